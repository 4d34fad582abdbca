// The learned start's network: a small CNN over the eps(r) grid with an MLP
// embedding of (k1, k2, polarisation, lattice), trained and run on the GPU by
// the kernels in cnn.hip.
#pragma once

#include <hip/hip_bf16.h>

#include <map>
#include <vector>

#include "common.h"
#include "dielectric.h"

namespace bh {

class Cnn {
  public:
    static constexpr int CH = 32;      // channels of every convolution
    static constexpr int LAYERS = 4;   // conv layers (stride 1, then 2, 2, 2)
    static constexpr int EMB_IN = 4;   // k1, k2, polarisation (0 TM, 1 TE), lattice (0 square, 1 hex)

    // h, w: input grid (multiples of 8); nb: predicted vectors (1..32);
    // pw_half: plane waves |n1|, |n2| <= pw_half (0..8); max_batch: samples per call
    Cnn(int h, int w, int nb, int pw_half, int max_batch, unsigned long long seed, hipStream_t st);
    ~Cnn();
    Cnn(const Cnn &) = delete;
    Cnn &operator=(const Cnn &) = delete;

    // x: [B][2][H][W] (eps / 16, 1 / eps), emb: [B][4]
    void upload_inputs(int B, const float *x, const float *emb);
    // the same inputs built on the device: the smoothed (or plain) eps of each
    // geometry on the network's H x W grid (the dielectric kernel), then
    // eps / 16 and 1 / eps
    void inputs_from_geometries(int B, const DielGeom *geoms, bool smoothing, const float *emb);
    void forward(int B);
    // One training step on inputs already uploaded: forward, projector loss and
    // its gradient, backward, and (lr > 0) an Adam update; returns the batch
    // mean loss.  targets [B][nb][ng] complex (interleaved doubles).  With
    // graphs on, the step's device work after the uploads is replayed from a
    // hipGraph captured once per (batch size, update or not); the step's
    // learning rate and bias corrections are read from device memory, so one
    // graph serves every step, and the eager path reads them the same way.
    double train_step(int B, const double *targets, float lr);
    void set_graphs(bool on) { use_graphs = on; }
    void dense_bwd_data(const float *dy, const __hip_bfloat16 *Wm, float *dx, int B, int K, int M);
    void get_output(int B, double *host) const;  // [B][nb][ng] complex
    void get_params(float *host) const;
    void set_params(const float *host);
    void get_grads(float *host) const;

    long param_count() const { return P; }
    int out_size() const { return OUT; }
    int max_batch_size() const { return max_batch; }

    const int H, W, nb, pw_half, max_batch;
    int ng = 0, F = 0, OUT = 0;

  private:
    hipStream_t st;
    long P = 0;
    int step = 0;
    // parameter offsets (floats) into params / grads
    long off_c[LAYERS]{}, off_cb[LAYERS]{}, off_e1 = 0, off_e1b = 0, off_e2 = 0, off_e2b = 0, off_h = 0, off_hb = 0;
    std::vector<void *> bufs;
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    __hip_bfloat16 *pbf = nullptr;
    float *x0 = nullptr, *e0 = nullptr, *eh = nullptr, *ev = nullptr;
    float *act[LAYERS]{}, *gact[LAYERS]{};
    float *yout = nullptr, *gy = nullptr, *geh = nullptr, *gev = nullptr;
    double *tgt = nullptr, *lossb = nullptr;
    float *wpart = nullptr;   // partial weight-gradient rows (one per workgroup)
    float *dpart = nullptr;   // partial sums of the dense layers' input gradients
    size_t dpart_floats = 0;
    double *d_eps = nullptr;  // dielectric scratch of inputs_from_geometries (allocated on first use)
    float *d_hyper = nullptr, *h_hyper = nullptr;  // [lr, 1 - b1^t, 1 - b2^t] of the step (device, pinned host)
    bool use_graphs = true;
    std::map<int, hipGraphExec_t> graphs;  // key 2 B + (Adam update)
    void step_launches(int B, bool update);  // the device work of one step (kernels only: capturable)
    void *d_geo = nullptr;
    size_t loss_lds = 0;
};

}  // namespace bh
