// Operator kernels: work descriptors and launch entry points.
#pragma once

#include "common.h"

#include <vector>

namespace bh {

enum OpMode {
    MODE_THETA_TM = 0,
    MODE_THETA_TE = 1,
    MODE_PREC_FD = 2,
    MODE_PREC_TP = 3,
    MODE_FFT_FWD = 4,
    MODE_FFT_INV = 5,
};

// One work item = a run of `count` consecutive columns of one slot.
template <typename T>
struct OpWork {
    const cplx<T> *in;    // first input column
    cplx<T> *out;         // first output column (may alias in)
    cplx<T> *bout;        // optional mass output (PREC_FD on TM: eps * out)
    const T *eps;         // eps(r) per point (TM mass, TP)
    const T *tensor;      // [N][4] inverse-eps tensor (TE)
    const OpParams *params;
    const int *dev_count; // if set, column count read on device
    long in_stride;       // elements between input columns
    long out_stride;
    int count;
    int pad;
    cplx<T> *tmp;         // multi-pass TE / TP: a second field per column (tmp_stride apart)
    long tmp_stride;
};

// Grids the operators run on: square 8..64 with the whole field in one
// workgroup's LDS, and any nx x ny with both sides in {8, 16, 24, 32, 48, 64,
// 96, 128, 192, 256, 384, 512} through the multi-pass kernels.
bool grid_supported(int nx, int ny);
// true when the grid takes the multi-pass path (its TE / TP applies need tmp)
bool grid_multipass(int nx, int ny);
// frequency index held at each digit-reversed position of an n-point line
std::vector<int> digit_reversal(int n);

// prec_fp32: the FD and TP preconditioners of an FP64 launch compute in FP32
// (FP64 columns in and out)
template <typename T>
void launch_op(int mode, int nx, int ny, const OpWork<T> *dwork, int nwork, int maxcols, const cplx<T> *twx,
               const cplx<T> *twy, hipStream_t st, bool prec_fp32 = false, const int *revx = nullptr,
               const int *revy = nullptr);

// Multi-pass operators: column passes (y transforms over tiles of 8 columns)
// and row passes (x transforms of whole rows with the Fourier-space step), the
// field passing through HBM (L2) between passes.  revx / revy: digit_reversal
// tables of nx and ny on the device.
template <typename T>
void launch_op_big(int mode, int nx, int ny, const OpWork<T> *dwork, int nwork, int maxcols, const cplx<T> *twx,
                   const cplx<T> *twy, const int *revx, const int *revy, hipStream_t st);

}  // namespace bh
