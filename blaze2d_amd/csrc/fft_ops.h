// Operator kernels: work descriptors and launch entry points.
#pragma once

#include "common.h"

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
};

bool grid_supported(int nx, int ny);

// prec_fp32: the FD and TP preconditioners of an FP64 launch compute in FP32
// (FP64 columns in and out)
template <typename T>
void launch_op(int mode, int nx, int ny, const OpWork<T> *dwork, int nwork, int maxcols, const cplx<T> *twx,
               const cplx<T> *twy, hipStream_t st, bool prec_fp32 = false);

}  // namespace bh
