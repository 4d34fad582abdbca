// Dielectric tables (smoothed eps, <1/eps>, inverse-eps tensor) on the GPU.
#pragma once

#include "common.h"

namespace bh {

constexpr int DIEL_MAX_ATOMS = 64;

struct DielGeom {
    double a1[2], a2[2];
    double eps_bg;
    int n_atoms;
    double atoms[DIEL_MAX_ATOMS][4];  // fractional x, y, radius (units of |a1|), eps_inside
};

// Tables of ngeom geometries on one nx x ny grid into device arrays laid out
// [geometry][node] (tensor [geometry][node][4]); any output may be null.
// d_geoms: device scratch of dielectric_geom_bytes(ngeom).  ev0 / ev1 (may be
// null) are recorded around the kernel.  Returns once the tables are written.
void launch_dielectric(const DielGeom *host_geoms, int ngeom, int nx, int ny, bool smoothing, double *d_eps,
                       double *d_inv_eps, double *d_tensor, void *d_geoms, hipStream_t st, hipEvent_t ev0 = nullptr,
                       hipEvent_t ev1 = nullptr);
size_t dielectric_geom_bytes(int ngeom);

}  // namespace bh
