// Fused face flux + conserved update of the mixed-precision Adams-Bashforth
// step (kernel in mhd_fused.hip): arguments and launcher.
#pragma once
#include "mhd_launch.hpp"

namespace mhd {

// z planes one workgroup marches through
constexpr int kFusedPlanes = 32;

struct FusedArgs {
    Dims d;
    Coef<float> c;
    Prim<float> q;                   // fp32 stencil inputs (the extrapolated state)
    float n[3][3];                   // unit face normals (affine metric)
    const double *rho, *vx, *vy, *vz, *p;   // binary64 state the corrector starts from
    double *orho, *ovx, *ovy, *ovz, *op;    // and the state it writes
    const double *dx, *dy, *dz;     // spacings (affine metric: spacing / g)
    const double* tau_p;
    double gm1;
    // CFL partials of the new state (DT): constrained transport's new faces
    const double *bi, *bj, *bk;
    double M[9];                     // affine metric's cell-B map (met_on != 0)
    int met_on;
    double gam, cfl;
    double* part;                    // one partial per workgroup
};

unsigned fused_blocks(const Dims& d);
void launch_flux_update(mhd_ctx* c, const FusedArgs& a, bool dt);

}  // namespace mhd
