// binary64 stencil kernels (compiled with -ffp-contract=off: same operation
// order as the reference, no fused multiply-add)
#include "mhd_ops.hpp"
#include "mhd_yee.hpp"

using namespace mhd;

void launch_flux_fp64(mhd_ctx* c, Src<double> s) { flux_launch<double, double, double>(c, s); }

void launch_emf_fp64(mhd_ctx* c, Src<double> s) { emf_launch<double, double, double>(c, s); }

void launch_recon(mhd_ctx* c, int dir, const double* f, int fn1, int fn2, int lnx, int lny, int lnz, double* fl,
                  double* fr) {
    const Coef<double> co = make_coef<double>(c);
    const long n = (long)lnx * lny * lnz;
    const int ng = c->p.ng;
    switch (c->p.reconop) {
    case PDM:
        launch(c, MHD_K_FLUX, blocks(n), dim3(kBlock), k_recon<double, PDM>, ng, dir, f, fn1, fn2, lnx, lny, lnz,
               co, fl, fr);
        break;
    case TVD:
        launch(c, MHD_K_FLUX, blocks(n), dim3(kBlock), k_recon<double, TVD>, ng, dir, f, fn1, fn2, lnx, lny, lnz,
               co, fl, fr);
        break;
    case WENO5:
        launch(c, MHD_K_FLUX, blocks(n), dim3(kBlock), k_recon<double, WENO5>, ng, dir, f, fn1, fn2, lnx, lny,
               lnz, co, fl, fr);
        break;
    default:
        launch(c, MHD_K_FLUX, blocks(n), dim3(kBlock), k_recon<double, MPWENO5>, ng, dir, f, fn1, fn2, lnx, lny,
               lnz, co, fl, fr);
        break;
    }
}

void launch_vface_fp64(mhd_ctx* c, long n, const double* lam, const double* v, double* v0p, double* v0n, double* v1p,
                     double* v1n) {
    vface_launch<double>(c, n, lam, v, v0p, v0n, v1p, v1n);
}

void launch_yee_split(mhd_ctx* c, long n, Prim<double> q, double* const* e) {
    YeeSplit o;
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 4; ++s) o.e[a][s] = e[4 * a + s];
    launch(c, MHD_K_EMF, blocks(n), dim3(kBlock), k_yee_split, n, q, make_coef<double>(c).rspi, o);
}

void launch_yee_sum(mhd_ctx* c, int o1, int o2, int o3, int ax, int n1, int n2, const double* const* e4,
                    double* out) {
    launch(c, MHD_K_EMF, blocks((long)o1 * o2 * o3), dim3(kBlock), k_yee_sum, o1, o2, o3, ax, c->p.ng, n1, n2,
           e4[0], e4[1], e4[2], e4[3], out);
}
