// fp32 stencil kernels of the mixed-precision step: reconstruction, face
// flux and edge EMF in fp32 (FMA contraction allowed), stored as fp32; the
// conserved accumulation and the CT face update stay binary64.
#include "mhd_ops.hpp"

using namespace mhd;

void launch_flux_fp32(mhd_ctx* c, Src<float> s) { flux_launch<float, float, float>(c, s); }

void launch_emf_fp32(mhd_ctx* c, Src<float> s) { emf_launch<float, float, float>(c, s); }

void launch_vface_fp32(mhd_ctx* c, long n, const double* lam, const double* v, double* v0p, double* v0n, double* v1p,
                     double* v1n) {
    vface_launch<float>(c, n, lam, v, v0p, v0n, v1p, v1n);
}
