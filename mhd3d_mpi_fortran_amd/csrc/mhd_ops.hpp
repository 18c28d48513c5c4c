// Launchers of the stencil kernels (face flux, edge EMF), instantiated once
// per arithmetic type in its own translation unit so that each can carry its
// own floating-point contract: binary64 without contraction (reference
// parity), fp32 with FMA contraction (mixed precision).
#pragma once
#include "mhd_emf.hpp"
#include "mhd_launch.hpp"

namespace mhd {

template <typename R, typename TI, typename TF, int M>
void flux_launch_m(mhd_ctx* c, Src<TI> s) {
    const Coef<R> co = make_coef<R>(c);
    const Prim<TI> q{s.rho, s.vx, s.vy, s.vz, s.p, s.bx, s.by, s.bz};
    launch(c, MHD_K_FLUX, blocks(c->nf[0]), dim3(kBlock), k_flux<R, M, 1, TI, TF>, c->d, co, q, (TF*)c->flux[0],
           c->nf[0]);
    launch(c, MHD_K_FLUX, blocks(c->nf[1]), dim3(kBlock), k_flux<R, M, 2, TI, TF>, c->d, co, q, (TF*)c->flux[1],
           c->nf[1]);
    launch(c, MHD_K_FLUX, blocks(c->nf[2]), dim3(kBlock), k_flux<R, M, 3, TI, TF>, c->d, co, q, (TF*)c->flux[2],
           c->nf[2]);
}

template <typename R, typename TI, typename TF>
void flux_launch(mhd_ctx* c, Src<TI> s) {
    switch (c->p.reconop) {
    case PDM: flux_launch_m<R, TI, TF, PDM>(c, s); break;
    case TVD: flux_launch_m<R, TI, TF, TVD>(c, s); break;
    case WENO5: flux_launch_m<R, TI, TF, WENO5>(c, s); break;
    default: flux_launch_m<R, TI, TF, MPWENO5>(c, s); break;
    }
}

// edges handled per thread along the march axis
template <int M>
struct EmfSteps {
    static constexpr int T = M == TVD ? 16 : 8;
};

template <typename R, int M, int K, typename TI, typename TF>
void emf_launch_k(mhd_ctx* c, EmfIn<TI> in) {
    constexpr int T = EmfSteps<M>::T;
    const Dims& d = c->d;
    const long ii = d.nx0 + (K != 1);
    const long no = K == 3 ? d.nz0 : (K == 1 ? d.ny0 + 1 : d.ny0);
    const long nm = K == 3 ? d.ny0 + 1 : d.nz0 + 1;
    const long n = ii * no * ((nm + T - 1) / T);
    launch(c, MHD_K_EMF, blocks(n), dim3(kBlock), k_emf_march<R, M, K, T, TI, TF>, d, make_coef<R>(c), in,
           (TF*)c->emf[K - 1], n);
}

// edge k: (v1, v2, b1, b2) = (vy, vz, bj, bk), (vz, vx, bk, bi), (vx, vy, bi, bj)
template <typename R, typename TI, typename TF, int M>
void emf_launch_m(mhd_ctx* c, Src<TI> s) {
    emf_launch_k<R, M, 1, TI, TF>(c, EmfIn<TI>{s.vy, s.vz, s.bj, s.bk, s.vf});
    emf_launch_k<R, M, 2, TI, TF>(c, EmfIn<TI>{s.vz, s.vx, s.bk, s.bi, s.vf});
    emf_launch_k<R, M, 3, TI, TF>(c, EmfIn<TI>{s.vx, s.vy, s.bi, s.bj, s.vf});
}

template <typename R, typename TI, typename TF>
void emf_launch(mhd_ctx* c, Src<TI> s) {
    switch (c->p.reconop) {
    case PDM: emf_launch_m<R, TI, TF, PDM>(c, s); break;
    case TVD: emf_launch_m<R, TI, TF, TVD>(c, s); break;
    case WENO5: emf_launch_m<R, TI, TF, WENO5>(c, s); break;
    default: emf_launch_m<R, TI, TF, MPWENO5>(c, s); break;
    }
}

}  // namespace mhd
