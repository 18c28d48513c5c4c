// Launchers of the stencil kernels (face flux, edge EMF), instantiated once
// per arithmetic type in its own translation unit so that each can carry its
// own floating-point contract: binary64 without contraction (reference
// parity), fp32 with FMA contraction (mixed precision).
#pragma once
#include "mhd_launch.hpp"

namespace mhd {

template <typename R, typename TF, int M>
void flux_launch_m(mhd_ctx* c, Prim<double> q) {
    const Coef<R> co = make_coef<R>(c);
    launch(c, MHD_K_FLUX, blocks(c->nf[0]), dim3(kBlock), k_flux<R, M, 1, TF>, c->d, co, q, (TF*)c->flux[0],
           c->nf[0]);
    launch(c, MHD_K_FLUX, blocks(c->nf[1]), dim3(kBlock), k_flux<R, M, 2, TF>, c->d, co, q, (TF*)c->flux[1],
           c->nf[1]);
    launch(c, MHD_K_FLUX, blocks(c->nf[2]), dim3(kBlock), k_flux<R, M, 3, TF>, c->d, co, q, (TF*)c->flux[2],
           c->nf[2]);
}

template <typename R, typename TF>
void flux_launch(mhd_ctx* c, Prim<double> q) {
    switch (c->p.reconop) {
    case PDM: flux_launch_m<R, TF, PDM>(c, q); break;
    case TVD: flux_launch_m<R, TF, TVD>(c, q); break;
    case WENO5: flux_launch_m<R, TF, WENO5>(c, q); break;
    default: flux_launch_m<R, TF, MPWENO5>(c, q); break;
    }
}

// edge k: (v1, v2, b1, b2) = (vy, vz, bj, bk), (vz, vx, bk, bi), (vx, vy, bi, bj)
template <typename R, typename TF, int M>
void emf_launch_m(mhd_ctx* c, Prim<double> q, const double* bi, const double* bj, const double* bk,
                  const double* vf) {
    const Coef<R> co = make_coef<R>(c);
    launch(c, MHD_K_EMF, blocks(c->ne[0]), dim3(kBlock), k_emf<R, M, 1, TF>, c->d, co, q.vy, q.vz, bj, bk, vf,
           (TF*)c->emf[0], c->ne[0]);
    launch(c, MHD_K_EMF, blocks(c->ne[1]), dim3(kBlock), k_emf<R, M, 2, TF>, c->d, co, q.vz, q.vx, bk, bi, vf,
           (TF*)c->emf[1], c->ne[1]);
    launch(c, MHD_K_EMF, blocks(c->ne[2]), dim3(kBlock), k_emf<R, M, 3, TF>, c->d, co, q.vx, q.vy, bi, bj, vf,
           (TF*)c->emf[2], c->ne[2]);
}

template <typename R, typename TF>
void emf_launch(mhd_ctx* c, Prim<double> q, const double* bi, const double* bj, const double* bk,
                const double* vf) {
    switch (c->p.reconop) {
    case PDM: emf_launch_m<R, TF, PDM>(c, q, bi, bj, bk, vf); break;
    case TVD: emf_launch_m<R, TF, TVD>(c, q, bi, bj, bk, vf); break;
    case WENO5: emf_launch_m<R, TF, WENO5>(c, q, bi, bj, bk, vf); break;
    default: emf_launch_m<R, TF, MPWENO5>(c, q, bi, bj, bk, vf); break;
    }
}

}  // namespace mhd
