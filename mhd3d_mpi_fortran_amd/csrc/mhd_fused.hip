// Fused face flux and conserved update of the mixed-precision
// Adams-Bashforth step: the fp32 PDM reconstruction and Gaussian-split face
// flux (get_flux_stress, mhd3d3.f90:2385-2560) and the binary64 corrector
// (advance_corrector1/2, 3842-3907) in one kernel, so the 24 face products
// per cell are formed on the chip and consumed there instead of making a
// round trip through HBM (written by the flux kernels, read back by the
// update).
//
// A workgroup owns a tile of TX x TY cells (64 along x, one wavefront per
// row, TY rows) and marches along z over KZ planes.  Per plane:
//   * the plane's 8 stencil inputs over the tile and its PDM halo (4 cells on
//     each side in x and y) are staged in LDS, each row one coalesced load;
//   * every thread forms the lower x face and the lower y face of its cell
//     from LDS and the upper z face from a register window of 8 planes of
//     its own column, which moves one plane per step (one new load per
//     variable), so no z face is formed twice;
//   * the faces the tile shares with its neighbours -- the upper y face of
//     the last row and the upper x face of each row's last column -- are
//     formed as well (by the last and the first wavefront), the lower z face
//     comes from the previous plane;
//   * x and y face products wait in LDS for the cells on their upper side,
//     and each thread updates its cell in binary64 from its six faces.
// The arithmetic of every face is the unfused kernels' (same reconstruction
// and flux functions), so fused and unfused steps agree to the rounding of
// the fp32 flux; the binary64 update keeps the unfused kernel's operation
// order.  With DT, the CFL partials of the new state are taken as the
// unfused update takes them (cell B of constrained transport's new faces).
//
// Cost: 225 VGPRs (the z window alone holds 64), 2 wavefronts per SIMD.
// At 256^3 the kernel takes 1.61 ms against 1.54 ms for the three unfused
// flux launches and the update: it moves 3.2 GB per step less, but with two
// wavefronts per SIMD and a barrier per plane its loads and its flux
// arithmetic do not overlap, and capping it at 168 VGPRs (three
// wavefronts) costs more than it gains (2.0 ms).  It is therefore not the
// default (mhd_set_fusion).
#include "mhd_ops.hpp"
#include "mhd_fused.hpp"

namespace mhd {
namespace fz {

constexpr int TX = kTileX, TY = kBlock / kTileX, KZ = kFusedPlanes;
constexpr int NT = TY * TX;  // threads: one wavefront per row
constexpr int H = 4;               // PDM reconstruction reads cells -4..+3
constexpr int PW = TX + 2 * H;     // columns of the staged plane
constexpr int NROW = TY + 2 * H;   // rows of the staged plane

// face products of one face along D; tap(v, m) = tap m (cells -4..+3
// around the face) of variable v, reconstructed one variable at a time
template <int D, int MET, typename Tap>
__device__ __forceinline__ void flux_vals(const FusedArgs& a, Tap tap, float* out) {
    float l[8], r[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        float s[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) s[m] = tap(v, m);
        Stencil<float, PDM>::run(s, a.c, l[v], r[v]);
    }
    if constexpr (MET == 1) face_flux_met_vals<float>(a.c, l, r, a.n[D - 1], out);
    else face_flux_vals<float, D>(a.c, l, r, out);
}

template <int MET, bool DT>
__global__ void __launch_bounds__(NT) k_flux_update(FusedArgs a) {
    __shared__ float tile[8][NROW][PW];
    __shared__ float fxs[8][TY][TX + 1];   // x face products of the plane (+ the tile's last upper face)
    __shared__ float fys[8][TY + 1][TX];   // y face products (+ the tile's upper row)
    const Dims& d = a.d;
    const int lane = threadIdx.x & (TX - 1), w = threadIdx.x / TX;
    const unsigned ntx = (d.nx0 + TX - 1) / TX, nty = (d.ny0 + TY - 1) / TY;
    const unsigned b = xcd_block<kXcdRun / 4>();
    const int x0 = (int)(b % ntx) * TX, y0 = (int)(b / ntx % nty) * TY, z0 = (int)(b / ntx / nty) * KZ;
    const int ng = d.ng;
    const float* f[8] = {a.q.rho, a.q.vx, a.q.vy, a.q.vz, a.q.p, a.q.bx, a.q.by, a.q.bz};
    // own column (1-based, ghost-inclusive), clamped onto the block for the
    // lanes past its edge (their results are not stored)
    const int cx = min(ng + 1 + x0 + lane, d.nx), cy = min(ng + 1 + y0 + w, d.ny);
    const bool col_live = x0 + lane < d.nx0 && y0 + w < d.ny0;
    const unsigned plane = (unsigned)d.nx * (unsigned)d.ny;
    const unsigned colo = (unsigned)ix(d.nx, d.ny, cx, cy, 1);
    auto zat = [&](int v, int kc) {  // own column at cell kc (1-based), clamped
        const int kk = max(1, min(kc, d.nz));
        return f[v][colo + (unsigned)(kk - 1) * plane];
    };
    // z window of the upper face of cell z0 - 1 (cells z0 - 4 .. z0 + 3)
    const bool row_wave = w < TY;   // (every wavefront owns a row)
    float zw[8][8], fzlo[8];
    if (row_wave) {
#pragma unroll
        for (int v = 0; v < 8; ++v)
#pragma unroll
            for (int m = 0; m < 8; ++m) zw[v][m] = zat(v, ng + 1 + z0 - 4 + m);
        flux_vals<3, MET>(a, [&](int v, int m) { return zw[v][m]; }, fzlo);
    }
    double dtmin = __builtin_huge_val();
    const int kend = min(z0 + KZ, d.nz0);
    for (int k = z0; k < kend; ++k) {
        const int ck = ng + 1 + k;
        __syncthreads();  // the previous plane's LDS reads are done
        for (int e = threadIdx.x; e < NROW * PW; e += NT) {
            const int row = e / PW, col = e - row * PW;
            const int gx = max(1, min(ng + 1 + x0 - H + col, d.nx));
            const int gy = max(1, min(ng + 1 + y0 - H + row, d.ny));
            const unsigned at = (unsigned)ix(d.nx, d.ny, gx, gy, ck);
#pragma unroll
            for (int v = 0; v < 8; ++v) tile[v][row][col] = f[v][at];
        }
        // the z window moves to the upper face of cell k (cells k-3 .. k+4)
        if (row_wave) {
#pragma unroll
            for (int v = 0; v < 8; ++v) {
#pragma unroll
                for (int m = 0; m < 7; ++m) zw[v][m] = zw[v][m + 1];
                zw[v][7] = zat(v, ck + 4);
            }
        }
        __syncthreads();
        float o[8], fzhi[8];
        if (row_wave) {
            // lower x face of the cell: cells x-4 .. x+3 of its row
            flux_vals<1, MET>(a, [&](int v, int m) { return tile[v][H + w][lane + m]; }, o);
#pragma unroll
            for (int v = 0; v < 8; ++v) fxs[v][w][lane] = o[v];
            // lower y face: rows y-4 .. y+3 of its column
            flux_vals<2, MET>(a, [&](int v, int m) { return tile[v][w + m][H + lane]; }, o);
#pragma unroll
            for (int v = 0; v < 8; ++v) fys[v][w][lane] = o[v];
            // upper z face from the window
            flux_vals<3, MET>(a, [&](int v, int m) { return zw[v][m]; }, fzhi);
        }
        if (w == TY - 1) {
            flux_vals<2, MET>(a, [&](int v, int m) { return tile[v][TY + m][H + lane]; }, o);
#pragma unroll
            for (int v = 0; v < 8; ++v) fys[v][TY][lane] = o[v];
        }
        if (w == 0 && lane < TY) {
            flux_vals<1, MET>(a, [&](int v, int m) { return tile[v][H + lane][TX + m]; }, o);
#pragma unroll
            for (int v = 0; v < 8; ++v) fxs[v][lane][TX] = o[v];
        }
        __syncthreads();
        if (row_wave && col_live) {
#pragma clang fp contract(off)
            // advance_corrector1/2 of cell (x0+lane, y0+w, k), Adams-Bashforth
            // form, as the unfused update: U = U0 - tau div F, p from the
            // new momentum before the stress, then the stress, then v
            const int i = x0 + lane + 1, j = y0 + w + 1, kk = k + 1;   // 1-based interior
            const ptrdiff_t q = ix(d.nx, d.ny, ng + i, ng + j, ng + kk);
            const double tau = *a.tau_p;
            const double dxi = a.dx[ng + i - 1], dyj = a.dy[ng + j - 1], dzk = a.dz[ng + kk - 1];
            auto div = [&](int c) -> double {
                return (double(fxs[c][w][lane + 1]) - double(fxs[c][w][lane])) / dxi +
                       (double(fys[c][w + 1][lane]) - double(fys[c][w][lane])) / dyj +
                       (double(fzhi[c]) - double(fzlo[c])) / dzk;
            };
            const double r0 = a.rho[q], u = a.vx[q], vv = a.vy[q], ww = a.vz[q];
            const double mx0 = r0 * u, my0 = r0 * vv, mz0 = r0 * ww;
            const double e0 = .5 * r0 * (u * u + vv * vv + ww * ww) + a.p[q] / a.gm1;
            const double rho = r0 - tau * div(0);
            const double eng = e0 - tau * div(4);
            double mx = mx0 - tau * div(1);
            double my = my0 - tau * div(2);
            double mz = mz0 - tau * div(3);
            const double p = a.gm1 * (eng - .5 * (mx * mx + my * my + mz * mz) / rho);
            mx = mx - tau * div(5);
            my = my - tau * div(6);
            mz = mz - tau * div(7);
            const double un = mx / rho, vn = my / rho, wn = mz / rho;
            a.orho[q] = rho;
            a.op[q] = p;
            a.ovx[q] = un;
            a.ovy[q] = vn;
            a.ovz[q] = wn;
            if constexpr (DT) {
                const ptrdiff_t gi = ix(d.nx + 1, d.ny, ng + i, ng + j, ng + kk);
                const ptrdiff_t gj = ix(d.nx, d.ny + 1, ng + i, ng + j, ng + kk);
                double x = .5 * (a.bi[gi + 1] + a.bi[gi]);
                double y = .5 * (a.bj[gj + d.nx] + a.bj[gj]);
                double z = .5 * (a.bk[q + (ptrdiff_t)d.nx * d.ny] + a.bk[q]);
                if (a.met_on) {
                    const double fx = x, fy = y, fzz = z;
                    x = a.M[0] * fx + a.M[1] * fy + a.M[2] * fzz;
                    y = a.M[3] * fx + a.M[4] * fy + a.M[5] * fzz;
                    z = a.M[6] * fx + a.M[7] * fy + a.M[8] * fzz;
                }
                const double dt = cfl_dt(a.gam, a.cfl, un, vn, wn, x, y, z, rho, p, dxi, dyj, dzk);
                dtmin = dt < dtmin ? dt : dtmin;
            }
        }
        if (row_wave) {
#pragma unroll
            for (int v = 0; v < 8; ++v) fzlo[v] = fzhi[v];
        }
    }
    if constexpr (DT) {
        __shared__ double red[TY];
        dtmin = wave_min(dtmin);
        if (lane == 0) red[w] = dtmin;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = red[0];
            for (int q = 1; q < TY; ++q) m = red[q] < m ? red[q] : m;
            a.part[blockIdx.x] = m;
        }
    }
}

}  // namespace fz

unsigned fused_blocks(const Dims& d) {
    const unsigned ntx = (d.nx0 + fz::TX - 1) / fz::TX, nty = (d.ny0 + fz::TY - 1) / fz::TY;
    const unsigned ntz = (d.nz0 + fz::KZ - 1) / fz::KZ;
    return ntx * nty * ntz;
}

void launch_flux_update(mhd_ctx* c, const FusedArgs& a, bool dt) {
    const dim3 g(fused_blocks(c->d)), b(fz::NT);
    if (c->met.on) {
        if (dt) launch(c, MHD_K_UPDATE, g, b, fz::k_flux_update<1, true>, a);
        else launch(c, MHD_K_UPDATE, g, b, fz::k_flux_update<1, false>, a);
    } else {
        if (dt) launch(c, MHD_K_UPDATE, g, b, fz::k_flux_update<0, true>, a);
        else launch(c, MHD_K_UPDATE, g, b, fz::k_flux_update<0, false>, a);
    }
}

}  // namespace mhd
