// Cell/face/ghost/reduction kernels of the step: the conserved update and
// stage combine, constrained transport, ghost fill, get_bxyz, get_dt, the
// time-stepper bookkeeping and the state diagnostics.  Included by the
// library's main translation unit only (binary64, no contraction).
#pragma once
#include "mhd_kernels.hpp"

namespace mhd {

// ------------------------------------------------------------------------
// conserved update of one interior cell: get_conserve + advance_corrector1
// (+ stage combine + pressure) + advance_corrector2 (+ combine + velocity)
// (mhd3d3.f90:1641-1650, 3842-3907, 3648-3792).
//   NST = 0: Adams-Bashforth corrector (no combine, p from the new momentum)
//   NST = 2/3: TVD RK stage ST
// ------------------------------------------------------------------------
struct Cells {  // cell primitives written by the update
    double *rho, *vx, *vy, *vz, *p;
};
struct CCells {  // cell primitives the update starts from
    const double *rho, *vx, *vy, *vz, *p;
};
struct PrevRK {  // advance_previous, RK flavour
    const double *rho, *mx, *my, *mz, *eng;
};

template <typename TF>
struct FluxSet {
    const TF *x, *y, *z;
    long nfx, nfy, nfz;
};


// CFL partials of the updated state, computed by the update itself (get_dt
// on the new state, mhd3d3.f90:1608-1639): cell B from the stage's new
// faces (CT has written them before the update runs; the affine metric's
// cell-B map when on), the CFL step of every owned cell, one minimum per
// workgroup.  Owned cells only: the caller takes this where every ghost
// cell repeats an owned cell (periodic axes whose ghost spacings repeat the
// owned ones), so the minimum is get_dt's over the ghost-inclusive block.
struct UpdDt {
    const double *bi, *bj, *bk;  // the new faces
    double M[9];                 // affine metric's cell-B map (on != 0)
    int on;
    double gam, cfl;
    double* part;                // one partial per workgroup
};

// GEN: general metric (gm), div = (sum_f F_f |S_f| out) / V; a template
// parameter so that the other grids' kernel carries none of its registers.
// DT: CFL partials of the new state as well (UpdDt)
template <int NST, int ST, typename TF, bool GEN, bool DT = false>
__global__ void __launch_bounds__(kBlock) k_update(Dims d, double gm1, const double* __restrict__ tau_p, CCells s,
                                                  Cells so, PrevRK pv, FluxSet<TF> F, const double* __restrict__ dx,
                                                  const double* __restrict__ dy, const double* __restrict__ dz,
                                                  GMet gm, long ncell, UpdDt udt) {
    const long t = gtid_xcd();
    if constexpr (!DT) {
        if (t >= ncell) return;
    }
    // with DT every thread reaches the workgroup minimum: threads past the
    // end work on cell 1 and store nothing
    const bool live = t < ncell;
    int i, j, k;
    unflat(live ? t : 0, d.nx0, d.ny0, i, j, k);
    const int ng = d.ng;
    const ptrdiff_t q = ix(d.nx, d.ny, ng + i, ng + j, ng + k);
    const double tau = *tau_p;
    const double dxi = dx[ng + i - 1], dyj = dy[ng + j - 1], dzk = dz[ng + k - 1];
    const ptrdiff_t fx0 = ix(d.nx0 + 1, d.ny0, i, j, k);
    const ptrdiff_t fy0 = ix(d.nx0, d.ny0 + 1, i, j, k), fy1 = fy0 + d.nx0;
    const ptrdiff_t fz0 = ix(d.nx0, d.ny0, i, j, k), fz1 = fz0 + (ptrdiff_t)d.nx0 * d.ny0;
    double ar[6] = {0., 0., 0., 0., 0., 0.}, iv = 0.;
    if constexpr (GEN) {
        const ptrdiff_t gi = ix(d.nx + 1, d.ny, ng + i, ng + j, ng + k), gj = ix(d.nx, d.ny + 1, ng + i, ng + j, ng + k);
        ar[0] = gm.af[0][gi];
        ar[1] = gm.af[0][gi + 1];
        ar[2] = gm.af[1][gj];
        ar[3] = gm.af[1][gj + d.nx];
        ar[4] = gm.af[2][q];
        ar[5] = gm.af[2][q + (ptrdiff_t)d.nx * d.ny];
        iv = gm.ivol[q];
    }
    auto div = [&](int f) -> double {
        if constexpr (GEN)
            return ((double(F.x[f * F.nfx + fx0 + 1]) * ar[1] - double(F.x[f * F.nfx + fx0]) * ar[0]) +
                    (double(F.y[f * F.nfy + fy1]) * ar[3] - double(F.y[f * F.nfy + fy0]) * ar[2]) +
                    (double(F.z[f * F.nfz + fz1]) * ar[5] - double(F.z[f * F.nfz + fz0]) * ar[4])) * iv;
        return (double(F.x[f * F.nfx + fx0 + 1]) - double(F.x[f * F.nfx + fx0])) / dxi +
               (double(F.y[f * F.nfy + fy1]) - double(F.y[f * F.nfy + fy0])) / dyj +
               (double(F.z[f * F.nfz + fz1]) - double(F.z[f * F.nfz + fz0])) / dzk;
    };
    // get_conserve from the stage's primitives
    const double r0 = s.rho[q], u = s.vx[q], v = s.vy[q], w = s.vz[q];
    const double mx0 = r0 * u, my0 = r0 * v, mz0 = r0 * w;
    const double e0 = .5 * r0 * (u * u + v * v + w * w) + s.p[q] / gm1;
    double rho = r0 - tau * div(0);
    double eng = e0 - tau * div(4);
    double mx = mx0 - tau * div(1);
    double my = my0 - tau * div(2);
    double mz = mz0 - tau * div(3);
    double hx = mx, hy = my, hz = mz;
    if (NST == 2 && ST == 2) {
        rho = (pv.rho[q] + rho) / 2.;
        hx = (pv.mx[q] + mx) / 2.;
        hy = (pv.my[q] + my) / 2.;
        hz = (pv.mz[q] + mz) / 2.;
        eng = (pv.eng[q] + eng) / 2.;
    } else if (NST == 3 && ST == 2) {
        rho = (3. * pv.rho[q] + rho) / 4.;
        hx = (3. * pv.mx[q] + mx) / 4.;
        hy = (3. * pv.my[q] + my) / 4.;
        hz = (3. * pv.mz[q] + mz) / 4.;
        eng = (3. * pv.eng[q] + eng) / 4.;
    } else if (NST == 3 && ST == 3) {
        rho = (pv.rho[q] + 2. * rho) / 3.;
        hx = (pv.mx[q] + 2. * mx) / 3.;
        hy = (pv.my[q] + 2. * my) / 3.;
        hz = (pv.mz[q] + 2. * mz) / 3.;
        eng = (pv.eng[q] + 2. * eng) / 3.;
    }
    // pressure from the hydro-only momentum (3685, 3876)
    const double p = gm1 * (eng - .5 * (hx * hx + hy * hy + hz * hz) / rho);
    mx = mx - tau * div(5);
    my = my - tau * div(6);
    mz = mz - tau * div(7);
    if (NST == 2 && ST == 2) {
        mx = (pv.mx[q] + mx) / 2.;
        my = (pv.my[q] + my) / 2.;
        mz = (pv.mz[q] + mz) / 2.;
    } else if (NST == 3 && ST == 2) {
        mx = (3. * pv.mx[q] + mx) / 4.;
        my = (3. * pv.my[q] + my) / 4.;
        mz = (3. * pv.mz[q] + mz) / 4.;
    } else if (NST == 3 && ST == 3) {
        mx = (pv.mx[q] + 2. * mx) / 3.;
        my = (pv.my[q] + 2. * my) / 3.;
        mz = (pv.mz[q] + 2. * mz) / 3.;
    }
    const double un = mx / rho, vn = my / rho, wn = mz / rho;
    if (live) {
        so.rho[q] = rho;
        so.p[q] = p;
        so.vx[q] = un;
        so.vy[q] = vn;
        so.vz[q] = wn;
    }
    if constexpr (DT) {
        double dtc = __builtin_huge_val();
        if (live) {
            // get_bxyz of the new faces, as k_bxyz
            const ptrdiff_t gi = ix(d.nx + 1, d.ny, ng + i, ng + j, ng + k), gj = ix(d.nx, d.ny + 1, ng + i, ng + j, ng + k);
            double x = .5 * (udt.bi[gi + 1] + udt.bi[gi]);
            double y = .5 * (udt.bj[gj + d.nx] + udt.bj[gj]);
            double z = .5 * (udt.bk[q + (ptrdiff_t)d.nx * d.ny] + udt.bk[q]);
            if (udt.on) {
                const double fx = x, fy = y, fz = z;
                x = udt.M[0] * fx + udt.M[1] * fy + udt.M[2] * fz;
                y = udt.M[3] * fx + udt.M[4] * fy + udt.M[5] * fz;
                z = udt.M[6] * fx + udt.M[7] * fy + udt.M[8] * fz;
            }
            dtc = cfl_dt(udt.gam, udt.cfl, un, vn, wn, x, y, z, rho, p, dxi, dyj, dzk);
        }
        dtc = block_min(dtc);
        if (threadIdx.x == 0) udt.part[blockIdx.x] = dtc;
    }
}

// ------------------------------------------------------------------------
// constrained transport: advance_bijk (mhd3d3.f90:3909-3933) + face combine
// one thread per owned face of bi, then bj, then bk
// ------------------------------------------------------------------------
struct Faces {
    double *bi, *bj, *bk;
};
struct CFaces {
    const double *bi, *bj, *bk;
};

// ctf[d] = 1/|C e_d| of a metric per face family, exactly 1 on the
// reference's tensor-product grid (then tau * ctf == tau bit for bit)
struct CtFactor {
    double f[3];
};

// GEN: general metric (gm), the EMFs are E . dl and the face update divides
// by the face's area
template <int NST, int ST, typename TF, bool GEN>
__global__ void __launch_bounds__(kBlock) k_ct(Dims d, const double* __restrict__ tau_p, CtFactor ctf, CFaces b,
                                              Faces bo, CFaces bp, const TF* __restrict__ ei,
                                              const TF* __restrict__ ej, const TF* __restrict__ ek,
                                              const double* __restrict__ dx, const double* __restrict__ dy,
                                              const double* __restrict__ dz, GMet gm) {
    const long t = gtid();
    const int nx0 = d.nx0, ny0 = d.ny0, nz0 = d.nz0, ng = d.ng;
    const long nbi = (long)(nx0 + 1) * ny0 * nz0, nbj = (long)nx0 * (ny0 + 1) * nz0,
               nbk = (long)nx0 * ny0 * (nz0 + 1);
    if (t >= nbi + nbj + nbk) return;
    const double tau = *tau_p;
    int i, j, k;
    double val, old;
    ptrdiff_t q;
    double* out;
    const double* prev;
    if (t < nbi) {
        unflat(t, nx0 + 1, ny0, i, j, k);
        // ek (nx0+1,ny0+1,nz0), ej (nx0+1,ny0,nz0+1)
        const ptrdiff_t a = ix(nx0 + 1, ny0 + 1, i, j, k), e = ix(nx0 + 1, ny0, i, j, k);
        const double dek = double(ek[a + (nx0 + 1)]) - double(ek[a]);
        const double dej = double(ej[e + (ptrdiff_t)(nx0 + 1) * ny0]) - double(ej[e]);
        q = ix(d.nx + 1, d.ny, ng + i, ng + j, ng + k);
        old = b.bi[q];
        val = GEN ? old - tau * gm.iaf[0][q] * (dek - dej)
                      : old - tau * ctf.f[0] * (dek / dy[ng + j - 1] - dej / dz[ng + k - 1]);
        out = bo.bi;
        prev = bp.bi;
    } else if (t < nbi + nbj) {
        unflat(t - nbi, nx0, ny0 + 1, i, j, k);
        // ei (nx0,ny0+1,nz0+1), ek (nx0+1,ny0+1,nz0)
        const ptrdiff_t a = ix(nx0, ny0 + 1, i, j, k), e = ix(nx0 + 1, ny0 + 1, i, j, k);
        const double dei = double(ei[a + (ptrdiff_t)nx0 * (ny0 + 1)]) - double(ei[a]);
        const double dek = double(ek[e + 1]) - double(ek[e]);
        q = ix(d.nx, d.ny + 1, ng + i, ng + j, ng + k);
        old = b.bj[q];
        val = GEN ? old - tau * gm.iaf[1][q] * (dei - dek)
                      : old - tau * ctf.f[1] * (dei / dz[ng + k - 1] - dek / dx[ng + i - 1]);
        out = bo.bj;
        prev = bp.bj;
    } else {
        unflat(t - nbi - nbj, nx0, ny0, i, j, k);
        // ej (nx0+1,ny0,nz0+1), ei (nx0,ny0+1,nz0+1)
        const ptrdiff_t a = ix(nx0 + 1, ny0, i, j, k), e = ix(nx0, ny0 + 1, i, j, k);
        const double dej = double(ej[a + 1]) - double(ej[a]);
        const double dei = double(ei[e + nx0]) - double(ei[e]);
        q = ix(d.nx, d.ny, ng + i, ng + j, ng + k);
        old = b.bk[q];
        val = GEN ? old - tau * gm.iaf[2][q] * (dej - dei)
                      : old - tau * ctf.f[2] * (dej / dx[ng + i - 1] - dei / dy[ng + j - 1]);
        out = bo.bk;
        prev = bp.bk;
    }
    (void)old;
    if (NST == 2 && ST == 2) val = (prev[q] + val) / 2.;
    else if (NST == 3 && ST == 2) val = (3. * prev[q] + val) / 4.;
    else if (NST == 3 && ST == 3) val = (prev[q] + 2. * val) / 3.;
    out[q] = val;
}

// ------------------------------------------------------------------------
// ghost fill of the whole block (fill_hydro_mpi + fill_mag_mpi,
// mhd3d3.f90:346-616, 781-819).  The reference's x, then y, then z fill over
// full extents composes to one map per axis: a ghost (i,j,k) copies the owned
// element (map_x(i), map_y(j), map_z(k)).  One thread per ghost element; every
// source is owned in all three axes, so no thread reads what another writes.
// ------------------------------------------------------------------------
struct FillArr {
    double* a;
    int n[3];     // extents
    int face;     // 0 cells, 1..3 normal axis of a face array
};
struct FillSet {
    FillArr arr[8];
    int narr;
    int n0[3];
    int bc[3];    // 0 even reflection, 2 periodic
};

__device__ __forceinline__ int fill_map(int i, int n0, int ng, bool face, int bc) {
    const int hi = ng + n0 + (face ? 1 : 0);
    if (i <= ng) return bc == 2 ? i + n0 : (face ? 2 * ng + 2 - i : 2 * ng + 1 - i);
    if (i > hi) return bc == 2 ? i - n0 : (face ? 2 * (n0 + ng) + 2 - i : 2 * (n0 + ng) + 1 - i);
    return i;
}

__device__ __forceinline__ int ghost_coord(int g, int ng, int own) { return g < ng ? g + 1 : ng + own + (g - ng) + 1; }

__global__ void __launch_bounds__(kBlock) k_fill(FillSet fs, int ng) {
    const FillArr A = fs.arr[blockIdx.y];
    const long t = gtid();
    int own[3], len[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        len[a] = A.n[a];
        own[a] = fs.n0[a] + (A.face == a + 1 ? 1 : 0);
    }
    const int g2 = 2 * ng;
    const long c1 = (long)g2 * len[1] * len[2];
    const long c2 = (long)own[0] * g2 * len[2];
    const long c3 = (long)own[0] * own[1] * g2;
    int i, j, k;
    if (t < c1) {
        long r = t / g2;
        i = ghost_coord((int)(t - r * g2), ng, own[0]);
        j = (int)(r % len[1]) + 1;
        k = (int)(r / len[1]) + 1;
    } else if (t < c1 + c2) {
        long u = t - c1;
        long r = u / own[0];
        i = ng + 1 + (int)(u - r * own[0]);
        j = ghost_coord((int)(r % g2), ng, own[1]);
        k = (int)(r / g2) + 1;
    } else if (t < c1 + c2 + c3) {
        long u = t - c1 - c2;
        long r = u / own[0];
        i = ng + 1 + (int)(u - r * own[0]);
        j = ng + 1 + (int)(r % own[1]);
        k = ghost_coord((int)(r / own[1]), ng, own[2]);
    } else {
        return;
    }
    const int si = fill_map(i, fs.n0[0], ng, A.face == 1, fs.bc[0]);
    const int sj = fill_map(j, fs.n0[1], ng, A.face == 2, fs.bc[1]);
    const int sk = fill_map(k, fs.n0[2], ng, A.face == 3, fs.bc[2]);
    A.a[ix(len[0], len[1], i, j, k)] = A.a[ix(len[0], len[1], si, sj, sk)];
}

// ------------------------------------------------------------------------
// get_bxyz (mhd3d3.f90:2118-2128) over the ghost-inclusive block, fused with
//   * the EMF's fast speed vf=|v|+|B|/sqrt(rho) (mhd3d3.f90:1766) when vf != 0,
//   * fp32 copies of the stencil inputs (mixed precision) when sh.rho != 0,
//   * per-workgroup minima of the CFL dt (get_dt, mhd3d3.f90:1608-1628)
//     when part != 0 -- the dt the next step starts with.
// Threads cover max(ncell, nbi, nbj, nbk) so the face copies ride along.
// ------------------------------------------------------------------------
struct Shadow {  // fp32 stencil inputs of one state
    float *rho, *vx, *vy, *vz, *p, *bx, *by, *bz, *vf, *bi, *bj, *bk;
};

// metric part of get_bxyz: cell B = M (face averages), M = A diag(g), and
// the face-normal velocities w_d = v . n_d (binary64 and/or fp32 copies)
struct BMet {
    int on;
    double M[9];
    double n[9];
    double* w[3];
    float* ws[3];
    GMet g;   // g.ivol != null: general metric (cell B from the 6 faces, per-cell normals and CFL spacing)
};

// cell B of cell t from its 6 faces through a general metric's coefficients
__device__ __forceinline__ void gen_cellb(const GMet& g, long t, const double* beta, double* B) {
    const long n = g.ncell;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double acc = 0.;
#pragma unroll
        for (int f = 0; f < 6; ++f) acc += beta[f] * g.cm[(f * 3 + r) * n + t];
        B[r] = acc;
    }
}

__device__ __forceinline__ double cfl_dt_gen(double gam, double cfl, double u, double v, double w, double x, double y,
                                             double z, double r, double pp, const GMet& g, long t) {
    const double vv = sqrt(u * u + v * v + w * w);
    const double btot = sqrt(x * x + y * y + z * z);
    const double valf = btot / sqrt(r);
    const double vs = sqrt(gam * pp / r);
    const double vcfl = vv + sqrt(valf * valf + vs * vs);
    return cfl / (vcfl * g.ih[t] + vcfl * g.ih[g.ncell + t] + vcfl * g.ih[2 * g.ncell + t]);
}

// GEN: general metric (met.g)
template <bool GEN>
__global__ void __launch_bounds__(kBlock) k_bxyz(Dims d, double gam, double cfl, CFaces b, CCells s,
                                                double* __restrict__ bx, double* __restrict__ by,
                                                double* __restrict__ bz, double* __restrict__ vf, Shadow sh, BMet met,
                                                const double* __restrict__ dx, const double* __restrict__ dy,
                                                const double* __restrict__ dz, double* __restrict__ part, long ncell,
                                                long nbi, long nbj, long nbk) {
    const long t = gtid_xcd();
    double dt = __builtin_huge_val();
    if (t < ncell) {
        int i, j, k;
        unflat(t, d.nx, d.ny, i, j, k);
        const double bi0 = b.bi[ix(d.nx + 1, d.ny, i, j, k)], bi1 = b.bi[ix(d.nx + 1, d.ny, i + 1, j, k)];
        const double bj0 = b.bj[ix(d.nx, d.ny + 1, i, j, k)], bj1 = b.bj[ix(d.nx, d.ny + 1, i, j + 1, k)];
        const double bk0 = b.bk[t], bk1 = b.bk[t + (ptrdiff_t)d.nx * d.ny];
        double x = .5 * (bi1 + bi0);
        double y = .5 * (bj1 + bj0);
        double z = .5 * (bk1 + bk0);
        const double u = s.vx[t], v = s.vy[t], w = s.vz[t], r = s.rho[t], pp = s.p[t];
        if constexpr (GEN) {
            const double beta[6] = {bi0, bi1, bj0, bj1, bk0, bk1};
            double B[3];
            gen_cellb(met.g, t, beta, B);
            x = B[0];
            y = B[1];
            z = B[2];
            const long n = met.g.ncell;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double wn = met.g.nc[(3 * a) * n + t] * u + met.g.nc[(3 * a + 1) * n + t] * v +
                                  met.g.nc[(3 * a + 2) * n + t] * w;
                if (met.w[a]) met.w[a][t] = wn;
                if (met.ws[a]) met.ws[a][t] = float(wn);
            }
        } else if (met.on) {
            const double fx = x, fy = y, fz = z;
            x = met.M[0] * fx + met.M[1] * fy + met.M[2] * fz;
            y = met.M[3] * fx + met.M[4] * fy + met.M[5] * fz;
            z = met.M[6] * fx + met.M[7] * fy + met.M[8] * fz;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double wn = met.n[3 * a] * u + met.n[3 * a + 1] * v + met.n[3 * a + 2] * w;
                if (met.w[a]) met.w[a][t] = wn;
                if (met.ws[a]) met.ws[a][t] = float(wn);
            }
        }
        if (bx) {  // null: the caller keeps cell B implicit (mixed Adams-Bashforth)
            bx[t] = x;
            by[t] = y;
            bz[t] = z;
        }
        if (vf || sh.rho) {
            const double f = sqrt(u * u + v * v + w * w) + sqrt((x * x + y * y + z * z) / r);
            if (vf) vf[t] = f;
            if (sh.rho) {
                sh.rho[t] = float(r);
                sh.vx[t] = float(u);
                sh.vy[t] = float(v);
                sh.vz[t] = float(w);
                sh.p[t] = float(pp);
                sh.bx[t] = float(x);
                sh.by[t] = float(y);
                sh.bz[t] = float(z);
                sh.vf[t] = float(f);
            }
        }
        if (part)
            dt = GEN ? cfl_dt_gen(gam, cfl, u, v, w, x, y, z, r, pp, met.g, t)
                            : cfl_dt(gam, cfl, u, v, w, x, y, z, r, pp, dx[i - 1], dy[j - 1], dz[k - 1]);
    }
    if (sh.bi) {
        if (t < nbi) sh.bi[t] = float(b.bi[t]);
        if (t < nbj) sh.bj[t] = float(b.bj[t]);
        if (t < nbk) sh.bk[t] = float(b.bk[t]);
    }
    if (part) {
        dt = block_min(dt);
        if (threadIdx.x == 0) part[blockIdx.x] = dt;
    }
}

// vf of an arbitrary state (get_eijk_LFM on a given state, mhd3d3.f90:1766)
__global__ void __launch_bounds__(kBlock) k_vf(Prim<double> q, double* __restrict__ vf, long ncell) {
    const long t = gtid();
    if (t >= ncell) return;
    const double u = q.vx[t], v = q.vy[t], w = q.vz[t], x = q.bx[t], y = q.by[t], z = q.bz[t];
    vf[t] = sqrt(u * u + v * v + w * w) + sqrt((x * x + y * y + z * z) / q.rho[t]);
}

// ------------------------------------------------------------------------
// CFL timestep: get_dt (mhd3d3.f90:1608-1639), min over the ghost-inclusive
// block.  Pass 1: per-workgroup min (wavefront shuffles + LDS); pass 2: one
// workgroup reduces the partials.  Deterministic, no atomics, no memset.
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_dt_partial(Dims d, double gam, double cfl, Prim<double> q,
                                                      const double* __restrict__ dx, const double* __restrict__ dy,
                                                      const double* __restrict__ dz, GMet gm,
                                                      double* __restrict__ part, long ncell) {
    const long t = gtid();
    double dt = __builtin_huge_val();
    if (t < ncell) {
        int i, j, k;
        unflat(t, d.nx, d.ny, i, j, k);
        dt = gm.ivol ? cfl_dt_gen(gam, cfl, q.vx[t], q.vy[t], q.vz[t], q.bx[t], q.by[t], q.bz[t], q.rho[t], q.p[t],
                                  gm, t)
                     : cfl_dt(gam, cfl, q.vx[t], q.vy[t], q.vz[t], q.bx[t], q.by[t], q.bz[t], q.rho[t], q.p[t],
                              dx[i - 1], dy[j - 1], dz[k - 1]);
    }
    dt = block_min(dt);
    if (threadIdx.x == 0) part[blockIdx.x] = dt;
}

// out[0] = min(part); if tau != null also tau[0] = out[0] (AB: tau from CFL)
// divergence monitor: -|div B| of each interior cell, as the update forms a
// flux divergence (tensor or affine metric: differences over the metric
// spacing; general metric: face fluxes times areas over the volume); the
// block minimum of the negated values goes to part[], k_dt_final then gives
// -max |div B|
template <bool GEN>
__global__ void __launch_bounds__(kBlock) k_divb_partial(Dims d, CFaces b, const double* __restrict__ dx,
                                                        const double* __restrict__ dy,
                                                        const double* __restrict__ dz, GMet gm,
                                                        double* __restrict__ part, long ncell) {
    const long t = gtid();
    double v = 0.;
    if (t < ncell) {
        int i, j, k;
        unflat(t, d.nx0, d.ny0, i, j, k);
        const int ng = d.ng;
        const ptrdiff_t q = ix(d.nx, d.ny, ng + i, ng + j, ng + k);
        const ptrdiff_t fi = ix(d.nx + 1, d.ny, ng + i, ng + j, ng + k), fj = ix(d.nx, d.ny + 1, ng + i, ng + j, ng + k);
        const ptrdiff_t fk1 = q + (ptrdiff_t)d.nx * d.ny;
        double div;
        if constexpr (GEN)
            div = ((b.bi[fi + 1] * gm.af[0][fi + 1] - b.bi[fi] * gm.af[0][fi]) +
                   (b.bj[fj + d.nx] * gm.af[1][fj + d.nx] - b.bj[fj] * gm.af[1][fj]) +
                   (b.bk[fk1] * gm.af[2][fk1] - b.bk[q] * gm.af[2][q])) * gm.ivol[q];
        else
            div = (b.bi[fi + 1] - b.bi[fi]) / dx[ng + i - 1] + (b.bj[fj + d.nx] - b.bj[fj]) / dy[ng + j - 1] +
                  (b.bk[fk1] - b.bk[q]) / dz[ng + k - 1];
        v = -fabs(div);
    }
    v = block_min(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void __launch_bounds__(1024) k_dt_final(const double* __restrict__ part, long npart,
                                                   double* __restrict__ out, double* __restrict__ tau) {
    __shared__ double red[16];
    // eight loads in flight per thread (one after the other, the 64 k partials
    // of a 256^3 block took 32 us); the minimum does not depend on the order
    constexpr int U = 8;
    double mu[U];
#pragma unroll
    for (int u = 0; u < U; ++u) mu[u] = __builtin_huge_val();
    const long bd = blockDim.x;
    long q = threadIdx.x;
    for (; q + (U - 1) * bd < npart; q += U * bd) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = part[q + u * bd];
#pragma unroll
        for (int u = 0; u < U; ++u) mu[u] = v[u] < mu[u] ? v[u] : mu[u];
    }
    for (; q < npart; q += bd) mu[0] = part[q] < mu[0] ? part[q] : mu[0];
    double m = mu[0];
#pragma unroll
    for (int u = 1; u < U; ++u) m = mu[u] < m ? mu[u] : m;
    m = wave_min(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) r = red[q] < r ? red[q] : r;
        out[0] = r;
        if (tau) tau[0] = r;
    }
}

__global__ void k_set2(double* __restrict__ p, double a, double b) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        p[0] = a;
        p[1] = b;
    }
}
template <int Z = 0>
__global__ void k_zero_int(int* __restrict__ p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = Z;
}
__global__ void k_tau0_from_tau(double* __restrict__ p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) p[1] = p[0];
}

// ------------------------------------------------------------------------
// advance_previous (mhd3d3.f90:3794-3819), RK flavour, full arrays
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_prev_rk(double gm1, Prim<double> q, double* __restrict__ rhop,
                                                   double* __restrict__ mxp, double* __restrict__ myp,
                                                   double* __restrict__ mzp, double* __restrict__ engp,
                                                   const double* __restrict__ bi, const double* __restrict__ bj,
                                                   const double* __restrict__ bk, double* __restrict__ bip,
                                                   double* __restrict__ bjp, double* __restrict__ bkp, long ncell,
                                                   long nbi, long nbj, long nbk) {
    const long t = gtid();
    if (t < ncell) {
        const double r = q.rho[t], u = q.vx[t], v = q.vy[t], w = q.vz[t];
        rhop[t] = r;
        mxp[t] = r * u;
        myp[t] = r * v;
        mzp[t] = r * w;
        engp[t] = .5 * r * (u * u + v * v + w * w) + q.p[t] / gm1;
    }
    if (t < nbi) bip[t] = bi[t];
    if (t < nbj) bjp[t] = bj[t];
    if (t < nbk) bkp[t] = bk[t];
}

// ------------------------------------------------------------------------
// Adams-Bashforth predictor (mhd3d3.f90:3821-3840):
//   xh = x + tau/tau0/2*(x - xp)   for the 11 state arrays, plus the h-state vf
// computed in binary64 and stored as TH (fp32 in mixed precision).  The
// reference's advance_previous copy xp = x is a buffer rotation on the host.
// ------------------------------------------------------------------------
template <typename TH>
struct AB11 {
    const double* x[11];   // rho vx vy vz p bx by bz bi bj bk
    const double* xp[11];
    TH* xh[11];
};

// face-normal velocities of the extrapolated state (metric only)
template <typename TH>
struct Proj {
    int on;
    double n[9];
    TH* w[3];
    const double* nc;  // general metric: per-cell normals (9 per cell, SoA), else null
    long ncell;
};

template <typename TH>
__device__ __forceinline__ void project(const Proj<TH>& pj, long t, double u, double v, double w) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double n0 = pj.nc ? pj.nc[(3 * a) * pj.ncell + t] : pj.n[3 * a];
        const double n1 = pj.nc ? pj.nc[(3 * a + 1) * pj.ncell + t] : pj.n[3 * a + 1];
        const double n2 = pj.nc ? pj.nc[(3 * a + 2) * pj.ncell + t] : pj.n[3 * a + 2];
        pj.w[a][t] = TH(n0 * u + n1 * v + n2 * w);
    }
}

template <typename TH>
__global__ void __launch_bounds__(kBlock) k_predict_ab(AB11<TH> s, const double* __restrict__ tau_p,
                                                      TH* __restrict__ vfh, Proj<TH> pj, long ncell, long nbi,
                                                      long nbj, long nbk) {
    const long t = gtid();
    const double c = tau_p[0] / tau_p[1] / 2.;
    if (t < ncell) {
        double h[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const double x = s.x[a][t], xp = s.xp[a][t];
            h[a] = x + c * (x - xp);
            s.xh[a][t] = TH(h[a]);
        }
        vfh[t] = TH(sqrt(h[1] * h[1] + h[2] * h[2] + h[3] * h[3]) + sqrt((h[5] * h[5] + h[6] * h[6] + h[7] * h[7]) / h[0]));
        if (pj.on) project(pj, t, h[1], h[2], h[3]);
    }
    const long nf[3] = {nbi, nbj, nbk};
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (t < nf[a]) {
            const double x = s.x[8 + a][t], xp = s.xp[8 + a][t];
            s.xh[8 + a][t] = TH(x + c * (x - xp));
        }
}

// Mixed-precision Adams-Bashforth: the same extrapolation, with the cell B
// of the extrapolated state taken from its extrapolated faces
// (M (face averages), as get_bxyz does) instead of from the cell B of the
// current and previous states -- equal in exact arithmetic, and the cell B
// arrays need not be stored each step.  Cell threads read their faces'
// neighbours (f+1 along each axis), which the neighbouring threads also
// read: the extra taps hit the caches.
struct FaceB {
    const double *bi, *bj, *bk, *bip, *bjp, *bkp;
};

// threads: one per cell (ghost-inclusive), then one per face that is no
// cell's lower face (i = nx+1 of bi, j = ny+1 of bj, k = nz+1 of bk).  A
// cell's thread extrapolates its lower faces once and uses them both for
// the face copies and for its cell B, so every face is read once, plus the
// upper neighbours of cell B (row and plane neighbours, mostly from L2).
__global__ void __launch_bounds__(kBlock) k_predict_ab_mixed(Dims d, AB11<float> s, FaceB fb, BMet met,
                                                            const double* __restrict__ tau_p,
                                                            float* __restrict__ vfh, Proj<float> pj, long ncell,
                                                            long nbi, long nbj, long nbk) {
    const long t = gtid_xcd();
    const double c = tau_p[0] / tau_p[1] / 2.;
    auto ext = [&](const double* x, const double* xp, ptrdiff_t f) { return x[f] + c * (x[f] - xp[f]); };
    if (t < ncell) {
        double h[8];
#pragma unroll
        for (int a = 0; a < 5; ++a) {
            const double x = s.x[a][t], xp = s.xp[a][t];
            h[a] = x + c * (x - xp);
            s.xh[a][t] = float(h[a]);
        }
        int i, j, k;
        unflat(t, d.nx, d.ny, i, j, k);
        const ptrdiff_t fi = ix(d.nx + 1, d.ny, i, j, k), fj = ix(d.nx, d.ny + 1, i, j, k);
        const double bi0 = ext(fb.bi, fb.bip, fi), bi1 = ext(fb.bi, fb.bip, fi + 1);
        const double bj0 = ext(fb.bj, fb.bjp, fj), bj1 = ext(fb.bj, fb.bjp, fj + d.nx);
        const double bk0 = ext(fb.bk, fb.bkp, t), bk1 = ext(fb.bk, fb.bkp, t + (ptrdiff_t)d.nx * d.ny);
        s.xh[8][fi] = float(bi0);
        s.xh[9][fj] = float(bj0);
        s.xh[10][t] = float(bk0);
        const double ax = .5 * (bi1 + bi0);
        const double ay = .5 * (bj1 + bj0);
        const double az = .5 * (bk1 + bk0);
        if (met.g.ivol) {
            const double beta[6] = {bi0, bi1, bj0, bj1, bk0, bk1};
            gen_cellb(met.g, t, beta, h + 5);
        } else if (met.on) {
            h[5] = met.M[0] * ax + met.M[1] * ay + met.M[2] * az;
            h[6] = met.M[3] * ax + met.M[4] * ay + met.M[5] * az;
            h[7] = met.M[6] * ax + met.M[7] * ay + met.M[8] * az;
        } else {
            h[5] = ax;
            h[6] = ay;
            h[7] = az;
        }
#pragma unroll
        for (int a = 5; a < 8; ++a) s.xh[a][t] = float(h[a]);
        vfh[t] = float(sqrt(h[1] * h[1] + h[2] * h[2] + h[3] * h[3]) + sqrt((h[5] * h[5] + h[6] * h[6] + h[7] * h[7]) / h[0]));
        if (pj.on) project(pj, t, h[1], h[2], h[3]);
        return;
    }
    // the faces on the upper boundary of each face family
    long e = t - ncell;
    const long nyz = (long)d.ny * d.nz, nxz = (long)d.nx * d.nz, nxy = (long)d.nx * d.ny;
    if (e < nyz) {
        const ptrdiff_t f = ix(d.nx + 1, d.ny, d.nx + 1, (int)(e % d.ny) + 1, (int)(e / d.ny) + 1);
        s.xh[8][f] = float(ext(fb.bi, fb.bip, f));
        return;
    }
    e -= nyz;
    if (e < nxz) {
        const ptrdiff_t f = ix(d.nx, d.ny + 1, (int)(e % d.nx) + 1, d.ny + 1, (int)(e / d.nx) + 1);
        s.xh[9][f] = float(ext(fb.bj, fb.bjp, f));
        return;
    }
    e -= nxz;
    if (e < nxy) {
        const ptrdiff_t f = ix(d.nx, d.ny, (int)(e % d.nx) + 1, (int)(e / d.nx) + 1, d.nz + 1);
        s.xh[10][f] = float(ext(fb.bk, fb.bkp, f));
    }
}

// ------------------------------------------------------------------------
// state diagnostics (diag_rho_neg, diag_overflow, mhd3d3.f90:3143-3160) plus
// a non-positive pressure check: flags |= 1 (rho<0 anywhere), 2 (NaN in
// interior rho), 4 (p<=0 or NaN in interior p)
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_check(Dims d, const double* __restrict__ rho,
                                                 const double* __restrict__ p, int* __restrict__ flags,
                                                 long ncell) {
    const long t = gtid();
    int f = 0;
    if (t < ncell) {
        int i, j, k;
        unflat(t, d.nx, d.ny, i, j, k);
        const double r = rho[t];
        if (r < 0.) f |= 1;
        const bool in = i > d.ng && i <= d.ng + d.nx0 && j > d.ng && j <= d.ng + d.ny0 && k > d.ng &&
                        k <= d.ng + d.nz0;
        if (in && r != r) f |= 2;
        const double pp = p[t];
        if (in && !(pp > 0.)) f |= 4;
    }
    // wave-level OR, then one atomic per wave
    for (int o = 32; o > 0; o >>= 1) f |= __shfl_xor(f, o, 64);
    if ((threadIdx.x & 63) == 0 && f) atomicOr(flags, f);
}

}  // namespace mhd
