// get_eijk_Yee (mhd3d3.f90:1654-1752): the reference's alternative edge EMF,
// which splits the cell-centred electric field by the Gaussian half-moments
// of get_v_center (2562-2592) and reconstructs each of the twelve parts
// twice, face then edge, with reconstruct_3d's untruncated directions 4-6.
// Binary64, compiled without contraction in the same unit as the flux: the
// split is elementwise here, the reconstructions are k_recon (the bit-exact
// reconstruct_3d), the four quadrant sums a last elementwise pass.
#pragma once
#include "mhd_launch.hpp"

namespace mhd {

struct YeeSplit {
    double* e[3][4];  // [edge direction][YEE_*], each a ghost-inclusive cell array
};

__global__ void __launch_bounds__(kBlock) k_yee_split(long n, Prim<double> q, double rspi, YeeSplit o) {
    const long t = gtid();
    if (t >= n) return;
    const double bx = q.bx[t], by = q.by[t], bz = q.bz[t];
    // bsq=lbx**2+lby**2+lbz**2; lam=lrho/(2.*lp+bsq) (1669-1670)
    const double bsq = bx * bx + by * by + bz * bz;
    const double lam = q.rho[t] / (2. * q.p[t] + bsq);
    // get_v_center: the same moments as get_v_face (vface)
    double m0[3][2], m1[3][2];  // [velocity][side p, n]
    const double v[3] = {q.vx[t], q.vy[t], q.vz[t]};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        vface<double>(lam, v[a], +1, rspi, m0[a][0], m1[a][0]);
        vface<double>(lam, v[a], -1, rspi, m0[a][1], m1[a][1]);
    }
    const double B[3] = {bx, by, bz};
    // ezpp=lbx*vx0p*vy1p-lby*vx1p*vy0p and its cyclic kin (1673-1691)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int b = (a + 1) % 3, c = (a + 2) % 3;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int s1 = s & 1, s2 = s >> 1;  // YEE_PP, NP, PN, NN
            o.e[a][s][t] = B[b] * m0[b][s1] * m1[c][s2] - B[c] * m1[b][s1] * m0[c][s2];
        }
    }
}

// e<a>(i,j,k) = epp + enp + epn + enn at the edge's owned index, the sums'
// arrays untruncated along axis `ax` (offset ng there): ei(i,:,:) =
// epp(ng+i,:,:)+enp(ng+i,:,:)+epn(ng+i,:,:)+enn(ng+i,:,:) (1717) and its kin
__global__ void __launch_bounds__(kBlock) k_yee_sum(int o1, int o2, int o3, int ax, int ng, int n1, int n2,
                                                   const double* __restrict__ epp, const double* __restrict__ enp,
                                                   const double* __restrict__ epn, const double* __restrict__ enn,
                                                   double* __restrict__ out) {
    const long t = gtid();
    if (t >= (long)o1 * o2 * o3) return;
    const int i = (int)(t % o1), r = (int)(t / o1), j = r % o2, k = r / o2;
    const long s = (long)(i + (ax == 0 ? ng : 0)) +
                   (long)n1 * ((long)(j + (ax == 1 ? ng : 0)) + (long)n2 * (k + (ax == 2 ? ng : 0)));
    out[t] = epp[s] + enp[s] + epn[s] + enn[s];
}

}  // namespace mhd
