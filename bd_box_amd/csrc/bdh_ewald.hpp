// Beenakker-Ewald RPY blocks on the device (src/diff_alg/diff_tensor.c:265-422).
//
// alpha = sqrt(pi)/L (forced by init_tensor, diff_tensor.c:551-554); real and reciprocal
// lattices share the sphere h^2 + f^2 + g^2 <= R^2 (ewald_real == ewald_recip,
// src/input.c:402-409).  The lattice vectors and every pair-independent factor of the
// reciprocal term are tabulated once on the host in FP64 (with the reference's float pi):
//   G_k = 6 pi (1 + (k^2/4a^2)(1 + k^2/2a^2)) exp(-k^2/4a^2) / k^2,  Rk = k^2 / 3,
//   k = 2 pi (h, f, g) / L  (in units where the box edge is folded into preQ = 1/L^3)
// so a pair's reciprocal sum is sum_k (1 - sigma^2 Rk) G_k cos(2 pi v.hfg) (I - e_k e_k^T),
// with v = r / L.  The reciprocal and self sums run in FP64 in both precisions; the
// real-space image sum runs in T.
#pragma once

#include <hip/hip_runtime.h>

#include "bdh_common.hpp"

namespace bdh {

constexpr int EW_MAXV = 700;  // lattice vectors with |hfg| <= 5

struct EwaldTables {
    int nv;                  // vectors in the sphere (including 0)
    int R;                   // lattice sphere radius
    double box, inv_L, alpha, alpha2, alpha3, inv_alpha2, preQ;
    double self_a, self_b;   // self term: -6 alpha/sqrt(pi) and 40 alpha^3/(3 sqrt(pi))
};

extern __constant__ EwaldTables c_ew;
// per lattice vector, h then f then g ascending over the sphere (the reference's order):
extern __constant__ double4 c_recip[EW_MAXV];    // (G_k, Rk, e_x, e_y); e_z in c_recip_z
extern __constant__ double c_recip_z[EW_MAXV];

// Upload the tables for box edge L and radius R; returns the vector count.
int ewald_upload(double box, int R);

struct Blk6 {
    double xx, yy, zz, xy, xz, yz;
};

// (re, im) *= (c + i s) and (re, im) *= (c - i s)
__device__ __forceinline__ void cmul_step(double& re, double& im, double c, double s)
{
    const double r = re * c - im * s;
    im = re * s + im * c;
    re = r;
}
__device__ __forceinline__ void cmul_conj_step(double& re, double& im, double c, double s)
{
    const double r = re * c + im * s;
    im = im * c - re * s;
    re = r;
}

// floor(sqrt(m)) for 0 <= m <= 25 (lattice radius <= 5)
__device__ __forceinline__ int isqrt_small(int m)
{
    int r = 0;
    while ((r + 1) * (r + 1) <= m) ++r;
    return r;
}

// Beenakker block for the pair (i, j): r = r_j - r_i (no minimum image, dist_vec), radii
// ai, aj; self = (i == j) adds the self term and clears the off-diagonal entries.
template <typename T>
__device__ __forceinline__ Blk6 beenakker_block(double rx, double ry, double rz, double ai,
                                                double aj, bool self, double preDii,
                                                double preDij)
{
    const EwaldTables& E = c_ew;
    const double sigma_sq = 0.5 * (ai * ai + aj * aj);
    double O[6] = {0, 0, 0, 0, 0, 0};  // xx yy zz xy xz yz
    double Qr[6] = {0, 0, 0, 0, 0, 0};
    Blk6 out{0, 0, 0, 0, 0, 0};
    const double r2 = rx * rx + ry * ry + rz * rz;
    const double s2 = (ai + aj) * (ai + aj);
    if (!self && r2 < s2) {
        // overlap correction: + overlapping RPY - non-overlapping RPY (diff_tensor.c:279-313)
        const double rn = sqrt(r2), inv = 1.0 / rn;
        const double ex = rx * inv, ey = ry * inv, ez = rz * inv;
        const double sig = 0.5 * (ai + aj);
        const double pre1 = preDii / sig;
        const double part1 = 1.0 - 9.0 * rn / (32.0 * sig);
        const double c1 = 3.0 * rn / (32.0 * sig);
        const double pre2 = preDij / rn;
        const double sr = (ai * ai + aj * aj) / r2;
        const double d2 = 1.0 + sr / 3.0, o2 = 1.0 - sr;
        out.xx += pre1 * (part1 + c1 * ex * ex) - pre2 * (d2 + ex * ex * o2);
        out.yy += pre1 * (part1 + c1 * ey * ey) - pre2 * (d2 + ey * ey * o2);
        out.zz += pre1 * (part1 + c1 * ez * ez) - pre2 * (d2 + ez * ez * o2);
        out.xy += pre1 * c1 * ex * ey - pre2 * ex * ey * o2;
        out.xz += pre1 * c1 * ex * ez - pre2 * ex * ez * o2;
        out.yz += pre1 * c1 * ey * ez - pre2 * ey * ez * o2;
    }
    const double vx = rx * E.inv_L, vy = ry * E.inv_L, vz = rz * E.inv_L;
    const T al = (T)E.alpha, al2 = (T)E.alpha2, al3 = (T)E.alpha3, sg = (T)sigma_sq;
    const T inv_sqrtpi = (T)(1.0 / SQRTPI);
    // Reciprocal phases cos 2 pi v.(h, f, g) = Re(e^{i h tx} e^{i f ty} e^{i g tz}): one
    // sincos per axis, the exponentials stepped along the lattice loops, which visit the
    // vectors of the sphere in the table's order (h, then f, then g ascending).
    double sx, cx, sy, cy, sz, cz;
    sincos(2.0 * PI_F * vx, &sx, &cx);
    sincos(2.0 * PI_F * vy, &sy, &cy);
    sincos(2.0 * PI_F * vz, &sz, &cz);
    const int R = E.R, RR = R * R;
    double hr = 1.0, hi = 0.0;
    for (int k = 0; k < R; ++k) cmul_conj_step(hr, hi, cx, sx);
    int q = 0;
    for (int h = -R; h <= R; ++h) {
        const int nf = isqrt_small(RR - h * h);
        double fr = 1.0, fi = 0.0;
        for (int k = 0; k < nf; ++k) cmul_conj_step(fr, fi, cy, sy);
        for (int f = -nf; f <= nf; ++f) {
            const int ng = isqrt_small(RR - h * h - f * f);
            const double hfr = hr * fr - hi * fi, hfi = hr * fi + hi * fr;
            double gr = 1.0, gi = 0.0;
            for (int k = 0; k < ng; ++k) cmul_conj_step(gr, gi, cz, sz);
            for (int g = -ng; g <= ng; ++g, ++q) {
                // real space image r + L (h, f, g)
                const T x = (T)(rx + E.box * h), y = (T)(ry + E.box * f), z = (T)(rz + E.box * g);
                const T rr2 = x * x + y * y + z * z;
                if (rr2 != (T)0) {
                    const T rn = sqrt(rr2);
                    const T inv_r = (T)1 / rn, inv_r2 = (T)1 / rr2, inv_r3 = inv_r * inv_r2;
                    const T ex = x * inv_r, ey = y * inv_r, ez = z * inv_r;
                    const T eg = exp(-al2 * rr2) * inv_sqrtpi;
                    const T ec = erfc(rn * al);
                    T sc = al * sg * inv_r2 + (T)14 * al3 * sg - (T)4.5 * al - (T)20 * al2 * al3 * sg * rr2 +
                           (T)3 * al3 * rr2 + (T)4 * al3 * al2 * al2 * sg * rr2 * rr2;
                    sc = sc * eg + ec * ((T)0.75 * inv_r + (T)0.5 * sg * inv_r3);
                    T sc2 = (T)-3 * al * sg * inv_r2 - (T)2 * al3 * sg + (T)1.5 * al + (T)16 * al2 * al3 * sg * rr2 -
                            (T)3 * al3 * rr2 - (T)4 * al3 * al2 * al2 * sg * rr2 * rr2;
                    sc2 = sc2 * eg + ec * ((T)0.75 * inv_r - (T)1.5 * sg * inv_r3);
                    O[0] += (double)(sc + sc2 * ex * ex);
                    O[1] += (double)(sc + sc2 * ey * ey);
                    O[2] += (double)(sc + sc2 * ez * ez);
                    O[3] += (double)(sc2 * ex * ey);
                    O[4] += (double)(sc2 * ex * ez);
                    O[5] += (double)(sc2 * ey * ez);
                }
                // reciprocal space (FP64)
                if (h != 0 || f != 0 || g != 0) {
                    const double4 gk = c_recip[q];
                    const double ekz = c_recip_z[q];
                    const double sc = (1.0 - sigma_sq * gk.y) * gk.x * (hfr * gr - hfi * gi);
                    Qr[0] += sc - sc * gk.z * gk.z;
                    Qr[1] += sc - sc * gk.w * gk.w;
                    Qr[2] += sc - sc * ekz * ekz;
                    Qr[3] -= sc * gk.z * gk.w;
                    Qr[4] -= sc * gk.z * ekz;
                    Qr[5] -= sc * gk.w * ekz;
                }
                cmul_step(gr, gi, cz, sz);
            }
            cmul_step(fr, fi, cy, sy);
        }
        cmul_step(hr, hi, cx, sx);
    }
    out.xx += preDii * (O[0] + E.preQ * Qr[0]);
    out.yy += preDii * (O[1] + E.preQ * Qr[1]);
    out.zz += preDii * (O[2] + E.preQ * Qr[2]);
    out.xy += preDii * (O[3] + E.preQ * Qr[3]);
    out.xz += preDii * (O[4] + E.preQ * Qr[4]);
    out.yz += preDii * (O[5] + E.preQ * Qr[5]);
    if (self) {
        // diff_tensor.c:403-422
        const double sc3 = preDii * (E.self_a + E.self_b * ai * ai);
        out.xx = preDii / ai + sc3 + out.xx;
        out.yy = preDii / ai + sc3 + out.yy;
        out.zz = preDii / ai + sc3 + out.zz;
        out.xy = out.xz = out.yz = 0.0;
    }
    return out;
}

}  // namespace bdh
