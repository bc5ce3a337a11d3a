// Ewald-summed RPY blocks on the device: Beenakker (src/diff_alg/diff_tensor.c:265-422)
// and Smith (:102-263).
//
// alpha = sqrt(pi)/L (forced by init_tensor, diff_tensor.c:551-554); real and reciprocal
// lattices share the sphere h^2 + f^2 + g^2 <= R^2 (ewald_real == ewald_recip,
// src/input.c:402-409), around the pair's separation as the wrapped positions give it (no
// minimum image).  An extension beyond the reference (bdh_set_ewald_split) scales alpha,
// gives the real and reciprocal sums spheres of their own radii (and the real-space images
// a distance cutoff, bdh_set_ewald_cutoff) and centres the image sum on the minimum-image
// separation, so the omitted images of every pair lie at least (sqrt(R_real^2 + 1) -
// sqrt(3)/2) L away, and forms the wave vectors k = 2 pi (h, f, g) / L of the box (the
// reference's k = 2 pi (h, f, g) is in box units against alpha and sigma in A, so its
// reciprocal terms underflow to zero).  The infinite sums do not depend on alpha or on the
// image chosen as the centre (the reciprocal phases are periodic in the separation); the
// truncated ones do.  The lattice vectors and every pair-independent factor of the
// reciprocal term are tabulated once on the host in FP64 (with the reference's float pi;
// the extension's with the double pi):
//   G_k = 6 pi (1 + (k^2/4a^2)(1 + k^2/2a^2)) exp(-k^2/4a^2) / k^2,  Rk = k^2 / 3,
//   k = 2 pi (h, f, g)  (reference),  k = 2 pi (h, f, g) / L  (extension)
// so a pair's reciprocal sum is sum_k (1 - sigma^2 Rk) G_k cos(2 pi v.hfg) (I - e_k e_k^T),
// with v = r / L, over the half of the sphere with (h, f, g) > 0 lexicographically and 2 G_k
// in the table (k and -k give the same term); the real-space images and the reciprocal
// vectors have loops of their own.  The reciprocal and self sums run in FP64 in both precisions; the
// real-space image sum runs in T.
#pragma once

#include <hip/hip_runtime.h>

#include "bdh_common.hpp"

namespace bdh {

constexpr int EW_MAXV = 700;  // lattice vectors with |hfg| <= 5

enum EwaldMethod { EW_SMITH = 0, EW_BEENAKKER = 1 };

struct EwaldTables {
    int nv;                  // vectors in the sphere (including 0)
    int R;                   // lattice sphere radius of the loops
    int RR_real;             // real-space images with h^2 + f^2 + g^2 <= RR_real
    int min_image;           // image sum around the minimum-image separation (extension)
    double rc2;              // extension: real-space images within this squared distance only
    int RR_recip;            // Beenakker: reciprocal half sphere |(h, f, g)|^2 <= RR_recip
    int nrecip;              // Beenakker: vectors of that half sphere in c_recip (0: none)
    int method;              // EwaldMethod
    double box, inv_L, alpha, alpha2, alpha3, inv_alpha2, preQ;
    double self_a, self_b;   // Beenakker self term: -6 alpha/sqrt(pi), 40 alpha^3/(3 sqrt(pi))
    double preO;             // Smith: 0.75 / L
};

extern __constant__ EwaldTables c_ew;
// per lattice vector, h then f then g ascending over the sphere (the reference's order):
//   Beenakker: (G_k, Rk, e_x, e_y);  Smith: (2 exp(m k^2) / (pi k^2), -1 + m k^2, e_x, e_y) with
//   m = -pi^2 / alpha^2 and k^2 = h^2 + f^2 + g^2, and 4 pi exp(m k^2) in c_recip_w
extern __constant__ double4 c_recip[EW_MAXV];
extern __constant__ double c_recip_z[EW_MAXV];   // e_z
extern __constant__ double c_recip_w[EW_MAXV];

// Upload the tables of one method for box edge L, lattice radius R and (Smith) the
// splitting parameter alpha in units of 1/L; returns the vector count.  Beenakker with
// alpha_scale > 0 (the extension): alpha = alpha_scale sqrt(pi) / L, spheres of radius
// R_real (images) and R_recip (reciprocal vectors) instead of R, minimum-image centres, and
// with r_cut > 0 only the images within r_cut (A) of the pair.
int ewald_upload(double box, int R, int method, double alpha, double alpha_scale = 0.0, int R_real = -1,
                 int R_recip = -1, double r_cut = 0.0);

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

// erfc(x) for x >= 0 given ex = exp(-x^2).  FP64: the library erfc.  FP32 (the mixed
// step's real-space images): Abramowitz & Stegun 7.1.26, erfc(x) = t (a1 + t (a2 + t (a3 +
// t (a4 + t a5)))) exp(-x^2), t = 1 / (1 + p x), absolute error <= 1.5e-7 -- the exponential
// is the one the image term needs anyway, so no second transcendental.
template <typename T>
__device__ __forceinline__ T erfc_with_exp(T x, T ex);
template <>
__device__ __forceinline__ double erfc_with_exp<double>(double x, double) { return erfc(x); }
template <>
__device__ __forceinline__ float erfc_with_exp<float>(float x, float ex)
{
    const float t = 1.0f / (1.0f + 0.3275911f * x);
    const float poly = t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f +
                       t * (-1.453152027f + t * 1.061405429f))));
    return poly * ex;
}

// r, 1/r, 1/r^2 and 1/r^3 of an image at squared distance r2: FP64 with IEEE sqrt and
// divisions; FP32 from one hardware reciprocal square root (about 1 ulp).
template <typename T>
struct InvPowers {
    T r, inv_r, inv_r2, inv_r3;
};
__device__ __forceinline__ InvPowers<double> inv_powers(double r2)
{
    const double r = sqrt(r2), inv_r = 1.0 / r, inv_r2 = 1.0 / r2;
    return {r, inv_r, inv_r2, inv_r * inv_r2};
}
__device__ __forceinline__ InvPowers<float> inv_powers(float r2)
{
    const float inv_r = __builtin_amdgcn_rsqf(r2), inv_r2 = inv_r * inv_r;
    return {r2 * inv_r, inv_r, inv_r2, inv_r * inv_r2};
}

// floor(sqrt(m)) for 0 <= m <= 25 (lattice radius <= 5)
__device__ __forceinline__ int isqrt_small(int m)
{
    int r = 0;
    while ((r + 1) * (r + 1) <= m) ++r;
    return r;
}

// Overlapping beads: + overlapping RPY - non-overlapping RPY, which the lattice sums count
// (diff_tensor.c:115-148 Smith, :279-313 Beenakker; the same terms).
__device__ __forceinline__ void overlap_correction(double rx, double ry, double rz, double r2,
                                                   double ai, double aj, double preDii,
                                                   double preDij, Blk6& out)
{
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

// Beenakker block for the pair (i, j): r = r_j - r_i (no minimum image, dist_vec), radii
// ai, aj; self = (i == j) adds the self term and clears the off-diagonal entries.
template <typename T>
__device__ __forceinline__ Blk6 beenakker_block(double rx, double ry, double rz, double ai,
                                                double aj, bool self, double preDii,
                                                double preDij)
{
    const EwaldTables& E = c_ew;
    if (E.min_image) {
        rx -= E.box * rint(rx * E.inv_L);
        ry -= E.box * rint(ry * E.inv_L);
        rz -= E.box * rint(rz * E.inv_L);
    }
    const double sigma_sq = 0.5 * (ai * ai + aj * aj);
    double O[6] = {0, 0, 0, 0, 0, 0};  // xx yy zz xy xz yz
    double Qr[6] = {0, 0, 0, 0, 0, 0};
    Blk6 out{0, 0, 0, 0, 0, 0};
    const double r2 = rx * rx + ry * ry + rz * rz;
    const double s2 = (ai + aj) * (ai + aj);
    if (!self && r2 < s2) overlap_correction(rx, ry, rz, r2, ai, aj, preDii, preDij, out);
    const T al = (T)E.alpha, al2 = (T)E.alpha2, al3 = (T)E.alpha3, sg = (T)sigma_sq;
    const T inv_sqrtpi = (T)(1.0 / SQRTPI);
    // real space: images r + L (h, f, g) of the sphere h^2 + f^2 + g^2 <= RR_real (h, then f,
    // then g ascending), within sqrt(rc2) of the pair, summed in T and added to O once
    T Ot[6] = {0, 0, 0, 0, 0, 0};
    const int RRr = E.RR_real, Rr = isqrt_small(RRr);
    for (int h = -Rr; h <= Rr; ++h) {
        const int nf = isqrt_small(RRr - h * h);
        for (int f = -nf; f <= nf; ++f) {
            const int ng = isqrt_small(RRr - h * h - f * f);
            for (int g = -ng; g <= ng; ++g) {
                const T x = (T)(rx + E.box * h), y = (T)(ry + E.box * f), z = (T)(rz + E.box * g);
                const T rr2 = x * x + y * y + z * z;
                if (rr2 != (T)0 && rr2 <= (T)E.rc2) {
                    const InvPowers<T> ip = inv_powers(rr2);
                    const T rn = ip.r, inv_r = ip.inv_r, inv_r2 = ip.inv_r2, inv_r3 = ip.inv_r3;
                    const T ex = x * inv_r, ey = y * inv_r, ez = z * inv_r;
                    const T ex2 = exp(-al2 * rr2);
                    const T eg = ex2 * inv_sqrtpi;
                    const T ec = erfc_with_exp<T>(rn * al, ex2);
                    T sc = al * sg * inv_r2 + (T)14 * al3 * sg - (T)4.5 * al - (T)20 * al2 * al3 * sg * rr2 +
                           (T)3 * al3 * rr2 + (T)4 * al3 * al2 * al2 * sg * rr2 * rr2;
                    sc = sc * eg + ec * ((T)0.75 * inv_r + (T)0.5 * sg * inv_r3);
                    T sc2 = (T)-3 * al * sg * inv_r2 - (T)2 * al3 * sg + (T)1.5 * al + (T)16 * al2 * al3 * sg * rr2 -
                            (T)3 * al3 * rr2 - (T)4 * al3 * al2 * al2 * sg * rr2 * rr2;
                    sc2 = sc2 * eg + ec * ((T)0.75 * inv_r - (T)1.5 * sg * inv_r3);
                    Ot[0] += sc + sc2 * ex * ex;
                    Ot[1] += sc + sc2 * ey * ey;
                    Ot[2] += sc + sc2 * ez * ez;
                    Ot[3] += sc2 * ex * ey;
                    Ot[4] += sc2 * ex * ez;
                    Ot[5] += sc2 * ey * ez;
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) O[k] += (double)Ot[k];
    // reciprocal space (FP64): a vector and its negative carry the same term (the phase's
    // cosine is even), so the loops visit the half sphere (h, f, g) > 0 lexicographically,
    // |.|^2 <= RR_recip, in the table's order, with the factor 2 in G.  The phases
    // cos 2 pi v.(h, f, g) = Re(e^{i h tx} e^{i f ty} e^{i g tz}) come from one sincos per
    // axis, the exponentials stepped along the loops.  None when every G is zero (the
    // reference's split, whose G underflow).
    if (E.nrecip > 0) {
        const double vx = rx * E.inv_L, vy = ry * E.inv_L, vz = rz * E.inv_L;
        double sx, cx, sy, cy, sz, cz;
        sincos(2.0 * PI_F * vx, &sx, &cx);
        sincos(2.0 * PI_F * vy, &sy, &cy);
        sincos(2.0 * PI_F * vz, &sz, &cz);
        const int RRk = E.RR_recip, Rk = isqrt_small(RRk);
        double hr = 1.0, hi = 0.0;
        int q = 0;
        for (int h = 0; h <= Rk; ++h) {
            const int nf = isqrt_small(RRk - h * h);
            const int f0 = h == 0 ? 0 : -nf;
            double fr = 1.0, fi = 0.0;
            for (int k = 0; k < -f0; ++k) cmul_conj_step(fr, fi, cy, sy);
            for (int f = f0; f <= nf; ++f) {
                const int ng = isqrt_small(RRk - h * h - f * f);
                const int g0 = (h == 0 && f == 0) ? 1 : -ng;
                const double hfr = hr * fr - hi * fi, hfi = hr * fi + hi * fr;
                double gr = 1.0, gi = 0.0;
                if (g0 > 0) cmul_step(gr, gi, cz, sz);
                for (int k = 0; k < -g0; ++k) cmul_conj_step(gr, gi, cz, sz);
                for (int g = g0; g <= ng; ++g, ++q) {
                    const double4 gk = c_recip[q];
                    const double ekz = c_recip_z[q];
                    const double sc = (1.0 - sigma_sq * gk.y) * gk.x * (hfr * gr - hfi * gi);
                    Qr[0] += sc - sc * gk.z * gk.z;
                    Qr[1] += sc - sc * gk.w * gk.w;
                    Qr[2] += sc - sc * ekz * ekz;
                    Qr[3] -= sc * gk.z * gk.w;
                    Qr[4] -= sc * gk.z * ekz;
                    Qr[5] -= sc * gk.w * ekz;
                    cmul_step(gr, gi, cz, sz);
                }
                cmul_step(fr, fi, cy, sy);
            }
            cmul_step(hr, hi, cx, sx);
        }
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

// Smith-Ewald block (src/diff_alg/diff_tensor.c:102-263) in reduced units v = r / L with
// the .prm alpha (EwaldTables.alpha): per lattice vector (h, f, g) of the sphere,
//   real space, x = v + (h, f, g), |x| = r, e = x / r (in T):
//     Op += (erfc(a r)/r) I + (erfc(a r)/r + 2 a exp(-a^2 r^2)/sqrt(pi)) e e^T
//     Qp += s I + (-3 s - 4 a^3 exp(-a^2 r^2)/sqrt(pi)) e e^T,
//           s = (erfc(a r) + 2 a r exp(-a^2 r^2)/sqrt(pi)) / r^3
//   reciprocal, (h, f, g) != 0, k^2 = h^2 + f^2 + g^2 (FP64, tabulated factors):
//     Op += G1 cos I + (-1 + m k^2) G1 cos e e^T,  Qp += G2 cos e e^T,
//     cos = cos 2 pi v.(h, f, g)
//   D_ij += preDii (0.75/L Op + sigma^2/(2 L^3) Qp); self: + preDii/a - self term.
// 1/sqrt(pi) is the reference's M1_SQRTPI = 0.5641895835 (src/data.h:228).
template <typename T>
__device__ __forceinline__ Blk6 smith_block(double rx, double ry, double rz, double ai,
                                            double aj, bool self, double preDii, double preDij)
{
    const EwaldTables& E = c_ew;
    const double sigma_sq = 0.5 * (ai * ai + aj * aj);
    double Op[6] = {0, 0, 0, 0, 0, 0};  // xx yy zz xy xz yz
    double Qp[6] = {0, 0, 0, 0, 0, 0};
    Blk6 out{0, 0, 0, 0, 0, 0};
    const double r2 = rx * rx + ry * ry + rz * rz;
    const double s2 = (ai + aj) * (ai + aj);
    if (!self && r2 < s2) overlap_correction(rx, ry, rz, r2, ai, aj, preDii, preDij, out);
    const double vx = rx * E.inv_L, vy = ry * E.inv_L, vz = rz * E.inv_L;
    const T al = (T)E.alpha, al2 = (T)E.alpha2, al3 = (T)E.alpha3;
    const T m1sp = (T)0.5641895835;
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
                const T x = (T)(vx + h), y = (T)(vy + f), z = (T)(vz + g);
                const T rr2 = x * x + y * y + z * z;
                if (rr2 != (T)0) {
                    const InvPowers<T> ip = inv_powers(rr2);
                    const T rn = ip.r, inv_r = ip.inv_r, inv_r3 = ip.inv_r3;
                    const T ex = x * inv_r, ey = y * inv_r, ez = z * inv_r;
                    const T eg = exp(-al2 * rr2);
                    const T ec = erfc_with_exp<T>(rn * al, eg);
                    const T so = ec * inv_r;
                    const T so2 = so + (T)2 * al * m1sp * eg;
                    Op[0] += (double)(so + so2 * ex * ex);
                    Op[1] += (double)(so + so2 * ey * ey);
                    Op[2] += (double)(so + so2 * ez * ez);
                    Op[3] += (double)(so2 * ex * ey);
                    Op[4] += (double)(so2 * ex * ez);
                    Op[5] += (double)(so2 * ey * ez);
                    const T sq = inv_r3 * (ec + (T)2 * al * rn * m1sp * eg);
                    const T sq2 = (T)-3 * sq - (T)4 * al3 * m1sp * eg;
                    Qp[0] += (double)(sq + sq2 * ex * ex);
                    Qp[1] += (double)(sq + sq2 * ey * ey);
                    Qp[2] += (double)(sq + sq2 * ez * ez);
                    Qp[3] += (double)(sq2 * ex * ey);
                    Qp[4] += (double)(sq2 * ex * ez);
                    Qp[5] += (double)(sq2 * ey * ez);
                }
                if (h != 0 || f != 0 || g != 0) {
                    const double4 gk = c_recip[q];
                    const double ekz = c_recip_z[q];
                    const double cosv = hfr * gr - hfi * gi;
                    const double so = gk.x * cosv;
                    const double so2 = gk.y * so;
                    Op[0] += so + so2 * gk.z * gk.z;
                    Op[1] += so + so2 * gk.w * gk.w;
                    Op[2] += so + so2 * ekz * ekz;
                    Op[3] += so2 * gk.z * gk.w;
                    Op[4] += so2 * gk.z * ekz;
                    Op[5] += so2 * gk.w * ekz;
                    const double sq = c_recip_w[q] * cosv;
                    Qp[0] += sq * gk.z * gk.z;
                    Qp[1] += sq * gk.w * gk.w;
                    Qp[2] += sq * ekz * ekz;
                    Qp[3] += sq * gk.z * gk.w;
                    Qp[4] += sq * gk.z * ekz;
                    Qp[5] += sq * gk.w * ekz;
                }
                cmul_step(gr, gi, cz, sz);
            }
            cmul_step(fr, fi, cy, sy);
        }
        cmul_step(hr, hi, cx, sx);
    }
    const double preQ = 0.5 * sigma_sq * E.inv_L * E.inv_L * E.inv_L;
    out.xx += preDii * (E.preO * Op[0] + preQ * Qp[0]);
    out.yy += preDii * (E.preO * Op[1] + preQ * Qp[1]);
    out.zz += preDii * (E.preO * Op[2] + preQ * Qp[2]);
    out.xy += preDii * (E.preO * Op[3] + preQ * Qp[3]);
    out.xz += preDii * (E.preO * Op[4] + preQ * Qp[4]);
    out.yz += preDii * (E.preO * Op[5] + preQ * Qp[5]);
    if (self) {
        // diff_tensor.c:241-263
        const double m1 = 0.5641895835;
        double sc3 = E.alpha * ai * E.inv_L;
        sc3 = sc3 * sc3 * sc3 * m1 / 3;
        sc3 = 0.75 * E.inv_L * 1.5 * E.alpha * ai * m1 * E.inv_L +
              0.5 * ai * ai * E.inv_L * E.inv_L * E.inv_L * sc3;
        sc3 *= preDii;
        out.xx = preDii / ai - sc3 + out.xx;
        out.yy = preDii / ai - sc3 + out.yy;
        out.zz = preDii / ai - sc3 + out.zz;
        out.xy = out.xz = out.yz = 0.0;
    }
    return out;
}

}  // namespace bdh
