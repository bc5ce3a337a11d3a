// Per-pair and per-bond force terms shared by the brute-force and cell-list force kernels,
// so both evaluate every pair with the same instructions.
//
//   Debye-Hueckel  F = (gamma/eps_r) Q_i Q_j (1 + kappa r) e^{-kappa r} / r^2   r <= cutoff_c
//                  (electro.c:36-78)
//   Lennard-Jones  F = alpha eps_ij 6 (s6 (2 s6 - lj6)) / r, s6 = (sigma_ij/r)^6,
//                  sigma_ij = R_i + R_j, eps_ij = sqrt(eps_i eps_j),
//                  r <= cutoff_lj (or 1.122462 sigma_ij when cutoff_lj <= 0)   (LJ.c:37-99)
//   overlap flag   sigma_ij > r (LJ_force / LJ_force_check, LJ.c:58, :101-107)
//   bonds          F = H (r0 - r) / (1 - r^2/rmax^2)                           (bond.c:43-85)
// The branch structure is force_single's (calc_func.c:370-409).
#pragma once

#include "bdh_common.hpp"
#include "bdh_kernels.hpp"

namespace bdh {

struct PairAcc {
    double fx, fy, fz;  // force on the row bead
    double e_lj, e_c;   // energies seen from the row bead (each pair is seen twice)
    int rep;            // bit 0: overlap, bit 1: overstretched bond
};

// Pair (i, j) seen from i: d = r_j - r_i (minimum image applied), r2d = |d|^2 from pair_r2.
template <typename T>
__device__ __forceinline__ void pair_term(double dx, double dy, double dz, double r2d, double qi,
                                          double qj, double2 lji, double2 ljj, bool conn,
                                          const ForceParams& fp, PairAcc& a)
{
    const T gam = (T)fp.gam, kap = (T)fp.kappa_c, alj = (T)fp.alpha_lj, lj6 = (T)fp.lj_6_term;
    const T r2 = (T)r2d;
    const T r = sqrt(r2);
    const T sig = (T)(lji.x + ljj.x);
    T F = 0;
    int ov = 0;
    if (fp.elec && r <= (T)fp.cutoff_c && (!conn || fp.bond_c_scale != 0.0)) {
        const T ex = exp(-kap * r);
        T f = gam * (T)qi * (T)qj * ((T)1 + kap * r) * ex / r2;
        T en = gam * (T)qi * (T)qj * ex / r;
        if (conn) { f *= (T)fp.bond_c_scale; en *= (T)fp.bond_c_scale; }
        F += f;
        a.e_c += (double)en;
    }
    const T lj_cut = fp.cutoff_lj > 0.0 ? (T)fp.cutoff_lj : (T)1.122462f * sig;
    if (fp.alpha_lj != 0.0 && r <= lj_cut) {
        if (!conn || fp.bond_lj_scale != 0.0) {
            const T eps = (T)sqrt(lji.y * ljj.y);
            T s6 = sig / r;
            s6 *= s6;
            s6 *= s6 * s6;
            T f = alj * eps * (T)6 * (s6 * ((T)2 * s6 - lj6)) / r;
            T en = alj * eps * (s6 * (s6 - lj6));
            if (conn) { f *= (T)fp.bond_lj_scale; en *= (T)fp.bond_lj_scale; }
            F += f;
            a.e_lj += (double)en;
            ov = (sig > r) || !isfinite(f);
        } else {
            ov = sig * sig > r2;
        }
    } else {
        ov = sig * sig > r2;
    }
    a.rep |= ov;
    // force on i: -e_ij F with e_ij = (r_j - r_i)/r
    const double s = (double)F / (double)r;
    a.fx -= dx * s;
    a.fy -= dy * s;
    a.fz -= dz * s;
}

__device__ __forceinline__ bool bonded(const Topology& topo, int nb0, int nb1, int j)
{
    bool conn = false;
    for (int b = nb0; b < nb1; ++b) conn |= (topo.conn[b] == j);
    return conn;
}

// All bonds of bead i (both ends carry each bond), returns the bond energy seen from i.
__device__ __forceinline__ double bond_terms(const double4* __restrict__ pos, int i, double4 pi,
                                             const Topology& topo, const ForceParams& fp,
                                             PairAcc& a)
{
    double e_b = 0.0;
    for (int b = topo.bond_off[i]; b < topo.bond_off[i + 1]; ++b) {
        const int p = topo.bond_p[b];
        const double r0 = topo.bond_parm[3 * b], rmax = topo.bond_parm[3 * b + 1];
        const double H = topo.bond_parm[3 * b + 2];
        double dx = pos[p].x - pi.x, dy = pos[p].y - pi.y, dz = pos[p].z - pi.z;
        const double r2 = pair_r2(dx, dy, dz, fp.pbc, fp.box);
        const double r = sqrt(r2);
        if (!(r < rmax)) a.rep |= 2;  // overstretched bond: fatal in the reference
        const double Ff = H * (r0 - r) / (1.0 - r2 / (rmax * rmax));
        a.fx -= Ff * dx / r;
        a.fy -= Ff * dy / r;
        a.fz -= Ff * dz / r;
        e_b += -0.5 * H * (rmax * rmax * log((rmax * rmax - r2) / (rmax * rmax - r0 * r0)) +
                           r0 * rmax * log((rmax + r) * (rmax - r0) / ((rmax - r) * (rmax + r0))));
    }
    return e_b;
}

}  // namespace bdh
