/*
 * FP64 CPU restatement of BD_BOX's Ermak-McCammon step with hydrodynamic interactions.
 *
 * TEST INFRASTRUCTURE ONLY -- the checker for the HIP path (see bdbox_oracle.h).
 *
 * Arithmetic is written in the same expression order as the reference's double-precision
 * CPU path so that, compiled without FMA contraction (gcc -O2 on x86-64), the RPY blocks,
 * the assembled D, the Gaussian stream and the single-thread forces agree with the
 * reference to the last bit; float literals of the reference (M_PIF, 1.122462f, 1e-7f)
 * are kept as float literals for the same reason.
 *
 * Pinning (tests/test_oracle.py): golden vectors written by the compiled reference
 * (the .npz files under tests/golden) for RPY/D (free space, overlapping unequal radii, Beenakker),
 * Cholesky factor, forces/energies, Gaussian stream, first-step X/Rf/Rs for Cholesky and
 * TEA noise, multi-step trajectories; plus closed forms (D_ii = kT/6 pi eta a, Oseen limit,
 * continuity of the RPY branches at r = a_i + a_j, Beenakker alpha-independence is not
 * testable with alpha fixed to sqrt(pi)/L and is left to the golden vectors), LL^T = D,
 * Newton's third law of the forces, central differences of the energies.
 */
#include "bdbox_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define KB        0.001986248532617          /* kB, data.h:224 */
#define KB_1_8PI  0.00007903031804378703     /* data.h:225 */
#define KB_1_6PI  0.00010537375739171604     /* data.h:226 */
#define SQRTPI_D  1.77245385090551602729     /* data.h:227 */
#define M_PIF_D   ((double)3.14159265358979323846f) /* float pi, data.h:254 */

void orc_prefactors(const orc_params* p, double* preDii, double* preDij)
{
    /* diff_decl.c:93 and :100 */
    *preDii = KB_1_6PI * p->T / (p->visc * p->vfactor);
    *preDij = KB_1_8PI * p->T / (p->visc * p->vfactor);
}

/* unit vector and norms of v (math_help.c:83-99: e = v * (1/|v|)) */
static void ev2_norms(double x, double y, double z, double* e, double* r_norm, double* r_sq)
{
    double inv;
    e[0] = x; e[1] = y; e[2] = z;
    *r_sq = x * x + y * y + z * z;
    *r_norm = sqrt(*r_sq);
    inv = 1 / *r_norm;
    e[0] *= inv; e[1] *= inv; e[2] *= inv;
}

/* diff_tensor.c:45-92 with a = i (columns), b = j (rows): sigma_i = a_j, sigma_j = a_i,
   e from coord_i - coord_j without minimum image (math_help.c:68-81). */
void orc_rpy_block(const double* ci, const double* cj, double preDii, double preDij, double* tD)
{
    double si = cj[3], sj = ci[3];
    double e[3], r_norm, r_sq;
    int a, b;
    ev2_norms(ci[0] - cj[0], ci[1] - cj[1], ci[2] - cj[2], e, &r_norm, &r_sq);
    if (r_norm < si + sj)
    {
        double sigma = (si + sj) / 2;
        double pre = preDii / sigma;
        double part1 = (1 - (9 * r_norm / (32 * sigma)));
        tD[0] = part1; tD[4] = part1; tD[8] = part1;
        for (a = 0; a < 3; ++a)
        {
            tD[a + 3 * a] += 3 * e[a] * e[a] * r_norm / (32 * sigma);
            tD[a + 3 * a] *= pre;
            for (b = 0; b < a; ++b)
            {
                tD[b + 3 * a] = pre * 3 * e[a] * e[b] * r_norm / (32 * sigma);
                tD[a + 3 * b] = tD[b + 3 * a];
            }
        }
    }
    else
    {
        double pre = preDij / r_norm;
        double sigm_r = (si * si + sj * sj) / r_sq;
        tD[0] = 1 + sigm_r / 3; tD[4] = 1 + sigm_r / 3; tD[8] = 1 + sigm_r / 3;
        for (a = 0; a < 3; ++a)
        {
            tD[a + 3 * a] += e[a] * e[a] * (1 - sigm_r);
            tD[a + 3 * a] *= pre;
            for (b = 0; b < a; ++b)
            {
                tD[b + 3 * a] = pre * e[a] * e[b] * (1 - sigm_r);
                tD[a + 3 * b] = tD[b + 3 * a];
            }
        }
    }
}

/* Beenakker-Ewald block (diff_tensor.c:265-422); alpha = sqrt(pi)/L as forced by
   init_tensor (diff_tensor.c:551-554); the real and reciprocal lattices share the sphere
   h^2+f^2+g^2 <= rr^2 (diff_tensor.c:319-327; ewald_real == ewald_recip, input.c:402). */
void orc_beenakker_block(const double* coord, int i, int j, double box, int rr,
                         double preDii, double preDij, double* out)
{
    const double inv_L = 1 / box;
    const double alpha = SQRTPI_D * inv_L;
    const double alpha2 = alpha * alpha;
    const double alpha3 = alpha * alpha2;
    const double inv_alpha2 = 1.0f / alpha2;
    const double si = coord[4 * j + 3], sj = coord[4 * i + 3];
    const double sigma_sq = 0.5 * (si * si + sj * sj);
    const double preQ = inv_L * inv_L * inv_L;
    const int rr_sq = rr * rr;
    double Op[9] = {0}, Qp[9] = {0}, tD[9] = {0};
    double r[3], v[3];
    int k, l, h, f, g;
    r[0] = coord[4 * j + 0] - coord[4 * i + 0];
    r[1] = coord[4 * j + 1] - coord[4 * i + 1];
    r[2] = coord[4 * j + 2] - coord[4 * i + 2];

    if ((r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) < (si + sj) * (si + sj) && i != j)
    {
        double r_norm, r_sq, e[3];
        ev2_norms(r[0], r[1], r[2], e, &r_norm, &r_sq);
        {   /* overlapping RPY added */
            double sigma = (si + sj) / 2;
            double pre = preDii / sigma;
            double part1 = (1 - (9 * r_norm / (32 * sigma)));
            for (k = 0; k < 3; ++k)
            {
                tD[k + 3 * k] += pre * (part1 + 3 * e[k] * e[k] * r_norm / (32 * sigma));
                for (l = 0; l < k; ++l)
                {
                    double vval = pre * 3 * e[k] * e[l] * r_norm / (32 * sigma);
                    tD[l + 3 * k] += vval;
                    tD[k + 3 * l] += vval;
                }
            }
        }
        {   /* non-overlapping RPY removed */
            double pre = preDij / r_norm;
            double sigm_r = (si * si + sj * sj) / r_sq;
            double part1 = 1 + sigm_r / 3;
            for (k = 0; k < 3; ++k)
            {
                tD[k + 3 * k] -= pre * (part1 + e[k] * e[k] * (1 - sigm_r));
                for (l = 0; l < k; ++l)
                {
                    double vval = pre * e[k] * e[l] * (1 - sigm_r);
                    tD[l + 3 * k] -= vval;
                    tD[k + 3 * l] -= vval;
                }
            }
        }
    }
    v[0] = r[0] * inv_L; v[1] = r[1] * inv_L; v[2] = r[2] * inv_L;

    for (h = -rr; h <= rr; ++h)
    {
        int nf = (int)sqrt((double)(rr_sq - h * h));
        for (f = -nf; f <= nf; ++f)
        {
            int ng = (int)sqrt((double)(rr_sq - h * h - f * f));
            for (g = -ng; g <= ng; ++g)
            {
                double r_norm, r_sq, e[3], inv_r, inv_r2, inv_r3, sc, sc2;
                ev2_norms(r[0] + box * h, r[1] + box * f, r[2] + box * g, e, &r_norm, &r_sq);
                if (r_norm != 0.0f)
                {
                    inv_r = 1 / r_norm;
                    inv_r2 = 1 / r_sq;
                    inv_r3 = inv_r * inv_r2;
                    sc = alpha * sigma_sq * inv_r2;
                    sc += 14 * alpha3 * sigma_sq;
                    sc += -4.5f * alpha;
                    sc += -20 * alpha2 * alpha3 * sigma_sq * r_sq;
                    sc += 3 * alpha3 * r_sq;
                    sc += 4 * alpha3 * alpha2 * alpha2 * sigma_sq * r_sq * r_sq;
                    sc *= exp(-alpha2 * r_sq);
                    sc /= SQRTPI_D;
                    sc += erfc(r_norm * alpha) * (0.75f * inv_r + 0.5f * sigma_sq * inv_r3);

                    sc2 = -3 * alpha * sigma_sq * inv_r2;
                    sc2 += -2 * alpha3 * sigma_sq;
                    sc2 += 1.5f * alpha;
                    sc2 += 16 * alpha2 * alpha3 * sigma_sq * r_sq;
                    sc2 += -3 * alpha3 * r_sq;
                    sc2 += -4 * alpha3 * alpha2 * alpha2 * sigma_sq * r_sq * r_sq;
                    sc2 *= exp(-alpha2 * r_sq);
                    sc2 /= SQRTPI_D;
                    sc2 += erfc(r_norm * alpha) * (0.75f * inv_r - 1.5f * sigma_sq * inv_r3);

                    for (k = 0; k < 3; ++k)
                    {
                        double sc_ek = e[k] * sc2;
                        Op[k * 4] += sc + sc_ek * e[k];
                        for (l = 0; l < k; ++l) Op[k + l * 3] += e[l] * sc_ek;
                    }
                }
                ev2_norms(h, f, g, e, &r_norm, &r_sq);
                if (r_norm != 0.0f)
                {
                    r_norm *= 2 * M_PIF_D;
                    r_sq *= 4 * M_PIF_D * M_PIF_D;
                    sc = 1 - sigma_sq * r_sq / 3;
                    sc *= 1 + (inv_alpha2 * r_sq / 4) * (1 + inv_alpha2 * r_sq / 2);
                    sc *= 6 * M_PIF_D * exp(-inv_alpha2 * r_sq / 4) / r_sq;
                    sc *= cos(2 * M_PIF_D * (v[0] * h + v[1] * f + v[2] * g));
                    for (k = 0; k < 3; ++k)
                    {
                        double sc_ek = sc * e[k];
                        Qp[k * 4] += sc - sc_ek * e[k];
                        for (l = 0; l < k; ++l) Qp[k + l * 3] -= e[l] * sc_ek;
                    }
                }
            }
        }
    }
    for (k = 0; k < 3; ++k)
    {
        tD[k * 4] += preDii * (Op[k * 4] + preQ * Qp[k * 4]);
        for (l = 0; l < k; ++l)
        {
            tD[k + l * 3] += preDii * (Op[k + l * 3] + preQ * Qp[k + l * 3]);
            tD[l + k * 3] += preDii * (Op[k + l * 3] + preQ * Qp[k + l * 3]);
        }
    }
    if (i == j)
    {
        /* self term (diff_tensor.c:403-422) */
        double aw = coord[4 * i + 3];
        double sc3 = -6 * alpha / SQRTPI_D + 40 * alpha3 * aw * aw / (3 * SQRTPI_D);
        sc3 *= preDii;
        for (k = 0; k < 3; ++k)
        {
            tD[k * 4] = preDii / aw + sc3 + tD[k * 4];
            for (l = 0; l < k; ++l) { tD[l + 3 * k] = 0; tD[k + 3 * l] = 0; }
        }
    }
    memcpy(out, tD, sizeof(tD));
}

/* Smith-Ewald block (diff_tensor.c:102-239, self term :241-263), in reduced units v / L
   with the .prm ewald_alpha (float-rounded) and the sphere h^2+f^2+g^2 <= rr^2 shared by
   both lattices.  Constants as the reference spells them: M1_SQRTPI = 0.5641895835
   (data.h:228), float pi, M_PI_ALPHA_SQ = -(pi_f pi_f in float) / alpha^2 (init_tensor,
   diff_tensor.c:546-558). */
void orc_smith_block(const double* coord, int i, int j, double box, int rr, double alpha,
                     double preDii, double preDij, double* out)
{
    const double M1_SQRTPI_D = 0.5641895835;
    const float pif = 3.14159265358979323846f;
    const float mpp = -pif * pif;
    const double inv_L = 1 / box;
    const double preO = 0.75 * inv_L;
    const double alpha2 = alpha * alpha;
    const double alpha3 = alpha * alpha2;
    const double m_pi_alpha_sq = mpp / alpha2;
    const double si = coord[4 * j + 3], sj = coord[4 * i + 3];
    const double sigma_sq = 0.5 * (si * si + sj * sj);
    const double preQ = 0.5 * sigma_sq * inv_L * inv_L * inv_L;
    const int rr_sq = rr * rr;
    double Op[9] = {0}, Qp[9] = {0}, tD[9] = {0};
    double v[3];
    int k, l, h, f, g;
    v[0] = coord[4 * i + 0] - coord[4 * j + 0];   /* dist_vec(c_j, c_i), trans.c:123-128 */
    v[1] = coord[4 * i + 1] - coord[4 * j + 1];
    v[2] = coord[4 * i + 2] - coord[4 * j + 2];
    if ((v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) < (si + sj) * (si + sj) && i != j)
    {
        double r_norm, r_sq, e[3];
        ev2_norms(v[0], v[1], v[2], e, &r_norm, &r_sq);
        {   /* overlapping RPY added */
            double sigma = (si + sj) / 2;
            double pre = preDii / sigma;
            double part1 = (1 - (9 * r_norm / (32 * sigma)));
            for (k = 0; k < 3; ++k)
            {
                tD[k + 3 * k] += pre * (part1 + 3 * e[k] * e[k] * r_norm / (32 * sigma));
                for (l = 0; l < k; ++l)
                {
                    double vval = pre * 3 * e[k] * e[l] * r_norm / (32 * sigma);
                    tD[l + 3 * k] += vval;
                    tD[k + 3 * l] += vval;
                }
            }
        }
        {   /* non-overlapping RPY removed */
            double pre = preDij / r_norm;
            double sigm_r = (si * si + sj * sj) / r_sq;
            double part1 = 1 + sigm_r / 3;
            for (k = 0; k < 3; ++k)
            {
                tD[k + 3 * k] -= pre * (part1 + e[k] * e[k] * (1 - sigm_r));
                for (l = 0; l < k; ++l)
                {
                    double vval = pre * e[k] * e[l] * (1 - sigm_r);
                    tD[l + 3 * k] -= vval;
                    tD[k + 3 * l] -= vval;
                }
            }
        }
    }
    v[0] *= inv_L; v[1] *= inv_L; v[2] *= inv_L;
    for (h = -rr; h <= rr; ++h)
    {
        int next_f = (int)sqrt((double)(rr_sq - h * h));
        for (f = -next_f; f <= next_f; ++f)
        {
            int next_g = (int)sqrt((double)(rr_sq - h * h - f * f));
            for (g = -next_g; g <= next_g; ++g)
            {
                double r_norm, r_sq, e[3], sc, sc2;
                ev2_norms(v[0] + h, v[1] + f, v[2] + g, e, &r_norm, &r_sq);
                if (r_norm != 0.0f)
                {
                    double inv_r = 1 / r_norm;
                    double inv_r3 = 1 / (r_norm * r_sq);
                    sc = erfc(r_norm * alpha) * inv_r;
                    sc2 = 2 * alpha * M1_SQRTPI_D * exp(-alpha2 * r_sq);
                    for (k = 0; k < 3; ++k)
                    {
                        double sc_ek = e[k] * (sc + sc2);
                        Op[k * 4] += sc + sc_ek * e[k];
                        for (l = 0; l < k; ++l) Op[k + l * 3] += e[l] * sc_ek;
                    }
                    sc = inv_r3 * (erfc(r_norm * alpha) + 2 * alpha * r_norm * M1_SQRTPI_D * exp(-alpha2 * r_sq));
                    sc2 = -4 * alpha3 * M1_SQRTPI_D * exp(-alpha2 * r_sq);
                    for (k = 0; k < 3; ++k)
                    {
                        double sc_ek = (-3 * sc + sc2) * e[k];
                        Qp[k * 4] += sc + sc_ek * e[k];
                        for (l = 0; l < k; ++l) Qp[k + l * 3] += e[l] * sc_ek;
                    }
                }
                ev2_norms(h, f, g, e, &r_norm, &r_sq);
                if (r_norm != 0.0f)
                {
                    sc = 2 * exp(m_pi_alpha_sq * r_sq) / (pif * r_sq);
                    sc2 = -1 + m_pi_alpha_sq * r_sq;
                    sc *= cos(2 * pif * (v[0] * h + v[1] * f + v[2] * g));
                    sc2 *= sc;
                    for (k = 0; k < 3; ++k)
                    {
                        double sc_ek = sc2 * e[k];
                        Op[k * 4] += sc + sc_ek * e[k];
                        for (l = 0; l < k; ++l) Op[k + l * 3] += e[l] * sc_ek;
                    }
                    sc = 4 * pif * exp(m_pi_alpha_sq * r_sq);
                    sc *= cos(2 * pif * (v[0] * h + v[1] * f + v[2] * g));
                    for (k = 0; k < 3; ++k)
                    {
                        double sc_ek = sc * e[k];
                        Qp[k * 4] += sc_ek * e[k];
                        for (l = 0; l < k; ++l) Qp[k + l * 3] += e[l] * sc_ek;
                    }
                }
            }
        }
    }
    for (k = 0; k < 3; ++k)
    {
        tD[k * 4] += preDii * (preO * Op[k * 4] + preQ * Qp[k * 4]);
        for (l = 0; l < k; ++l)
        {
            tD[k + l * 3] += preDii * (preO * Op[k + l * 3] + preQ * Qp[k + l * 3]);
            tD[l + k * 3] += preDii * (preO * Op[k + l * 3] + preQ * Qp[k + l * 3]);
        }
    }
    if (i == j)
    {
        double aw = coord[4 * i + 3];
        double sc3 = alpha * aw * inv_L;
        sc3 = sc3 * sc3 * sc3 * M1_SQRTPI_D / 3;
        sc3 = 0.75f * inv_L * 1.5f * alpha * aw * M1_SQRTPI_D * inv_L +
              0.5f * aw * aw * inv_L * inv_L * inv_L * sc3;
        sc3 *= preDii;
        for (k = 0; k < 3; ++k)
        {
            tD[k * 4] = preDii / aw - sc3 + tD[k * 4];
            for (l = 0; l < k; ++l) { tD[l + 3 * k] = 0; tD[k + 3 * l] = 0; }
        }
    }
    memcpy(out, tD, sizeof(tD));
}

/* the periodic block of the configured Ewald method */
static void ewald_block(const double* coord, int i, int j, const orc_params* p, double preDii,
                        double preDij, double* tD)
{
    if (p->ewald_method == 0)
        orc_smith_block(coord, i, j, p->box, p->ewald_real, p->ewald_alpha, preDii, preDij, tD);
    else
        orc_beenakker_block(coord, i, j, p->box, p->ewald_real, preDii, preDij, tD);
}

/* block (i, j), j < i, in the orientation compute_D stores it (rows of j, columns of i) */
static void pair_block(int N, const double* coord, const orc_params* p, double preDii,
                       double preDij, int i, int j, double* tD)
{
    (void)N;
    if (p->ewald)
        ewald_block(coord, i, j, p, preDii, preDij, tD);
    else
        orc_rpy_block(coord + 4 * i, coord + 4 * j, preDii, preDij, tD);
}

void orc_compute_D(int N, const double* coord, const orc_params* p, double* D, int ld)
{
    double preDii, preDij;
    int i;
    orc_prefactors(p, &preDii, &preDij);
#pragma omp parallel for schedule(dynamic)
    for (i = 0; i < N; ++i)
    {
        int j, k, l;
        double tD[9];
        for (j = 0; j < i; ++j)
        {
            pair_block(N, coord, p, preDii, preDij, i, j, tD);
            /* insert9 (diff_tensor.c:425-435): D(3j+k, 3i+l) = tD[k + 3l]; mirrored */
            for (k = 0; k < 3; ++k)
                for (l = 0; l < 3; ++l)
                {
                    D[(3 * j + k) + (size_t)(3 * i + l) * ld] = tD[k + 3 * l];
                    D[(3 * i + l) + (size_t)(3 * j + k) * ld] = tD[k + 3 * l];
                }
        }
        if (p->ewald)
            ewald_block(coord, i, i, p, preDii, preDij, tD);
        else
        {
            memset(tD, 0, sizeof(tD));
            tD[0] = tD[4] = tD[8] = preDii / coord[4 * i + 3];
        }
        for (k = 0; k < 3; ++k)
            for (l = 0; l < 3; ++l)
                D[(3 * i + k) + (size_t)(3 * i + l) * ld] = tD[k + 3 * l];
    }
}

int orc_cholesky(int n, double* A, int ld)
{
    /* left-looking column Cholesky, lower triangle, column-major */
    int j;
    for (j = 0; j < n; ++j)
    {
        double* cj = A + (size_t)j * ld;
        int k, i;
        for (k = 0; k < j; ++k)
        {
            const double* ck = A + (size_t)k * ld;
            double ljk = ck[j];
            for (i = j; i < n; ++i) cj[i] -= ljk * ck[i];
        }
        if (!(cj[j] > 0.0)) return j + 1;
        cj[j] = sqrt(cj[j]);
        {
            double inv = 1.0 / cj[j];
            for (i = j + 1; i < n; ++i) cj[i] *= inv;
        }
    }
    return 0;
}

void orc_trmv_lower(int n, const double* L, int ld, const double* x, double* y)
{
    int i, j;
    for (i = 0; i < n; ++i) y[i] = 0.0;
    for (j = 0; j < n; ++j)
    {
        const double* c = L + (size_t)j * ld;
        double xj = x[j];
        for (i = j; i < n; ++i) y[i] += c[i] * xj;
    }
}

void orc_symv_lower(int n, const double* D, int ld, const double* x, double* y)
{
    int i, j;
    for (i = 0; i < n; ++i) y[i] = 0.0;
    for (j = 0; j < n; ++j)
    {
        const double* c = D + (size_t)j * ld;
        double xj = x[j], acc = 0.0;
        y[j] += c[j] * xj;
        for (i = j + 1; i < n; ++i)
        {
            y[i] += c[i] * xj;
            acc += c[i] * x[i];
        }
        y[j] += acc;
    }
}

/* ------------------------------------------------------------------------------ forces */
static void dist_vec_pbc(const orc_params* p, const double* a, const double* b, double* v)
{
    /* trans_dist_vec (trans.c:111-121): b - a, minimum image for bc periodic */
    v[0] = b[0] - a[0]; v[1] = b[1] - a[1]; v[2] = b[2] - a[2];
    if (p->bc == 1)
    {
        double inv = 1 / p->box;
        v[0] -= p->box * floor(v[0] * inv + 0.5f);
        v[1] -= p->box * floor(v[1] * inv + 0.5f);
        v[2] -= p->box * floor(v[2] * inv + 0.5f);
    }
}

static int bad(double x) { return isnan(x) || isinf(x); }

int orc_forces(int N, const double* coord, const double* Q, const double* LJ,
               int nbonds, const int* bonds, const double* bond_parms,
               const orc_params* p, double* F, double* E)
{
    int i, j, b, rep = 0;
    double E_lj = 0, E_c = 0, E_b = 0;
    unsigned char* conn = NULL;
    memset(F, 0, sizeof(double) * 3 * N);

    /* bonded pairs: bond.c:43-85, in bond order (pre_func.c:125-131) */
    for (b = 0; b < nbonds; ++b)
    {
        int b0 = bonds[2 * b], b1 = bonds[2 * b + 1];
        double r_0 = bond_parms[3 * b], r_max = bond_parms[3 * b + 1], H = bond_parms[3 * b + 2];
        double v[3], e[3], r_norm, r_sq, Ff;
        dist_vec_pbc(p, coord + 4 * b0, coord + 4 * b1, v);
        ev2_norms(v[0], v[1], v[2], e, &r_norm, &r_sq);
        if (!(r_norm < r_max)) { free(conn); return -1; } /* "Overstretched bond" UNERR */
        E_b += -0.5 * H * (r_max * r_max * log((r_max * r_max - r_sq) / (r_max * r_max - r_0 * r_0))
                           + r_0 * r_max * log((r_max + r_norm) * (r_max - r_0) / ((r_max - r_norm) * (r_max + r_0))));
        Ff = H * (r_0 - r_norm) / (1 - r_sq / (r_max * r_max));
        F[3 * b0 + 0] -= Ff * e[0]; F[3 * b0 + 1] -= Ff * e[1]; F[3 * b0 + 2] -= Ff * e[2];
        F[3 * b1 + 0] += Ff * e[0]; F[3 * b1 + 1] += Ff * e[1]; F[3 * b1 + 2] += Ff * e[2];
        rep |= bad(Ff) || bad(e[0]) || bad(e[1]) || bad(e[2]);
    }
    /* bonded-partner table exists only when a scale differs from 1 (bond.c:113-130) */
    if (nbonds && ((p->elec && p->bond_c_scale != 1) || (p->alpha_lj != 0.0 && p->bond_lj_scale != 1)))
    {
        conn = (unsigned char*)calloc((size_t)N * N, 1);
        for (b = 0; b < nbonds; ++b)
        {
            conn[(size_t)bonds[2 * b] * N + bonds[2 * b + 1]] = 1;
            conn[(size_t)bonds[2 * b + 1] * N + bonds[2 * b]] = 1;
        }
    }
    if (p->elec || p->alpha_lj != 0.0)
    {
        const double gam = p->gamma_c / p->epsilon_c;
        for (i = 0; i < N; ++i)
        {
            for (j = i + 1; j < N && !rep; ++j)
            {
                /* force_single (calc_func.c:370-409) */
                double v[3], e[3], r_norm, r_sq;
                int c = conn ? conn[(size_t)i * N + j] : 0;
                int r = 0;
                double lj_cut;
                double* Fi = F + 3 * i;
                double* Fj = F + 3 * j;
                dist_vec_pbc(p, coord + 4 * i, coord + 4 * j, v);
                ev2_norms(v[0], v[1], v[2], e, &r_norm, &r_sq);
                if (p->elec && r_norm <= p->cutoff_c && (!c || p->bond_c_scale != 0.0))
                {
                    /* electro.c:36-78 */
                    double Fc = gam * Q[i] * Q[j] * (1 + p->kappa_c * r_norm);
                    Fc *= exp(-p->kappa_c * r_norm) / r_sq;
                    if (c) Fc *= p->bond_c_scale;
                    E_c += (c ? p->bond_c_scale : 1.0) * gam * Q[i] * Q[j] * exp(-p->kappa_c * r_norm) / r_norm;
                    Fi[0] -= e[0] * Fc; Fi[1] -= e[1] * Fc; Fi[2] -= e[2] * Fc;
                    Fj[0] += e[0] * Fc; Fj[1] += e[1] * Fc; Fj[2] += e[2] * Fc;
                }
                lj_cut = p->cutoff_lj > 0.0 ? p->cutoff_lj
                                            : 1.122462f * (LJ[2 * i] + LJ[2 * j]);
                if (p->alpha_lj != 0.0 && r_norm <= lj_cut)
                {
                    if (!c || p->bond_lj_scale != 0.0)
                    {
                        /* LJ.c:37-93 */
                        double sigma_ij = LJ[2 * i] + LJ[2 * j];
                        double eps_ij = sqrt(LJ[2 * i + 1] * LJ[2 * j + 1]);
                        double s6 = sigma_ij / r_norm, Fl;
                        double scale = c ? p->bond_lj_scale : 1.0;
                        s6 *= s6;
                        s6 *= s6 * s6;
                        E_lj += scale * p->alpha_lj * eps_ij * (s6 * (s6 - p->lj_6_term));
                        Fl = p->alpha_lj * eps_ij * 6 * (s6 * (2 * s6 - p->lj_6_term)) / r_norm;
                        if (c) Fl = scale * Fl;
                        Fi[0] -= e[0] * Fl; Fi[1] -= e[1] * Fl; Fi[2] -= e[2] * Fl;
                        Fj[0] += e[0] * Fl; Fj[1] += e[1] * Fl; Fj[2] += e[2] * Fl;
                        r = (sigma_ij > r_norm) || bad(Fl) || bad(e[0]) || bad(e[1]) || bad(e[2]);
                    }
                    else
                    {
                        double s = LJ[2 * i] + LJ[2 * j];
                        r = s * s > r_sq;
                    }
                }
                else
                {
                    double s = LJ[2 * i] + LJ[2 * j];
                    r = s * s > r_sq;
                }
                rep = rep || (p->check_overlap && r);
            }
        }
    }
    else if (p->check_overlap)
    {
        for (i = 0; i < N; ++i)
            for (j = i + 1; j < N && !rep; ++j)
            {
                double v[3], r_sq, s = LJ[2 * i] + LJ[2 * j];
                dist_vec_pbc(p, coord + 4 * i, coord + 4 * j, v);
                r_sq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                rep = rep || (s * s > r_sq);
            }
    }
    free(conn);
    if (E) { E[0] = E_lj; E[1] = E_c; E[2] = E_b; }
    return rep;
}

/* ----------------------------------------------------------------------- random stream */
uint64_t orc_srand48(uint32_t seed)
{
    /* srand48_r: X0 = (seed << 16) | 0x330E (glibc, called from rand_move.c:70) */
    return (((uint64_t)seed) << 16) | 0x330Eu;
}

double orc_drand48(uint64_t* state)
{
    /* X <- (a X + c) mod 2^48, value X / 2^48 (exact in double) */
    *state = (0x5DEECE66DULL * *state + 0xBULL) & ((1ULL << 48) - 1);
    return ldexp((double)*state, -48);
}

void orc_gauss(uint64_t* state, double* out, int count)
{
    /* polar Box-Muller of gauss_rand (rand_move.c:132-168): pair state local to the call */
    int i, iset = 0;
    double gset = 0.0, fac, rsq, v1 = 0.0, v2 = 0.0;
    for (i = 0; i < count; ++i)
    {
        if (iset == 0)
        {
            rsq = 0.0;
            while ((rsq >= 1) || (rsq == 0.0))
            {
                v1 = 2.0f * orc_drand48(state) - 1.0f;
                v2 = 2.0f * orc_drand48(state) - 1.0f;
                rsq = v1 * v1 + v2 * v2;
            }
            fac = sqrt(-2.0f * log(rsq) / rsq);
            gset = v1 * fac;
            iset = 1;
            out[i] = v2 * fac;
        }
        else
        {
            iset = 0;
            out[i] = gset;
        }
    }
}

/* ------------------------------------------------------------------- TEA and on-the-fly */
void orc_rf_onfly(int N, const double* coord, const orc_params* p, double dt,
                  const double* F, double* Rf)
{
    /* compute_Rf_geyer_in_fly (diff_equation.c:55-117), single thread order */
    double preDii, preDij, pre = dt / (KB * p->T);
    int i, k, a, b;
    orc_prefactors(p, &preDii, &preDij);
    memset(Rf, 0, sizeof(double) * 3 * N);
    for (i = 0; i < N; ++i)
    {
        double tD[9];
        for (k = 0; k < i; ++k)
        {
            pair_block(N, coord, p, preDii, preDij, i, k, tD);
            for (a = 0; a < 3; ++a)
                for (b = 0; b < 3; ++b)
                {
                    Rf[i * 3 + a] += F[k * 3 + b] * tD[a * 3 + b];
                    Rf[k * 3 + b] += F[i * 3 + a] * tD[a * 3 + b];
                }
        }
        if (p->ewald) ewald_block(coord, i, i, p, preDii, preDij, tD);
        else { memset(tD, 0, sizeof(tD)); tD[0] = tD[4] = tD[8] = preDii / coord[4 * i + 3]; }
        Rf[i * 3] += F[i * 3] * tD[0];
        Rf[i * 3 + 1] += F[i * 3 + 1] * tD[4];
        Rf[i * 3 + 2] += F[i * 3 + 2] * tD[8];
    }
    for (i = 0; i < 3 * N; ++i) Rf[i] *= pre;
}

static double diag_of(int N, const double* coord, const orc_params* p, double preDii,
                      double preDij, int dof)
{
    /* diag_D (diff_decl.c:94-98): preDii / a_i for all three components */
    (void)N; (void)p; (void)preDij;
    return preDii / coord[4 * (dof / 3) + 3];
}

/* The close of compute_r's DIFF_GEYER branch (diff_rand.c:165-224): eps from the sum of the
   off-diagonal entries, beta, then Rs = sqrt(1 / (1 + beta^2 C / D_rr)) (beta Feff + D_rr X);
   Feff arrives in Rs. */
static void geyer_close(int N, const double* coord, const orc_params* p, double eps,
                        const double* C, const double* X, double* Rs)
{
    const int n = 3 * N;
    double preDii, preDij, a, beta;
    int i;
    orc_prefactors(p, &preDii, &preDij);
    eps /= n;
    eps /= (n - 1);
    eps *= 2;
    if (eps > 1.0) eps = 1.0;
    a = (3 * N - 1) * eps * eps - (3 * N - 2) * eps;
    if (fabs(a) < 0.0000001f) beta = 0.5;
    else beta = (1 - sqrt(1 - a)) / a;
    for (i = 0; i < n; ++i)
    {
        double dii = diag_of(N, coord, p, preDii, preDij, i);
        double c = C[i];
        c *= beta * beta * (1 / dii);
        c = 1 / (1 + c);
        c = sqrt(c);
        Rs[i] = c * (beta * Rs[i] + dii * X[i]);
    }
}

void orc_geyer_rs(int N, const double* coord, const orc_params* p, double dt,
                  double* X, double* Rs)
{
    /* compute_r, DIFF_GEYER branch, on the fly (diff_rand.c:86-135, 165-224) */
    const int n = 3 * N;
    double preDii, preDij, eps = 0.0;
    double* C = (double*)calloc((size_t)n, sizeof(double));
    double* Feff = Rs;
    int i, k, jj, l;
    orc_prefactors(p, &preDii, &preDij);
    memset(Rs, 0, sizeof(double) * n);
    for (i = 0; i < n; ++i)
    {
        double inv_d = 1 / diag_of(N, coord, p, preDii, preDij, i);
        X[i] *= sqrt(2 * dt * inv_d);
    }
    for (i = 0; i < N; ++i)
    {
        double e_loc = 0.0, tD[9];
        for (k = 0; k < i; ++k)
        {
            pair_block(N, coord, p, preDii, preDij, i, k, tD);
            for (jj = 0; jj < 3; ++jj)
                for (l = 0; l < 3; ++l)
                {
                    double d = tD[jj * 3 + l];
                    C[i * 3 + jj] += d * d * (1 / diag_of(N, coord, p, preDii, preDij, k * 3 + l));
                    C[k * 3 + l] += d * d * (1 / diag_of(N, coord, p, preDii, preDij, i * 3 + jj));
                    Feff[i * 3 + jj] += X[k * 3 + l] * d;
                    Feff[k * 3 + l] += X[i * 3 + jj] * d;
                    e_loc += d;
                }
        }
        eps += e_loc;
    }
    geyer_close(N, coord, p, eps, C, X, Rs);
    free(C);
}

/* compute_r, DIFF_GEYER branch over a stored tensor (geyer_on_the_fly no, diff_rand.c:136-164,
   then :165-224): the sums run over the entries D[r][k], k < r, of the lower triangle of D
   (ld, column-major lower = row-major upper), D symmetric. */
void orc_geyer_rs_D(int N, const double* D, int ld, const double* coord, const orc_params* p,
                    double dt, double* X, double* Rs)
{
    const int n = 3 * N;
    double preDii, preDij, eps = 0.0;
    double* C = (double*)calloc((size_t)n, sizeof(double));
    double* Feff = Rs;
    int r, k;
    orc_prefactors(p, &preDii, &preDij);
    memset(Rs, 0, sizeof(double) * n);
    for (r = 0; r < n; ++r)
    {
        double inv_d = 1 / diag_of(N, coord, p, preDii, preDij, r);
        X[r] *= sqrt(2 * dt * inv_d);
    }
    for (r = 0; r < n; ++r)
    {
        double e_loc = 0.0;
        for (k = 0; k < r; ++k)
        {
            const double d = D[(size_t)k * ld + r];
            C[r] += d * d * (1 / diag_of(N, coord, p, preDii, preDij, k));
            C[k] += d * d * (1 / diag_of(N, coord, p, preDii, preDij, r));
            Feff[r] += X[k] * d;
            Feff[k] += X[r] * d;
            e_loc += d;
        }
        eps += e_loc;
    }
    geyer_close(N, coord, p, eps, C, X, Rs);
    free(C);
}

/* --------------------------------------------------------------------------- integrator */
struct orc_sim {
    int N, nb, refactor_every, steps_done;
    orc_params p;
    double *coord, *Q, *LJ, *bond_parms, *F, *D, *L, *Rf, *Rs, *X;
    int* bonds;
    double *coord_prev, *F_prev;
    double *D1, *L2, *Rf0;          /* IG-T: D at the predictor, the averaged factor, Rf0 */
    uint64_t rng;
    int have_L;
    int algorithm, move_attempts;   /* 0 ermak_const, 1 ermak_var; rejections allowed */
    double last_dt;
    long long rejections;
    int flow;                       /* vel_grad_tensor set: G (row, column) and FS */
    double G[9];
    double* FS;
};

orc_sim* orc_sim_create(int N, const double* coord, const double* Q, const double* LJ,
                        int nbonds, const int* bonds, const double* bond_parms,
                        const orc_params* p, uint32_t seed, int refactor_every)
{
    orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
    size_t n = 3 * (size_t)N;
    s->N = N; s->nb = nbonds; s->p = *p;
    s->refactor_every = refactor_every < 1 ? 1 : refactor_every;
    s->coord = (double*)malloc(sizeof(double) * 4 * N);
    s->Q = (double*)malloc(sizeof(double) * N);
    s->LJ = (double*)malloc(sizeof(double) * 2 * N);
    s->bonds = (int*)malloc(sizeof(int) * 2 * (nbonds + 1));
    s->bond_parms = (double*)malloc(sizeof(double) * 3 * (nbonds + 1));
    memcpy(s->coord, coord, sizeof(double) * 4 * N);
    memcpy(s->Q, Q, sizeof(double) * N);
    memcpy(s->LJ, LJ, sizeof(double) * 2 * N);
    if (nbonds) { memcpy(s->bonds, bonds, sizeof(int) * 2 * nbonds);
                  memcpy(s->bond_parms, bond_parms, sizeof(double) * 3 * nbonds); }
    s->F = (double*)calloc(n, sizeof(double));
    s->Rf = (double*)calloc(n, sizeof(double));
    s->Rs = (double*)calloc(n, sizeof(double));
    s->X = (double*)calloc(n, sizeof(double));
    s->coord_prev = (double*)malloc(sizeof(double) * 4 * N);
    s->F_prev = (double*)malloc(sizeof(double) * n);
    s->move_attempts = 10000;   /* diff_decl.c:71 */
    s->last_dt = p->dt;
    if (p->hydro == 1)
    {
        s->D = (double*)malloc(sizeof(double) * n * n);
        s->L = (double*)malloc(sizeof(double) * n * n);
    }
    s->rng = orc_srand48(seed);
    /* initial forces (main_loop.c:91) */
    orc_forces(N, s->coord, s->Q, s->LJ, s->nb, s->bonds, s->bond_parms, &s->p, s->F, NULL);
    return s;
}

/* The move of compute_hydro_ermak in a flow field (diff_equation.c:265-285): with
   t = dt G r at the step's start, r' = ((a + b) + r) + t, and FS = t + dt G r' (r' before
   the box wrap), which the IG-T corrector adds half of (sim_move_fs). */
static void sim_move_flow(orc_sim* s, const double* a, const double* b, double dt)
{
    const double* G = s->G;
    int i, k;
    for (i = 0; i < s->N; ++i)
    {
        double* c = s->coord + 4 * i;
        double t[3], o[3];
        for (k = 0; k < 3; ++k) t[k] = dt * (G[3 * k] * c[0] + G[3 * k + 1] * c[1] + G[3 * k + 2] * c[2]);
        for (k = 0; k < 3; ++k) o[k] = a[3 * i + k] + b[3 * i + k] + c[k] + t[k];
        for (k = 0; k < 3; ++k)
            s->FS[3 * i + k] = t[k] + dt * (G[3 * k] * o[0] + G[3 * k + 1] * o[1] + G[3 * k + 2] * o[2]);
        for (k = 0; k < 3; ++k)
        {
            c[k] = o[k];
            if (s->p.bc == 1)
            {
                double inv = 1 / s->p.box;
                c[k] -= s->p.box * floor(c[k] * inv + 0.5f);
            }
        }
    }
}

/* The IG-T corrector in a flow field (diff_equation.c:304-312): r += (a + b) + FS / 2. */
static void sim_move_fs(orc_sim* s, const double* a, const double* b)
{
    int i, k;
    for (i = 0; i < s->N; ++i)
        for (k = 0; k < 3; ++k)
        {
            s->coord[4 * i + k] += a[3 * i + k] + b[3 * i + k] + 0.5 * s->FS[3 * i + k];
            if (s->p.bc == 1)
            {
                double inv = 1 / s->p.box;
                s->coord[4 * i + k] -= s->p.box * floor(s->coord[4 * i + k] * inv + 0.5f);
            }
        }
}

/* one attempt of a step of length dt from the current state: Rf, Rs, move, forces;
   returns orc_forces' status (> 0 repeat, < 0 overstretched bond) */
static int sim_attempt(orc_sim* s, double dt, const double* X_in, double preDii)
{
    const int N = s->N, n = 3 * N;
    const double pre = dt / (KB * s->p.T);
    int i;
    /* deterministic part Rf = dt/kT D F (diff_equation.c:119-214) */
    if (s->p.hydro == 1)
    {
        orc_symv_lower(n, s->D, n, s->F, s->Rf);
        for (i = 0; i < n; ++i) s->Rf[i] *= pre;
    }
    else if (s->p.hydro == 2)
        orc_rf_onfly(N, s->coord, &s->p, dt, s->F, s->Rf);
    else
        for (i = 0; i < n; ++i) s->Rf[i] = pre * (preDii / s->coord[4 * (i / 3) + 3]) * s->F[i];
    /* random part (diff_rand.c:47-236) */
    if (X_in) memcpy(s->X, X_in, sizeof(double) * n);
    else orc_gauss(&s->rng, s->X, n);
    if (s->p.hydro == 1)
    {
        double a = sqrt(2 * dt);
        for (i = 0; i < n; ++i) s->X[i] *= a;
        orc_trmv_lower(n, s->L, n, s->X, s->Rs);
    }
    else if (s->p.hydro == 2)
        orc_geyer_rs(N, s->coord, &s->p, dt, s->X, s->Rs);
    else
        for (i = 0; i < n; ++i) s->Rs[i] = s->X[i] * sqrt((preDii / s->coord[4 * (i / 3) + 3]) * 2 * dt);
    /* move (diff_equation.c:254-285) */
    if (s->flow)
        sim_move_flow(s, s->Rf, s->Rs, dt);
    else
        for (i = 0; i < N; ++i)
        {
            int k;
            for (k = 0; k < 3; ++k)
                s->coord[4 * i + k] = s->Rf[3 * i + k] + s->Rs[3 * i + k] + s->coord[4 * i + k];
            if (s->p.bc == 1)
            {
                double inv = 1 / s->p.box;
                for (k = 0; k < 3; ++k)
                    s->coord[4 * i + k] -= s->p.box * floor(s->coord[4 * i + k] * inv + 0.5f);
            }
        }
    return orc_forces(N, s->coord, s->Q, s->LJ, s->nb, s->bonds, s->bond_parms, &s->p, s->F, NULL);
}

/* comp_step_ermak (diff_comp.c:152-179): D and its factor once per step at the step's
   starting positions; an attempt whose forces report an overlap (check_overlap) is undone
   (positions and forces restored, load_coord / load_forces) and redrawn with the next
   Gaussians of the stream, ermak_var halving dt for the first 16 rejections; more than
   move_attempts rejections fail. */
/* Rf = pre D F (hydro cholesky and geyer: D stored in Dm; none: diag_D) */
static void sim_rf(orc_sim* s, const double* Dm, double pre, double preDii, double* out)
{
    const int n = 3 * s->N;
    int i;
    if (s->p.hydro == 1 || s->p.hydro == 2)
    {
        orc_symv_lower(n, Dm, n, s->F, out);
        for (i = 0; i < n; ++i) out[i] *= pre;
    }
    else
        for (i = 0; i < n; ++i) out[i] = pre * (preDii / s->coord[4 * (i / 3) + 3]) * s->F[i];
}

/* Rs from the next Gaussians of the stream: L sqrt(2 dt) X, the TEA displacement over the
   stored tensor in Lm (hydro geyer), or X sqrt(2 dt diag_D) */
static void sim_rs(orc_sim* s, const double* Lm, double dt, double preDii)
{
    const int n = 3 * s->N;
    int i;
    orc_gauss(&s->rng, s->X, n);
    if (s->p.hydro == 2)
        orc_geyer_rs_D(s->N, Lm, n, s->coord, &s->p, dt, s->X, s->Rs);
    else if (s->p.hydro == 1)
    {
        double a = sqrt(2 * dt);
        for (i = 0; i < n; ++i) s->X[i] *= a;
        orc_trmv_lower(n, Lm, n, s->X, s->Rs);
    }
    else
        for (i = 0; i < n; ++i) s->Rs[i] = s->X[i] * sqrt((preDii / s->coord[4 * (i / 3) + 3]) * 2 * dt);
}

/* coord = coord + (a + b), then the box wrap (diff_equation.c:254-264, :296-306) */
static void sim_move(orc_sim* s, const double* a, const double* b)
{
    int i, k;
    for (i = 0; i < s->N; ++i)
        for (k = 0; k < 3; ++k)
        {
            s->coord[4 * i + k] = (a[3 * i + k] + b[3 * i + k]) + s->coord[4 * i + k];
            if (s->p.bc == 1)
            {
                double inv = 1 / s->p.box;
                s->coord[4 * i + k] -= s->p.box * floor(s->coord[4 * i + k] * inv + 0.5f);
            }
        }
}

/* One IG-T predictor-corrector step, comp_step_IGT_const / _var (diff_comp.c:76-144) over
   compute_hydro_IGT and its second phase (diff_equation.c:288-374):
     predictor  r* = r0 + Rf(r0) + L(r0) sqrt(2 dt) X1, forces at r*;
     corrector  r  = r0 + (Rf(r0) + Rf(r*)) / 2 + L_avg sqrt(2 dt) X2,
                L_avg the factor of (D(r0) + D(r*)) / 2, forces at r.
   Rejections: a rejected predictor is redrawn (igt_var halves dt); a rejected corrector is
   redrawn alone with L_avg kept (igt_const; its first rejection is not counted) or undoes
   the whole step (igt_var, dt halved, L(r0) kept).  hydro none uses diag_D throughout. */
static int sim_step_igt(orc_sim* s, double preDii)
{
    const int N = s->N, n = 3 * N;
    const int var = s->algorithm == 4;
    double dt = s->p.dt;
    int count = 0, i;
    memcpy(s->coord_prev, s->coord, sizeof(double) * 4 * N);
    memcpy(s->F_prev, s->F, sizeof(double) * n);
    if (s->p.hydro == 1)
    {
        orc_compute_D(N, s->coord, &s->p, s->D, n);
        memcpy(s->L, s->D, sizeof(double) * (size_t)n * n);
        if (orc_cholesky(n, s->L, n)) return 1;
    }
    else if (s->p.hydro == 2)
    {
        /* TEA noise reads the stored tensor itself (input.c:497-501 turns the on-the-fly
           sums off for IG-T) */
        orc_compute_D(N, s->coord, &s->p, s->D, n);
        memcpy(s->L, s->D, sizeof(double) * (size_t)n * n);
    }
    for (;;)
    {
        const double pre = dt / (KB * s->p.T);
        int rep, fails = 0;
        /* predictor */
        sim_rf(s, s->D, pre, preDii, s->Rf);
        sim_rs(s, s->L, dt, preDii);
        if (s->flow) sim_move_flow(s, s->Rf, s->Rs, dt);
        else sim_move(s, s->Rf, s->Rs);
        rep = orc_forces(N, s->coord, s->Q, s->LJ, s->nb, s->bonds, s->bond_parms, &s->p, s->F, NULL);
        if (rep < 0) return 2;
        if (rep && s->p.check_overlap)
        {
            memcpy(s->coord, s->coord_prev, sizeof(double) * 4 * N);
            memcpy(s->F, s->F_prev, sizeof(double) * n);
            if (var && count < 16) dt /= 2;
            if (++count > s->move_attempts) return 3;
            continue;
        }
        /* averaged deterministic move and tensor at (r0, r*) */
        if (s->p.hydro == 1 || s->p.hydro == 2) orc_compute_D(N, s->coord, &s->p, s->D1, n);
        sim_rf(s, s->D1, pre, preDii, s->Rs);
        for (i = 0; i < n; ++i) s->Rf0[i] = 0.5 * (s->Rf[i] + s->Rs[i]);
        if (s->p.hydro == 1 || s->p.hydro == 2)
        {
            size_t e;
            for (e = 0; e < (size_t)n * n; ++e) s->L2[e] = (s->D[e] + s->D1[e]) / 2;
            if (s->p.hydro == 1 && orc_cholesky(n, s->L2, n)) return 1;
        }
        memcpy(s->coord, s->coord_prev, sizeof(double) * 4 * N);
        /* corrector */
        for (;;)
        {
            sim_rs(s, s->L2, dt, preDii);
            if (s->flow) sim_move_fs(s, s->Rf0, s->Rs);
            else sim_move(s, s->Rf0, s->Rs);
            rep = orc_forces(N, s->coord, s->Q, s->LJ, s->nb, s->bonds, s->bond_parms, &s->p, s->F, NULL);
            if (rep < 0) return 2;
            if (!(rep && s->p.check_overlap)) break;
            memcpy(s->coord, s->coord_prev, sizeof(double) * 4 * N);
            if (var) break;
            if (fails++ > 0 && ++count > s->move_attempts) return 3;
        }
        if (!(rep && s->p.check_overlap)) break;
        /* igt_var: the whole step is undone */
        memcpy(s->F, s->F_prev, sizeof(double) * n);
        if (count < 16) dt /= 2;
        if (++count > s->move_attempts) return 3;
    }
    memcpy(s->Rf, s->Rf0, sizeof(double) * n);
    s->last_dt = dt;
    s->rejections += count;
    s->steps_done++;
    return 0;
}

int orc_sim_run(orc_sim* s, int steps, const double* xi)
{
    const int N = s->N, n = 3 * N;
    double preDii, preDij;
    int st;
    orc_prefactors(&s->p, &preDii, &preDij);
    if (s->algorithm >= 3)
    {
        if (xi) return 4;   /* IG-T: stream noise */
        for (st = 0; st < steps; ++st)
        {
            int rc = sim_step_igt(s, preDii);
            if (rc) return rc;
        }
        return 0;
    }
    for (st = 0; st < steps; ++st)
    {
        double dt = s->p.dt;
        int count = 0;
        if (s->p.hydro == 1)
        {
            orc_compute_D(N, s->coord, &s->p, s->D, n);
            if (!s->have_L || (s->steps_done % s->refactor_every) == 0)
            {
                memcpy(s->L, s->D, sizeof(double) * (size_t)n * n);
                if (orc_cholesky(n, s->L, n)) return 1;
                s->have_L = 1;
            }
        }
        memcpy(s->coord_prev, s->coord, sizeof(double) * 4 * N);
        memcpy(s->F_prev, s->F, sizeof(double) * n);
        for (;;)
        {
            /* supplied Gaussians serve the first attempt; redraws come from the stream */
            const double* X = (xi && count == 0) ? xi + (size_t)st * n : NULL;
            int rep = sim_attempt(s, dt, X, preDii);
            if (rep < 0) return 2;
            if (rep == 0 || !s->p.check_overlap) break;
            if (s->algorithm == 1 && count < 16) dt /= 2;
            memcpy(s->coord, s->coord_prev, sizeof(double) * 4 * N);
            memcpy(s->F, s->F_prev, sizeof(double) * n);
            if (++count > s->move_attempts) return 3;
        }
        s->last_dt = dt;
        s->rejections += count;
        s->steps_done++;
    }
    return 0;
}

void orc_sim_set_algorithm(orc_sim* s, int algorithm, int move_attempts)
{
    const size_t n = 3 * (size_t)s->N;
    s->algorithm = algorithm;
    s->move_attempts = move_attempts;
    if (algorithm >= 3 && !s->Rf0)
    {
        s->Rf0 = (double*)calloc(n, sizeof(double));
        if (s->p.hydro == 2 && !s->D)
        {
            s->D = (double*)malloc(sizeof(double) * n * n);
            s->L = (double*)malloc(sizeof(double) * n * n);
        }
        if (s->p.hydro == 1 || s->p.hydro == 2)
        {
            s->D1 = (double*)malloc(sizeof(double) * n * n);
            s->L2 = (double*)malloc(sizeof(double) * n * n);
        }
    }
}

void orc_sim_set_flow(orc_sim* s, const double* G)
{
    s->flow = G != NULL;
    if (!G) return;
    memcpy(s->G, G, sizeof(s->G));
    if (!s->FS) s->FS = (double*)calloc(3 * (size_t)s->N, sizeof(double));
}

double orc_sim_last_dt(const orc_sim* s) { return s->last_dt; }
long long orc_sim_rejections(const orc_sim* s) { return s->rejections; }

void orc_sim_coords(const orc_sim* s, double* c) { memcpy(c, s->coord, sizeof(double) * 4 * s->N); }
void orc_sim_forces(const orc_sim* s, double* F) { memcpy(F, s->F, sizeof(double) * 3 * s->N); }
void orc_sim_last(const orc_sim* s, double* Rf, double* Rs)
{
    memcpy(Rf, s->Rf, sizeof(double) * 3 * s->N);
    memcpy(Rs, s->Rs, sizeof(double) * 3 * s->N);
}
uint64_t orc_sim_rng_state(const orc_sim* s) { return s->rng; }

void orc_sim_destroy(orc_sim* s)
{
    if (!s) return;
    free(s->coord); free(s->Q); free(s->LJ); free(s->bonds); free(s->bond_parms);
    free(s->F); free(s->D); free(s->L); free(s->Rf); free(s->Rs); free(s->X);
    free(s->coord_prev); free(s->F_prev);
    free(s->D1); free(s->L2); free(s->Rf0); free(s->FS);
    free(s);
}
