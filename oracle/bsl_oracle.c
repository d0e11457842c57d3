/*
 * bsl_oracle.c -- CPU oracle (test infrastructure only, see bsl_oracle.h).
 *
 * Plain C99 + OpenMP, FP64 throughout, restated from the reference sources.
 * Loop nests, summation order and OpenMP placement follow the reference so
 * that (a) results agree with the reference to the last bit where the
 * reference is deterministic and (b) the CPU baseline times the same work
 * (serial kinetic term, OpenMP potential term).  Every function cites the
 * reference file:line it restates.  Build without -ffast-math and without
 * FMA contraction (-ffp-contract=off) so the operation order is kept.
 */
#include "bsl_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* basis                                                                     */
/* ------------------------------------------------------------------------ */

/* linRotCartSphHvRoutines.cpp:8-17 */
static double factorial(int num)
{
    double fact = 1.0;
    for (int i = 2; i <= num; i++) fact *= (double)i;
    return fact;
}

/* linRotCartSphHvRoutines.cpp:24-68: m outer (-lmax..lmax), l inner (|m|..lmax) */
int bso_gen_indices_lm(int lmax, int *qnum)
{
    int i = 0;
    for (int m = -lmax; m <= lmax; m++)
        for (int l = abs(m); l <= lmax; l++) {
            qnum[2 * i] = l;
            qnum[2 * i + 1] = m;
            i++;
        }
    return i;
}

/* linRotCartSphHvRoutines.cpp:219-255 (Newton iteration to DBL_EPSILON) */
void bso_gauleg(double x1, double x2, double *x, double *w, int n)
{
    int m = (n + 1) / 2;
    double xm = 0.5 * (x2 + x1);
    double xl = 0.5 * (x2 - x1);
    for (int i = 1; i <= m; i++) {
        double z = cos(M_PI * ((double)i - 0.25) / ((double)n + 0.5));
        double z1, pp, p1, p2, p3;
        do {
            p1 = 1.0;
            p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                p3 = p2;
                p2 = p1;
                p1 = ((2.0 * (double)j - 1.0) * z * p2 - ((double)j - 1.0) * p3) / (double)j;
            }
            pp = (double)n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
        } while (fabs(z - z1) > DBL_EPSILON);
        x[i - 1] = xm - xl * z;
        x[n - i] = xm + xl * z;
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
        w[n - i] = w[i - 1];
    }
}

/* linRotCartSphHvRoutines.cpp:450-463 */
void bso_gauss_chebyshev(int n, double *x, double *w)
{
    for (int i = 0; i < n; i++) {
        x[i] = cos(BSO_PI * (2.0 * (double)(i + 1) - 1.0) / (2.0 * (double)n));
        w[i] = BSO_PI / (double)n;
    }
}

/* linRotCartSphHvRoutines.cpp:293-403: P_l^m by the diagonal and column
 * recurrences, times (-1)^m sqrt((l-m)!(2l+1)/(2(l+m)!)); |m| for m<0 */
void bso_norm_assoc_legendre(int lmax, const int *qnum, int nn, double x, double *out)
{
    int L = lmax + 1;
    double *P = (double *)calloc((size_t)L * L, sizeof(double));
#define PA(l, m) P[(l) * L + (m)]
    double rec = sqrt(1.0 - x * x);
    PA(0, 0) = 1.0;
    for (int l = 1; l <= lmax; l++) {
        double ld = (double)l;
        PA(l, l) = (2.0 * ld - 1.0) * PA(l - 1, l - 1) * rec;
    }
    for (int m = 0; m <= lmax - 1; m++)
        for (int l = m; l <= lmax - 1; l++) {
            double ld = (double)l, md = (double)m;
            double prev = ((l - 1) < m) ? 0.0 : PA(l - 1, m);
            PA(l + 1, m) = (((2.0 * ld + 1.0) * x * PA(l, m)) - ((ld + md) * prev)) / (ld - md + 1.0);
        }
    for (int l = 0; l <= lmax; l++) {
        double phase = 1.0;
        for (int m = 0; m <= l; m++) {
            double ld = (double)l, md = (double)m;
            double pnf = factorial(l - m) * (2.0 * ld + 1.0) / (2.0 * factorial(l + m));
            (void)md;
            PA(l, m) *= phase * sqrt(pnf);
            phase *= -1.0;
        }
    }
    for (int n = 0; n < nn; n++) {
        int l = qnum[2 * n], m = qnum[2 * n + 1];
        out[n] = PA(l, abs(m));
    }
#undef PA
    free(P);
}

/* linRotCartSphHvRoutines.cpp:405-434 */
void bso_tesseral_trig(const int *qnum, int nn, double phi, double *out)
{
    double spi = 1.0 / sqrt(BSO_PI);
    for (int n = 0; n < nn; n++) {
        int m = qnum[2 * n + 1];
        if (m == 0)
            out[n] = spi / sqrt(2.0);
        else if (m > 0)
            out[n] = spi * cos((double)m * phi);
        else
            out[n] = spi * sin((double)(-m) * phi);
    }
}

/* linRotCartSphHvRoutines.cpp:484-517 (Colbert-Miller sinc DVR) */
void bso_cart_kin_grid(double x_max, int n, double mass, double *kin, double *grid)
{
    double dx = x_max / (double)(n - 1);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double v = (BSO_H_BAR * BSO_H_BAR) / (2.0 * mass * dx * dx) * pow(-1.0, (double)((i + 1) - (j + 1)));
            if (i == j)
                v *= (BSO_PI * BSO_PI) / 3.0;
            else
                v *= 2.0 / (double)(((i + 1) - (j + 1)) * ((i + 1) - (j + 1)));
            kin[i * n + j] = v;
        }
    for (int i = 0; i < n; i++) grid[i] = (double)i * dx - x_max / 2.0;
}

/* linRotCartSphHvRoutines.cpp:522-538 */
void bso_rot_kin(const int *qnum, int nn, double B, double *out)
{
    for (int i = 0; i < nn; i++) {
        double l = (double)qnum[2 * i];
        out[i] = B * l * (l + 1.0);
    }
}

/* ------------------------------------------------------------------------ */
/* sites and pair energies                                                   */
/* ------------------------------------------------------------------------ */

/* rotmat.cpp:6-32 (ZY'Z'' Euler matrix), stored row-major here */
static void rot_zyz(const double *ea, double *mat)
{
    double cp = cos(ea[0]), sp = sin(ea[0]);
    double ct = cos(ea[1]), st = sin(ea[1]);
    double cx = cos(ea[2]), sx = sin(ea[2]);
    mat[0] = (cp * ct * cx) - (sp * sx);
    mat[1] = (-1 * cp * ct * sx) - (sp * cx);
    mat[2] = cp * st;
    mat[3] = (sp * ct * cx) + (cp * sx);
    mat[4] = (-1 * sp * ct * sx) + (cp * cx);
    mat[5] = sp * st;
    mat[6] = -1 * st * cx;
    mat[7] = st * sx;
    mat[8] = ct;
}

/* SPCE_AH2_EngRoutines.h:17-32 and TIP4P_AH2_EngRoutines.h:16-35 body-frame
 * sites (nm) and charges; getSPCEatoms SPCE_AH2_EngRoutines.cpp:5-108,
 * getTIP4Patoms TIP4P_AH2_EngRoutines.cpp:5-117 */
int bso_build_sites(int model, int nmol, const double *cm_ea, double *pos, double *q, char *type)
{
    static const double spce[3][3] = {{0., 0., 0.006461408220},
                                      {-8.164965810E-2, 0., -0.05127361870},
                                      {8.164965810E-2, 0., -0.05127361870}};
    static const double spce_q[3] = {-0.8476, 0.4238, 0.4238};
    static const double tip4p[4][3] = {{0., 0., 0.00655538},
                                       {-0.075695, 0., -0.0520326},
                                       {0.075695, 0., -0.0520326},
                                       {0., 0., -0.00844462}};
    static const double tip4p_q[4] = {0.0, 0.520, 0.520, -1.040};
    static const char types[4] = {'O', 'H', 'H', 'M'};
    int ns = model == 0 ? 3 : 4;
    for (int i = 0; i < nmol; i++) {
        const double *r = cm_ea + 6 * i;
        double cm[3] = {r[0], r[2], r[4]};
        double ea[3] = {r[1], r[3], r[5]};
        double mat[9];
        rot_zyz(ea, mat);
        for (int s = 0; s < ns; s++) {
            const double *b = model == 0 ? spce[s] : tip4p[s];
            for (int c = 0; c < 3; c++) {
                /* matvec2.cpp:133-158: prod = 0; prod += M(c,k) b(k) */
                double prod = 0.;
                for (int k = 0; k < 3; k++) prod += mat[c * 3 + k] * b[k];
                /* O_mol + NM_PER_ANG*CM (matvec2.cpp:125-131, :531-539) */
                pos[(ns * i + s) * 3 + c] = prod + BSO_NM_PER_ANG * cm[c];
            }
            q[ns * i + s] = model == 0 ? spce_q[s] : tip4p_q[s];
            type[ns * i + s] = types[s];
        }
    }
    return nmol * ns;
}

/* energyRoutines.cpp:5-25 CoulombEng: K q1 q2 / |q2 - q1| (eNorm matvec2.cpp:1324) */
static inline double coulomb(const double *p, double q, const double *a, double qa)
{
    double d0 = a[0] - p[0], d1 = a[1] - p[1], d2 = a[2] - p[2];
    double s = 0.;
    s += d0 * d0;
    s += d1 * d1;
    s += d2 * d2;
    double dist = sqrt(s);
    return BSO_K_CONST * q * qa / dist;
}

/* Q_SPCE_Eng SPCE_AH2_EngRoutines.cpp:110-132 (site order, charge per site).
 * TIP4P sites with zero charge ('O' in Q_TIP4P_Eng) are skipped there; here a
 * zero-charge site is skipped the same way. */
double bso_coulomb_sum(const double *pos3, double q, int nsites, const double *site_pos, const double *site_q)
{
    double e = 0;
    for (int i = 0; i < nsites; i++) {
        if (site_q[i] == 0.0) continue;
        e += coulomb(pos3, q, site_pos + 3 * i, site_q[i]);
    }
    return e;
}

/* LJ_SPCE_Eng_Fast SPCE_AH2_EngRoutines.cpp:157-178; LJEngFast energyRoutines.cpp:42-54 */
double bso_lj_sum(const double *pos3, int nsites, const double *site_pos, const char *site_type, double A, double B)
{
    double e = 0;
    for (int i = 0; i < nsites; i++) {
        if (site_type[i] != 'O') continue;
        const double *a = site_pos + 3 * i;
        double d0 = a[0] - pos3[0], d1 = a[1] - pos3[1], d2 = a[2] - pos3[2];
        double s = 0.;
        s += d0 * d0;
        s += d1 * d1;
        s += d2 * d2;
        double dist = sqrt(s);
        e += (A / pow(dist, 12)) - (B / pow(dist, 6));
    }
    return e;
}

/* ------------------------------------------------------------------------ */
/* prepared state                                                            */
/* ------------------------------------------------------------------------ */

struct bso_state {
    bso_problem p;
    int nx, ny, nz, nn, lmax, na, nb;
    long long P, N;
    double *kx, *ky, *kz, *gx, *gy, *gz;
    int *qnum;
    double *rot, *sym;
    double *ct, *wt, *th, *cp, *wp, *ph1, *ph2;
    double *L;        /* [n][a] */
    double *S1, *S2;  /* [m+lmax][b] */
    double *v1, *v2;  /* [P(i slowest)][a][b] */
    double *vblk;     /* [P(i slowest)][n][n'] */
};

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* IsoHarmOsc_Routines.cpp:6-31 */
static double iho_point(const bso_problem *p, const double *cm)
{
    double d2 = 0.;
    d2 += cm[0] * cm[0];
    d2 += cm[1] * cm[1];
    d2 += cm[2] * cm[2];
    double v = p->iho_factor * d2;
    if (fabs(v) > fabs(p->iho_ceil)) v = p->iho_ceil;
    return v;
}

/* Grid-mapped Alavi cage (Alavi_H2_Routines.cpp:5-317, gridFcns.cpp:6-101):
 * centre and H-atom potentials tabulated on the universe grid (point
 * u = i0 + n0 (i1 + n1 i2) at i_d dx_d - max_d/2, dx_d = max_d/(n_d - 1)),
 * looked up at the nearest universe point of the centre and of the two H
 * atoms (round half away from zero, kept inside the universe). */
static int universe_index(double c, double maxd, double dd, int n)
{
    double cmin = (-0.5) * maxd;
    int i = (int)round((c - cmin) / dd);
    if (i <= 0) i = 0;
    else if (i >= n) i = n - 1;
    return i;
}

static void build_vgrid_universe(bso_state *s)
{
    const bso_problem *p = &s->p;
    int nx = s->nx, ny = s->ny, nz = s->nz, na = s->na, nb = s->nb;
    int un[3] = {p->pnx, p->pny, p->pnz};
    double um[3] = {p->px_max, p->py_max, p->pz_max}, ud[3];
    for (int d = 0; d < 3; d++) ud[d] = um[d] / ((double)(un[d] - 1));
    long long U = (long long)un[0] * un[1] * un[2];
    double *cmu = dalloc((size_t)U), *hu = dalloc((size_t)U);
#pragma omp parallel for schedule(guided)
    for (long long u = 0; u < U; u++) {
        int i0 = (int)(u % un[0]), i1 = (int)((u / un[0]) % un[1]), i2 = (int)(u / ((long long)un[0] * un[1]));
        double g[3] = {((double)i0 * ud[0]) - (um[0] / 2.0), ((double)i1 * ud[1]) - (um[1] / 2.0),
                       ((double)i2 * ud[2]) - (um[2] / 2.0)};
        cmu[u] = bso_coulomb_sum(g, p->q_h2_cm, p->nsites, p->site_pos, p->site_q);
        cmu[u] += bso_lj_sum(g, p->nsites, p->site_pos, p->site_type, p->lj_a, p->lj_b);
        hu[u] = bso_coulomb_sum(g, p->q_h2_h, p->nsites, p->site_pos, p->site_q);
    }
#pragma omp parallel for schedule(guided) collapse(3)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                double c[3] = {s->gx[i], s->gy[j], s->gz[k]};
                size_t ind = (((size_t)i * ny + j) * nz + k) * na;
                for (int a = 0; a < na; a++)
                    for (int b = 0; b < nb; b++)
                        for (int half = 0; half < 2; half++) {
                            double th = s->th[a], ph = half == 0 ? s->ph1[b] : s->ph2[b];
                            double e[3] = {sin(th) * cos(ph), sin(th) * sin(ph), cos(th)};
                            int ic[3], i1[3], i2[3];
                            for (int d = 0; d < 3; d++) {
                                ic[d] = universe_index(c[d], um[d], ud[d], un[d]);
                                i1[d] = universe_index(c[d] + e[d] * p->h2_half_bond, um[d], ud[d], un[d]);
                                i2[d] = universe_index(c[d] + e[d] * (-p->h2_half_bond), um[d], ud[d], un[d]);
                            }
                            double v = 0.0;
                            v += cmu[ic[0] + (long long)un[0] * (ic[1] + (long long)un[1] * ic[2])];
                            v += hu[i1[0] + (long long)un[0] * (i1[1] + (long long)un[1] * i1[2])];
                            v += hu[i2[0] + (long long)un[0] * (i2[1] + (long long)un[1] * i2[2])];
                            if (v >= p->ceiling) v = p->ceiling;
                            if (half == 0)
                                s->v1[(ind + a) * nb + b] = v;
                            else
                                s->v2[(ind + a) * nb + b] = v;
                        }
            }
    free(cmu);
    free(hu);
}

/* Potential grids.
 *  pot_kind 0: calc_Vlmlpmp_NoQuad linRotCartSphHvFullVPrepRoutines.cpp:134-184
 *              (point function per (P,a,b), clamp V >= ceiling)
 *  pot_kind 1: Alavi_SiteSite_Eng_contGrid Alavi_ContGrid_H2_Routines.cpp:6-147 */
static void build_vgrid(bso_state *s)
{
    const bso_problem *p = &s->p;
    int nx = s->nx, ny = s->ny, nz = s->nz, na = s->na, nb = s->nb;
    double ceil_v = p->ceiling;
    if (p->pot_kind == 0) {
#pragma omp parallel for schedule(guided) collapse(3)
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                for (int k = 0; k < nz; k++) {
                    double cm[3] = {s->gx[i], s->gy[j], s->gz[k]};
                    for (int a = 0; a < na; a++)
                        for (int b = 0; b < nb; b++) {
                            size_t ind = (((size_t)i * ny + j) * nz + k) * na;
                            double v1 = iho_point(p, cm);
                            if (v1 >= ceil_v) v1 = ceil_v;
                            s->v1[(ind + a) * nb + b] = v1;
                            double v2 = iho_point(p, cm);
                            if (v2 >= ceil_v) v2 = ceil_v;
                            s->v2[(ind + a) * nb + b] = v2;
                        }
                }
        return;
    }
    if (p->pot_kind == 3) {
        /* CoulombPotential Coulomb_Routines.cpp:6-31: K q1 q2 / |CM - 0| */
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                for (int k = 0; k < nz; k++) {
                    double x = s->gx[i], y = s->gy[j], z = s->gz[k];
                    double d2 = 0.;
                    d2 += x * x;
                    d2 += y * y;
                    d2 += z * z;
                    double v = BSO_K_CONST * p->coul_q_centre * p->coul_q_mol / sqrt(d2);
                    if (fabs(v) > fabs(p->coul_ceil)) v = p->coul_ceil;
                    if (v >= ceil_v) v = ceil_v;
                    size_t ind = (((size_t)i * ny + j) * nz + k) * na;
                    for (int a = 0; a < na; a++)
                        for (int b = 0; b < nb; b++) {
                            s->v1[(ind + a) * nb + b] = v;
                            s->v2[(ind + a) * nb + b] = v;
                        }
                }
        return;
    }
    if (p->pot_kind == 2) {
        build_vgrid_universe(s);
        return;
    }
    long long P = s->P;
    double *cmpot = dalloc((size_t)P);
    double *off = dalloc((size_t)na * nb * 3);
    double *offpi = dalloc((size_t)na * nb * 3);
    /* grid[ind], ind = (i*nj + j)*nk + k: Alavi_ContGrid_H2_Routines.cpp:212-226 */
#pragma omp parallel for schedule(guided)
    for (long long ip = 0; ip < P; ip++) {
        int k = (int)(ip % nz), j = (int)((ip / nz) % ny), i = (int)(ip / ((long long)nz * ny));
        double g[3] = {s->gx[i], s->gy[j], s->gz[k]};
        cmpot[ip] = bso_coulomb_sum(g, p->q_h2_cm, p->nsites, p->site_pos, p->site_q);
    }
#pragma omp parallel for schedule(guided)
    for (long long ip = 0; ip < P; ip++) {
        int k = (int)(ip % nz), j = (int)((ip / nz) % ny), i = (int)(ip / ((long long)nz * ny));
        double g[3] = {s->gx[i], s->gy[j], s->gz[k]};
        cmpot[ip] += bso_lj_sum(g, p->nsites, p->site_pos, p->site_type, p->lj_a, p->lj_b);
    }
    /* :74-90 H-H axis offsets for phi in [0,pi) and the mirrored [pi,2pi) */
    for (int a = 0; a < na; a++)
        for (int b = 0; b < nb; b++) {
            int ind = a * nb + b;
            off[3 * ind + 0] = p->h2_half_bond * sin(s->th[a]) * s->cp[b];
            off[3 * ind + 1] = p->h2_half_bond * sin(s->th[a]) * sin(s->ph1[b]);
            off[3 * ind + 2] = p->h2_half_bond * s->ct[a];
            offpi[3 * ind + 0] = off[3 * ind + 0];
            offpi[3 * ind + 1] = -1.0 * (off[3 * ind + 1]);
            offpi[3 * ind + 2] = off[3 * ind + 2];
        }
    /* :101-139 */
#pragma omp parallel for schedule(guided) collapse(3)
    for (long long ip = 0; ip < P; ip++)
        for (int a = 0; a < na; a++)
            for (int b = 0; b < nb; b++) {
                int k = (int)(ip % nz), j = (int)((ip / nz) % ny), i = (int)(ip / ((long long)nz * ny));
                double g[3] = {s->gx[i], s->gy[j], s->gz[k]};
                int ind = a * nb + b;
                size_t ind2 = ((size_t)ip * na + a) * nb + b;
                for (int half = 0; half < 2; half++) {
                    const double *o = (half == 0 ? off : offpi) + 3 * ind;
                    double h1[3] = {g[0] + o[0], g[1] + o[1], g[2] + o[2]};
                    double h2[3] = {g[0] - o[0], g[1] - o[1], g[2] - o[2]};
                    double pot = bso_coulomb_sum(h1, p->q_h2_h, p->nsites, p->site_pos, p->site_q);
                    pot += bso_coulomb_sum(h2, p->q_h2_h, p->nsites, p->site_pos, p->site_q);
                    pot += cmpot[ip];
                    if (pot >= ceil_v) pot = ceil_v;
                    if (half == 0)
                        s->v1[ind2] = pot;
                    else
                        s->v2[ind2] = pot;
                }
            }
    free(cmpot);
    free(off);
    free(offpi);
}

/* <lm|V|l'm'>(P) = sum_a sum_b (h1 V1 + h2 V2):
 * linRotCartSphHvFullVPrepRoutines.cpp:113-131 (harmonic factors) and
 * :187-218 / linRotCartSphHvContGridRoutines.cpp:92-143 (contraction) */
static void build_vblocks(bso_state *s)
{
    int nn = s->nn, na = s->na, nb = s->nb, lmax = s->lmax;
    size_t hsz = (size_t)nn * nn * na * nb;
    double *h1 = dalloc(hsz), *h2 = dalloc(hsz);
#pragma omp parallel for schedule(guided) collapse(4)
    for (int n = 0; n < nn; n++)
        for (int np = 0; np < nn; np++)
            for (int a = 0; a < na; a++)
                for (int b = 0; b < nb; b++) {
                    size_t ind2 = (((size_t)n * nn + np) * na + a) * nb + b;
                    int m = s->qnum[2 * n + 1], mp = s->qnum[2 * np + 1];
                    h1[ind2] = s->wt[a] * s->wp[b] * s->L[n * na + a] * s->S1[(m + lmax) * nb + b] *
                               s->L[np * na + a] * s->S1[(mp + lmax) * nb + b];
                    h2[ind2] = s->wt[a] * s->wp[b] * s->L[n * na + a] * s->S2[(m + lmax) * nb + b] *
                               s->L[np * na + a] * s->S2[(mp + lmax) * nb + b];
                }
    long long P = s->P;
#pragma omp parallel for schedule(guided) collapse(3)
    for (long long ip = 0; ip < P; ip++)
        for (int n = 0; n < nn; n++)
            for (int np = 0; np < nn; np++) {
                size_t ind = ((size_t)ip * nn + n) * nn + np;
                size_t ind2 = ((size_t)n * nn + np) * na;
                size_t ind3 = (size_t)ip * na;
                double acc = 0.0;
                for (int a = 0; a < na; a++)
                    for (int b = 0; b < nb; b++)
                        acc += h1[(ind2 + a) * nb + b] * s->v1[(ind3 + a) * nb + b] +
                               h2[(ind2 + a) * nb + b] * s->v2[(ind3 + a) * nb + b];
                s->vblk[ind] = acc;
            }
    free(h1);
    free(h2);
}

/* HvPrep_Internal_NoQuad linRotCartSphHvFullVPrepRoutines.cpp:229-528, with
 * QuadraturePrep linRotCartSphHvRoutines.cpp:1480-1532, TesseralPrep
 * :1534-1735 and calcSym linRotCartSphHvFullVPrepRoutines.cpp:530-573 */
static bso_state *prepare(const bso_problem *p, const double *vblk_ext);

bso_state *bso_prepare(const bso_problem *p) { return prepare(p, NULL); }

/* Same operator with the V blocks supplied by the caller (reference order
 * [i][j][k][n][n']); the potential grid and the contraction are skipped.
 * Used to time the CPU Hv on operators whose preparation would take hours
 * on the CPU. */
bso_state *bso_prepare_external(const bso_problem *p, const double *vblk) { return prepare(p, vblk); }

static bso_state *prepare(const bso_problem *p, const double *vblk_ext)
{
    if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0 || p->lmax < 0 || p->na <= 0 || p->nb <= 0) return NULL;
    if (p->sym < 0 || p->sym > 2 || p->pot_kind < 0 || p->pot_kind > 3) return NULL;
    bso_state *s = (bso_state *)calloc(1, sizeof(bso_state));
    s->p = *p;
    s->nx = p->nx;
    s->ny = p->ny;
    s->nz = p->nz;
    s->lmax = p->lmax;
    s->na = p->na;
    s->nb = p->nb;
    s->nn = (p->lmax + 1) * (p->lmax + 1);
    s->P = (long long)p->nx * p->ny * p->nz;
    s->N = s->P * s->nn;
    s->kx = dalloc((size_t)p->nx * p->nx);
    s->gx = dalloc(p->nx);
    s->ky = dalloc((size_t)p->ny * p->ny);
    s->gy = dalloc(p->ny);
    s->kz = dalloc((size_t)p->nz * p->nz);
    s->gz = dalloc(p->nz);
    bso_cart_kin_grid(p->x_max, p->nx, p->mass, s->kx, s->gx);
    bso_cart_kin_grid(p->y_max, p->ny, p->mass, s->ky, s->gy);
    bso_cart_kin_grid(p->z_max, p->nz, p->mass, s->kz, s->gz);
    s->qnum = (int *)calloc((size_t)2 * s->nn, sizeof(int));
    bso_gen_indices_lm(p->lmax, s->qnum);
    s->rot = dalloc(s->nn);
    bso_rot_kin(s->qnum, s->nn, p->rot_const, s->rot);
    s->sym = dalloc(s->nn);
    for (int n = 0; n < s->nn; n++) {
        int l = s->qnum[2 * n];
        s->sym[n] = p->sym == 0 ? 1.0 : (p->sym == 1 ? (l % 2 == 0 ? 1.0 : 0.0) : (l % 2 == 1 ? 1.0 : 0.0));
    }
    int na = p->na, nb = p->nb;
    s->ct = dalloc(na);
    s->wt = dalloc(na);
    s->th = dalloc(na);
    s->cp = dalloc(nb);
    s->wp = dalloc(nb);
    s->ph1 = dalloc(nb);
    s->ph2 = dalloc(nb);
    bso_gauleg(-1.0, 1.0, s->ct, s->wt, na);
    bso_gauss_chebyshev(nb, s->cp, s->wp);
    for (int a = 0; a < na; a++) s->th[a] = acos(s->ct[a]);
    for (int b = 0; b < nb; b++) {
        s->ph1[b] = acos(s->cp[b]);
        s->ph2[b] = 2 * BSO_PI - acos(s->cp[b]);
    }
    int nm = 2 * p->lmax + 1;
    s->L = dalloc((size_t)s->nn * na);
    s->S1 = dalloc((size_t)nm * nb);
    s->S2 = dalloc((size_t)nm * nb);
    double *tmp = dalloc(s->nn);
    for (int a = 0; a < na; a++) {
        bso_norm_assoc_legendre(p->lmax, s->qnum, s->nn, s->ct[a], tmp);
        for (int n = 0; n < s->nn; n++) s->L[n * na + a] = tmp[n];
    }
    /* S_m[m][b] taken from the l = lmax entry of each m (TesseralPrep :1590-1606) */
    for (int b = 0; b < nb; b++) {
        bso_tesseral_trig(s->qnum, s->nn, acos(s->cp[b]), tmp);
        for (int m = -p->lmax; m <= p->lmax; m++) {
            int n = 0;
            for (int q = 0; q < s->nn; q++)
                if (s->qnum[2 * q] == p->lmax && s->qnum[2 * q + 1] == m) n = q;
            s->S1[(m + p->lmax) * nb + b] = tmp[n];
        }
        bso_tesseral_trig(s->qnum, s->nn, 2.0 * BSO_PI - acos(s->cp[b]), tmp);
        for (int m = -p->lmax; m <= p->lmax; m++) {
            int n = 0;
            for (int q = 0; q < s->nn; q++)
                if (s->qnum[2 * q] == p->lmax && s->qnum[2 * q + 1] == m) n = q;
            s->S2[(m + p->lmax) * nb + b] = tmp[n];
        }
    }
    free(tmp);
    if (vblk_ext) {
        s->v1 = dalloc(1);
        s->v2 = dalloc(1);
        s->vblk = (double *)malloc((size_t)s->P * s->nn * s->nn * sizeof(double));
        memcpy(s->vblk, vblk_ext, (size_t)s->P * s->nn * s->nn * sizeof(double));
        return s;
    }
    s->v1 = dalloc((size_t)s->P * na * nb);
    s->v2 = dalloc((size_t)s->P * na * nb);
    build_vgrid(s);
    s->vblk = dalloc((size_t)s->P * s->nn * s->nn);
    build_vblocks(s);
    return s;
}

void bso_free(bso_state *s)
{
    if (!s) return;
    free(s->kx); free(s->ky); free(s->kz); free(s->gx); free(s->gy); free(s->gz);
    free(s->qnum); free(s->rot); free(s->sym);
    free(s->ct); free(s->wt); free(s->th); free(s->cp); free(s->wp); free(s->ph1); free(s->ph2);
    free(s->L); free(s->S1); free(s->S2); free(s->v1); free(s->v2); free(s->vblk);
    free(s);
}

long long bso_basis_size(const bso_state *s) { return s->N; }
int bso_nn(const bso_state *s) { return s->nn; }
void bso_get_vgrid(const bso_state *s, double *v1, double *v2)
{
    size_t n = (size_t)s->P * s->na * s->nb;
    memcpy(v1, s->v1, n * sizeof(double));
    memcpy(v2, s->v2, n * sizeof(double));
}
void bso_get_vblocks(const bso_state *s, double *out)
{
    memcpy(out, s->vblk, (size_t)s->P * s->nn * s->nn * sizeof(double));
}
void bso_get_kin(const bso_state *s, int axis, double *out)
{
    int n = axis == 0 ? s->nx : axis == 1 ? s->ny : s->nz;
    const double *k = axis == 0 ? s->kx : axis == 1 ? s->ky : s->kz;
    memcpy(out, k, (size_t)n * n * sizeof(double));
}
void bso_get_grid(const bso_state *s, int axis, double *out)
{
    int n = axis == 0 ? s->nx : axis == 1 ? s->ny : s->nz;
    const double *g = axis == 0 ? s->gx : axis == 1 ? s->gy : s->gz;
    memcpy(out, g, (size_t)n * sizeof(double));
}
void bso_get_rot(const bso_state *s, double *out) { memcpy(out, s->rot, s->nn * sizeof(double)); }
void bso_get_sym(const bso_state *s, double *out) { memcpy(out, s->sym, s->nn * sizeof(double)); }
void bso_get_qnum(const bso_state *s, int *out) { memcpy(out, s->qnum, 2 * s->nn * sizeof(int)); }
void bso_get_quad(const bso_state *s, double *ct, double *wt, double *cp, double *wp)
{
    memcpy(ct, s->ct, s->na * sizeof(double));
    memcpy(wt, s->wt, s->na * sizeof(double));
    memcpy(cp, s->cp, s->nb * sizeof(double));
    memcpy(wp, s->wp, s->nb * sizeof(double));
}
void bso_get_tess(const bso_state *s, double *L, double *S1, double *S2)
{
    int nm = 2 * s->lmax + 1;
    memcpy(L, s->L, (size_t)s->nn * s->na * sizeof(double));
    memcpy(S1, s->S1, (size_t)nm * s->nb * sizeof(double));
    memcpy(S2, s->S2, (size_t)nm * s->nb * sizeof(double));
}

/* ------------------------------------------------------------------------ */
/* Hv                                                                        */
/* ------------------------------------------------------------------------ */

/* Hv_linRotCartSph_NoQuad linRotCartSphHvFullVPrepRoutines.cpp:807-822 ->
 * Hv_5D_oneCompositeIndex_NoQuad :736-796.
 *   Tv: Tv_5D_oneCompositeIndex linRotCartSphHvRoutines.cpp:1872-1969 (serial):
 *       Mv_5D sums sum_ip T[i][ip] v[..ip] from 0.0 in ip order (:1807-1829);
 *       the partial results are added as ((Tx + Ty) + Tz) + Trot.
 *   Vv: Vv_5D_oneCompositeIndex_NoQuad :675-734 (OpenMP), sum over n' in order.
 *   Hv = (Tv + Vv) * S_n; then uec += Hv (serial, :816-818). */
void bso_hv(const bso_state *s, const double *vec, double *uec)
{
    int nx = s->nx, ny = s->ny, nz = s->nz, nn = s->nn;
    long long N = s->N;
    double *tv = dalloc((size_t)N), *vv = dalloc((size_t)N);
    /* kinetic + rotor, serial as in the reference */
    for (int n = 0; n < nn; n++)
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    size_t base = (((size_t)n * nz + k) * ny + j) * nx;
                    double tx = 0.0, ty = 0.0, tz = 0.0;
                    for (int ip = 0; ip < nx; ip++) tx += s->kx[i * nx + ip] * vec[base + ip];
                    for (int jp = 0; jp < ny; jp++)
                        ty += s->ky[j * ny + jp] * vec[(((size_t)n * nz + k) * ny + jp) * nx + i];
                    for (int kp = 0; kp < nz; kp++)
                        tz += s->kz[k * nz + kp] * vec[(((size_t)n * nz + kp) * ny + j) * nx + i];
                    tv[base + i] = ((tx + ty) + tz) + s->rot[n] * vec[base + i];
                }
    /* potential (reshuffle + block GEMV) */
    double *vnp = dalloc((size_t)N);
#pragma omp parallel
    {
#pragma omp for schedule(guided) collapse(4)
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                for (int k = 0; k < nz; k++)
                    for (int np = 0; np < nn; np++)
                        vnp[(((size_t)k * ny + j) * nx + i) * nn + np] =
                            vec[(((size_t)np * nz + k) * ny + j) * nx + i];
#pragma omp for schedule(guided) collapse(4)
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                for (int k = 0; k < nz; k++)
                    for (int n = 0; n < nn; n++) {
                        size_t ind = (((size_t)n * nz + k) * ny + j) * nx + i;
                        size_t ind2 = ((((size_t)i * ny + j) * nz + k) * nn + n) * nn;
                        size_t ind3 = (((size_t)k * ny + j) * nx + i) * nn;
                        double acc = 0.0;
                        for (int np = 0; np < nn; np++) acc += s->vblk[ind2 + np] * vnp[ind3 + np];
                        vv[ind] = acc;
                    }
    }
    free(vnp);
#pragma omp parallel for schedule(guided) collapse(4)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++)
                for (int n = 0; n < nn; n++) {
                    size_t ind = (((size_t)n * nz + k) * ny + j) * nx + i;
                    double h = tv[ind] + vv[ind];
                    h *= s->sym[n];
                    tv[ind] = h;
                }
    for (long long i = 0; i < N; i++) uec[i] += tv[i];
    free(tv);
    free(vv);
}

/* ------------------------------------------------------------------------ */
/* Lanczos recurrence                                                        */
/* ------------------------------------------------------------------------ */

/* Counter-based U[-1,1) generator (SplitMix64 finaliser on seed and index);
 * the reference hard-codes the uniform vector instead
 * (lanczosEigenValuesVectors.cpp:177-181). */
static inline double start_value(uint64_t seed, uint64_t idx)
{
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + (idx + 1) * 0xBF58476D1CE4E5B9ull + 0x94D049BB133111EBull;
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    return 2.0 * u - 1.0;
}

void bso_start_vector(long long n, uint64_t seed, double *v)
{
    double s = 0.0;
    for (long long i = 0; i < n; i++) {
        v[i] = start_value(seed, (uint64_t)i);
        s += v[i] * v[i];
    }
    double nrm = sqrt(s);
    for (long long i = 0; i < n; i++) v[i] = v[i] / nrm;
}

/* Loop 1, lanczosEigenValuesVectors.cpp:170-242 */
void bso_lanczos(const bso_state *s, int niter, const double *v0, double *alpha, double *beta, double *beta2)
{
    long long N = s->N;
    double *rvec = dalloc((size_t)N), *vec = dalloc((size_t)N), *uec = dalloc((size_t)N);
    for (long long ib = 0; ib < N; ib++) {
        uec[ib] = 0.;
        vec[ib] = v0 ? v0[ib] : 1. / sqrt((double)N);
        rvec[ib] = 0.;
    }
    beta[0] = 0.;
    beta2[0] = 0.;
    for (int j = 1; j <= niter; j++) {
        for (long long ib = 0; ib < N; ib++) uec[ib] = 0.;
        bso_hv(s, vec, uec);
        for (long long ib = 0; ib < N; ib++) rvec[ib] += uec[ib];
        double a = 0.;
        for (long long ib = 0; ib < N; ib++) a += vec[ib] * rvec[ib];
        alpha[j - 1] = a;
        for (long long ib = 0; ib < N; ib++) rvec[ib] -= (a * vec[ib]);
        double b2 = 0.;
        for (long long ib = 0; ib < N; ib++) b2 += rvec[ib] * rvec[ib];
        beta2[j] = b2;
        beta[j] = sqrt(b2);
        for (long long ib = 0; ib < N; ib++) rvec[ib] = (1. / beta[j]) * rvec[ib];
        for (long long ib = 0; ib < N; ib++) vec[ib] = (-beta[j]) * vec[ib];
        for (long long ib = 0; ib < N; ib++) {
            double t = vec[ib];
            vec[ib] = rvec[ib];
            rvec[ib] = t;
        }
    }
    free(rvec);
    free(vec);
    free(uec);
}

/* ------------------------------------------------------------------------ */
/* Cullum-Willoughby tridiagonal routines (1-based indexing as in F77)       */
/* ------------------------------------------------------------------------ */

#define CW_EPS (2.0 * 1.1102230246251565e-16) /* 2 * dlamch('e') */

/* GENRAN genran.f:4-32 with GNU SRAND/RAND: Park-Miller minimal standard
 * (a = 16807, m = 2^31-1); SRAND(0) selects 123459876.  RAND returns the
 * REAL*4 value ((s - 1) with its low 9 bits cleared) / 2^31, which is exact.
 * Pinned bitwise against the reference's GENRAN compiled with the GNU
 * runtime (oracle/_ref, tests/test_oracle_pins.py). */
void bso_genran(int seed, double *r, int n)
{
    uint64_t st = seed ? (uint64_t)seed : 123459876ull;
    for (int i = 0; i < n; i++) {
        st = (16807ull * st) % 2147483647ull;
        r[i] = (double)((st - 1) & ~(uint64_t)511) * (1.0 / 2147483648.0);
    }
}

/* BISEC bisec.f:4-225.  ALPHA(i)=al[i-1], BETA(i)=be[i-1] (be[0] = 0),
 * BETA2 = b2 (overwritten with BETA**2 as in :91-92).  VB workspace 2*nmax. */
static int cw_bisec(int mev, const double *al, double *be, double *b2, double *vs, int *mp, double *vb,
                    int nmax, double bb, double ub, int *err)
{
#define ALPHA(i) al[(i) - 1]
#define BETA(i) be[(i) - 1]
#define BETA2(i) b2[(i) - 1]
#define VS(i) vs[(i) - 1]
#define MP(i) mp[(i) - 1]
#define VB(i) vb[(i) - 1]
    (void)nmax;
    double eps = CW_EPS;
    double tkmax = 0.0;
    for (int i = 1; i <= mev; i++) tkmax = fmax(fmax(fabs(ALPHA(i)), BETA(i)), tkmax);
    double ttol = eps * tkmax;
    int mxstur = 100 * mev;
    int ndis = 0;
    double betam = BETA(mev + 1);
    BETA(mev + 1) = 0.0;
    for (int i = 1; i <= mev + 1; i++) BETA2(i) = BETA(i) * BETA(i);
    double temp = (double)(mev + 1000);
    double ep0 = temp * ttol;
    double ep1 = sqrt(temp) * ttol;
    int na = 0, md = 0, ng = 0, nev = 0, mt = 0, k = 0, nu = 0;
    int kev = 0, jev = 0, kl = 0, jl = 0, ju = 0, ku = 0, mpev = 0, knew = 0;
    double x1 = ub, xl = 0, xu = 0;
    int isturm = 1, jsturm = 0;
    *err = 0;
    goto L330;
L50:
    na = nev;
    x1 = bb;
    isturm = 2;
    goto L330;
L60:
    mt = nev;
    if (30 * mt >= mxstur) { *err = 1; goto L410; }
    if (mt == 0) goto L410;
    for (int i = 1; i <= mt; i++) {
        VB(i) = bb;
        VB(mt + i) = ub;
    }
    k = mt;
L150:
    xl = VB(k);
    xu = VB(mt + k);
    isturm = 3;
    x1 = xu;
    goto L330;
L160:
    nu = nev;
    isturm = 4;
L170:
    x1 = (xl + xu) * 0.5;
    {
        double xs = fabs(xl) + fabs(xu);
        double x0 = xu - xl;
        double ept = eps * xs + ep1;
        if (x0 <= ept) goto L230;
    }
    goto L330;
L180:
    if (nev < k) goto L190;
    xl = x1;
    goto L170;
L190:
    xu = x1;
    nu = nev;
    for (int i = 1; i <= nev; i++) VB(i) = fmax(x1, VB(i));
    for (int ii = nev + 1; ii <= k; ii++) VB(mt + ii) = fmin(x1, VB(mt + ii));
    goto L170;
L230:
    ndis++;
    md++;
    VS(ndis) = x1;
    jsturm = 1;
    x1 = xl - ep0;
    goto L370;
L240:
    kl = kev;
    jl = jev;
    jsturm = 2;
    x1 = xu + ep0;
    goto L370;
L250:
    ju = jev;
    ku = kev;
    if (kl - ku - 1 == 0) goto L290;
    if (ku == nu) goto L280;
L260:
    isturm = 5;
    x1 = x1 + ep0;
    goto L330;
L270: {
    int kne = ku - nev;
    ku = nev;
    if (kne != 0) goto L260;
}
L280:
    mpev = kl - ku;
    knew = ku;
    goto L300;
L290:
    mpev = 1;
    if (ju < jl) mpev = 0;
    knew = k - 1;
L300:
    k = knew;
    MP(ndis) = mpev;
    if (mpev >= 1) ng++;
    if (k <= 0) goto L410;
    for (int i = 1; i <= knew; i++) VB(i) = fmax(x1, VB(i));
    goto L150;
L330: {
    /* forward Sturm count :187-198 */
    nev = -na;
    double yu = 1.0;
    for (int i = 1; i <= mev; i++) {
        double yv = (yu == 0.0) ? BETA(i) / eps : BETA2(i) / yu;
        yu = x1 - ALPHA(i) - yv;
        if (yu < 0.0) nev++;
    }
    switch (isturm) {
    case 1: goto L50;
    case 2: goto L60;
    case 3: goto L160;
    case 4: goto L180;
    default: goto L270;
    }
}
L370: {
    /* backward Sturm count on T-hat :199-215 */
    kev = -na;
    double yu = 1.0;
    for (int i = mev; i >= 1; i--) {
        double yv = (yu == 0.0) ? BETA(i + 1) / eps : BETA2(i + 1) / yu;
        yu = x1 - ALPHA(i) - yv;
        jev = 0;
        if (yu < 0.0) {
            kev++;
            jev = 1;
        }
    }
    jev = kev - jev;
    if (jsturm == 1) goto L240;
    goto L250;
}
L410:
    BETA(mev + 1) = betam;
    (void)md;
    (void)ng;
    return ndis;
#undef ALPHA
#undef BETA
#undef BETA2
#undef VS
#undef MP
#undef VB
}

/* ISOEV isoev.f:4-55.  A single computed value has no neighbour; the F77
 * routine then reads outside VS/MP, here it is left isolated. */
static void cw_isoev(double *vs, int *mp, int ndis)
{
    const double gaptol = 1.e-8;
    if (ndis < 2) return;
    for (int j = 1; j <= ndis; j++) {
        if (mp[j - 1] != 1) continue;
        double t0 = 0, t1 = 0, g;
        if (j != ndis) t1 = vs[j] - vs[j - 1];
        if (j != 1) t0 = vs[j - 1] - vs[j - 2];
        if (j != 1 && j != ndis) {
            g = (t0 < t1) ? -t0 : t1;
        } else {
            g = 0;
            if (j == 1) g = vs[1] - vs[0];
            if (j == ndis) g = vs[ndis - 2] - vs[ndis - 1];
        }
        int i = (g >= 0.0) ? j + 1 : j - 1;
        if (mp[i - 1] == 0) {
            double gap = fabs(g);
            double temp = fmax(gaptol, gaptol * fabs(vs[j - 1]));
            if (gap <= temp) mp[j - 1] = -mp[j - 1];
        }
    }
}

/* inverse iteration sweep shared by INVERR inverr.f:173-215 and INVERM
 * inverm.f:168-210; BETA(i) = be[i-1], v1/v2 1-based.  Returns IT. */
static int cw_inverse_iteration(int mev, const double *al, double *be, double *v1, double *v2, double x1,
                                double eps3, double eps4, int iter)
{
#define ALPHA(i) al[(i) - 1]
#define BETA(i) be[(i) - 1]
#define V1(i) v1[(i) - 1]
#define V2(i) v2[(i) - 1]
    int it = 1;
    int mm1 = mev - 1;
    for (;;) {
        double u = ALPHA(1) - x1;
        double z = BETA(2);
        for (int i = 2; i <= mev; i++) {
            if (BETA(i) > fabs(u)) {
                double ratio = u / BETA(i);
                BETA(i) = -BETA(i);
                V1(i - 1) = ALPHA(i) - x1;
                u = z - ratio * V1(i - 1);
                z = -ratio * BETA(i + 1);
                double temp = V2(i - 1);
                V2(i - 1) = V2(i);
                V2(i) = temp - ratio * V2(i);
            } else {
                V1(i - 1) = z / u;
                V2(i - 1) = V2(i - 1) / u;
                V2(i) = V2(i) - BETA(i) * V2(i - 1);
                double ratio = BETA(i) / u;
                u = ALPHA(i) - x1 - z * ratio;
                z = BETA(i + 1);
            }
        }
        if (u == 0.0) u = eps3;
        V2(mev) = V2(mev) / u;
        for (int ii = 1; ii <= mm1; ii++) {
            int i = mev - ii;
            if (BETA(i + 1) < 0.0) {
                BETA(i + 1) = -BETA(i + 1);
                /* BETA(MEV+1) is zero here, so V2(MEV+1) (outside the
                 * order-MEV vector) only ever enters multiplied by zero */
                double far = (i + 2 <= mev) ? BETA(i + 2) * V2(i + 2) : 0.0;
                V2(i) = (V2(i) - V1(i) * V2(i + 1) - far) / BETA(i + 1);
            } else {
                V2(i) = V2(i) - V1(i) * V2(i + 1);
            }
        }
        double xnorm = fabs(V2(mev));
        for (int ii = 1; ii <= mm1; ii++) {
            int i = mev - ii;
            xnorm = xnorm + fabs(V2(i));
        }
        if (xnorm >= 1.0) break;
        it = it + 1;
        if (it > iter) break;
        double xu = eps4 / xnorm;
        for (int i = 1; i <= mev; i++) V2(i) = V2(i) * xu;
    }
    /* SCALAR scalar.f:4-21 and normalisation */
    double sum = 0.0;
    for (int i = 1; i <= mev; i++) sum = sum + V2(i) * V2(i);
    sum = 1.0 / sqrt(sum);
    for (int ii = 1; ii <= mev; ii++) V2(ii) = sum * V2(ii);
    return it;
#undef ALPHA
#undef BETA
#undef V1
#undef V2
}

/* INVERR inverr.f:4-167 */
static void cw_inverr(int mev, const double *al, double *be, double *v1, double *v2, double *g, const double *vs,
                      const int *mp, double *err, int ndis)
{
    const int iter = 10;
    for (int i = 0; i < ndis; i++) err[i] = 0.0;
    double eps = CW_EPS;
    double betam = be[mev];
    be[mev] = 0.0;
    double tsum = fabs(al[0]);
    for (int i = 2; i <= mev; i++) tsum = tsum + fabs(al[i - 1]) + be[i - 1];
    double eps3 = eps * tsum;
    double eps4 = (double)mev * eps3;
    bso_genran(1, g, mev);
    double gsum = 0.0;
    for (int i = 0; i < mev; i++) gsum = gsum + fabs(g[i]);
    gsum = eps4 / gsum;
    for (int i = 0; i < mev; i++) g[i] = gsum * g[i];
    for (int jev = 0; jev < ndis; jev++) {
        if (mp[jev] != 1) continue;
        for (int i = 0; i < mev; i++) v2[i] = g[i];
        int it = cw_inverse_iteration(mev, al, be, v1, v2, vs[jev], eps3, eps4, iter);
        double est = betam * fabs(v2[mev - 1]);
        if (it > iter) est = -est;
        err[jev] = est;
    }
    be[mev] = betam;
}

/* lancbis lancbis.cpp:5-105 */
int bso_lancbis(int niter, const double *alpha, const double *beta, const double *beta2, double emin, double emax,
                double *eval, double *err)
{
    double *lal = dalloc(niter), *lbe = dalloc(niter + 1), *lb2 = dalloc(niter + 1);
    for (int j = 0; j < niter; j++) lal[j] = alpha[j];
    lbe[0] = 0.0;
    lb2[0] = 0.0;
    for (int j = 1; j <= niter; j++) {
        lbe[j] = beta[j];
        lb2[j] = beta2[j];
    }
    double *vs = dalloc(niter), *wrk2 = dalloc(2 * (size_t)niter);
    int *mp = (int *)calloc(niter ? niter : 1, sizeof(int));
    int e = 0;
    int ndis = cw_bisec(niter, lal, lbe, lb2, vs, mp, wrk2, niter, emin, emax, &e);
    int ngood = 0;
    if (e) {
        ngood = -1;
    } else {
        cw_isoev(vs, mp, ndis);
        double *w1 = dalloc(niter + 1), *w2 = dalloc(niter + 2), *g = dalloc(niter);
        double *er = dalloc(ndis);
        if (ndis > 0) cw_inverr(niter, lal, lbe, w1, w2, g, vs, mp, er, ndis);
        for (int i = 0; i < ndis; i++)
            if (mp[i] != 0) {
                eval[ngood] = vs[i];
                err[ngood] = er[i];
                ngood++;
            }
        free(w1);
        free(w2);
        free(g);
        free(er);
    }
    free(lal); free(lbe); free(lb2); free(vs); free(wrk2); free(mp);
    return ngood;
}

int bso_bisec(int niter, const double *alpha, const double *beta, double emin, double emax, double *vs, int *mp)
{
    double *lal = dalloc(niter), *lbe = dalloc(niter + 1), *lb2 = dalloc(niter + 1), *wrk2 = dalloc(2 * (size_t)niter);
    for (int j = 0; j < niter; j++) lal[j] = alpha[j];
    for (int j = 1; j <= niter; j++) lbe[j] = beta[j];
    int e = 0;
    int ndis = cw_bisec(niter, lal, lbe, lb2, vs, mp, wrk2, niter, emin, emax, &e);
    free(lal); free(lbe); free(lb2); free(wrk2);
    return e ? -1 : ndis;
}

/* STURMI sturmi.f:4-116; returns 0 or -1 ("eigenvalue not found") */
static int cw_sturmi(const double *al, double *be, const double *b2, int mmax, double x1, int *mk1, int *mk2)
{
    double epsm = CW_EPS;
    double tkmax = 0.0;
    for (int i = 1; i <= mmax; i++) tkmax = fmax(fmax(fabs(al[i - 1]), be[i - 1]), tkmax);
    double ttol = epsm * tkmax;
    double xmltol = (double)(mmax + 1000) * ttol;
    double stutol = 5.0 * xmltol;
    double toln = fmax(fabs(x1) * 1e-10, stutol);
    *mk1 = 0;
    *mk2 = 0;
    be[0] = 0.0;
    double evl = x1 - toln, evu = x1 + toln;
    double u1 = 1.0, u2 = 1.0;
    int ic0 = 0, ic1 = 0, ic2 = 0, icd = 0;
    int i;
    int broke = 0;
    for (i = 1; i <= mmax; i++) {
        double v1 = (u1 != 0.0) ? b2[i - 1] / u1 : be[i - 1] / epsm;
        u1 = evl - al[i - 1] - v1;
        if (u1 < 0.0) ic1++;
        double v2 = (u2 != 0.0) ? b2[i - 1] / u2 : be[i - 1] / epsm;
        u2 = evu - al[i - 1] - v2;
        if (u2 < 0.0) ic2++;
        icd = ic1 - ic2;
        int ic = icd - ic0;
        if (ic >= 1) {
            if (ic0 == 0) *mk1 = i;
            ic0++;
            if (ic0 > 1) {
                broke = 1;
                break;
            }
        }
    }
    if (!broke) {
        i = i - 1; /* DO loop index is MMAX+1 after completion */
        if (ic0 == 0) *mk1 = mmax;
    }
    *mk2 = i;
    if (icd == 0) return -1;
    return 0;
}

/* INVERM inverm.f:4-142 */
static void cw_inverm(const double *al, double *be, double *v1, double *v2, const double *g, int mev, double x1,
                      double *error, int *qconv)
{
    const int iter = 10;
    double eps = CW_EPS;
    double betam = be[mev];
    be[mev] = 0.0;
    double tsum = fabs(al[0]);
    for (int i = 2; i <= mev; i++) tsum = tsum + fabs(al[i - 1]) + be[i - 1];
    double eps3 = eps * tsum;
    double eps4 = (double)mev * eps3;
    double gsum = 0.0;
    for (int i = 0; i < mev; i++) gsum = gsum + fabs(g[i]);
    gsum = eps4 / gsum;
    for (int i = 0; i < mev; i++) v2[i] = gsum * g[i];
    int it = cw_inverse_iteration(mev, al, be, v1, v2, x1, eps3, eps4, iter);
    *error = fabs(v2[mev - 1]);
    *qconv = it > iter ? 0 : 1;
    be[mev] = betam;
}

/* TRIVEC trivec.f:4-95 called by lanczosvectors lanczosEigenValuesVectors.cpp:509-561.
 * The eigenvector column keeps entries beyond the chosen order from earlier
 * trial orders, as in the reference (INVERM only writes MEV entries). */
int bso_trivec(int niter, const double *alpha, const double *beta, const double *beta2, const double *eval,
               int ngood, double *evtr, int *mamin)
{
    double *lal = dalloc(niter), *lbe = dalloc(niter + 2), *lb2 = dalloc(niter + 1);
    for (int j = 0; j < niter; j++) lal[j] = alpha[j];
    lbe[0] = 0.;
    lb2[0] = 0.;
    for (int j = 1; j <= niter; j++) {
        lbe[j] = beta[j];
        lb2[j] = beta2[j];
    }
    double *wrk1 = dalloc(niter + 1), *wrk2 = dalloc(niter);
    for (size_t i = 0; i < (size_t)niter * ngood; i++) evtr[i] = 0.0;
    for (int i = 0; i < ngood; i++) mamin[i] = 0;
    bso_genran(1, wrk2, niter);
    int status = 0;
    for (int ig = 0; ig < ngood; ig++) {
        int mk1, mk2;
        if (cw_sturmi(lal, lbe, lb2, niter, eval[ig], &mk1, &mk2) != 0) {
            status = -1;
            break;
        }
        int ma, ml, idelta;
        if (mk2 == niter) {
            ma = (5 * mk1) / 4 + 1;
            if (ma > niter) ma = niter;
            ml = niter;
            int t = (13 * mk1) / 8 + 1;
            if (t > niter) t = niter;
            idelta = (t - ma) / 10 + 1;
        } else {
            ma = (3 * mk1 + mk2) / 4 + 1;
            ml = (mk1 + 3 * mk2) / 4;
            idelta = ((3 * mk1 + 5 * mk2) / 8 + 1 - ma) / 10 + 1;
        }
        int icount = 0, qfind = 0;
        double errmin = 1.0;
        double *col = evtr + (size_t)ig * niter;
        for (;;) {
            icount++;
            double error;
            int qconv;
            cw_inverm(lal, lbe, wrk1, col, wrk2, ma, eval[ig], &error, &qconv);
            qfind = qfind || qconv;
            if (error <= errmin) {
                errmin = error;
                mamin[ig] = ma;
            }
            ma = ma + idelta;
            if (!((ma < ml) && (icount <= 10))) break;
        }
        if (!qfind) {
            status = -2;
            break;
        }
        double error;
        int qconv;
        cw_inverm(lal, lbe, wrk1, col, wrk2, mamin[ig], eval[ig], &error, &qconv);
    }
    free(lal); free(lbe); free(lb2); free(wrk1); free(wrk2);
    return status;
}

/* Loop 2, lanczosEigenValuesVectors.cpp:371-469: the recurrence is re-run
 * from the same start vector and ARvL[row + N n] += evtr(j-1,n) v_j[row]
 * while the cumulative norm stays below 1 - 1e-16. */
void bso_ritz(const bso_state *s, int niter, const double *v0, const double *evtr, int ngood, double *ritz)
{
    long long N = s->N;
    double *rvec = dalloc((size_t)N), *vec = dalloc((size_t)N), *uec = dalloc((size_t)N);
    double *cum = dalloc(ngood);
    for (long long ib = 0; ib < N; ib++) {
        vec[ib] = v0 ? v0[ib] : 1. / sqrt((double)N);
        rvec[ib] = 0.;
    }
    for (size_t i = 0; i < (size_t)N * ngood; i++) ritz[i] = 0.0;
    for (int j = 1; j <= niter; j++) {
        for (int n = 0; n < ngood; n++) {
            double c = evtr[(j - 1) + (size_t)niter * n];
            cum[n] += pow(c, 2.);
            if (cum[n] < (1. - 1.e-16))
                for (long long row = 0; row < N; row++) ritz[row + N * n] += c * vec[row];
        }
        for (long long ib = 0; ib < N; ib++) uec[ib] = 0.;
        bso_hv(s, vec, uec);
        for (long long ib = 0; ib < N; ib++) rvec[ib] += uec[ib];
        double a = 0.;
        for (long long ib = 0; ib < N; ib++) a += vec[ib] * rvec[ib];
        for (long long ib = 0; ib < N; ib++) rvec[ib] -= (a * vec[ib]);
        double b2 = 0.;
        for (long long ib = 0; ib < N; ib++) b2 += rvec[ib] * rvec[ib];
        double b = sqrt(b2);
        for (long long ib = 0; ib < N; ib++) rvec[ib] = (1. / b) * rvec[ib];
        for (long long ib = 0; ib < N; ib++) vec[ib] = (-b) * vec[ib];
        for (long long ib = 0; ib < N; ib++) {
            double t = vec[ib];
            vec[ib] = rvec[ib];
            rvec[ib] = t;
        }
    }
    free(rvec);
    free(vec);
    free(uec);
    free(cum);
}

/* Hellmann-Feynman gradient: sum_k g_k <psi_k| dV/dtheta |psi_k>, with
 * <psi|dV|psi> = sum_{P,q} w_q (sum_n Y_n(q) psi(P,n))^2 dV(P,q)
 * (the contraction of build_vblocks) and V(P,q) the pair sums of
 * build_vgrid: V = sum_s K q_cm q_s/|c - s| + LJ(|c - s|) on 'O' sites
 *              + sum_{h=+-} sum_s K q_h q_s/|c + h off(q) - s|,
 * clamped at the ceiling (no dependence there). */
void bso_hf_gradient(const bso_state *s, int nvec, const double *psi, const double *g, double *grad)
{
    const bso_problem *p = &s->p;
    int nx = s->nx, ny = s->ny, nz = s->nz, nn = s->nn, na = s->na, nb = s->nb, lmax = s->lmax;
    int ns = p->nsites, nent = 4 * ns + 4;
    long long P = s->P;
    for (int e = 0; e < nent; e++) grad[e] = 0.0;
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    double *acc = dalloc((size_t)nthr * nent);
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *my = acc + (size_t)tid * nent;
#pragma omp for schedule(guided)
        for (long long ip = 0; ip < P; ip++) {
            int k = (int)(ip % nz), j = (int)((ip / nz) % ny), i = (int)(ip / ((long long)nz * ny));
            double c[3] = {s->gx[i], s->gy[j], s->gz[k]};
            long long pv = ((long long)k * ny + j) * nx + i; /* vector layout [n][k][j][i] */
            double wcm = 0.0;
            for (int a = 0; a < na; a++)
                for (int b = 0; b < nb; b++)
                    for (int half = 0; half < 2; half++) {
                        size_t ind2 = ((size_t)ip * na + a) * nb + b;
                        double v = half == 0 ? s->v1[ind2] : s->v2[ind2];
                        if (v >= p->ceiling) continue;
                        const double *S = half == 0 ? s->S1 : s->S2;
                        double w = 0.0;
                        for (int kv = 0; kv < nvec; kv++) {
                            double y = 0.0;
                            for (int n = 0; n < nn; n++)
                                y += s->L[n * na + a] * S[(s->qnum[2 * n + 1] + lmax) * nb + b] *
                                     psi[(size_t)kv * s->N + (size_t)n * P + pv];
                            w += g[kv] * y * y;
                        }
                        w *= s->wt[a] * s->wp[b];
                        wcm += w;
                        double o[3] = {p->h2_half_bond * sin(s->th[a]) * s->cp[b],
                                       p->h2_half_bond * sin(s->th[a]) * sin(s->ph1[b]),
                                       p->h2_half_bond * s->ct[a]};
                        if (half == 1) o[1] = -o[1];
                        for (int hh = 0; hh < 2; hh++) {
                            double sg = hh == 0 ? 1.0 : -1.0;
                            for (int st = 0; st < ns; st++) {
                                const double *sp = p->site_pos + 3 * st;
                                double e[3] = {c[0] + sg * o[0] - sp[0], c[1] + sg * o[1] - sp[1],
                                               c[2] + sg * o[2] - sp[2]};
                                double r = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
                                my[4 * st] += w * BSO_K_CONST * p->q_h2_h / r;
                                my[4 * ns] += w * BSO_K_CONST * p->site_q[st] / r;
                                double f = w * BSO_K_CONST * p->q_h2_h * p->site_q[st] / (r * r * r);
                                for (int d = 0; d < 3; d++) my[4 * st + 1 + d] += f * e[d];
                            }
                        }
                    }
            for (int st = 0; st < ns; st++) {
                const double *sp = p->site_pos + 3 * st;
                double e[3] = {c[0] - sp[0], c[1] - sp[1], c[2] - sp[2]};
                double r = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
                my[4 * st] += wcm * BSO_K_CONST * p->q_h2_cm / r;
                my[4 * ns + 1] += wcm * BSO_K_CONST * p->site_q[st] / r;
                double f = wcm * BSO_K_CONST * p->q_h2_cm * p->site_q[st] / (r * r * r);
                if (p->site_type[st] == 'O') {
                    my[4 * ns + 2] += wcm / pow(r, 12.0);
                    my[4 * ns + 3] -= wcm / pow(r, 6.0);
                    f += wcm * (12.0 * p->lj_a / pow(r, 14.0) - 6.0 * p->lj_b / pow(r, 8.0));
                }
                for (int d = 0; d < 3; d++) my[4 * st + 1 + d] += f * e[d];
            }
        }
    }
    for (int t = 0; t < nthr; t++)
        for (int e = 0; e < nent; e++) grad[e] += acc[(size_t)t * nent + e];
    free(acc);
}
