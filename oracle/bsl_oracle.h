/*
 * bsl_oracle.h -- CPU oracle for the Hv matvec and Lanczos recurrence of
 * jtcantin/boundstate-lanczos.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64 restatement of the
 * reference algorithm, written from the reference sources (cited per
 * function in bsl_oracle.c).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or as the
 * timed CPU baseline.  The product library (libbsl_hip.so) never links,
 * loads or calls anything here.
 *
 * Parity pins (what fixes each part; see tests/test_oracle_*.py):
 *   basis / quadrature / harmonics  closed forms, numpy leggauss, scipy lpmv,
 *                                   quadrature orthonormality, oracle/_ref
 *   kinetic + rotor + mask          closed form, HO spectrum, oracle/_ref Hv
 *   V blocks (NoQuad, ContGrid)     constant / radial / cos(theta) potentials,
 *                                   symmetry, oracle/_ref V blocks + Hv
 *   cage potential (SPC/E, Alavi)   point energies of the compiled reference
 *                                   sources (oracle/_ref; the -30.2331 kJ/mol
 *                                   of a reference comment is stale: those
 *                                   sources give -58.37227147344424),
 *                                   closed-form pair sums
 *   Lanczos loop 1                  numpy Lanczos on the dense operator,
 *                                   reference run-dir eigenvalues
 *   BISEC / ISOEV                   brute force (eigvals of T and T-hat),
 *                                   reference run-dir eigenvalues, and the
 *                                   reference's own lancbis.cpp + Fortran
 *                                   built from source (oracle/_ref)
 *   Hellmann-Feynman gradient       central finite differences of eigenvalues
 *                                   of this oracle (tests/test_adjoint.py)
 *   INVERR / TRIVEC / INVERM /
 *   STURMI / GENRAN                 residual checks; run-dir error column and
 *                                   lv / eigVecFile fixtures; the reference's
 *                                   own Fortran built from source with GNU
 *                                   RAND and LAPACK DLAMCH (oracle/_ref)
 */
#ifndef BSL_ORACLE_H
#define BSL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- units, lanczosUnits.h:5-14 ---- */
#define BSO_PI 3.141592653589793
#define BSO_K_CONST 138.9354846111096
#define BSO_H_BAR 0.0635077857002
#define BSO_NM_PER_ANG 0.1

/* ---- basis ---- */
int  bso_gen_indices_lm(int lmax, int *qnum /* [nn*2]: l, m */);
void bso_gauleg(double x1, double x2, double *x, double *w, int n);
void bso_gauss_chebyshev(int n, double *x, double *w);
void bso_norm_assoc_legendre(int lmax, const int *qnum, int nn, double x, double *out);
void bso_tesseral_trig(const int *qnum, int nn, double phi, double *out);
void bso_cart_kin_grid(double x_max, int n, double mass, double *kin, double *grid);
void bso_rot_kin(const int *qnum, int nn, double B, double *out);

/* ---- sites: water molecules from COM + ZYZ Euler angles ---- */
/* model: 0 = SPC/E (O,H,H), 1 = TIP4P (O,H,H,M).  cm_ea: per molecule
 * CMx phi CMy theta CMz chi (Angstrom / radian, interleaved as in the .xyz
 * files).  Writes nsites*3 positions (nm), charges and type chars. */
int  bso_build_sites(int model, int nmol, const double *cm_ea,
                     double *pos, double *q, char *type);
double bso_coulomb_sum(const double *pos3, double q, int nsites,
                       const double *site_pos, const double *site_q);
double bso_lj_sum(const double *pos3, int nsites, const double *site_pos,
                  const char *site_type, double A, double B);

/* ---- problem description ---- */
typedef struct bso_problem {
    int nx, ny, nz;
    double x_max, y_max, z_max;   /* full box lengths, nm */
    int lmax;
    int na, nb;                   /* theta (GL) and phi (GC) points */
    double mass;                  /* amu */
    double rot_const;             /* B, kJ/mol */
    double ceiling;               /* V >= ceiling -> ceiling, kJ/mol */
    int sym;                      /* 0 All, 1 Even, 2 Odd */
    int pot_kind;                 /* 0 isotropic harmonic oscillator, 1 Alavi cage (exact H
                                     positions), 2 Alavi cage on a potential universe grid,
                                     3 Coulomb point charge at the origin */
    double iho_factor;            /* 0.5 m w^2 */
    double iho_ceil;              /* |V| > |ceil| -> ceil */
    int nsites;
    const double *site_pos;       /* [nsites*3] nm */
    const double *site_q;         /* [nsites] e */
    const char *site_type;        /* 'O' carries LJ, others charge only */
    double q_h2_h, q_h2_cm;       /* Alavi charges (e) */
    double h2_half_bond;          /* nm, H offset from the H2 centre */
    double lj_a, lj_b;            /* A/r^12 - B/r^6 on the O sites, kJ nm^n/mol */
    /* pot_kind 2 (grid-mapped cage): potential universe */
    int pnx, pny, pnz;
    double px_max, py_max, pz_max;
    /* pot_kind 3 (Coulomb): charges at the origin and on the molecule, |V| cap */
    double coul_q_centre, coul_q_mol, coul_ceil;
} bso_problem;

typedef struct bso_state bso_state;

bso_state *bso_prepare(const bso_problem *p);
bso_state *bso_prepare_external(const bso_problem *p, const double *vblk);
void       bso_free(bso_state *s);
long long  bso_basis_size(const bso_state *s);
int        bso_nn(const bso_state *s);
/* V grids in reference order [i][j][k][a][b] (i slowest), both phi halves */
void bso_get_vgrid(const bso_state *s, double *v1, double *v2);
/* V blocks in reference order [i][j][k][n][n'] */
void bso_get_vblocks(const bso_state *s, double *out);
void bso_get_kin(const bso_state *s, int axis, double *out);
void bso_get_grid(const bso_state *s, int axis, double *out);
void bso_get_rot(const bso_state *s, double *out);
void bso_get_sym(const bso_state *s, double *out);
void bso_get_qnum(const bso_state *s, int *out);
/* quadrature + harmonic tables: cos(theta), w_theta, cos(phi), w_phi,
 * L[n][a], S1[m][b], S2[m][b] (m shifted by lmax) */
void bso_get_quad(const bso_state *s, double *ct, double *wt, double *cp, double *wp);
void bso_get_tess(const bso_state *s, double *L, double *S1, double *S2);

/* uec += H vec ; vectors in reference layout [n][k][j][i] */
void bso_hv(const bso_state *s, const double *vec, double *uec);

/* Lanczos loop 1.  v0 == NULL -> uniform 1/sqrt(N).  alpha[niter],
 * beta[niter+1], beta2[niter+1] (index 0 unused, zero). */
void bso_lanczos(const bso_state *s, int niter, const double *v0,
                 double *alpha, double *beta, double *beta2);

/* counter-based start vector U[-1,1) from seed, normalised */
void bso_start_vector(long long n, uint64_t seed, double *v);

/* Cullum-Willoughby post-processing (lancbis.cpp).  eval/err sized niter. */
int  bso_lancbis(int niter, const double *alpha, const double *beta,
                 const double *beta2, double emin, double emax,
                 double *eval, double *err);
/* raw BISEC: vs[niter], mp[niter]; returns ndis */
int  bso_bisec(int niter, const double *alpha, const double *beta,
               double emin, double emax, double *vs, int *mp);
/* TRIVEC: evtr column-major [niter][ngood] (entry i + j*niter), mamin[ngood] */
int  bso_trivec(int niter, const double *alpha, const double *beta,
                const double *beta2, const double *eval, int ngood,
                double *evtr, int *mamin);
/* Loop 2: Ritz vectors ritz[n*N + row] */
void bso_ritz(const bso_state *s, int niter, const double *v0,
              const double *evtr, int ngood, double *ritz);

/* Hellmann-Feynman gradient of sum_k g[k] E_k for the cage potential:
 * grad[4*s] = d/d site_q[s], grad[4*s+1+d] = d/d site_pos[s][d],
 * grad[4*nsites + 0..3] = d/d q_h2_h, q_h2_cm, lj_a, lj_b.  psi: nvec unit
 * eigenvectors in reference layout.  Derivatives of the pair sums of
 * Alavi_ContGrid_H2_Routines.cpp:6-147 (no reference counterpart: the
 * reference has no derivatives); clamped nodes (V >= ceiling) contribute
 * nothing.  Needs the V grids (bso_prepare, not bso_prepare_external). */
void bso_hf_gradient(const bso_state *s, int nvec, const double *psi, const double *g, double *grad);
/* GNU Fortran style SRAND/RAND stream (Park-Miller minimal standard) */
void bso_genran(int seed, double *r, int n);

#ifdef __cplusplus
}
#endif
#endif
