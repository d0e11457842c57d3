/*
 * bsl_hip.h -- C ABI of libbsl_hip.so, the MI355X (gfx950) implementation of
 * the Hv matvec and Lanczos recurrence of jtcantin/boundstate-lanczos.
 *
 * Plain C types only.  Every entry point returns 0 on success and a non-zero
 * BSL_E_* code on failure; bsl_last_error() then describes the failure (per
 * thread).  The entry points that return a count or a value instead
 * (bsl_get_potential_form, bsl_get_fusion, bsl_gaunt_table) return minus the BSL_E_* code on
 * failure, so that no error reads as a valid result.  A context is not
 * thread-safe: calls on one context must be serialised, as the reference
 * driver serialises its Hv calls.  Several contexts may coexist.
 *
 * Vector layout at this boundary is the reference's composite index
 * [n][k][j][i] (i fastest; n = lm index in the reference order, m outer from
 * -lmax..lmax, l inner from |m|..lmax), length nx*ny*nz*(lmax+1)^2, exactly
 * what the reference's Hv function pointer receives.
 */
#ifndef BSL_HIP_H
#define BSL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    BSL_OK = 0,
    BSL_E_ARG = 1,      /* invalid argument (bad size, flag or pointer) */
    BSL_E_HIP = 2,      /* HIP runtime error */
    BSL_E_NOMEM = 3,    /* device or host allocation failed */
    BSL_E_STURM = 4,    /* BISEC: estimated number of Sturm sequences exceeds its limit */
    BSL_E_NOTFOUND = 5, /* STURMI: eigenvalue not found */
    BSL_E_NOCONV = 6,   /* TRIVEC: inverse iteration never converged */
    BSL_E_STATE = 7     /* call not valid for this context */
};

enum { BSL_SYM_ALL = 0, BSL_SYM_EVEN = 1, BSL_SYM_ODD = 2 };
/* FULL: all (lmax+1)^2 lm, output masked by the symmetry (reference
 * semantics, calcSym linRotCartSphHvFullVPrepRoutines.cpp:530-573).
 * REDUCED: only the lm of the symmetry class are carried (para / ortho
 * blocks); the spurious zero level of the masked space does not appear. */
enum { BSL_BASIS_FULL = 0, BSL_BASIS_REDUCED = 1 };
/* NOQUAD: precomputed per-point <lm|V|l'm'> blocks streamed by a batched
 * GEMV (linRotCartSph_NoQuad).  QUAD: Y_lm <-> (theta, phi) quadrature
 * transforms around the resident potential grid, evaluated on MFMA
 * (the on-the-fly linRotCartSph calculator, with its summation corrected). */
enum { BSL_CALC_NOQUAD = 0, BSL_CALC_QUAD = 1 };
/* How a NOQUAD context holds its per-point potential.  BLOCKS: the packed
 * <lm|V|l'm'> blocks themselves (B(B+1)/2 values per point and parity
 * block).  HARMONICS: the even-L real-harmonic coefficients
 * vh_LM(P) = sum_q w_q Y_LM(q) V(P, q) on the same quadrature (66 values per
 * point for lmax 5), from which the potential kernel rebuilds each block as
 * <lm|V|l'm'> = sum_LM G(lm, l'm', LM) vh_LM with the Gaunt coefficients G of
 * the real harmonics.  The two are the same matrices, exactly: the product
 * of two harmonics of l and l' is a finite sum of harmonics of L <= l + l',
 * at every node of any quadrature; they differ by rounding only.  HARMONICS
 * is the default wherever it applies (every parity block of l <= 5). */
enum { BSL_VFORM_BLOCKS = 0, BSL_VFORM_HARMONICS = 1 };
/* Hv arithmetic; the recurrence, its dot products and norms are always FP64.
 * FP32: kinetic contractions and quadrature transforms on FP32 MFMA, the
 * NoQuad potential's coefficients in FP32.  BF16X3: the same with the MFMA
 * products taken as three bfloat16 split products (hi*hi + hi*lo + lo*hi,
 * FP32 accumulation) on the K = 32 BF16 MFMA. */
enum { BSL_PREC_FP64 = 0, BSL_PREC_FP32 = 1, BSL_PREC_BF16X3 = 2 };
/* IHO: isotropic harmonic oscillator (IsoHarmOsc_Routines.cpp:6-79).
 * CAGE: Alavi H2 pair sums at the exact H positions of every quadrature
 *   node (Alavi_ContGrid_H2_Routines.cpp:6-291, the ContGrid calculators).
 * CAGE_GRID: the same sites, with the centre and H-atom potentials tabulated
 *   on a separate "potential universe" grid and looked up at the nearest
 *   universe point (Alavi_H2_Routines.cpp:5-317, the grid-mapped calculators).
 * COULOMB: a point charge at the origin acting on the molecule's centre,
 *   |V| > |coul_ceil| -> coul_ceil (Coulomb_Routines.cpp:6-86). */
enum { BSL_POT_IHO = 0, BSL_POT_CAGE = 1, BSL_POT_CAGE_GRID = 2, BSL_POT_COULOMB = 3 };
enum { BSL_START_UNIFORM = 0, BSL_START_RANDOM = 1 };

typedef struct bsl_params {
    int32_t nx, ny, nz;          /* grid points per axis, 2..256 (MFMA kinetic kernels up to 64,
                                    FP64 line kernels above) */
    double x_max, y_max, z_max;  /* full box lengths (nm); points span [-x_max/2, x_max/2] */
    int32_t lmax;                /* 0..12 (harmonic form for parity blocks of l <= 5, packed
                                    blocks above; the quadrature calculator up to 7) */
    int32_t n_theta;             /* Gauss-Legendre points in cos(theta) */
    int32_t n_phi;               /* Gauss-Chebyshev points; 2*n_phi phi nodes on [0, 2 pi) */
    double mass;                 /* amu */
    double rot_const;            /* B, kJ/mol */
    double ceiling;              /* V >= ceiling is set to ceiling (kJ/mol) */
    int32_t symmetry;            /* BSL_SYM_* */
    int32_t basis;               /* BSL_BASIS_* */
    int32_t calculator;          /* BSL_CALC_* */
    int32_t precision;           /* BSL_PREC_* */
    int32_t potential;           /* BSL_POT_* */
    /* isotropic harmonic oscillator: V = iho_factor |r|^2, |V| > |iho_ceil| -> iho_ceil */
    double iho_factor, iho_ceil;
    /* cage: point charges (all sites) and the H2-centre LJ term (flagged sites) */
    int32_t n_sites;
    const double *site_pos;      /* [n_sites][3], nm */
    const double *site_q;        /* [n_sites], e */
    const int32_t *site_lj;      /* [n_sites], 1 = LJ centre (water O) */
    double q_h2_h, q_h2_cm;      /* Alavi H2 charges on each H and at the centre, e */
    double h2_half_bond;         /* nm */
    double lj_a, lj_b;           /* A/r^12 - B/r^6, kJ nm^12/mol and kJ nm^6/mol */
    /* CAGE_GRID: potential universe, points per axis and full lengths (nm);
     * must contain the system box (HvPrep_Internal_NoQuad :330-348) */
    int32_t pnx, pny, pnz;
    double px_max, py_max, pz_max;
    /* COULOMB: charges (e) at the origin and on the molecule, and |V| cap */
    double coul_q_centre, coul_q_mol, coul_ceil;
} bsl_params;

typedef struct bsl_ctx bsl_ctx;

/* Prepares a calculator (replaces HvPrep: HvPrep_Internal_NoQuad,
 * linRotCartSphHvFullVPrepRoutines.cpp:229-528, and calc_Vlmlpmp_NoQuad[_ContGrid],
 * linRotCartSphHvFullVPrepRoutines.cpp:7-227 / linRotCartSphHvContGridRoutines.cpp:7-152).
 * Builds the basis tables on the host, then the potential grid and the
 * per-point V blocks on the device, where they stay resident with the
 * Krylov vectors. */
int bsl_create(const bsl_params *p, bsl_ctx **out);
int bsl_destroy(bsl_ctx *ctx);
const char *bsl_last_error(void);

/* total_basis_size in the reference convention (lanczosStor,
 * interfaceContainers.h:6-10; set at linRotCartSphHvFullVPrepRoutines.cpp:500) */
int64_t bsl_basis_size(const bsl_ctx *ctx);
/* length of the device-resident Krylov vectors (differs from the above
 * only for BSL_BASIS_REDUCED) */
int64_t bsl_krylov_size(const bsl_ctx *ctx);

/* uec += H vec on host vectors in the reference layout.  Replaces the Hv
 * function pointer (Hv_linRotCartSph_NoQuad, linRotCartSphHvFullVPrepRoutines.cpp:807-822;
 * called at lanczosEigenValuesVectors.cpp:195 and :426). */
int bsl_hv(bsl_ctx *ctx, const double *vec, double *uec);
/* the same on device pointers (reference layout, length bsl_basis_size) */
int bsl_hv_device(bsl_ctx *ctx, const double *d_vec, double *d_uec);

/* Lanczos loop 1 with the vectors resident in HBM
 * (lanczosEigenValuesVectors.cpp:170-242).  Start vector: v0 (host, reference
 * layout, normalised by the caller) if not NULL, else BSL_START_UNIFORM
 * (1/sqrt(N), the reference's) or BSL_START_RANDOM (U[-1,1) from the
 * counter-based generator with `seed`, normalised).  alpha[niter],
 * beta[niter+1], beta2[niter+1]; index 0 of beta/beta2 is zero. */
int bsl_lanczos(bsl_ctx *ctx, int32_t niter, const double *v0, int32_t start, uint64_t seed,
                double *alpha, double *beta, double *beta2);

/* bsl_lanczos without any device allocation in the call.  No reference
 * counterpart: for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc), on a
 * context prepared beforehand by bsl_create.  Same arguments, same kernels
 * and the same coefficients, bitwise.  The recurrence scalars, the reduction
 * partials and the coefficient lists live in one static device array per
 * GPU (about 1.2 MiB, held until the process ends) instead of buffers the
 * library allocates; hence at most BSL_NOALLOC_MAX_STEPS steps and potential
 * kernels of at most BSL_NOALLOC_MAX_PARTIALS workgroups (BSL_E_ARG
 * otherwise).  All calls on one GPU share that array: calls that overlap on
 * one GPU race. */
#define BSL_NOALLOC_MAX_STEPS 32768
#define BSL_NOALLOC_MAX_PARTIALS 16384
int bsl_lanczos_noalloc(bsl_ctx *ctx, int32_t niter, const double *v0, int32_t start, uint64_t seed,
                        double *alpha, double *beta, double *beta2);

/* Cullum-Willoughby post-processing (lancbis, lancbis.cpp:5-105: BISEC
 * bisec.f:4-225, ISOEV isoev.f:4-55, INVERR inverr.f:4-167).  eval/err hold
 * niter entries; *ngood receives the number of good eigenvalues. */
int bsl_lancbis(int32_t niter, const double *alpha, const double *beta, const double *beta2, double emin,
                double emax, double *eval, double *err, int32_t *ngood);

/* The same as bsl_lancbis with the Sturm bisections and the inverse-iteration
 * error estimates on the GPU (current device), one eigenvalue per thread; the
 * cluster bookkeeping and ISOEV stay on the host.  Eigenvalues agree with
 * bsl_lancbis to BISEC's bisection tolerance.  No reference counterpart (the
 * reference's BISEC/INVERR are serial Fortran, bisec.f:4-225, inverr.f:4-167). */
int bsl_lancbis_device(int32_t niter, const double *alpha, const double *beta, const double *beta2, double emin,
                       double emax, double *eval, double *err, int32_t *ngood);

/* bsl_lancbis and bsl_lancbis_device with the half-width of BISEC's
 * multiplicity interval (EP0, bisec.f:66-95) raised to at least mult_tol
 * (energy units of alpha): eigenvalues of T_m closer than that are one
 * multiple, hence good, eigenvalue.  For the mixed-precision calculators,
 * whose operator rounding spreads the ghost copies of a converged level by
 * more than FP64's EP0, so that BISEC would find each copy in T-hat as well
 * and drop the level as spurious.  mult_tol = 0 is bsl_lancbis itself.  No
 * reference counterpart (the reference runs an FP64 operator only). */
/* The multiplicity tolerances the mixed-precision calculators use (kJ/mol).
 * FP32: its rounding spreads the ghost copies of a converged level by ~1e-8
 * relative; 1e-6 kJ/mol (8e-5 cm^-1) merges them, two orders below the
 * 0.01 cm^-1 accuracy of mixed levels.  BF16x3 (two-term bf16 splits, about
 * 16 significant bits): copies spread by ~1e-6 relative; 3e-5 kJ/mol
 * (2.5e-3 cm^-1), the level accuracy of that precision. */
#define BSL_MIXED_MULT_TOL 1.0e-6
#define BSL_BF16X3_MULT_TOL 3.0e-5
/* NoQuad BF16x3 (its kinetic contractions as BF16 split products): levels
 * shift by up to ~1.5e-2 cm^-1 at 54^3 and ghost copies spread by ~1e-4
 * kJ/mol; at this width TRIVEC converges for the lowest levels, whose FP64
 * Rayleigh refinement (bsl_ritz_refine) is then within 1e-4 cm^-1. */
#define BSL_BF16X3_KIN_MULT_TOL 1.0e-4
int bsl_lancbis_tol(int32_t niter, const double *alpha, const double *beta, const double *beta2, double emin,
                    double emax, double mult_tol, double *eval, double *err, int32_t *ngood);
int bsl_lancbis_device_tol(int32_t niter, const double *alpha, const double *beta, const double *beta2,
                           double emin, double emax, double mult_tol, double *eval, double *err, int32_t *ngood);

/* Tm eigenvectors (lanczosvectors, lanczosEigenValuesVectors.cpp:509-561 ->
 * TRIVEC trivec.f:4-95).  evtr column-major [niter][ngood]. */
int bsl_trivec(int32_t niter, const double *alpha, const double *beta, const double *beta2, const double *eval,
               int32_t ngood, double *evtr, int32_t *mamin);
/* bsl_trivec with STURMI's window (sturmi.f:4-116) widened to at least
 * mult_tol, for the values bsl_lancbis_tol returned with the same mult_tol:
 * the order search then starts where the first of the merged copies
 * converged.  mult_tol = 0 is bsl_trivec itself.  No reference counterpart. */
int bsl_trivec_tol(int32_t niter, const double *alpha, const double *beta, const double *beta2, const double *eval,
                   int32_t ngood, double mult_tol, double *evtr, int32_t *mamin);

/* Lanczos loop 2 (lanczosEigenValuesVectors.cpp:371-469): recomputes the
 * Krylov vectors from the same start and accumulates the Ritz vectors on
 * the device; ritz is host memory [ngood][bsl_basis_size] (reference layout).
 * The device accumulators are kept by the context and reused (they grow to
 * the largest ngood asked for, until bsl_destroy). */
int bsl_ritz(bsl_ctx *ctx, int32_t niter, const double *v0, int32_t start, uint64_t seed, const double *evtr,
             int32_t ngood, double *ritz);

/* FP64 Rayleigh quotients <y|H|y>/<y|y> of host vectors (reference layout,
 * [count][bsl_basis_size]) with an FP64 Hv, whatever the context precision;
 * also returns the residual norms |Hy - e y|/|y|.  Hy, the dot products and
 * the residual are computed on the device (only y crosses PCIe).  No
 * reference counterpart (the reference reports the Lanczos values). */
int bsl_rayleigh(bsl_ctx *ctx, int32_t count, const double *y, double *energy, double *resid);

/* Loop 2 followed by the FP64 refinement of every Ritz vector it builds,
 * with the vectors kept on the device: bsl_ritz (same arguments) and then
 * bsl_rayleigh of each result, without the vectors crossing PCIe.  ritz may
 * be NULL; otherwise it receives what bsl_ritz returns.  Loop 2
 * (lanczosEigenValuesVectors.cpp:371-469) plus a refinement with no
 * reference counterpart. */
int bsl_ritz_refine(bsl_ctx *ctx, int32_t niter, const double *v0, int32_t start, uint64_t seed,
                    const double *evtr, int32_t ngood, double *ritz, double *energy, double *resid);

/* Water sites from centres of mass and ZY'Z'' Euler angles, as read from
 * the reference's .xyz geometry files (getSPCEatoms,
 * SPCE_AH2_EngRoutines.cpp:5-108; rotMat3D rotmat.cpp:6-32).  cm_ea holds
 * per molecule CMx phi CMy theta CMz chi (Angstrom, radian) in file order.
 * model 0 = SPC/E (O, H, H; 3 sites per molecule), 1 = TIP4P (O, H, H, M;
 * getTIP4Patoms TIP4P_AH2_EngRoutines.cpp:5-117, 4 sites, O uncharged).
 * Writes the sites' positions (nm), charges (e) and LJ flags (1 on O). */
int bsl_water_sites(int32_t model, int32_t nmol, const double *cm_ea, double *pos, double *q, int32_t *lj);

/* Diagnostics: V blocks in the reference order [i][j][k][n][n'] (FULL basis
 * contexts only; cross-parity couplings the device drops are returned as 0),
 * and the potential grid [i][j][k][a][b] for both phi halves. */
int bsl_get_vblocks(bsl_ctx *ctx, double *out);
int bsl_get_vgrid(bsl_ctx *ctx, double *v1, double *v2);

/* The BSL_VFORM_* a NOQUAD context's potential kernel reads, and a switch
 * between the two (the potential is rebuilt from a regenerated grid and the
 * kernel variant re-tuned; the Krylov vectors are cleared, so call it before
 * a run).  BLOCKS is the representation HvPrep_Internal_NoQuad builds
 * (linRotCartSphHvFullVPrepRoutines.cpp:229-528); HARMONICS has no
 * reference counterpart.  For the quadrature calculator these act on the
 * FP64 NoQuad operator the context keeps for its FP64 passes. */
int bsl_get_potential_form(const bsl_ctx *ctx); /* BSL_VFORM_*, or -BSL_E_ARG */
int bsl_set_potential_form(bsl_ctx *ctx, int32_t form);

/* Kernel structure of a NOQUAD step: 0 = x-y kinetic, z kinetic, potential
 * (three kernels), 1 = x-y kinetic, then one kernel applying the z kinetic
 * term and the potential (one read of X and one read-modify-write of R
 * where the separate kernels make two).  Both compute Tv_5D + Vv_5D
 * (linRotCartSphHvRoutines.cpp:1872-1969, linRotCartSphHvFullVPrepRoutines.cpp:675-734)
 * to rounding; the fused kernel exists for the lmax-5 harmonic-form blocks
 * (the full basis pair or one reduced-basis block) with every axis of at
 * most 64 points, in FP64 or FP32 (not BF16X3).  Contexts start with (and
 * bsl_set_potential_form returns to) the three-kernel step, which is the
 * faster one on MI355X; bsl_set_fusion(ctx, 1) selects the fused kernel and
 * fails with BSL_E_ARG where it does not exist.  No reference counterpart. */
int bsl_get_fusion(const bsl_ctx *ctx); /* 0 or 1, or -BSL_E_ARG */
int bsl_set_fusion(bsl_ctx *ctx, int32_t on);

/* The harmonic-form term list of parity block (lb, par): the planes are
 * every l of parity par up to lb in the reference lm order, the columns the
 * even-L real harmonics (L = 0, 2, .., 2 lb; M = -L .. L).  Term t adds
 * g[t] * vh[ijc[3t + 2]] to <i|V|j> with i = ijc[3t], j = ijc[3t + 1]
 * (i <= j).  Returns the number of terms, or minus a BSL_E_* code; writes
 * them only if cap is at least that (host only, no device).  No reference counterpart
 * (diagnostics of BSL_VFORM_HARMONICS). */
int bsl_gaunt_table(int32_t lb, int32_t par, int32_t cap, int32_t *ijc, double *g);

/* Timing of the device recurrence: runs `steps` Lanczos steps from the
 * random start after `warmup` untimed ones.  ms[0] = total time of the timed
 * steps (HIP events on the context stream, no host synchronisation inside);
 * ms[1], ms[2], ms[3] = average per-step duration of the potential kernel,
 * the two kinetic kernels and the step finalize (alpha, beta and the scalars
 * of the next step), from the kernels' own start and end timestamps
 * (hipExtLaunchKernel events: the first launch's start to the last launch's
 * end of each group) in a further min(steps, 64) steps. */
int bsl_time_steps(bsl_ctx *ctx, int32_t warmup, int32_t steps, double *ms);

/* Bytes each kernel of one Lanczos step must move at minimum (algorithmic
 * bytes): out[0] potential kernel (V + read X, read and write R), out[1]
 * kinetic kernels (seven vector passes), out[2] recurrence update (0: it is
 * folded into the kinetic and potential kernels), out[3] minimum of a step
 * (V + six vector passes); out[4] FLOPs of one Hv. */
int bsl_step_bytes(const bsl_ctx *ctx, double *out);

// Hellmann-Feynman gradients: grad = sum_k g[k] <psi_k| dV/dtheta |psi_k>
// for the cage potential (Alavi H2 in point-charge/LJ sites,
// Alavi_ContGrid_H2_Routines.cpp:6-147; clamped nodes contribute nothing).
// psi: nvec eigenvectors of length bsl_basis_size, reference layout,
// unit norm; g: nvec weights (e.g. the gradient of a loss with respect to
// each eigenvalue).  grad (4*n_sites + 4 doubles): d/d site_q[s] at [4*s],
// d/d site_pos[s][d] at [4*s + 1 + d]; then d/d q_h2_h, q_h2_cm, lj_a, lj_b
// at [4*n_sites + 0..3].  The first call on a context evaluates the
// potential grid once (in a temporary P * Q doubles) to find the clamped
// nodes and keeps until bsl_destroy one bit per (point, node) (P * Q / 8
// bytes, Q = 2 * n_theta * n_phi), the list of kept node pairs (4 bytes
// each; a node and its inversion partner, the same H2 molecule with the
// atoms exchanged, count once) and the node tables; later calls skip that
// pass.  The call's own buffers (the eigenvectors, one weight per kept node
// pair and set) stay with the context and are reused by later calls of the
// same or smaller size.  No reference counterpart (the reference has no
// derivatives).
int bsl_hf_gradient(bsl_ctx *ctx, int32_t nvec, const double *psi, const double *g, double *grad);
// The gradients of nset weighted sets at once (e.g. every degenerate
// multiplet of the lowest levels): set m is the counts[m] eigenvectors that
// follow the previous sets' in psi, with their weights in g; grad receives
// nset blocks of 4*n_sites + 4, each equal (bitwise) to bsl_hf_gradient of
// that set alone.  Up to eight sets share one pass over the (point, node)
// pairs, whose geometry (the distances to every site) is computed once per
// pass.  No reference counterpart.
int bsl_hf_gradient_sets(bsl_ctx *ctx, int32_t nset, const int32_t *counts, const double *psi, const double *g,
                         double *grad);
// The same for the Ritz vectors that the context's last bsl_ritz_refine
// left on the device (vector i < its ngood is the one of evtr[i]), so that
// they never cross PCIe: set m is the counts[m] vectors index[...] that
// follow the previous sets' in index, with their weights in g.  Each vector
// y is taken as the unit vector y / |y| (|y|^2 from the refinement; the
// weight g / |y|^2 carries it), so the result equals bsl_hf_gradient_sets
// of the downloaded, normalised vectors up to the rounding of that
// normalisation.  BSL_E_STATE when the context holds no such vectors (a
// later bsl_ritz or bsl_ritz_refine replaces them).  No reference
// counterpart.
int bsl_hf_gradient_ritz(bsl_ctx *ctx, int32_t nset, const int32_t *counts, const int32_t *index, const double *g,
                         double *grad);

// Counts of the (point, quadrature node) pairs of a cage context's potential
// grid that lie below the ceiling clamp (the pairs bsl_hf_gradient sums
// over) and of all pairs; builds the clamp mask if no gradient call did.  No
// reference counterpart (measurement support).
int bsl_clamp_stats(bsl_ctx *ctx, int64_t *unclamped, int64_t *total);

// ---- batched recurrences (no reference counterpart: the reference runs one
// recurrence per process, lanczosEigenValuesVectors.cpp:158-242) ----
// A batch owns n contexts (cages, para/ortho blocks, grid sizes may differ)
// whose Loop-1 recurrences advance in lockstep; with use_graph != 0 each pair
// of lockstep iterations is one replayed hipGraph whose member chains run as
// independent branches.  Member i starts from seed + i.
typedef struct bsl_batch bsl_batch;
int bsl_batch_create(int32_t n, const bsl_params *params /* [n] */, bsl_batch **out);
int bsl_batch_destroy(bsl_batch *b);
int bsl_batch_size(const bsl_batch *b);
// the context of member i (owned by the batch: Ritz vectors, Rayleigh, Hv)
int bsl_batch_member(bsl_batch *b, int32_t i, bsl_ctx **out);
// alpha [n][niter], beta and beta2 [n][niter + 1] (index 0 zero), as bsl_lanczos
int bsl_batch_lanczos(bsl_batch *b, int32_t niter, int32_t start, uint64_t seed, int32_t use_graph,
                      double *alpha, double *beta, double *beta2);
// ms[0]: wall time (ms, HIP events) of `steps` lockstep iterations after `warmup`
int bsl_batch_time_steps(bsl_batch *b, int32_t warmup, int32_t steps, int32_t use_graph, double *ms);

// Name (with template arguments, as a profiler prints it) of the potential
// kernel the context launches: "k_vgaunt_split<double, 4>" (harmonic form of
// the lmax-5 pair of blocks, split over four waves per 64 points),
// "k_vgaunt<4, 0, 3, 1, double, 2>" (other harmonic-form shapes: largest l and
// parity of each block) or "k_vblock<15, 21, double, 1>" (block form: block
// sizes); the last argument is the occupancy (waves per SIMD) of the variant
// picked by the autotuner at bsl_create.
// No reference counterpart (measurement support).
int bsl_potential_kernel(const bsl_ctx *ctx, char *name, int32_t len);

#ifdef __cplusplus
}
#endif
#endif
