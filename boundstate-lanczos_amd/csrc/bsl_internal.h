// bsl_internal.h -- shared declarations of libbsl_hip.so (host tables,
// device buffers, kernel launchers).  Not part of the public ABI.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "bsl_gaunt.h"
#include "bsl_hip.h"

namespace bsl {

// Scalars of the recurrence, kept in device memory so that the Lanczos loop
// never synchronises with the host.  At the start of a step
// the first vector slot X holds r~ = H v_j - beta_j v_{j-1} (alpha not yet
// removed) and the second slot R holds v_j; the x-y kinetic kernel forms
// v_{j+1} = ca X + cb R (= (r~ - alpha v_j) / beta) and R = cr R (= -beta v_j)
// on the fly, so no separate update pass is needed.
struct Scal {
    double ca, cb, cr;
    double alpha, beta, beta2;
    long long step;  // completed Lanczos steps: alpha/beta outputs are indexed by it
    // first step (1-based) whose beta^2 = |R|^2 - 2 alpha v.R + alpha^2 |v|^2
    // kept no correct digit (at most 64 eps |R|^2: the residual vanished or
    // cancelled), 0 if none; the host turns it into an error after the run
    long long breakdown;
};
// BSL_E_STATE with the step of a recorded cancellation, or BSL_OK (reads the
// scalars after synchronising `s`)
int check_breakdown(const Scal *sc, hipStream_t s);

// error reporting (thread-local message + code)
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what, const char *file, int line);
#define BSL_HIP(call)                                                              \
    do {                                                                           \
        hipError_t _e = (call);                                                    \
        if (_e != hipSuccess) return ::bsl::hip_fail(_e, #call, __FILE__, __LINE__); \
    } while (0)

// ---------------------------------------------------------------------------
// host tables (bsl_tables.cpp)
// ---------------------------------------------------------------------------
struct Tables {
    int lmax = 0, nn = 0, na = 0, nb = 0;
    std::vector<int> l, m;                 // per reference lm index n
    std::vector<double> ct, wt, th;        // GL: cos(theta), weights, theta
    std::vector<double> cp, wp, ph1, ph2;  // GC: cos(phi), weights, phi in (0,pi), 2pi - phi
    std::vector<double> L;                 // [n][a] normalised associated Legendre
    std::vector<double> S1, S2;            // [m+lmax][b] tesseral trig for both phi halves
};
void build_tables(int lmax, int na, int nb, Tables &t);
// first row of the (symmetric Toeplitz) sinc-DVR kinetic matrix and the grid
void kinetic_row(double x_max, int n, double mass, std::vector<double> &row, std::vector<double> &grid);
// harmonic-form angular table hh[q][c] = w_a w_b Yh_c(node q), q = (half, a, b)
// as in the V-block table, c over the ncoef(lb) even-L real harmonics
void harmonic_table(const Tables &t, int lb, std::vector<double> &hh);
// values of the harmonic term list of parity block (lb, par), in list order,
// from an exact quadrature; *excluded = largest |G| among the (i, j, c) the
// selection rules leave out (rounding level when the rules hold)
void gaunt_values(int lb, int par, std::vector<double> &g, double *excluded);

// ---------------------------------------------------------------------------
// Cullum-Willoughby post-processing (bsl_tridiag.cpp)
// ---------------------------------------------------------------------------
// device Sturm bisection and error estimates (bsl_tridiag_dev.hip)
// per-index output of k_tri_bisect: x1, xl, xu, nu, kl, jl, ku, ju, klm, kum
constexpr int kTriOut = 10;
int launch_tri_bisect(const double *al, const double *be, int m, double lo, double hi, int above_hi, int mt,
                      double ep0, double ep1, double epm, double *out, hipStream_t s);
int launch_tri_inverr(const double *al, const double *be, int m, const double *xs, int nx, const double *g,
                      double eps3, double eps4, double beta_m1, int maxit, double *sv, double *sm, unsigned char *sf,
                      double *err, hipStream_t s);
// mtol: lower bound of BISEC's multiplicity interval half-width (0 = BISEC)
int cw_lancbis_device(int niter, const double *alpha, const double *beta, double emin, double emax, double mtol,
                      hipStream_t stream, double *eval, double *err, int *ngood);
int cw_lancbis(int niter, const double *alpha, const double *beta, const double *beta2, double emin, double emax,
               double mtol, double *eval, double *err, int *ngood);
int cw_trivec(int niter, const double *alpha, const double *beta, const double *beta2, const double *eval,
              int ngood, double mtol, double *evtr, int *mamin);

// ---------------------------------------------------------------------------
// device-side launchers (bsl_kernels.hip)
// ---------------------------------------------------------------------------
// Storage order of the packed upper triangle of a B x B potential block:
// row r then row B-1-r, for r = 0 .. (B-1)/2 (rows as columns c >= row).
inline std::vector<std::pair<int, int>> block_pair_order(int B)
{
    std::vector<std::pair<int, int>> out;
    for (int r = 0; r <= (B - 1) / 2; r++) {
        const int r2 = B - 1 - r;
        for (int c = r; c < B; c++) out.emplace_back(r, c);
        if (r2 != r)
            for (int c = r2; c < B; c++) out.emplace_back(r2, c);
    }
    return out;
}

// Packed V layout.  The grid points are cut into tiles of kVTile points (the
// potential kernel's workgroup, one point per thread); the pairs of parity
// block b are stored tile by tile, [tile][pair of b][kVTile], block 0 first,
// so that a workgroup streams one contiguous region instead of one stripe
// per pair.  npair0 = pairs of block 0, ntile = ceil(P / kVTile); the last
// tile's points beyond P are padding and never read.
constexpr int kVTile = 256;
__host__ __device__ inline int64_t vpk_ntile(int64_t P) { return (P + kVTile - 1) / kVTile; }
__host__ __device__ inline size_t vpk_index(int64_t pair, int64_t p, int64_t npair0, int64_t npair, int64_t ntile)
{
    const bool b1 = pair >= npair0;
    const int64_t off = b1 ? npair0 : 0, nb = b1 ? npair - npair0 : npair0;
    return (size_t)(off * ntile * kVTile + ((p / kVTile) * nb + (pair - off)) * kVTile + p % kVTile);
}

// Harmonic-form potential kernel of an operator: 1 = the lmax-5 pair of
// blocks split over four waves per 64 points (k_vgaunt_split), 2 = a single
// lmax-5 block of the reduced basis (l = 0, 2, 4 or 1, 3, 5) split over three
// (k_vgaunt_split1), 0 = one point per thread and block (k_vgaunt).  The
// workgroups (= partial sums) are 64 points for the split kernels and
// kVTile points otherwise.
inline int harmonic_split(int nblk, int lb0, int par0, int lb1, int par1, int plane1, int size0)
{
    if (nblk == 2 && lb0 == 4 && par0 == 0 && lb1 == 5 && par1 == 1 && plane1 == size0) return 1;
    if (nblk == 1 && ((lb0 == 5 && par0 == 1) || (lb0 == 4 && par0 == 0))) return 2;
    return 0;
}
inline int64_t harmonic_workgroups(int64_t P, int split) { return split ? (P + 63) / 64 : vpk_ntile(P); }

// a parity block of the potential: `size` consecutive carried planes from `plane0`
struct Block {
    int size = 0;
    int plane0 = 0;
    int64_t pair_offset = 0;      // first packed pair of this block
    double mask = 1.0;            // symmetry mask of its planes (uniform within a block)
    int lb = 0, par = 0;          // largest l and the l parity of its planes
};

struct DevOperator {
    int nx = 0, ny = 0, nz = 0;
    int64_t P = 0;                // grid points
    int nk = 0;                   // carried lm planes
    int vprec = 0;                // BSL_PREC_* of the potential blocks
    int kprec = 0;                // BSL_PREC_*: kinetic contractions (FP64, FP32 or BF16x3 MFMA)
    double *tx = nullptr, *ty = nullptr, *tz = nullptr;  // Toeplitz rows
    double *rot = nullptr, *mask = nullptr;              // per carried plane
    void *vpk = nullptr;          // potential of the points (double or float): packed V blocks in the
                                  // tiled layout of vpk_index (vform 0) or harmonic coefficients
                                  // [tile][ncoef][kVTile] (vform 1)
    int vform = 0;                // BSL_VFORM_*
    int ncoef = 0;                // harmonic columns per point (vform 1)
    double *gc[2] = {nullptr, nullptr};  // per block: the values of its harmonic term list (vform 1)
    int nblk = 0;
    Block blk[2];
    int *plane_ref = nullptr;     // reference lm index of each carried plane
    int64_t npair = 0;
    int vb_variant = 1;           // occupancy variant of the potential kernel (autotuned)
    int fused = 0;                // 1: the z kinetic term runs inside the potential kernel (k_zpot*);
                                  // nalpha is then that kernel's workgroup count
    // quadrature calculator (calc == BSL_CALC_QUAD)
    int calc = 0;
    void *wgrid = nullptr;        // weighted potential grid in MFMA accumulator layout
    int wprec = 0;
    double *qL = nullptr;         // [nk][na] normalised Legendre of the carried planes
    double *qS = nullptr;         // [nm][nphi] tesseral trig factors, nphi = 2*nb
    int *mplane = nullptr;        // m + lmax of each carried plane
    int na = 0, nphi = 0, nm = 0;  // na: theta nodes of the folded hemisphere, ceil(n_theta / 2)
    int qmode = 0;                // 0 FP64 MFMA, 1 FP32 MFMA, 2 BF16x3 MFMA
    int q_npts = 0;               // grid points per group of the quadrature kernel
    int q_ne = 0;                 // carried planes of even l (the first q_ne)
    int q_grid = 0;               // its persistent workgroups (resident ones, at most the groups)
    // reduction scratch: the potential kernel writes nalpha per-workgroup
    // partials of v.R (partial), |R|^2 (partial_b) and |v|^2 (partial_c)
    double *partial = nullptr, *partial_b = nullptr, *partial_c = nullptr;
    int npartial = 0;
    int64_t nalpha = 0;
};

// Kernel-timestamp timing of a group of launches: while a group is armed,
// its first launch records `start` at the kernel's own start and every
// launch records `stop` at its end (hipExtLaunchKernel events), so that the
// interval is the group's execution as a profiler reports it, without the
// gaps an ordinary event record adds around it.  Thread-local; disarmed
// after the group.
struct LaunchTimer {
    hipEvent_t start = nullptr, stop = nullptr;
};
LaunchTimer &launch_timer();
template <typename F, typename... Args>
inline void timed_launch(F kernel, const dim3 &grid, const dim3 &block, uint32_t lds, hipStream_t s, Args... args)
{
    LaunchTimer &t = launch_timer();
    if (!t.stop) {  // not armed: the ordinary launch (also under stream capture)
        hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
        return;
    }
    hipEvent_t a = t.start;
    t.start = nullptr;
    hipExtLaunchKernelGGL(kernel, grid, block, lds, s, a, t.stop, 0u, args...);
}

// one Hv: v = ca X + cb R is formed by the x-y kinetic kernel and written
// to X; R = cr R + mask*(H v); the potential kernel leaves the nalpha
// partial sums of v.R, |R|^2, |v|^2 in op.partial, partial_b, partial_c.
// ev (or null): the start/stop events of the kinetic and of the potential
// launches (ev[0..1], ev[2..3])
int launch_hv(const DevOperator &op, double *X, double *R, Scal *sc, hipStream_t s,
              hipEvent_t *ev /* 4 events or null */);
// one Lanczos step: launch_hv, then launch_finalize (its start/stop in ev[4..5])
int launch_step(const DevOperator &op, double *X, double *R, Scal *sc, double *alpha_out, double *beta_out,
                double *beta2_out, hipStream_t s, hipEvent_t *ev /* 6 events or null */);
// the potential kernel alone (used by the autotuner)
int launch_potential(const DevOperator &op, const double *X, double *R, Scal *sc, hipStream_t s);
// The fused z-kinetic + potential kernel (k_zpot2 / k_zpot1): the lmax-5
// harmonic-form blocks (full basis pair, or one reduced-basis block), every
// axis of at most 64 points, FP64 or FP32 throughout (zpot_fits); one
// workgroup (= partial sum) per 4-column tile and block (zpot_workgroups).
bool zpot_fits(const DevOperator &op);
int64_t zpot_workgroups(const DevOperator &op);
int launch_zpot(const DevOperator &op, const double *X, double *R, hipStream_t s);
// what follows the kinetic launches of a step: the fused kernel when
// op.fused, else the potential kernel
int launch_potential_part(const DevOperator &op, const double *X, double *R, Scal *sc, hipStream_t s);
// the kinetic kernels alone (the MFMA contractions where every axis has at
// most 64 points, else the FP64 line kernels); forms v as launch_hv does
int launch_kinetic(const DevOperator &op, double *X, double *R, Scal *sc, hipStream_t s);
bool kinetic_mfma_fits(const DevOperator &op);
// a compiled register kernel for these parity block sizes (else the
// row-parallel block kernel, 64-point workgroups)
bool vblock_has_kernel(int b0, int b1);
// alpha = v.R, beta^2 = |R|^2 - 2 alpha v.R + alpha^2 |v|^2 from the
// potential kernel's three partial lists, and the scalars of the next step;
// alpha_out[step], beta_out[step + 1], beta2_out[step + 1] with step the
// device counter of completed steps, which advances by one
int launch_finalize(const DevOperator &op, Scal *sc, double *alpha_out, double *beta_out, double *beta2_out,
                    hipStream_t s);
// ca = scale of X in the vector the next Hv reads, cb = 0, cr = scale of R
int launch_init_scal(Scal *sc, double ca, double cr, hipStream_t s);
int launch_start_vector(double *X, int64_t n, int64_t P, const int *plane_ref, int kind, uint64_t seed, Scal *sc,
                        double *partial, int npartial, hipStream_t s);
int launch_zero(double *p, int64_t n, hipStream_t s);
// ritz[g] += coef[k][g] w[k] for the nw <= kRitzWindow Lanczos vectors w[k]
// of consecutive steps, in step order
constexpr int kRitzWindow = 16;
int launch_ritz_flush(double *ritz, const double *w, const double *coef, int nw, int64_t n, int ngood, hipStream_t s);

// per-workgroup pieces of a Rayleigh quotient (mode 0: y.y and y.hy; mode 1:
// |hy - e y|^2) into partial[2 b], partial[2 b + 1], b < *groups
int launch_rayleigh_dots(const double *y, const double *hy, int64_t n, double e, int mode, double *partial,
                         int npartial, int *groups, hipStream_t s);

// ---------------------------------------------------------------------------
// batches: one launch per kernel variant for all members that use it
// ---------------------------------------------------------------------------
// One member's operator for one launch of a table-driven batch kernel: its
// vectors for the current step parity, tables, scalars, potential and
// coefficient outputs, and its range [first, first + count) of the launch's
// virtual workgroups (the member's own grid, linearised).
struct BatchEntry {
    double *X, *R;
    const double *tx, *ty, *tz, *rot, *mask;
    Scal *sc;
    const void *H;
    const double *G0, *G1;
    double *pa, *pb, *pc;
    double *alpha, *beta, *beta2;
    int64_t P;
    int nx, ny, nz, nk;
    int nc, plane1;
    double mask0, mask1;
    int first, count;
    int nalpha, pad;
};
// the kernel variants a batch step is made of
enum BatchKind {
    kBXyHalf = 0, kBZHalf, kBPotSplitPara, kBPotSplitOrtho, kBZPotPara, kBZPotOrtho, kBFinalize, kBKinds
};
constexpr int kBatchMaxTable = 64;  // members per launch
// Device layout of one launch's table: kBatchMaxTable entries (the first nm
// used) followed by kBatchMaxTable ints, the entries' `first` values
// ascending, INT_MAX in the unused slots.
constexpr size_t kBatchTableBytes = sizeof(BatchEntry) * kBatchMaxTable + sizeof(int) * kBatchMaxTable;
// launch `kind` over the nm <= kBatchMaxTable entries of d_tab (device);
// prec: BSL_PREC_FP64 or FP32 (kinetic T, or potential VT); total = the
// entries' workgroups (for kBFinalize one per entry); G: the Gaunt values
// of the potential kinds' block (the same for every member of the launch);
// lds_nz: the longest z axis among the members (sizes the fused kinds' LDS)
int launch_batch(int kind, int prec, const BatchEntry *d_tab, int nm, int total, const double *G, int lds_nz,
                 hipStream_t s);

// prep kernels
struct CageDev {
    int nsites = 0;
    double *site = nullptr;  // [nsites][4]: x, y, z, q
    int *lj = nullptr;
};
// the clamped grid into vg, or (mask non-null, vg unused) only its clamp
// mask: bit idx = [vg[idx] < ceiling], 64 per word
int launch_vgrid_cage(double *vg, uint64_t *mask, int64_t P, int Q, const double *gx, const double *gy,
                      const double *gz, int nx, int ny, const double *off /*[Q][3]*/, const CageDev &cage,
                      double q_h, double q_cm, double lj_a, double lj_b, double ceiling, hipStream_t s);
int launch_vgrid_iho(double *vg, int64_t P, int Q, const double *gx, const double *gy, const double *gz, int nx,
                     int ny, double factor, double iho_ceil, double ceiling, hipStream_t s);
// potential universe of the grid-mapped cage potential
struct Universe {
    int n[3];
    double d[3], mx[3];  // spacing max/(n-1) and full length per axis
};
int launch_universe_cage(double *cm, double *hp, const Universe &u, const CageDev &cage, double q_cm, double q_h,
                         double lj_a, double lj_b, hipStream_t s);
int launch_vgrid_mapped(double *vg, int64_t P, int Q, const double *gx, const double *gy, const double *gz, int nx,
                        int ny, const double *dir, const double *cm, const double *hp, const Universe &u, double dist,
                        double ceiling, hipStream_t s);
int launch_vgrid_coulomb(double *vg, int64_t P, int Q, const double *gx, const double *gy, const double *gz, int nx,
                         int ny, double q1, double q2, double v_ceil, double ceiling, hipStream_t s);
// Vpk[pair][P] = sum_q Vg[P][q] * H[q][pair]
int launch_vblock_gemm(const double *vg, const double *h, void *vpk, int vprec, int64_t P, int Q, int64_t npair0,
                       int64_t npair, hipStream_t s);
// quadrature calculator (bsl_quad.hip)
// weighted quadrature grid in the MFMA accumulator layout of the quadrature
// kernel (npts grid points per workgroup)
int launch_weight_grid(const double *vg, void *W, int wprec, int mode, const double *wt, const double *wp, int64_t P,
                       int na, int nb, int npts, hipStream_t s);
// context run helpers (bsl_capi.cpp) used by the batch driver: coefficient
// capacity for niter steps, start vector, one step on the context stream
// (swaps the context's X/R; enqueues kernels only, so it can be captured)
int ctx_reserve_steps(bsl_ctx *c, int niter);
int ctx_start(bsl_ctx *c, const double *v0, int start, uint64_t seed);
int ctx_step(bsl_ctx *c);

// Hellmann-Feynman adjoint (bsl_adjoint.hip)
constexpr int kHfPoints = 4;       // points per workgroup of the gradient kernels
constexpr int kHfDensitySets = 4;  // most weight sets of a density launch
constexpr int kHfGradSets = 8;     // most weight sets of a gradient pass
int hf_grad_groups(int64_t P, int npb);
// the kept (point, node pair) entries of each group of npb points (npb <=
// 256, Q even and <= 2^24; node q < Q/2 stands for q and its inversion
// partner Q - 1 - q, kept if either is unclamped): count[g] per group, then,
// from first[g] (exclusive prefix sums of the counts, first[groups] =
// total), entries (pl << 24) | q in pair order
int launch_pair_count(const uint64_t *unclamped, int64_t P, int Q, int npb, int *count, hipStream_t s);
int launch_pair_fill(const uint64_t *unclamped, int64_t P, int Q, int npb, const int64_t *first, uint32_t *pairs,
                     hipStream_t s);
// nset (1 to kHfDensitySets) sets of counts[m] consecutive eigenvectors (psi, g): the
// folded weights of set m at om + m ne for the ne entries of the pair list
// (pairs / first, groups of kHfPoints points) and the H2-centre weights of
// its points at wcm + m P
int launch_hf_density(const double *psi, int nset, const int *counts, const double *g, const double *Y,
                      const double *wq, const uint64_t *unclamped, const uint32_t *pairs, const int64_t *first,
                      int64_t ne, double *om, double *wcm, int64_t P, int nk, int Q, hipStream_t s);
// the gradients of nset (1 to kHfGradSets) sets whose weights are at om +
// m ne and wcm + m P (npb = kHfPoints); partial:
// hf_grad_groups(P, npb) * nset * (4 ns + 4) doubles; out: nset * (4 ns + 4),
// set after set
int launch_hf_grad(const double *om, int64_t ne, const double *wcm, int nset, int64_t P, const double *gx,
                   const double *gy, const double *gz, int nx, int ny, const double *off, const double *site4,
                   const int *lj, int ns, double q_h, double q_cm, double lj_a, double lj_b, int npb,
                   const uint32_t *pairs, const int64_t *first, double *partial, double *out, hipStream_t s);
int quad_points_per_group(int na, int nk);
// workgroups of the persistent quadrature kernel: its resident workgroups
// per CU times the CUs of `device`, at most the number of point groups
int quad_grid(const DevOperator &op, int device, int *grid);
int64_t quad_tiles(int64_t P, int na, int npts);
int quad_phi_tiles(int nphi);
// the points-per-group candidates of the quadrature kernel's autotuner
// (op.q_npts the smallest), in increasing order
std::vector<int> quad_group_candidates(const DevOperator &op);
int launch_quad_potential(const DevOperator &op, const double *X, double *R, Scal *sc, hipStream_t s);

}  // namespace bsl

struct bsl_ctx {
    bsl_params p{};
    std::vector<double> site_pos, site_q;
    std::vector<int32_t> site_lj;
    bsl::Tables tab;
    int nn = 0;                        // full lm count
    int64_t P = 0, n_full = 0, n_kry = 0;
    std::vector<int> carried;          // carried plane -> reference lm index (grouped by parity block)
    std::vector<double> rot, mask;     // per carried plane
    bsl::DevOperator op;               // the operator in the context precision
    bsl::DevOperator op64;             // FP64 operator (same as op when FP64)
    void *vpk64 = nullptr;             // FP64 blocks kept for refinement when op is mixed
    double *d_gx = nullptr, *d_gy = nullptr, *d_gz = nullptr;
    double *d_vgrid = nullptr;         // [P][Q] potential grid (kept for the quadrature calculator)
    uint64_t *d_unclamped = nullptr;   // bit p*Q + q: V(p, q) < ceiling; built by the first gradient call
    uint32_t *d_pairs = nullptr;       // with it: the kept node pairs of each group of kHfPoints points
    int64_t *d_pair_first = nullptr;   // (launch_pair_fill), and each group's first entry
    // gradient workspace, kept between calls: the node tables (harmonics at
    // the nodes, node weights, H offsets, sites) built with the pair list,
    // and buffers that grow to the largest request
    double *d_hfY = nullptr, *d_hfwq = nullptr, *d_hfoff = nullptr, *d_hfsite = nullptr;
    int *d_hflj = nullptr;
    double *d_hfpsi = nullptr, *d_hfg = nullptr, *d_hfom = nullptr, *d_hfwcm = nullptr, *d_hfpart = nullptr,
           *d_hfout = nullptr;
    size_t hf_cap[6] = {0, 0, 0, 0, 0, 0};  // capacities (doubles) of d_hfpsi .. d_hfout
    int64_t hf_ne = 0;                        // entries of the pair list
    int Q = 0;
    double *d_X = nullptr, *d_R = nullptr, *d_T = nullptr;
    // Loop-2 workspace, kept between calls (grows to the largest request):
    // Ritz accumulators [ngood][n_kry] and the per-step coefficients
    double *d_ritz = nullptr, *d_rcoef = nullptr, *d_rwin = nullptr;  // d_rwin: staged Lanczos vectors
    size_t ritz_cap = 0, rcoef_cap = 0, rwin_cap = 0;
    // squared norms of the Ritz vectors the last bsl_ritz_refine left in
    // d_ritz (empty when they are gone: bsl_ritz and Loop-2 calls reuse it)
    std::vector<double> ritz_vv;
    bsl::Scal *d_scal = nullptr;
    double *d_alpha = nullptr, *d_beta = nullptr, *d_beta2 = nullptr;
    int niter_cap = 0;
    hipStream_t stream = nullptr;
    int device = 0;
};
