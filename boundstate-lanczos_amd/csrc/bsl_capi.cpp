// bsl_capi.cpp -- the C ABI of libbsl_hip.so: context preparation (host
// tables, device potential grid and V blocks), Hv on host or device vectors,
// the device-resident Lanczos loops, and the host Cullum-Willoughby
// post-processing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "bsl_internal.h"

namespace bsl {

namespace {
thread_local std::string g_err;
}

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    char buf[512];
    std::snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    g_err = buf;
    return e == hipErrorOutOfMemory ? BSL_E_NOMEM : BSL_E_HIP;
}

namespace {

template <typename T>
int dev_alloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    BSL_HIP(hipMalloc((void **)p, sizeof(T) * count));
    return BSL_OK;
}

template <typename T>
int dev_upload(T **p, const std::vector<T> &v)
{
    int rc = dev_alloc(p, v.size());
    if (rc) return rc;
    BSL_HIP(hipMemcpy(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return BSL_OK;
}

void dev_free(void *p)
{
    if (p) (void)hipFree(p);
}

int check_params(const bsl_params &p)
{
    if (p.nx < 2 || p.ny < 2 || p.nz < 2) return fail(BSL_E_ARG, "grid needs at least 2 points per axis");
    if (p.nx > 256 || p.ny > 256 || p.nz > 256) return fail(BSL_E_ARG, "at most 256 grid points per axis");
    if (!(p.x_max > 0) || !(p.y_max > 0) || !(p.z_max > 0)) return fail(BSL_E_ARG, "box lengths must be positive");
    if (p.lmax < 0 || p.lmax > 12) return fail(BSL_E_ARG, "lmax must be in 0..12");
    if (p.n_theta < 1 || p.n_phi < 1) return fail(BSL_E_ARG, "quadrature sizes must be positive");
    if (!(p.mass > 0)) return fail(BSL_E_ARG, "mass must be positive");
    if (p.symmetry < 0 || p.symmetry > 2) return fail(BSL_E_ARG, "symmetry flag not recognised");
    if (p.basis < 0 || p.basis > 1) return fail(BSL_E_ARG, "basis flag not recognised");
    if (p.calculator < 0 || p.calculator > 1) return fail(BSL_E_ARG, "calculator flag not recognised");
    if (p.precision < 0 || p.precision > 2) return fail(BSL_E_ARG, "precision flag not recognised");
    if (p.potential < 0 || p.potential > 3) return fail(BSL_E_ARG, "potential flag not recognised");
    const bool sites = p.potential == BSL_POT_CAGE || p.potential == BSL_POT_CAGE_GRID;
    if (sites && (p.n_sites <= 0 || !p.site_pos || !p.site_q || !p.site_lj))
        return fail(BSL_E_ARG, "cage potential needs sites");
    if (sites && p.n_sites > 4096) return fail(BSL_E_ARG, "at most 4096 sites");
    if (p.potential == BSL_POT_CAGE_GRID) {
        if (p.pnx < 2 || p.pny < 2 || p.pnz < 2) return fail(BSL_E_ARG, "potential universe needs 2 points per axis");
        if ((double)p.pnx * p.pny * p.pnz > 2.0e9) return fail(BSL_E_ARG, "potential universe too large");
        // HvPrep_Internal_NoQuad linRotCartSphHvFullVPrepRoutines.cpp:330-348
        if (p.x_max > p.px_max || p.y_max > p.py_max || p.z_max > p.pz_max)
            return fail(BSL_E_ARG, "the system's dimensions must not exceed the potential universe's");
    }
    return BSL_OK;
}

// Does every parity block fit the harmonic-form kernels (l <= kMaxLb, planes
// in the order of its term list)?
bool harmonic_fits(const bsl_ctx *c, const DevOperator &op)
{
    for (int g = 0; g < op.nblk; g++) {
        const Block &B = op.blk[g];
        if (B.lb > gaunt::kMaxLb || B.size != gaunt::block_size(B.lb, B.par)) return false;
        for (int k = 0; k < B.size; k++) {
            const int n = c->carried[B.plane0 + k];
            if (c->tab.l[n] != gaunt::block_l(B.lb, B.par, k) || c->tab.m[n] != gaunt::block_m(B.lb, B.par, k))
                return false;
        }
    }
    return true;
}

// Angular table of a potential form, [q][column] with ncol columns:
// BLOCKS H[q][pair] = w_a w_b L_n(a) S_m(b) L_n'(a) S_m'(b), q = (half, a, b)
// (calc_Vlmlpmp_NoQuad linRotCartSphHvFullVPrepRoutines.cpp:116-131);
// HARMONICS H[q][c] = w_a w_b Yh_c(a, b) over the even-L harmonics of the
// largest block.
void potential_table(const bsl_ctx *c, const DevOperator &op, int form, std::vector<double> &h, int64_t *ncol)
{
    const Tables &t = c->tab;
    const int na = t.na, nb = t.nb, lmax = t.lmax;
    if (form == BSL_VFORM_HARMONICS) {
        int lb = 0;
        for (int g = 0; g < op.nblk; g++) lb = std::max(lb, op.blk[g].lb);
        harmonic_table(t, lb, h);
        *ncol = gaunt::ncoef(lb);
        return;
    }
    const int64_t pairs = op.npair;
    h.assign((size_t)c->Q * pairs, 0.0);
    for (int half = 0; half < 2; half++) {
        const std::vector<double> &S = half == 0 ? t.S1 : t.S2;
        for (int a = 0; a < na; a++)
            for (int b = 0; b < nb; b++) {
                const size_t q = (size_t)half * na * nb + (size_t)a * nb + b;
                for (int g = 0; g < op.nblk; g++) {
                    const Block &B = op.blk[g];
                    int64_t pr = B.pair_offset;
                    for (const auto &rc : block_pair_order(B.size)) {
                        const int n = c->carried[B.plane0 + rc.first], np = c->carried[B.plane0 + rc.second];
                        const int m = t.m[n], mp = t.m[np];
                        h[q * pairs + pr] = t.wt[a] * t.wp[b] * t.L[(size_t)n * na + a] *
                                            S[(size_t)(m + lmax) * nb + b] * t.L[(size_t)np * na + a] *
                                            S[(size_t)(mp + lmax) * nb + b];
                        pr++;
                    }
                }
            }
    }
    *ncol = pairs;
}

// Workgroups (= partial sums) of the kernel that ends a NoQuad step: the
// fused z-kinetic + potential kernel when op.fused, else the potential kernel
// of the operator's form and blocks.
int64_t potential_workgroups(const DevOperator &op)
{
    if (op.fused) return zpot_workgroups(op);
    const Block &b0 = op.blk[0], &b1 = op.blk[op.nblk > 1 ? 1 : 0];
    const int split = harmonic_split(op.nblk, b0.lb, b0.par, op.nblk > 1 ? b1.lb : -1, b1.par, b1.plane0, b0.size);
    const int bs1 = op.nblk > 1 ? b1.size : 0;
    return op.vform == BSL_VFORM_HARMONICS      ? harmonic_workgroups(op.P, split)
           : vblock_has_kernel(b0.size, bs1) ? vpk_ntile(op.P) * op.nblk
                                             : ((op.P + 63) / 64) * op.nblk;
}

// Switch a NoQuad operator between the three-kernel step and the fused one
// (on is ignored where zpot_fits does not hold); keeps nalpha in step.
int set_fused(DevOperator &op, int on)
{
    if (op.calc == BSL_CALC_QUAD) return BSL_OK;
    op.fused = on && zpot_fits(op) ? 1 : 0;
    op.nalpha = potential_workgroups(op);
    if (op.nalpha > op.npartial) {
        op.fused = 0;
        op.nalpha = potential_workgroups(op);
        return fail(BSL_E_STATE, "partial buffer too small for the fused kernel");
    }
    return BSL_OK;
}

// Build a DevOperator's potential (form, precision) from the grid d_vg and
// its angular table d_h of ncol columns
int build_vpk(bsl_ctx *c, DevOperator &op, int form, int vprec, const double *d_vg, const double *d_h, int64_t ncol)
{
    op.vprec = vprec;
    op.vform = form;
    const size_t es = vprec == BSL_PREC_FP64 ? sizeof(double) : sizeof(float);
    const int64_t ntile = vpk_ntile(c->P);
    BSL_HIP(hipMalloc(&op.vpk, es * (size_t)ncol * (size_t)(ntile * kVTile)));
    int64_t col0 = ncol;  // columns of the first layout region
    if (form == BSL_VFORM_HARMONICS) {
        op.ncoef = (int)ncol;
        for (int g = 0; g < op.nblk; g++) {
            std::vector<double> gv;
            double excluded = 0.0;
            gaunt_values(op.blk[g].lb, op.blk[g].par, gv, &excluded);
            if ((int64_t)gv.size() != gaunt::count_terms(op.blk[g].lb, op.blk[g].par) || excluded > 1e-12)
                return fail(BSL_E_STATE, "harmonic term list does not match the block's couplings");
            int rc = dev_upload(&op.gc[g], gv);
            if (rc) return rc;
        }
    } else {
        op.ncoef = 0;
        col0 = op.nblk > 1 ? op.blk[1].pair_offset : op.npair;
    }
    if (op.calc != BSL_CALC_QUAD) {
        op.fused = 0;
        op.nalpha = potential_workgroups(op);
        if (op.nalpha > op.npartial) return fail(BSL_E_STATE, "partial buffer too small for the potential kernel");
    }
    return launch_vblock_gemm(d_vg, d_h, op.vpk, vprec, c->P, c->Q, col0, ncol, c->stream);
}

// release the potential buffers of one operator
void free_potential(DevOperator &o)
{
    dev_free(o.vpk);
    dev_free(o.gc[0]);
    dev_free(o.gc[1]);
    o.vpk = nullptr;
    o.gc[0] = o.gc[1] = nullptr;
}

// release the potential of op (and of op64 unless it shares op's)
void free_potentials(bsl_ctx *c)
{
    const bool shared = c->op64.vpk == c->op.vpk;
    free_potential(c->op);
    if (shared) {
        c->op64.vpk = nullptr;
        c->op64.gc[0] = c->op64.gc[1] = nullptr;
    } else {
        free_potential(c->op64);
    }
}

// op reads the potential of src (its buffers, form and partial count)
void share_potential(DevOperator &op, const DevOperator &src)
{
    op.vpk = src.vpk;
    op.vprec = src.vprec;
    op.vform = src.vform;
    op.ncoef = src.ncoef;
    op.gc[0] = src.gc[0];
    op.gc[1] = src.gc[1];
    op.fused = src.fused;
    op.nalpha = src.nalpha;
}

// The potential of both operators in `form` from the grid d_vg: op64 in
// FP64, op in the context precision (shared when that is FP64).  A
// quadrature context's op has no potential blocks.  The new potential is
// built into copies of the operators and replaces the old one only when
// every part of it was built, so that a failure (out of memory, a term list
// that does not match) leaves the context as it was.
int build_potentials(bsl_ctx *c, int form, const double *d_vg)
{
    std::vector<double> h;
    int64_t ncol = 0;
    potential_table(c, c->op64, form, h, &ncol);
    const bool quad = c->op.calc == BSL_CALC_QUAD;
    const bool shared = !quad && c->p.precision == BSL_PREC_FP64;
    DevOperator n64 = c->op64, nop = c->op;
    n64.vpk = nop.vpk = nullptr;
    n64.gc[0] = n64.gc[1] = nop.gc[0] = nop.gc[1] = nullptr;
    double *d_h = nullptr;
    int rc = dev_upload(&d_h, h);
    if (!rc) rc = build_vpk(c, n64, form, BSL_PREC_FP64, d_vg, d_h, ncol);
    if (!rc && !quad) {
        if (shared)
            share_potential(nop, n64);
        else
            rc = build_vpk(c, nop, form, BSL_PREC_FP32, d_vg, d_h, ncol);
    }
    const hipError_t e = hipStreamSynchronize(c->stream);
    dev_free(d_h);
    if (!rc && e != hipSuccess) rc = hip_fail(e, "build_potentials", __FILE__, __LINE__);
    if (rc) {
        free_potential(n64);
        if (!shared) free_potential(nop);
        return rc;
    }
    free_potentials(c);
    share_potential(c->op64, n64);
    if (!quad) share_potential(c->op, nop);
    return BSL_OK;
}

// H offsets of the quadrature nodes (q = half*na*nb + a*nb + b; the second
// half mirrors y), site [x, y, z, q] rows and LJ flags of a cage context
void cage_arrays(const bsl_ctx *c, std::vector<double> &off, std::vector<double> &site4, std::vector<int> &lj)
{
    const bsl_params &p = c->p;
    const Tables &t = c->tab;
    const int na = t.na, nb = t.nb;
    off.assign((size_t)c->Q * 3, 0.0);
    for (int a = 0; a < na; a++)
        for (int b = 0; b < nb; b++) {
            const size_t q0 = (size_t)a * nb + b, q1 = (size_t)na * nb + q0;
            off[3 * q0 + 0] = p.h2_half_bond * std::sin(t.th[a]) * t.cp[b];
            off[3 * q0 + 1] = p.h2_half_bond * std::sin(t.th[a]) * std::sin(t.ph1[b]);
            off[3 * q0 + 2] = p.h2_half_bond * t.ct[a];
            off[3 * q1 + 0] = off[3 * q0 + 0];
            off[3 * q1 + 1] = -1.0 * (off[3 * q0 + 1]);
            off[3 * q1 + 2] = off[3 * q0 + 2];
        }
    site4.assign((size_t)4 * p.n_sites, 0.0);
    for (int s = 0; s < p.n_sites; s++) {
        for (int d = 0; d < 3; d++) site4[4 * s + d] = c->site_pos[3 * s + d];
        site4[4 * s + 3] = c->site_q[s];
    }
    lj.assign(c->site_lj.begin(), c->site_lj.end());
}

// the cage potential's clamped grid [P][Q] (device point order) into d_vg,
// or (mask non-null) only its clamp mask
int cage_vgrid(bsl_ctx *c, double *d_vg, uint64_t *mask)
{
    const bsl_params &p = c->p;
    std::vector<double> off, site4;
    std::vector<int> lj;
    cage_arrays(c, off, site4, lj);
    CageDev cage;
    cage.nsites = p.n_sites;
    double *d_off = nullptr;
    int rc = dev_upload(&cage.site, site4);
    if (!rc) rc = dev_upload(&cage.lj, lj);
    if (!rc) rc = dev_upload(&d_off, off);
    if (!rc)
        rc = launch_vgrid_cage(d_vg, mask, c->P, c->Q, c->d_gx, c->d_gy, c->d_gz, p.nx, p.ny, d_off, cage, p.q_h2_h,
                               p.q_h2_cm, p.lj_a, p.lj_b, p.ceiling, c->stream);
    (void)hipStreamSynchronize(c->stream);
    dev_free(cage.site);
    dev_free(cage.lj);
    dev_free(d_off);
    return rc;
}

// the clamped potential grid [P][Q] (device point order) into d_vg
int make_vgrid(bsl_ctx *c, double *d_vg)
{
    const bsl_params &p = c->p;
    int rc = BSL_OK;
    if (p.potential == BSL_POT_CAGE) {
        rc = cage_vgrid(c, d_vg, nullptr);
    } else if (p.potential == BSL_POT_CAGE_GRID) {
        std::vector<double> off, site4;
        std::vector<int> lj;
        cage_arrays(c, off, site4, lj);
        // H-H axis directions of the nodes, as the reference evaluates them
        // (theta = acos(x_a), phi = acos(x_b) and 2 pi - acos(x_b))
        const Tables &t = c->tab;
        const int na = t.na, nb = t.nb;
        std::vector<double> dir((size_t)c->Q * 3);
        for (int half = 0; half < 2; half++)
            for (int a = 0; a < na; a++)
                for (int b = 0; b < nb; b++) {
                    const size_t q = (size_t)half * na * nb + (size_t)a * nb + b;
                    const double th = std::acos(t.ct[a]);
                    const double ph = half == 0 ? std::acos(t.cp[b]) : 2 * M_PI - std::acos(t.cp[b]);
                    dir[3 * q + 0] = std::sin(th) * std::cos(ph);
                    dir[3 * q + 1] = std::sin(th) * std::sin(ph);
                    dir[3 * q + 2] = std::cos(th);
                }
        Universe u;
        const int un[3] = {p.pnx, p.pny, p.pnz};
        const double um[3] = {p.px_max, p.py_max, p.pz_max};
        for (int d = 0; d < 3; d++) {
            u.n[d] = un[d];
            u.mx[d] = um[d];
            u.d[d] = um[d] / (double(un[d] - 1));
        }
        const size_t nu = (size_t)p.pnx * p.pny * p.pnz;
        CageDev cage;
        cage.nsites = p.n_sites;
        double *d_dir = nullptr, *d_cm = nullptr, *d_hp = nullptr;
        if (!rc) rc = dev_upload(&cage.site, site4);
        if (!rc) rc = dev_upload(&cage.lj, lj);
        if (!rc) rc = dev_upload(&d_dir, dir);
        if (!rc) rc = dev_alloc(&d_cm, nu);
        if (!rc) rc = dev_alloc(&d_hp, nu);
        if (!rc) rc = launch_universe_cage(d_cm, d_hp, u, cage, p.q_h2_cm, p.q_h2_h, p.lj_a, p.lj_b, c->stream);
        if (!rc)
            rc = launch_vgrid_mapped(d_vg, c->P, c->Q, c->d_gx, c->d_gy, c->d_gz, p.nx, p.ny, d_dir, d_cm, d_hp, u,
                                     p.h2_half_bond, p.ceiling, c->stream);
        (void)hipStreamSynchronize(c->stream);
        dev_free(cage.site);
        dev_free(cage.lj);
        dev_free(d_dir);
        dev_free(d_cm);
        dev_free(d_hp);
    } else if (p.potential == BSL_POT_COULOMB) {
        rc = launch_vgrid_coulomb(d_vg, c->P, c->Q, c->d_gx, c->d_gy, c->d_gz, p.nx, p.ny, p.coul_q_centre,
                                  p.coul_q_mol, p.coul_ceil, p.ceiling, c->stream);
    } else {
        rc = launch_vgrid_iho(d_vg, c->P, c->Q, c->d_gx, c->d_gy, c->d_gz, p.nx, p.ny, p.iho_factor, p.iho_ceil,
                              p.ceiling, c->stream);
    }
    return rc;
}

// Picks the occupancy variant of the potential-block kernel by timing each
// on the context's own operator (the variants compute identical results).
int time_potential(bsl_ctx *c, DevOperator &op, float *ms)
{
    hipEvent_t e0, e1;
    BSL_HIP(hipEventCreate(&e0));
    BSL_HIP(hipEventCreate(&e1));
    int rc = launch_init_scal(c->d_scal, 1.0, 1.0, c->stream);
    for (int i = 0; i < 2 && !rc; i++) rc = launch_potential(op, c->d_X, c->d_R, c->d_scal, c->stream);
    if (!rc) {
        BSL_HIP(hipEventRecord(e0, c->stream));
        for (int i = 0; i < 5 && !rc; i++) rc = launch_potential(op, c->d_X, c->d_R, c->d_scal, c->stream);
        BSL_HIP(hipEventRecord(e1, c->stream));
        BSL_HIP(hipEventSynchronize(e1));
        BSL_HIP(hipEventElapsedTime(ms, e0, e1));
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// The quadrature kernel's tile shape: points per workgroup group (each
// candidate a multiple of the smallest whole-tile group whose shared memory
// fits), with the persistent grid the occupancy query gives for it; the
// fastest on the context's own operator is kept.
int tune_quad(bsl_ctx *c)
{
    DevOperator &op = c->op;
    int rc = launch_zero(c->d_X, c->n_kry, c->stream);
    if (!rc) rc = launch_zero(c->d_R, c->n_kry, c->stream);
    float best = 1e30f;
    int pick = op.q_npts, pick_grid = op.q_grid;
    for (int npts : quad_group_candidates(op)) {
        if (rc) break;
        op.q_npts = npts;
        if ((rc = quad_grid(op, c->device, &op.q_grid))) break;
        if (op.q_grid + 1 > op.npartial) continue;
        op.nalpha = op.q_grid;
        float ms = 0;
        rc = time_potential(c, op, &ms);
        if (!rc && ms < best) {
            best = ms;
            pick = npts;
            pick_grid = op.q_grid;
        }
    }
    op.q_npts = pick;
    op.q_grid = pick_grid;
    op.nalpha = pick_grid;
    if (!rc) rc = launch_init_scal(c->d_scal, 0.0, 0.0, c->stream);
    if (!rc) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(e, "autotune", __FILE__, __LINE__);
    }
    return rc;
}

// Picks, per operator, the occupancy variant of the potential kernel (the
// variants compute identical results).  The step keeps its three kernels;
// the fused z-kinetic + potential kernel is selected only on request
// (bsl_set_fusion): it moves fewer bytes but runs slower on this part.
int autotune(bsl_ctx *c)
{
    if (c->op.calc == BSL_CALC_QUAD) return tune_quad(c);
    int rc = launch_zero(c->d_X, c->n_kry, c->stream);
    if (!rc) rc = launch_zero(c->d_R, c->n_kry, c->stream);
    DevOperator *ops[2] = {&c->op, &c->op64};
    for (int o = 0; o < 2 && !rc; o++) {
        if (o == 1 && c->op64.vpk == c->op.vpk) {
            c->op64.vb_variant = c->op.vb_variant;
            c->op64.fused = c->op.fused;
            c->op64.nalpha = c->op.nalpha;
            break;
        }
        if ((rc = set_fused(*ops[o], 0))) break;
        // each variant's best of three rounds, the rounds interleaved so that
        // a slow moment of the device does not decide the choice
        const int nvar = ops[o]->vform == BSL_VFORM_HARMONICS ? 2 : 3;
        float best[3] = {1e30f, 1e30f, 1e30f};
        for (int round = 0; round < 3 && !rc; round++)
            for (int v = 0; v < nvar && !rc; v++) {
                ops[o]->vb_variant = v;
                float ms = 0;
                rc = time_potential(c, *ops[o], &ms);
                if (!rc && ms < best[v]) best[v] = ms;
            }
        int pick = 1;
        for (int v = 0; v < nvar; v++)
            if (best[v] < best[pick]) pick = v;
        ops[o]->vb_variant = pick;
    }
    if (!rc) rc = launch_init_scal(c->d_scal, 0.0, 0.0, c->stream);
    if (!rc) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(e, "autotune", __FILE__, __LINE__);
    }
    return rc;
}

}  // namespace
}  // namespace bsl

using namespace bsl;

extern "C" const char *bsl_last_error(void) { return g_err.c_str(); }

namespace {
void free_hf_setup(bsl_ctx *c);
}  // namespace

extern "C" int bsl_destroy(bsl_ctx *c)
{
    if (!c) return BSL_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_potentials(c);
    DevOperator &o = c->op;
    dev_free(o.tx);
    if (o.ty != o.tx) dev_free(o.ty);
    if (o.tz != o.tx && o.tz != o.ty) dev_free(o.tz);
    dev_free(o.rot);
    dev_free(o.mask);
    dev_free(o.plane_ref);
    dev_free(o.wgrid);
    dev_free(o.qL);
    dev_free(o.qS);
    dev_free(o.mplane);
    dev_free(o.partial);
    dev_free(o.partial_b);
    dev_free(o.partial_c);
    dev_free(c->d_gx);
    dev_free(c->d_gy);
    dev_free(c->d_gz);
    dev_free(c->d_vgrid);
    free_hf_setup(c);
    for (void *p : {(void *)c->d_hfpsi, (void *)c->d_hfg, (void *)c->d_hfom, (void *)c->d_hfwcm, (void *)c->d_hfpart,
                    (void *)c->d_hfout})
        dev_free(p);
    dev_free(c->d_X);
    dev_free(c->d_R);
    dev_free(c->d_T);
    dev_free(c->d_ritz);
    dev_free(c->d_rwin);
    dev_free(c->d_rcoef);
    dev_free(c->d_scal);
    dev_free(c->d_alpha);
    dev_free(c->d_beta);
    dev_free(c->d_beta2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return BSL_OK;
}

extern "C" int bsl_create(const bsl_params *pp, bsl_ctx **out)
{
    if (!pp || !out) return fail(BSL_E_ARG, "null argument");
    *out = nullptr;
    int rc = check_params(*pp);
    if (rc) return rc;
    // every early return below releases what was allocated so far
    std::unique_ptr<bsl_ctx, int (*)(bsl_ctx *)> c(new bsl_ctx(), bsl_destroy);
    c->p = *pp;
    if (pp->potential == BSL_POT_CAGE || pp->potential == BSL_POT_CAGE_GRID) {
        c->site_pos.assign(pp->site_pos, pp->site_pos + 3 * pp->n_sites);
        c->site_q.assign(pp->site_q, pp->site_q + pp->n_sites);
        c->site_lj.assign(pp->site_lj, pp->site_lj + pp->n_sites);
        c->p.site_pos = c->site_pos.data();
        c->p.site_q = c->site_q.data();
        c->p.site_lj = c->site_lj.data();
    }
    const bsl_params &p = c->p;
    BSL_HIP(hipGetDevice(&c->device));
    BSL_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));

    // ---- host tables (HvPrep_Internal_NoQuad :356-455) ----
    build_tables(p.lmax, p.n_theta, p.n_phi, c->tab);
    const Tables &t = c->tab;
    c->nn = t.nn;
    c->P = (int64_t)p.nx * p.ny * p.nz;
    c->n_full = c->P * c->nn;
    std::vector<double> rowx, rowy, rowz, gx, gy, gz;
    kinetic_row(p.x_max, p.nx, p.mass, rowx, gx);
    kinetic_row(p.y_max, p.ny, p.mass, rowy, gy);
    kinetic_row(p.z_max, p.nz, p.mass, rowz, gz);

    // carried lm planes, grouped by parity block (even l first), each block
    // in reference lm order
    auto in_sym = [&](int l) { return p.symmetry == BSL_SYM_ALL || (l % 2 == 0) == (p.symmetry == BSL_SYM_EVEN); };
    DevOperator &op = c->op;
    op.nx = p.nx;
    op.ny = p.ny;
    op.nz = p.nz;
    op.P = c->P;
    int64_t pairs = 0;
    for (int parity = 0; parity < 2; parity++) {
        Block b;
        b.plane0 = (int)c->carried.size();
        for (int n = 0; n < t.nn; n++)
            if (t.l[n] % 2 == parity && (p.basis == BSL_BASIS_FULL || in_sym(t.l[n]))) c->carried.push_back(n);
        b.size = (int)c->carried.size() - b.plane0;
        if (b.size == 0) continue;
        b.pair_offset = pairs;
        b.mask = (p.basis == BSL_BASIS_FULL && !in_sym(parity)) ? 0.0 : 1.0;
        b.par = parity;
        b.lb = 0;
        for (int k = b.plane0; k < (int)c->carried.size(); k++) b.lb = std::max(b.lb, t.l[c->carried[k]]);
        pairs += (int64_t)b.size * (b.size + 1) / 2;
        op.blk[op.nblk++] = b;
    }
    op.npair = pairs;
    const int nk = (int)c->carried.size();
    op.nk = nk;
    c->n_kry = c->P * nk;
    for (int kidx = 0; kidx < nk; kidx++) {
        const int n = c->carried[kidx];
        const double l = (double)t.l[n];
        c->rot.push_back(p.rot_const * l * (l + 1.0));  // rotKinEng :522-538
        c->mask.push_back(p.basis == BSL_BASIS_FULL ? (in_sym(t.l[n]) ? 1.0 : 0.0) : 1.0);
    }

    const int na = t.na, nb = t.nb, lmax = t.lmax;
    c->Q = 2 * na * nb;

    // ---- device: tables ----
    if ((rc = dev_upload(&op.tx, rowx))) return rc;
    if ((rc = dev_upload(&op.ty, rowy))) return rc;
    // one zero past the row (tz[nz] = 0: the fused kernel clamps Toeplitz
    // offsets beyond the axis onto it)
    rowz.push_back(0.0);
    if ((rc = dev_upload(&op.tz, rowz))) return rc;
    if ((rc = dev_upload(&op.rot, c->rot))) return rc;
    if ((rc = dev_upload(&op.mask, c->mask))) return rc;
    if ((rc = dev_upload(&op.plane_ref, c->carried))) return rc;
    if ((rc = dev_upload(&c->d_gx, gx))) return rc;
    if ((rc = dev_upload(&c->d_gy, gy))) return rc;
    if ((rc = dev_upload(&c->d_gz, gz))) return rc;
    op.nalpha = ((c->P + 255) / 256) * op.nblk;
    // room for every potential kernel's workgroups (64 points the smallest)
    op.npartial = (int)std::max<int64_t>(((c->P + 63) / 64) * op.nblk, 4096);
    if ((rc = dev_alloc(&op.partial, (size_t)op.npartial))) return rc;
    if ((rc = dev_alloc(&op.partial_b, (size_t)op.npartial))) return rc;
    if ((rc = dev_alloc(&op.partial_c, (size_t)op.npartial))) return rc;

    // ---- device: potential grid and V blocks ----
    BSL_HIP(hipMalloc((void **)&c->d_vgrid, sizeof(double) * (size_t)c->P * c->Q));
    if ((rc = make_vgrid(c.get(), c->d_vgrid))) return rc;
    c->op64 = op;
    c->op64.vpk = nullptr;
    c->op64.kprec = BSL_PREC_FP64;
    // mixed precisions contract the kinetic term in FP32 as well
    // (BF16x3 quadrature contexts keep FP32 kinetics: BF16 split products
    // gain nothing on the bandwidth-bound kinetic kernels and cost accuracy)
    op.kprec = (p.precision == BSL_PREC_BF16X3 && p.calculator == BSL_CALC_QUAD) ? BSL_PREC_FP32 : p.precision;
    if (p.calculator == BSL_CALC_QUAD) op.calc = BSL_CALC_QUAD;
    const int form = harmonic_fits(c.get(), op) ? BSL_VFORM_HARMONICS : BSL_VFORM_BLOCKS;
    if ((rc = build_potentials(c.get(), form, c->d_vgrid))) return rc;
    if (p.calculator == BSL_CALC_QUAD) {
        // weighted grid W[P][a][half*nb + b] and the transform tables
        op.calc = BSL_CALC_QUAD;
        op.wprec = p.precision == BSL_PREC_FP64 ? BSL_PREC_FP64 : BSL_PREC_FP32;
        op.qmode = p.precision;
        // hemisphere fold: the kernel's columns are the nodes a < ceil(na/2)
        const int nah = (na + 1) / 2;
        op.na = nah;
        op.nphi = 2 * nb;
        op.nm = 2 * lmax + 1;
        op.q_ne = 0;
        for (int r = 0; r < nk; r++) op.q_ne += t.l[c->carried[r]] % 2 == 0 ? 1 : 0;
        for (int r = 0; r < op.q_ne; r++)
            if (t.l[c->carried[r]] % 2 != 0) return fail(BSL_E_STATE, "carried planes are not grouped by l parity");
        op.q_npts = quad_points_per_group(nah, nk);
        // the grid W is laid out by global column tile whatever the group
        // size, so one allocation (padded for the largest group the
        // autotuner tries) serves every candidate
        const int npts_max = quad_group_candidates(op).back();
        const size_t es = op.wprec == BSL_PREC_FP64 ? sizeof(double) : sizeof(float);
        const int64_t ntiles = quad_tiles(c->P, nah, npts_max);
        BSL_HIP(hipMalloc(&op.wgrid, es * (size_t)ntiles * quad_phi_tiles(op.nphi) * 256));
        double *d_wt = nullptr, *d_wp = nullptr;
        rc = dev_upload(&d_wt, t.wt);
        if (!rc) rc = dev_upload(&d_wp, t.wp);
        if (!rc)
            rc = launch_weight_grid(c->d_vgrid, op.wgrid, op.wprec, op.qmode, d_wt, d_wp, c->P, na, nb, npts_max,
                                    c->stream);
        const hipError_t ew = hipStreamSynchronize(c->stream);
        dev_free(d_wt);
        dev_free(d_wp);
        if (!rc && ew != hipSuccess) rc = hip_fail(ew, "weight grid", __FILE__, __LINE__);
        if (rc) return rc;
        std::vector<double> qL((size_t)nk * nah), qS((size_t)op.nm * op.nphi);
        std::vector<int> mpl(nk);
        for (int r = 0; r < nk; r++) {
            const int n = c->carried[r];
            for (int a = 0; a < nah; a++) qL[(size_t)r * nah + a] = t.L[(size_t)n * na + a];
            mpl[r] = t.m[n] + lmax;
        }
        for (int mi = 0; mi < op.nm; mi++)
            for (int b = 0; b < nb; b++) {
                qS[(size_t)mi * op.nphi + b] = t.S1[(size_t)mi * nb + b];
                qS[(size_t)mi * op.nphi + nb + b] = t.S2[(size_t)mi * nb + b];
            }
        if ((rc = dev_upload(&op.qL, qL))) return rc;
        if ((rc = dev_upload(&op.qS, qS))) return rc;
        if ((rc = dev_upload(&op.mplane, mpl))) return rc;
        if ((rc = quad_grid(op, c->device, &op.q_grid))) return rc;
        op.nalpha = op.q_grid;
        op.npartial = std::max<int>(op.npartial, (int)op.nalpha + 1);
        dev_free(op.partial);
        dev_free(op.partial_b);
        dev_free(op.partial_c);
        if ((rc = dev_alloc(&op.partial, (size_t)op.npartial))) return rc;
        if ((rc = dev_alloc(&op.partial_b, (size_t)op.npartial))) return rc;
        if ((rc = dev_alloc(&op.partial_c, (size_t)op.npartial))) return rc;
        c->op64.partial = op.partial;
        c->op64.partial_b = op.partial_b;
        c->op64.partial_c = op.partial_c;
        c->op64.npartial = op.npartial;
    }
    BSL_HIP(hipStreamSynchronize(c->stream));
    dev_free(c->d_vgrid);
    c->d_vgrid = nullptr;

    // ---- device: Krylov vectors and scalars ----
    if ((rc = dev_alloc(&c->d_X, (size_t)c->n_kry))) return rc;
    if ((rc = dev_alloc(&c->d_R, (size_t)c->n_kry))) return rc;
    if ((rc = dev_alloc(&c->d_scal, 1))) return rc;
    BSL_HIP(hipMemset(c->d_scal, 0, sizeof(Scal)));
    if ((rc = autotune(c.get()))) return rc;
    *out = c.release();
    return BSL_OK;
}

extern "C" int64_t bsl_basis_size(const bsl_ctx *c) { return c ? c->n_full : -1; }
extern "C" int64_t bsl_krylov_size(const bsl_ctx *c) { return c ? c->n_kry : -1; }

namespace {
// copy the carried planes of a reference-layout host vector into a device vector
int upload_carried(bsl_ctx *c, const double *host, double *dev)
{
    for (size_t k = 0; k < c->carried.size(); k++)
        BSL_HIP(hipMemcpyAsync(dev + k * c->P, host + (size_t)c->carried[k] * c->P, sizeof(double) * c->P,
                               hipMemcpyHostToDevice, c->stream));
    return BSL_OK;
}

int download_carried(bsl_ctx *c, const double *dev, double *host)
{
    for (size_t k = 0; k < c->carried.size(); k++)
        BSL_HIP(hipMemcpyAsync(host + (size_t)c->carried[k] * c->P, dev + k * c->P, sizeof(double) * c->P,
                               hipMemcpyDeviceToHost, c->stream));
    return BSL_OK;
}

int ensure_tmp(bsl_ctx *c)
{
    if (!c->d_T) {
        int rc = dev_alloc(&c->d_T, (size_t)2 * c->n_kry);
        if (rc) return rc;
    }
    return BSL_OK;
}

int ensure_coeffs(bsl_ctx *c, int niter)
{
    if (c->niter_cap >= niter) return BSL_OK;
    dev_free(c->d_alpha);
    dev_free(c->d_beta);
    dev_free(c->d_beta2);
    c->d_alpha = c->d_beta = c->d_beta2 = nullptr;
    int rc;
    if ((rc = dev_alloc(&c->d_alpha, (size_t)niter + 1))) return rc;
    if ((rc = dev_alloc(&c->d_beta, (size_t)niter + 1))) return rc;
    if ((rc = dev_alloc(&c->d_beta2, (size_t)niter + 1))) return rc;
    c->niter_cap = niter;
    return BSL_OK;
}

// host Hv on the context operator `op`: uec += H vec
int hv_host_with(bsl_ctx *c, const DevOperator &op, const double *vec, double *uec)
{
    int rc = ensure_tmp(c);
    if (rc) return rc;
    double *dX = c->d_T, *dR = c->d_T + c->n_kry;
    Scal *sc = c->d_scal;
    if ((rc = upload_carried(c, vec, dX))) return rc;
    if ((rc = upload_carried(c, uec, dR))) return rc;
    // keep the Lanczos scalars: stash them and restore afterwards
    Scal saved;
    BSL_HIP(hipMemcpyAsync(&saved, sc, sizeof(Scal), hipMemcpyDeviceToHost, c->stream));
    if ((rc = launch_init_scal(sc, 1.0, 1.0, c->stream))) return rc;
    if ((rc = launch_hv(op, dX, dR, sc, c->stream, nullptr))) return rc;
    if ((rc = download_carried(c, dR, uec))) return rc;
    BSL_HIP(hipStreamSynchronize(c->stream));
    BSL_HIP(hipMemcpyAsync(sc, &saved, sizeof(Scal), hipMemcpyHostToDevice, c->stream));
    BSL_HIP(hipStreamSynchronize(c->stream));
    return BSL_OK;
}

int start_lanczos(bsl_ctx *c, const double *v0, int start, uint64_t seed)
{
    int rc;
    if (v0) {
        if ((rc = upload_carried(c, v0, c->d_X))) return rc;
        if ((rc = launch_init_scal(c->d_scal, 1.0, 0.0, c->stream))) return rc;
    } else {
        if (start != BSL_START_UNIFORM && start != BSL_START_RANDOM) return fail(BSL_E_ARG, "unknown start vector");
        rc = launch_start_vector(c->d_X, c->n_kry, c->P, c->op.plane_ref, start, seed, c->d_scal, c->op.partial,
                                 c->op.npartial, c->stream);
        if (rc) return rc;
    }
    return launch_zero(c->d_R, c->n_kry, c->stream);
}

// one Lanczos step j (1-based) on the current X/R; swaps the buffers
int lanczos_step(bsl_ctx *c, const DevOperator &op, int j, hipEvent_t *ev)
{
    (void)j;  // the device step counter indexes the coefficient outputs
    int rc = launch_step(op, c->d_X, c->d_R, c->d_scal, c->d_alpha, c->d_beta, c->d_beta2, c->stream, ev);
    if (rc) return rc;
    std::swap(c->d_X, c->d_R);
    return BSL_OK;
}
}  // namespace

namespace bsl {
int ctx_reserve_steps(bsl_ctx *c, int niter) { return ensure_coeffs(c, niter); }
int ctx_start(bsl_ctx *c, const double *v0, int start, uint64_t seed) { return start_lanczos(c, v0, start, seed); }
int ctx_step(bsl_ctx *c) { return lanczos_step(c, c->op, 0, nullptr); }
}  // namespace bsl

extern "C" int bsl_hv(bsl_ctx *c, const double *vec, double *uec)
{
    if (!c || !vec || !uec) return fail(BSL_E_ARG, "null argument");
    BSL_HIP(hipSetDevice(c->device));
    return hv_host_with(c, c->op, vec, uec);
}

extern "C" int bsl_hv_device(bsl_ctx *c, const double *d_vec, double *d_uec)
{
    if (!c || !d_vec || !d_uec) return fail(BSL_E_ARG, "null argument");
    BSL_HIP(hipSetDevice(c->device));
    int rc = ensure_tmp(c);
    if (rc) return rc;
    double *dX = c->d_T, *dR = c->d_T + c->n_kry;
    for (size_t k = 0; k < c->carried.size(); k++) {
        BSL_HIP(hipMemcpyAsync(dX + k * c->P, d_vec + (size_t)c->carried[k] * c->P, sizeof(double) * c->P,
                               hipMemcpyDeviceToDevice, c->stream));
        BSL_HIP(hipMemcpyAsync(dR + k * c->P, d_uec + (size_t)c->carried[k] * c->P, sizeof(double) * c->P,
                               hipMemcpyDeviceToDevice, c->stream));
    }
    Scal saved;
    BSL_HIP(hipMemcpyAsync(&saved, c->d_scal, sizeof(Scal), hipMemcpyDeviceToHost, c->stream));
    if ((rc = launch_init_scal(c->d_scal, 1.0, 1.0, c->stream))) return rc;
    if ((rc = launch_hv(c->op, dX, dR, c->d_scal, c->stream, nullptr))) return rc;
    for (size_t k = 0; k < c->carried.size(); k++)
        BSL_HIP(hipMemcpyAsync(d_uec + (size_t)c->carried[k] * c->P, dR + k * c->P, sizeof(double) * c->P,
                               hipMemcpyDeviceToDevice, c->stream));
    BSL_HIP(hipStreamSynchronize(c->stream));
    BSL_HIP(hipMemcpyAsync(c->d_scal, &saved, sizeof(Scal), hipMemcpyHostToDevice, c->stream));
    BSL_HIP(hipStreamSynchronize(c->stream));
    return BSL_OK;
}

extern "C" int bsl_lanczos(bsl_ctx *c, int32_t niter, const double *v0, int32_t start, uint64_t seed,
                           double *alpha, double *beta, double *beta2)
{
    if (!c || niter <= 0 || !alpha || !beta || !beta2) return fail(BSL_E_ARG, "bad arguments");
    BSL_HIP(hipSetDevice(c->device));
    int rc = ensure_coeffs(c, niter);
    if (rc) return rc;
    if ((rc = start_lanczos(c, v0, start, seed))) return rc;
    for (int j = 1; j <= niter; j++)
        if ((rc = lanczos_step(c, c->op, j, nullptr))) return rc;
    BSL_HIP(hipMemcpyAsync(alpha, c->d_alpha, sizeof(double) * niter, hipMemcpyDeviceToHost, c->stream));
    BSL_HIP(hipMemcpyAsync(beta + 1, c->d_beta + 1, sizeof(double) * niter, hipMemcpyDeviceToHost, c->stream));
    BSL_HIP(hipMemcpyAsync(beta2 + 1, c->d_beta2 + 1, sizeof(double) * niter, hipMemcpyDeviceToHost, c->stream));
    BSL_HIP(hipStreamSynchronize(c->stream));
    beta[0] = 0.0;
    beta2[0] = 0.0;
    return check_breakdown(c->d_scal, c->stream);
}

extern "C" int bsl_lancbis_tol(int32_t niter, const double *alpha, const double *beta, const double *beta2,
                               double emin, double emax, double mult_tol, double *eval, double *err,
                               int32_t *ngood)
{
    if (!alpha || !beta || !beta2 || !eval || !err || !ngood) return fail(BSL_E_ARG, "null argument");
    if (!(mult_tol >= 0.0)) return fail(BSL_E_ARG, "lancbis: mult_tol must be >= 0");
    int ng = 0;
    int rc = cw_lancbis(niter, alpha, beta, beta2, emin, emax, mult_tol, eval, err, &ng);
    *ngood = ng;
    return rc;
}

extern "C" int bsl_lancbis(int32_t niter, const double *alpha, const double *beta, const double *beta2,
                           double emin, double emax, double *eval, double *err, int32_t *ngood)
{
    return bsl_lancbis_tol(niter, alpha, beta, beta2, emin, emax, 0.0, eval, err, ngood);
}

extern "C" int bsl_lancbis_device_tol(int32_t niter, const double *alpha, const double *beta, const double *beta2,
                                      double emin, double emax, double mult_tol, double *eval, double *err,
                                      int32_t *ngood)
{
    (void)beta2;
    if (!alpha || !beta || !eval || !err || !ngood) return fail(BSL_E_ARG, "null argument");
    if (!(mult_tol >= 0.0)) return fail(BSL_E_ARG, "lancbis: mult_tol must be >= 0");
    hipStream_t s = nullptr;
    BSL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int ng = 0;
    int rc = cw_lancbis_device(niter, alpha, beta, emin, emax, mult_tol, s, eval, err, &ng);
    (void)hipStreamDestroy(s);
    *ngood = ng;
    return rc;
}

extern "C" int bsl_lancbis_device(int32_t niter, const double *alpha, const double *beta, const double *beta2,
                                  double emin, double emax, double *eval, double *err, int32_t *ngood)
{
    return bsl_lancbis_device_tol(niter, alpha, beta, beta2, emin, emax, 0.0, eval, err, ngood);
}

extern "C" int bsl_trivec_tol(int32_t niter, const double *alpha, const double *beta, const double *beta2,
                              const double *eval, int32_t ngood, double mult_tol, double *evtr, int32_t *mamin)
{
    if (!alpha || !beta || !beta2 || (ngood > 0 && (!eval || !evtr || !mamin))) return fail(BSL_E_ARG, "null argument");
    if (!(mult_tol >= 0.0)) return fail(BSL_E_ARG, "trivec: mult_tol must be >= 0");
    std::vector<int> mm(std::max(ngood, 1));
    int rc = cw_trivec(niter, alpha, beta, beta2, eval, ngood, mult_tol, evtr, mm.data());
    for (int i = 0; i < ngood; i++) mamin[i] = mm[i];
    return rc;
}

extern "C" int bsl_trivec(int32_t niter, const double *alpha, const double *beta, const double *beta2,
                          const double *eval, int32_t ngood, double *evtr, int32_t *mamin)
{
    return bsl_trivec_tol(niter, alpha, beta, beta2, eval, ngood, 0.0, evtr, mamin);
}

namespace {
// Grow-only device workspace: *p holds at least `count` doubles afterwards
int ensure_workspace(double **p, size_t *cap, size_t count)
{
    if (*cap >= count) return BSL_OK;
    dev_free(*p);
    *p = nullptr;
    *cap = 0;
    int rc = dev_alloc(p, count);
    if (!rc) *cap = count;
    return rc;
}

// Loop 2 (lanczosEigenValuesVectors.cpp:371-469): recomputes the Krylov
// vectors from the same start and accumulates the ngood Ritz vectors into
// c->d_ritz ([ngood][n_kry], device order)
int ritz_run(bsl_ctx *c, int niter, const double *v0, int start, uint64_t seed, const double *evtr, int ngood)
{
    // coefficient of v_j for each vector, zero once the cumulative norm
    // reaches 1 - 1e-16 (lanczosEigenValuesVectors.cpp:404-419)
    std::vector<double> coef((size_t)niter * ngood);
    int jlast = 0;  // the last step with a non-zero coefficient: later steps add nothing
    for (int g = 0; g < ngood; g++) {
        double cum = 0.0;
        for (int j = 1; j <= niter; j++) {
            const double e = evtr[(j - 1) + (size_t)niter * g];
            cum += std::pow(e, 2.);
            coef[(size_t)(j - 1) * ngood + g] = (cum < (1. - 1.e-16)) ? e : 0.0;
            if (coef[(size_t)(j - 1) * ngood + g] != 0.0) jlast = std::max(jlast, j);
        }
    }
    int rc = ensure_workspace(&c->d_rcoef, &c->rcoef_cap, coef.size());
    if (!rc) rc = ensure_workspace(&c->d_ritz, &c->ritz_cap, (size_t)c->n_kry * ngood);
    if (rc) return rc;
    BSL_HIP(hipMemcpyAsync(c->d_rcoef, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice, c->stream));
    rc = launch_zero(c->d_ritz, c->n_kry * ngood, c->stream);
    if (!rc) rc = ensure_coeffs(c, niter);
    if (!rc) rc = start_lanczos(c, v0, start, seed);
    // Step j writes its Lanczos vector v_j over X, which the swap after the
    // step makes R.  Several Ritz vectors take the v_j of a window of
    // kRitzWindow steps, copied aside, in one read and write each; a single
    // one takes each v_j from R directly.  Either way ritz += coef_j v_j in
    // step order, with exactly the v_j the recurrence used.
    const int win = ngood > 1 ? std::min(kRitzWindow, std::max(jlast, 1)) : 1;
    if (!rc && win > 1) rc = ensure_workspace(&c->d_rwin, &c->rwin_cap, (size_t)c->n_kry * win);
    int j0 = 1;  // first step of the window
    for (int j = 1; j <= niter && !rc; j++) {
        rc = lanczos_step(c, c->op, j, nullptr);
        if (rc || j > jlast) continue;
        const double *wcoef = c->d_rcoef + (size_t)(j0 - 1) * ngood;
        if (win == 1) {
            rc = launch_ritz_flush(c->d_ritz, c->d_R, wcoef, 1, c->n_kry, ngood, c->stream);
            j0 = j + 1;
        } else {
            const hipError_t e = hipMemcpyAsync(c->d_rwin + (size_t)(j - j0) * c->n_kry, c->d_R,
                                                sizeof(double) * c->n_kry, hipMemcpyDeviceToDevice, c->stream);
            if (e != hipSuccess) rc = hip_fail(e, "stage Lanczos vector", __FILE__, __LINE__);
            if (!rc && (j - j0 + 1 == win || j == jlast)) {
                rc = launch_ritz_flush(c->d_ritz, c->d_rwin, wcoef, j - j0 + 1, c->n_kry, ngood, c->stream);
                j0 = j + 1;
            }
        }
    }
    // the host coefficient list must outlive the copy
    const hipError_t e = hipStreamSynchronize(c->stream);
    if (!rc && e != hipSuccess) rc = hip_fail(e, "Lanczos loop 2", __FILE__, __LINE__);
    if (!rc) rc = check_breakdown(c->d_scal, c->stream);
    return rc;
}

// sum of the first 2 g partial pairs in a fixed order
int sum_pairs(bsl_ctx *c, int g, double *s0, double *s1)
{
    std::vector<double> h((size_t)2 * g);
    BSL_HIP(hipMemcpyAsync(h.data(), c->op.partial, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    BSL_HIP(hipStreamSynchronize(c->stream));
    double a = 0.0, b = 0.0;
    for (int i = 0; i < g; i++) {
        a += h[2 * i];
        b += h[2 * i + 1];
    }
    *s0 = a;
    *s1 = b;
    return BSL_OK;
}

// FP64 Rayleigh quotient <y|H|y>/<y|y> and residual |Hy - e y|/|y| of the
// device vector y (n_kry, device order) with the context's FP64 operator,
// whatever its precision; Hy goes to the second half of d_T
int rayleigh_dev(bsl_ctx *c, const double *y, double *energy, double *resid, double *norm2 = nullptr)
{
    int rc = ensure_tmp(c);
    if (rc) return rc;
    double *dX = c->d_T, *dR = c->d_T + c->n_kry;
    if (y != dX) BSL_HIP(hipMemcpyAsync(dX, y, sizeof(double) * c->n_kry, hipMemcpyDeviceToDevice, c->stream));
    if ((rc = launch_zero(dR, c->n_kry, c->stream))) return rc;
    if ((rc = launch_init_scal(c->d_scal, 1.0, 1.0, c->stream))) return rc;
    if ((rc = launch_hv(c->op64, dX, dR, c->d_scal, c->stream, nullptr))) return rc;
    int g = 0;
    double vv = 0.0, vh = 0.0, r2 = 0.0, unused = 0.0;
    if ((rc = launch_rayleigh_dots(dX, dR, c->n_kry, 0.0, 0, c->op.partial, c->op.npartial, &g, c->stream)))
        return rc;
    if ((rc = sum_pairs(c, g, &vv, &vh))) return rc;
    if (!(vv > 0.0)) return fail(BSL_E_ARG, "Rayleigh quotient of a zero vector");
    const double e = vh / vv;
    if ((rc = launch_rayleigh_dots(dX, dR, c->n_kry, e, 1, c->op.partial, c->op.npartial, &g, c->stream)))
        return rc;
    if ((rc = sum_pairs(c, g, &r2, &unused))) return rc;
    *energy = e;
    *resid = std::sqrt(r2 / vv);
    if (norm2) *norm2 = vv;
    return BSL_OK;
}

// keeps the context's recurrence scalars across a refinement
struct ScalGuard {
    bsl_ctx *c;
    Scal saved{};
    bool ok = false;
    explicit ScalGuard(bsl_ctx *cc) : c(cc)
    {
        ok = hipMemcpyAsync(&saved, c->d_scal, sizeof(Scal), hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
             hipStreamSynchronize(c->stream) == hipSuccess;
    }
    ~ScalGuard()
    {
        if (!ok) return;
        (void)hipMemcpyAsync(c->d_scal, &saved, sizeof(Scal), hipMemcpyHostToDevice, c->stream);
        (void)hipStreamSynchronize(c->stream);
    }
};
}  // namespace

extern "C" int bsl_ritz(bsl_ctx *c, int32_t niter, const double *v0, int32_t start, uint64_t seed,
                        const double *evtr, int32_t ngood, double *ritz)
{
    if (!c || niter <= 0 || ngood < 0 || (ngood > 0 && (!evtr || !ritz))) return fail(BSL_E_ARG, "bad arguments");
    if (ngood == 0) return BSL_OK;
    BSL_HIP(hipSetDevice(c->device));
    c->ritz_vv.clear();
    int rc = ritz_run(c, niter, v0, start, seed, evtr, ngood);
    if (rc) return rc;
    std::memset(ritz, 0, sizeof(double) * (size_t)c->n_full * ngood);
    for (int g = 0; g < ngood && !rc; g++)
        rc = download_carried(c, c->d_ritz + (size_t)g * c->n_kry, ritz + (size_t)g * c->n_full);
    if (!rc) BSL_HIP(hipStreamSynchronize(c->stream));
    return rc;
}

extern "C" int bsl_ritz_refine(bsl_ctx *c, int32_t niter, const double *v0, int32_t start, uint64_t seed,
                               const double *evtr, int32_t ngood, double *ritz, double *energy, double *resid)
{
    if (!c || niter <= 0 || ngood < 0 || (ngood > 0 && (!evtr || !energy || !resid)))
        return fail(BSL_E_ARG, "bad arguments");
    if (ngood == 0) return BSL_OK;
    BSL_HIP(hipSetDevice(c->device));
    c->ritz_vv.clear();
    int rc = ritz_run(c, niter, v0, start, seed, evtr, ngood);
    std::vector<double> vv(ngood);
    for (int g = 0; g < ngood && !rc; g++)
        rc = rayleigh_dev(c, c->d_ritz + (size_t)g * c->n_kry, energy + g, resid + g, &vv[g]);
    if (!rc) c->ritz_vv = vv;
    if (!rc && ritz) {
        std::memset(ritz, 0, sizeof(double) * (size_t)c->n_full * ngood);
        for (int g = 0; g < ngood && !rc; g++)
            rc = download_carried(c, c->d_ritz + (size_t)g * c->n_kry, ritz + (size_t)g * c->n_full);
        if (!rc) BSL_HIP(hipStreamSynchronize(c->stream));
    }
    return rc;
}

extern "C" int bsl_rayleigh(bsl_ctx *c, int32_t count, const double *y, double *energy, double *resid)
{
    if (!c || count < 0 || (count > 0 && (!y || !energy || !resid))) return fail(BSL_E_ARG, "bad arguments");
    if (count == 0) return BSL_OK;
    BSL_HIP(hipSetDevice(c->device));
    int rc = ensure_tmp(c);
    if (rc) return rc;
    ScalGuard keep(c);
    for (int g = 0; g < count && !rc; g++) {
        rc = upload_carried(c, y + (size_t)g * c->n_full, c->d_T);
        if (!rc) rc = rayleigh_dev(c, c->d_T, energy + g, resid + g);
    }
    return rc;
}

extern "C" int bsl_water_sites(int32_t model, int32_t nmol, const double *cm_ea, double *pos, double *q,
                               int32_t *lj)
{
    if (model != 0 && model != 1) return fail(BSL_E_ARG, "water model must be 0 (SPC/E) or 1 (TIP4P)");
    if (nmol < 0 || (nmol > 0 && (!cm_ea || !pos || !q || !lj))) return fail(BSL_E_ARG, "bad arguments");
    // body-frame sites (nm) and charges: SPC/E (SPCE_AH2_EngRoutines.h:17-32)
    // and TIP4P with its massless M site (TIP4P_AH2_EngRoutines.h:16-35)
    static const double spce[3][3] = {{0., 0., 0.006461408220},
                                      {-8.164965810E-2, 0., -0.05127361870},
                                      {8.164965810E-2, 0., -0.05127361870}};
    static const double spce_q[3] = {-0.8476, 0.4238, 0.4238};
    static const double tip4p[4][3] = {{0., 0., 0.00655538},
                                       {-0.075695, 0., -0.0520326},
                                       {0.075695, 0., -0.0520326},
                                       {0., 0., -0.00844462}};
    static const double tip4p_q[4] = {0.0, 0.520, 0.520, -1.040};
    const int ns = model == 0 ? 3 : 4;
    for (int i = 0; i < nmol; i++) {
        const double *r = cm_ea + 6 * i;
        const double cp = std::cos(r[1]), sp = std::sin(r[1]);
        const double ct = std::cos(r[3]), st = std::sin(r[3]);
        const double cx = std::cos(r[5]), sx = std::sin(r[5]);
        const double rot[3][3] = {{(cp * ct * cx) - (sp * sx), (-1 * cp * ct * sx) - (sp * cx), cp * st},
                                  {(sp * ct * cx) + (cp * sx), (-1 * sp * ct * sx) + (cp * cx), sp * st},
                                  {-1 * st * cx, st * sx, ct}};
        const double cm[3] = {r[0], r[2], r[4]};
        for (int s = 0; s < ns; s++) {
            const double *body = model == 0 ? spce[s] : tip4p[s];
            for (int d = 0; d < 3; d++) {
                double acc = 0.;
                for (int e = 0; e < 3; e++) acc += rot[d][e] * body[e];
                pos[((size_t)ns * i + s) * 3 + d] = acc + 0.1 * cm[d];
            }
            q[(size_t)ns * i + s] = model == 0 ? spce_q[s] : tip4p_q[s];
            lj[(size_t)ns * i + s] = s == 0 ? 1 : 0;
        }
    }
    return BSL_OK;
}

extern "C" int bsl_get_vblocks(bsl_ctx *c, double *out)
{
    if (!c || !out) return fail(BSL_E_ARG, "null argument");
    if (c->p.basis != BSL_BASIS_FULL) return fail(BSL_E_STATE, "V blocks are reported for FULL-basis contexts");
    BSL_HIP(hipSetDevice(c->device));
    const DevOperator &op = c->op64;
    const int64_t ntile = vpk_ntile(c->P);
    const bool harm = op.vform == BSL_VFORM_HARMONICS;
    const int64_t ncol = harm ? op.ncoef : op.npair;
    const int64_t npair0 = op.nblk > 1 ? op.blk[1].pair_offset : op.npair;
    std::vector<double> vpk((size_t)ncol * ntile * kVTile);
    BSL_HIP(hipMemcpy(vpk.data(), op.vpk, sizeof(double) * vpk.size(), hipMemcpyDeviceToHost));
    const int nn = c->nn, nx = op.nx, ny = op.ny, nz = op.nz;
    std::memset(out, 0, sizeof(double) * (size_t)c->P * nn * nn);
    auto put = [&](int n, int np, const std::vector<double> &v /* [P], device point order */) {
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    const int64_t p = ((int64_t)k * ny + j) * nx + i;
                    const int64_t ref = (((int64_t)i * ny + j) * nz + k);
                    out[(ref * nn + n) * nn + np] = v[p];
                    out[(ref * nn + np) * nn + n] = v[p];
                }
    };
    std::vector<double> v((size_t)c->P);
    for (int g = 0; g < op.nblk; g++) {
        const Block &B = op.blk[g];
        if (harm) {
            // V[i][j] = sum over the pair's terms of G h[c], the kernel's sum
            std::vector<double> gv;
            double excluded = 0.0;
            gaunt_values(B.lb, B.par, gv, &excluded);
            size_t t = 0;
            for (int i = 0; i < B.size; i++)
                for (int j = i; j < B.size; j++) {
                    std::fill(v.begin(), v.end(), 0.0);
                    for (int cc = 0; cc < gaunt::ncoef(B.lb); cc++) {
                        if (!gaunt::coupled(gaunt::block_l(B.lb, B.par, i), gaunt::block_m(B.lb, B.par, i),
                                            gaunt::block_l(B.lb, B.par, j), gaunt::block_m(B.lb, B.par, j),
                                            gaunt::coef_L(cc), gaunt::coef_M(cc)))
                            continue;
                        const double gt = gv[t++];
                        for (int64_t p = 0; p < c->P; p++)
                            v[p] += gt * vpk[(size_t)((p / kVTile) * ncol + cc) * kVTile + p % kVTile];
                    }
                    put(c->carried[B.plane0 + i], c->carried[B.plane0 + j], v);
                }
        } else {
            int64_t pr = B.pair_offset;
            for (const auto &rc : block_pair_order(B.size)) {
                for (int64_t p = 0; p < c->P; p++) v[p] = vpk[vpk_index(pr, p, npair0, op.npair, ntile)];
                put(c->carried[B.plane0 + rc.first], c->carried[B.plane0 + rc.second], v);
                pr++;
            }
        }
    }
    return BSL_OK;
}

extern "C" int bsl_gaunt_table(int32_t lb, int32_t par, int32_t cap, int32_t *ijc, double *g)
{
    if (lb < 0 || lb > 6 || (par != 0 && par != 1) || lb % 2 != par)
        return -fail(BSL_E_ARG, "a parity block needs 0 <= lb <= 6 with lb of parity par");
    std::vector<double> gv;
    double excluded = 0.0;
    gaunt_values(lb, par, gv, &excluded);
    const int nt = (int)gv.size();
    if (!ijc || !g || cap < nt) return nt;
    int t = 0;
    for (int i = 0; i < gaunt::block_size(lb, par); i++)
        for (int j = i; j < gaunt::block_size(lb, par); j++)
            for (int c = 0; c < gaunt::ncoef(lb); c++)
                if (gaunt::coupled(gaunt::block_l(lb, par, i), gaunt::block_m(lb, par, i), gaunt::block_l(lb, par, j),
                                   gaunt::block_m(lb, par, j), gaunt::coef_L(c), gaunt::coef_M(c))) {
                    ijc[3 * t + 0] = i;
                    ijc[3 * t + 1] = j;
                    ijc[3 * t + 2] = c;
                    g[t] = gv[t];
                    t++;
                }
    return nt;
}

extern "C" int bsl_get_fusion(const bsl_ctx *c)
{
    if (!c) return -fail(BSL_E_ARG, "null argument");
    return c->op.fused;
}

extern "C" int bsl_set_fusion(bsl_ctx *c, int32_t on)
{
    if (!c) return fail(BSL_E_ARG, "null argument");
    if (on != 0 && on != 1) return fail(BSL_E_ARG, "fusion is 0 or 1");
    if (on && !zpot_fits(c->op))
        return fail(BSL_E_ARG, "no fused z-kinetic and potential kernel for this operator (it needs the lmax-5 "
                               "harmonic-form blocks, axes of at most 64 points and FP64 or FP32 throughout)");
    BSL_HIP(hipSetDevice(c->device));
    BSL_HIP(hipStreamSynchronize(c->stream));
    int rc = set_fused(c->op, on);
    if (!rc && c->op64.vpk == c->op.vpk) {
        c->op64.fused = c->op.fused;
        c->op64.nalpha = c->op.nalpha;
    } else if (!rc) {
        rc = set_fused(c->op64, on);
    }
    return rc;
}

extern "C" int bsl_get_potential_form(const bsl_ctx *c)
{
    if (!c) return -fail(BSL_E_ARG, "null argument");
    return c->op64.vform;
}

extern "C" int bsl_set_potential_form(bsl_ctx *c, int32_t form)
{
    if (!c) return fail(BSL_E_ARG, "null argument");
    if (form != BSL_VFORM_BLOCKS && form != BSL_VFORM_HARMONICS) return fail(BSL_E_ARG, "potential form not recognised");
    if (form == c->op64.vform) return BSL_OK;
    if (form == BSL_VFORM_HARMONICS && !harmonic_fits(c, c->op64))
        return fail(BSL_E_ARG, "the harmonic form needs parity blocks of l <= " + std::to_string(gaunt::kMaxLb));
    BSL_HIP(hipSetDevice(c->device));
    BSL_HIP(hipStreamSynchronize(c->stream));
    double *d_vg = nullptr;
    int rc = dev_alloc(&d_vg, (size_t)c->P * c->Q);
    if (!rc) rc = make_vgrid(c, d_vg);
    if (!rc) rc = build_potentials(c, form, d_vg);
    dev_free(d_vg);
    if (!rc) rc = autotune(c);
    return rc;
}

extern "C" int bsl_get_vgrid(bsl_ctx *c, double *v1, double *v2)
{
    if (!c || !v1 || !v2) return fail(BSL_E_ARG, "null argument");
    BSL_HIP(hipSetDevice(c->device));
    // the grid is regenerated on demand (it is not kept after preparation)
    const bsl_params &p = c->p;
    double *d_vg = nullptr;
    int rc = dev_alloc(&d_vg, (size_t)c->P * c->Q);
    if (rc) return rc;
    const Tables &t = c->tab;
    const int na = t.na, nb = t.nb;
    rc = make_vgrid(c, d_vg);
    std::vector<double> vg;
    if (!rc) {
        vg.resize((size_t)c->P * c->Q);
        hipError_t e = hipMemcpyAsync(vg.data(), d_vg, sizeof(double) * vg.size(), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(e, "download potential grid", __FILE__, __LINE__);
    }
    dev_free(d_vg);
    if (rc) return rc;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const int64_t pp = ((int64_t)k * ny + j) * nx + i;
                const int64_t ref = ((int64_t)i * ny + j) * nz + k;
                for (int q = 0; q < na * nb; q++) {
                    v1[ref * na * nb + q] = vg[(size_t)pp * c->Q + q];
                    v2[ref * na * nb + q] = vg[(size_t)pp * c->Q + na * nb + q];
                }
            }
    return BSL_OK;
}

namespace {
void free_hf_setup(bsl_ctx *c)
{
    for (void *p : {(void *)c->d_unclamped, (void *)c->d_pairs, (void *)c->d_pair_first, (void *)c->d_hfY,
                    (void *)c->d_hfwq, (void *)c->d_hfoff, (void *)c->d_hfsite, (void *)c->d_hflj})
        dev_free(p);
    c->d_unclamped = nullptr;
    c->d_pairs = nullptr;
    c->d_pair_first = nullptr;
    c->d_hfY = c->d_hfwq = c->d_hfoff = c->d_hfsite = nullptr;
    c->d_hflj = nullptr;
    c->hf_ne = 0;
}

// a workspace buffer of at least `need` doubles (contents not kept when it grows)
int grow(double *&buf, size_t &cap, size_t need)
{
    if (need <= cap) return BSL_OK;
    dev_free(buf);
    buf = nullptr;
    cap = 0;
    const int rc = dev_alloc(&buf, need);
    if (!rc) cap = need;
    return rc;
}

// What the gradients need of a cage context, fixed by its potential and
// built once: which (point, node) pairs are below the ceiling clamp (one bit
// each, from a pass of the grid kernel that keeps only the bits), the list
// of kept node pairs per group of kHfPoints points that the gradient kernels
// walk, and the node tables (carried harmonics and weights at the nodes, H
// offsets, sites).
int ensure_hf_setup(bsl_ctx *c)
{
    if (c->d_unclamped) return BSL_OK;
    const bsl_params &p = c->p;
    const Tables &t = c->tab;
    const int na = t.na, nb = t.nb, nk = (int)c->carried.size(), Q = c->Q;
    const int ng = hf_grad_groups(c->P, kHfPoints);
    int *d_count = nullptr;
    std::vector<int> count(ng);
    std::vector<int64_t> first(ng + 1, 0);
    int rc = dev_alloc(&c->d_unclamped, ((size_t)c->P * Q + 63) / 64);
    if (!rc) rc = cage_vgrid(c, nullptr, c->d_unclamped);
    if (!rc) rc = dev_alloc(&d_count, (size_t)ng);
    if (!rc) rc = launch_pair_count(c->d_unclamped, c->P, Q, kHfPoints, d_count, c->stream);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(count.data(), d_count, sizeof(int) * ng, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(e, "pair counts", __FILE__, __LINE__);
    }
    for (int g = 0; g < ng && !rc; g++) first[g + 1] = first[g] + count[g];
    if (!rc) rc = dev_upload(&c->d_pair_first, first);
    if (!rc) rc = dev_alloc(&c->d_pairs, (size_t)std::max<int64_t>(first[ng], 1));
    if (!rc) rc = launch_pair_fill(c->d_unclamped, c->P, Q, kHfPoints, c->d_pair_first, c->d_pairs, c->stream);
    if (!rc) c->hf_ne = first[ng];
    if (!rc) {
        // quadrature-node values of the carried harmonics and the node weights
        std::vector<double> Y((size_t)nk * Q), wq(Q);
        for (int a = 0; a < na; a++)
            for (int b = 0; b < nb; b++) {
                const size_t q0 = (size_t)a * nb + b, q1 = (size_t)na * nb + q0;
                wq[q0] = wq[q1] = t.wt[a] * t.wp[b];
                for (int r = 0; r < nk; r++) {
                    const int n = c->carried[r], mi = t.m[n] + p.lmax;
                    Y[(size_t)r * Q + q0] = t.L[(size_t)n * na + a] * t.S1[(size_t)mi * nb + b];
                    Y[(size_t)r * Q + q1] = t.L[(size_t)n * na + a] * t.S2[(size_t)mi * nb + b];
                }
            }
        std::vector<double> off, site4;
        std::vector<int> lj;
        cage_arrays(c, off, site4, lj);
        rc = dev_upload(&c->d_hfY, Y);
        if (!rc) rc = dev_upload(&c->d_hfwq, wq);
        if (!rc) rc = dev_upload(&c->d_hfoff, off);
        if (!rc) rc = dev_upload(&c->d_hfsite, site4);
        if (!rc) rc = dev_upload(&c->d_hflj, lj);
    }
    const hipError_t e = hipStreamSynchronize(c->stream);
    if (!rc && e != hipSuccess) rc = hip_fail(e, "clamp mask", __FILE__, __LINE__);
    dev_free(d_count);
    if (rc) free_hf_setup(c);
    return rc;
}

}  // namespace

extern "C" int bsl_clamp_stats(bsl_ctx *c, int64_t *unclamped, int64_t *total)
{
    if (!c || !unclamped || !total) return fail(BSL_E_ARG, "null argument");
    if (c->p.potential != BSL_POT_CAGE) return fail(BSL_E_STATE, "bsl_clamp_stats: cage potential only");
    BSL_HIP(hipSetDevice(c->device));
    const int rc = ensure_hf_setup(c);
    if (rc) return rc;
    const size_t n = (size_t)c->P * c->Q, words = (n + 63) / 64;
    std::vector<uint64_t> m(words);
    BSL_HIP(hipMemcpyAsync(m.data(), c->d_unclamped, sizeof(uint64_t) * words, hipMemcpyDeviceToHost, c->stream));
    BSL_HIP(hipStreamSynchronize(c->stream));
    int64_t cnt = 0;
    for (size_t w = 0; w < words; w++) {
        uint64_t x = m[w];
        if (w == words - 1 && n % 64) x &= (~0ull) >> (64 - n % 64);
        cnt += __builtin_popcountll(x);
    }
    *unclamped = cnt;
    *total = (int64_t)n;
    return BSL_OK;
}

namespace {
// The passes of bsl_hf_gradient_sets over vectors already in d_hfpsi with
// their weights in d_hfg (both filled by the caller after hf_prepare).
int hf_prepare(bsl_ctx *c, int nset, const int32_t *counts, int64_t *nvec)
{
    if (!c || nset <= 0 || !counts) return fail(BSL_E_ARG, "bsl_hf_gradient: bad arguments");
    *nvec = 0;
    for (int m = 0; m < nset; m++) {
        if (counts[m] <= 0) return fail(BSL_E_ARG, "bsl_hf_gradient: every set needs an eigenvector");
        *nvec += counts[m];
    }
    if (c->p.potential != BSL_POT_CAGE) return fail(BSL_E_STATE, "bsl_hf_gradient: cage potential only");
    BSL_HIP(hipSetDevice(c->device));
    const int nent = 4 * c->p.n_sites + 4, nwg = hf_grad_groups(c->P, kHfPoints);
    const int chunk = std::min(nset, kHfGradSets);  // sets of a gradient pass
    size_t *cap = c->hf_cap;
    int rc = ensure_hf_setup(c);
    if (!rc) rc = grow(c->d_hfpsi, cap[0], (size_t)*nvec * c->n_kry);
    if (!rc) rc = grow(c->d_hfg, cap[1], (size_t)*nvec);
    if (!rc) rc = grow(c->d_hfom, cap[2], (size_t)chunk * std::max<int64_t>(c->hf_ne, 1));
    if (!rc) rc = grow(c->d_hfwcm, cap[3], (size_t)chunk * c->P);
    if (!rc) rc = grow(c->d_hfpart, cap[4], (size_t)nwg * chunk * nent);
    if (!rc) rc = grow(c->d_hfout, cap[5], (size_t)nset * nent);
    return rc;
}

int hf_passes(bsl_ctx *c, int nset, const int32_t *counts, double *grad)
{
    const bsl_params &p = c->p;
    const int nk = (int)c->carried.size(), ns = p.n_sites, Q = c->Q;
    const int npb = kHfPoints, nent = 4 * ns + 4;
    const int64_t ne = c->hf_ne;
    // gradient passes of up to kHfGradSets sets; their weights come from
    // density launches of up to kHfDensitySets sets whose eigenvectors'
    // coefficients fit the density kernel's LDS (a set alone that does not
    // fit fails there), each writing its sets' slots of the pass
    const int vmax = std::max(4, (int)(65536 / (sizeof(double) * kHfPoints * nk)) & ~3);
    int rc = BSL_OK;
    int64_t v0 = 0;  // first eigenvector of the density launch
    for (int m0 = 0, np = 0; m0 < nset && !rc; m0 += np) {
        for (np = 0; np < kHfGradSets && m0 + np < nset && !rc;) {
            const int m1 = m0 + np;
            int nv = counts[m1], nm = 1;
            for (; nm < kHfDensitySets && np + nm < kHfGradSets && m1 + nm < nset && nv + counts[m1 + nm] <= vmax;
                 nm++)
                nv += counts[m1 + nm];
            rc = launch_hf_density(c->d_hfpsi + (size_t)v0 * c->n_kry, nm, counts + m1, c->d_hfg + v0, c->d_hfY,
                                   c->d_hfwq, c->d_unclamped, c->d_pairs, c->d_pair_first, ne,
                                   c->d_hfom + (size_t)np * ne, c->d_hfwcm + (size_t)np * c->P, c->P, nk, Q,
                                   c->stream);
            v0 += nv;
            np += nm;
        }
        if (!rc)
            rc = launch_hf_grad(c->d_hfom, ne, c->d_hfwcm, np, c->P, c->d_gx, c->d_gy, c->d_gz, p.nx, p.ny,
                                c->d_hfoff, c->d_hfsite, c->d_hflj, ns, p.q_h2_h, p.q_h2_cm, p.lj_a, p.lj_b, npb,
                                c->d_pairs, c->d_pair_first, c->d_hfpart, c->d_hfout + (size_t)m0 * nent, c->stream);
    }
    if (!rc) {
        hipError_t e = hipMemcpyAsync(grad, c->d_hfout, sizeof(double) * nset * nent, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(e, "download gradient", __FILE__, __LINE__);
    }
    (void)hipStreamSynchronize(c->stream);
    return rc;
}
}  // namespace

extern "C" int bsl_hf_gradient_sets(bsl_ctx *c, int32_t nset, const int32_t *counts, const double *psi,
                                    const double *g, double *grad)
{
    if (!psi || !g || !grad) return fail(BSL_E_ARG, "bsl_hf_gradient_sets: bad arguments");
    int64_t nvec = 0;
    int rc = hf_prepare(c, nset, counts, &nvec);
    for (int64_t k = 0; k < nvec && !rc; k++)
        rc = upload_carried(c, psi + (size_t)k * c->n_full, c->d_hfpsi + (size_t)k * c->n_kry);
    if (!rc) {
        const hipError_t e = hipMemcpyAsync(c->d_hfg, g, sizeof(double) * nvec, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) rc = hip_fail(e, "upload weights", __FILE__, __LINE__);
    }
    if (!rc) rc = hf_passes(c, nset, counts, grad);
    return rc;
}

extern "C" int bsl_hf_gradient_ritz(bsl_ctx *c, int32_t nset, const int32_t *counts, const int32_t *index,
                                    const double *g, double *grad)
{
    if (!index || !g || !grad) return fail(BSL_E_ARG, "bsl_hf_gradient_ritz: bad arguments");
    int64_t nvec = 0;
    int rc = hf_prepare(c, nset, counts, &nvec);
    if (rc) return rc;
    const int nr = (int)c->ritz_vv.size();
    if (nr == 0) return fail(BSL_E_STATE, "bsl_hf_gradient_ritz: no Ritz vectors of bsl_ritz_refine on the context");
    // unit vectors psi = y / |y|: weight g / |y|^2 on the kept vector y
    std::vector<double> gs(nvec);
    for (int64_t k = 0; k < nvec && !rc; k++) {
        if (index[k] < 0 || index[k] >= nr) return fail(BSL_E_ARG, "bsl_hf_gradient_ritz: Ritz vector index out of range");
        gs[k] = g[k] / c->ritz_vv[index[k]];
        BSL_HIP(hipMemcpyAsync(c->d_hfpsi + (size_t)k * c->n_kry, c->d_ritz + (size_t)index[k] * c->n_kry,
                               sizeof(double) * c->n_kry, hipMemcpyDeviceToDevice, c->stream));
    }
    BSL_HIP(hipMemcpyAsync(c->d_hfg, gs.data(), sizeof(double) * nvec, hipMemcpyHostToDevice, c->stream));
    return hf_passes(c, nset, counts, grad);
}

extern "C" int bsl_hf_gradient(bsl_ctx *c, int32_t nvec, const double *psi, const double *g, double *grad)
{
    if (!c || nvec <= 0 || !psi || !g || !grad) return fail(BSL_E_ARG, "bsl_hf_gradient: bad arguments");
    return bsl_hf_gradient_sets(c, 1, &nvec, psi, g, grad);
}

extern "C" int bsl_step_bytes(const bsl_ctx *c, double *out)
{
    if (!c || !out) return fail(BSL_E_ARG, "null argument");
    const DevOperator &op = c->op;
    const double vec = 8.0 * (double)c->n_kry;
    double pot_bytes, pot_flops;
    if (op.calc == BSL_CALC_QUAD) {
        const double es = op.wprec == BSL_PREC_FP64 ? 8.0 : 4.0;
        // the hemisphere's weighted grid (op.na = ceil(n_theta / 2) nodes)
        pot_bytes = es * (double)c->P * (double)op.na * (double)op.nphi;
        // two Legendre transforms, and per carried l parity two phi
        // transforms and the pointwise product
        const double npar = (op.q_ne > 0 && op.q_ne < op.nk) ? 2.0 : 1.0;
        pot_flops = (double)c->P * (4.0 * op.na * op.nk + npar * (4.0 * op.na * op.nphi * op.nm + op.na * op.nphi));
    } else {
        const double es = op.vprec == BSL_PREC_FP64 ? 8.0 : 4.0;
        const bool harm = op.vform == BSL_VFORM_HARMONICS;
        pot_bytes = es * (double)(harm ? op.ncoef : op.npair) * (double)c->P;
        // per point: B^2 multiply-adds of the block product, plus one per
        // harmonic term in the harmonic form
        double pair_fma = 0.0;
        for (int g = 0; g < op.nblk; g++) {
            pair_fma += (double)op.blk[g].size * op.blk[g].size;
            if (harm) pair_fma += (double)gaunt::count_terms(op.blk[g].lb, op.blk[g].par);
        }
        pot_flops = 2.0 * (double)c->P * pair_fma;
    }
    // potential kernel (or the fused z-kinetic + potential kernel): V + read
    // X, read and write R
    out[0] = pot_bytes + 3.0 * vec;
    // kinetic kernels: x-y reads X, R and writes X, R; z (unless fused)
    // reads X, R, writes R
    out[1] = (op.fused ? 4.0 : 7.0) * vec;
    out[2] = 0.0;                    // recurrence update: folded into the kinetic and potential kernels
    out[3] = pot_bytes + 6.0 * vec;  // fused-step minimum: V + the six vector passes of the recurrence
    out[4] = pot_flops + 2.0 * (double)(op.nx + op.ny + op.nz) * (double)c->n_kry + 4.0 * (double)c->n_kry;
    return BSL_OK;
}

extern "C" int bsl_potential_kernel(const bsl_ctx *c, char *name, int32_t len)
{
    if (!c || !name || len <= 0) return fail(BSL_E_ARG, "bsl_potential_kernel: null argument");
    const DevOperator &op = c->op;
    char buf[128];
    if (op.calc == BSL_CALC_QUAD)
        std::snprintf(buf, sizeof buf, "k_quad_mfma<%d> (%d points per group, %d workgroups)", op.qmode, op.q_npts,
                      op.q_grid);
    else if (op.fused && op.nblk > 1)
        std::snprintf(buf, sizeof buf, "k_zpot2<%s>", op.vprec == BSL_PREC_FP64 ? "double" : "float");
    else if (op.fused)
        std::snprintf(buf, sizeof buf, "k_zpot1<%d, %d, %s>", op.blk[0].lb, op.blk[0].par,
                      op.vprec == BSL_PREC_FP64 ? "double" : "float");
    else if (op.vform == BSL_VFORM_HARMONICS &&
             harmonic_split(op.nblk, op.blk[0].lb, op.blk[0].par, op.nblk > 1 ? op.blk[1].lb : -1,
                            op.nblk > 1 ? op.blk[1].par : 0, op.blk[op.nblk > 1 ? 1 : 0].plane0,
                            op.blk[0].size) == 1)
        std::snprintf(buf, sizeof buf, "k_vgaunt_split<%s, %d>", op.vprec == BSL_PREC_FP64 ? "double" : "float",
                      op.vb_variant + 3);
    else if (op.vform == BSL_VFORM_HARMONICS &&
             harmonic_split(op.nblk, op.blk[0].lb, op.blk[0].par, -1, 0, 0, op.blk[0].size) == 2)
        std::snprintf(buf, sizeof buf, "k_vgaunt_split1<%d, %d, %s, %d>", op.blk[0].lb, op.blk[0].par,
                      op.vprec == BSL_PREC_FP64 ? "double" : "float", op.vb_variant + 3);
    else if (op.vform == BSL_VFORM_HARMONICS)
        std::snprintf(buf, sizeof buf, "k_vgaunt<%d, %d, %d, %d, %s, %d>", op.blk[0].lb, op.blk[0].par,
                      op.nblk > 1 ? op.blk[1].lb : -1, op.nblk > 1 ? op.blk[1].par : 0,
                      op.vprec == BSL_PREC_FP64 ? "double" : "float", op.vb_variant + 1);
    else if (!vblock_has_kernel(op.blk[0].size, op.nblk > 1 ? op.blk[1].size : 0))
        std::snprintf(buf, sizeof buf, "k_vblock_rows<%s>", op.vprec == BSL_PREC_FP64 ? "double" : "float");
    else
        std::snprintf(buf, sizeof buf, "k_vblock<%d, %d, %s, %d>", op.blk[0].size, op.nblk > 1 ? op.blk[1].size : 0,
                      op.vprec == BSL_PREC_FP64 ? "double" : "float", op.vb_variant + 1);
    std::snprintf(name, (size_t)len, "%s", buf);
    return BSL_OK;
}

extern "C" int bsl_time_steps(bsl_ctx *c, int32_t warmup, int32_t steps, double *ms)
{
    if (!c || steps <= 0 || warmup < 0 || !ms) return fail(BSL_E_ARG, "bad arguments");
    BSL_HIP(hipSetDevice(c->device));
    // every step writes its coefficients at the device step counter
    const int ksteps = std::min(steps, 64);
    int rc = ensure_coeffs(c, warmup + steps + ksteps + 1);
    if (rc) return rc;
    if ((rc = start_lanczos(c, nullptr, BSL_START_RANDOM, 0))) return rc;
    int j = 1;
    for (int w = 0; w < warmup; w++, j++)
        if ((rc = lanczos_step(c, c->op, j, nullptr))) return rc;
    hipEvent_t e0, e1;
    BSL_HIP(hipEventCreate(&e0));
    BSL_HIP(hipEventCreate(&e1));
    BSL_HIP(hipStreamSynchronize(c->stream));
    BSL_HIP(hipEventRecord(e0, c->stream));
    const int j0 = j;
    for (int s = 0; s < steps; s++, j++)
        if ((rc = lanczos_step(c, c->op, j, nullptr))) return rc;
    BSL_HIP(hipEventRecord(e1, c->stream));
    BSL_HIP(hipEventSynchronize(e1));
    float total = 0;
    BSL_HIP(hipEventElapsedTime(&total, e0, e1));
    (void)j0;
    // per-kernel durations over fewer steps: the kernels' own start and end
    // timestamps of each group (kinetic, potential, finalize)
    std::vector<hipEvent_t> ev((size_t)6 * ksteps);
    for (auto &e : ev) BSL_HIP(hipEventCreate(&e));
    double kin = 0, pot = 0, upd = 0;
    for (int s = 0; s < ksteps; s++)
        if ((rc = lanczos_step(c, c->op, j + s, &ev[(size_t)6 * s]))) return rc;
    BSL_HIP(hipStreamSynchronize(c->stream));
    for (int s = 0; s < ksteps; s++) {
        float a = 0, b = 0, d = 0;
        BSL_HIP(hipEventElapsedTime(&a, ev[6 * s + 0], ev[6 * s + 1]));
        BSL_HIP(hipEventElapsedTime(&b, ev[6 * s + 2], ev[6 * s + 3]));
        BSL_HIP(hipEventElapsedTime(&d, ev[6 * s + 4], ev[6 * s + 5]));
        kin += a;
        pot += b;
        upd += d;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ms[0] = total;
    ms[1] = pot / ksteps;
    ms[2] = kin / ksteps;
    ms[3] = upd / ksteps;
    return BSL_OK;
}
