// bsl_batch.cpp -- many independent Lanczos recurrences (cages, para/ortho
// blocks, grid sizes) advanced in lockstep on one GPU.
//
// Each member is a full context (its own operator, Krylov vectors, scalars
// and coefficient outputs).  A lockstep iteration is one launch per kernel
// variant over every member that uses it (x-y kinetic, z kinetic, potential,
// finalize; split by potential block and precision): each
// launch takes a device table of the members' pointers and shapes, and its
// workgroups are the members' own grids laid end to end.  A member whose
// step fuses the z kinetic term into the potential kernel takes the fused
// table kernel instead of the z and potential ones.  A member whose
// potential has no table kernel (quadrature calculator, the full basis,
// other block shapes) gets its own potential launch inside the iteration;
// one with an axis of more than 64 points its own kinetic and potential
// launches.
// Two tables are kept, one per step parity (the members' X and R swap roles
// every step), so two iterations leave every buffer where it began; with
// graphs on they are captured once and replayed.  No reference counterpart: the reference runs
// one recurrence per process (lanczosEigenValuesVectors.cpp:158-242).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#include "bsl_hip.h"
#include "bsl_internal.h"

using namespace bsl;

namespace {
struct Launch {
    int kind, prec, offset, nm, total;
    const double *G;  // the Gaunt values of a potential launch's block (members share them)
    int nz;           // the longest z axis among the members
};
}  // namespace

struct bsl_batch {
    std::vector<bsl_ctx *> ctx;
    hipStream_t main = nullptr;
    int device = 0;
    std::vector<BatchEntry> h_tab;      // both parities' entries, launch after launch
    std::vector<unsigned char> h_blob;  // the launches' device tables (kBatchTableBytes each)
    unsigned char *d_tab = nullptr;
    size_t tab_cap = 0;
    std::vector<Launch> launches[2];    // per step parity, in launch order
    std::vector<int> direct;            // members with a potential launch of their own
};

namespace {

void batch_free(bsl_batch *b)
{
    if (!b) return;
    if (b->main) (void)hipStreamSynchronize(b->main);
    for (bsl_ctx *c : b->ctx) bsl_destroy(c);
    if (b->d_tab) (void)hipFree(b->d_tab);
    if (b->main) (void)hipStreamDestroy(b->main);
    delete b;
}

// the potential kind of a member's operator (the para or ortho lmax-5 block
// of the reduced basis), or -1 if it has no table kernel
int potential_kind(const DevOperator &op)
{
    // (a member whose kinetic term has no table kernel runs all of its
    // step on its own launches, in order)
    if (op.calc == BSL_CALC_QUAD || op.vform != BSL_VFORM_HARMONICS || !kinetic_mfma_fits(op)) return -1;
    const Block &b0 = op.blk[0], &b1 = op.nblk > 1 ? op.blk[1] : op.blk[0];
    const int split = harmonic_split(op.nblk, b0.lb, b0.par, op.nblk > 1 ? b1.lb : -1, b1.par, b1.plane0, b0.size);
    if (split == 2 && op.fused) return b0.par == 0 ? kBZPotPara : kBZPotOrtho;
    if (split == 2) return b0.par == 0 ? kBPotSplitPara : kBPotSplitOrtho;
    return -1;
}

BatchEntry entry_of(const bsl_ctx *c, int parity)
{
    const DevOperator &op = c->op;
    BatchEntry e{};
    e.X = parity == 0 ? c->d_X : c->d_R;
    e.R = parity == 0 ? c->d_R : c->d_X;
    e.tx = op.tx;
    e.ty = op.ty;
    e.tz = op.tz;
    e.rot = op.rot;
    e.mask = op.mask;
    e.sc = c->d_scal;
    e.H = op.vpk;
    e.G0 = op.gc[0];
    e.G1 = op.gc[1];
    e.pa = op.partial;
    e.pb = op.partial_b;
    e.pc = op.partial_c;
    e.alpha = c->d_alpha;
    e.beta = c->d_beta;
    e.beta2 = c->d_beta2;
    e.P = op.P;
    e.nx = op.nx;
    e.ny = op.ny;
    e.nz = op.nz;
    e.nk = op.nk;
    e.nc = op.ncoef;
    e.plane1 = op.nblk > 1 ? op.blk[1].plane0 : 0;
    e.mask0 = op.blk[0].mask;
    e.mask1 = op.nblk > 1 ? op.blk[1].mask : 0.0;
    e.nalpha = (int)op.nalpha;
    return e;
}

// The per-parity launch lists and their tables, from the members' current
// buffers (parity 0 = the contexts' X and R as they are now), uploaded to
// the batch's device table.
int build_tables(bsl_batch *b)
{
    b->h_tab.clear();
    b->direct.clear();
    for (size_t i = 0; i < b->ctx.size(); i++)
        if (potential_kind(b->ctx[i]->op) < 0) b->direct.push_back((int)i);
    for (int parity = 0; parity < 2; parity++) {
        std::vector<Launch> &L = b->launches[parity];
        L.clear();
        // members grouped by (kind, precision), in member order, at most
        // kBatchMaxTable per launch
        auto add = [&](int kind, int prec, auto count_of, auto uses) {
            Launch cur{kind, prec, (int)b->h_tab.size(), 0, 0, nullptr, 0};
            for (bsl_ctx *c : b->ctx) {
                if (!uses(c)) continue;
                if (cur.nm == kBatchMaxTable) {
                    L.push_back(cur);
                    cur = Launch{kind, prec, (int)b->h_tab.size(), 0, 0, nullptr, 0};
                }
                BatchEntry e = entry_of(c, parity);
                if (!cur.G) cur.G = c->op.gc[0];
                cur.nz = std::max(cur.nz, c->op.nz);
                e.first = cur.total;
                e.count = count_of(c);
                cur.total += e.count;
                cur.nm++;
                b->h_tab.push_back(e);
            }
            if (cur.nm > 0) L.push_back(cur);
        };
        for (int prec : {BSL_PREC_FP64, BSL_PREC_FP32, BSL_PREC_BF16X3})
            add(kBXyHalf, prec, [](const bsl_ctx *c) { return c->op.nz * c->op.nk; },
                [&](const bsl_ctx *c) { return c->op.kprec == prec && kinetic_mfma_fits(c->op); });
        for (int prec : {BSL_PREC_FP64, BSL_PREC_FP32, BSL_PREC_BF16X3})
            add(kBZHalf, prec, [](const bsl_ctx *c) { return ((c->op.nx * c->op.ny + 63) / 64) * c->op.nk; },
                [&](const bsl_ctx *c) { return c->op.kprec == prec && kinetic_mfma_fits(c->op) && !c->op.fused; });
        for (int kind : {(int)kBPotSplitPara, (int)kBPotSplitOrtho})
            for (int prec : {BSL_PREC_FP64, BSL_PREC_FP32})
                add(kind, prec, [](const bsl_ctx *c) { return (int)((c->op.P + 63) / 64); },
                    [&](const bsl_ctx *c) { return potential_kind(c->op) == kind && c->op.vprec == prec; });
        for (int kind : {(int)kBZPotPara, (int)kBZPotOrtho})
            for (int prec : {BSL_PREC_FP64, BSL_PREC_FP32})
                add(kind, prec, [](const bsl_ctx *c) { return (int)zpot_workgroups(c->op); },
                    [&](const bsl_ctx *c) { return potential_kind(c->op) == kind && c->op.vprec == prec; });
        add(kBFinalize, BSL_PREC_FP64, [](const bsl_ctx *) { return 1; }, [](const bsl_ctx *) { return true; });
    }
    // every count must match the grids the member's own kernels use
    for (const BatchEntry &e : b->h_tab)
        if (e.count <= 0) return fail(BSL_E_STATE, "empty batch member grid");
    // one device table per launch: its entries, then their first[] list
    size_t nblk = b->launches[0].size() + b->launches[1].size();
    b->h_blob.assign(nblk * kBatchTableBytes, 0);
    size_t blk = 0;
    for (int parity = 0; parity < 2; parity++)
        for (Launch &l : b->launches[parity]) {
            unsigned char *base = b->h_blob.data() + blk * kBatchTableBytes;
            std::memcpy(base, b->h_tab.data() + l.offset, sizeof(BatchEntry) * l.nm);
            int *first = reinterpret_cast<int *>(base + sizeof(BatchEntry) * kBatchMaxTable);
            for (int i = 0; i < kBatchMaxTable; i++)
                first[i] = i < l.nm ? b->h_tab[l.offset + i].first : std::numeric_limits<int>::max();
            l.offset = (int)blk++;
        }
    if (b->h_blob.size() > b->tab_cap) {
        if (b->d_tab) (void)hipFree(b->d_tab);
        b->d_tab = nullptr;
        b->tab_cap = 0;
        BSL_HIP(hipMalloc((void **)&b->d_tab, b->h_blob.size()));
        b->tab_cap = b->h_blob.size();
    }
    BSL_HIP(hipMemcpyAsync(b->d_tab, b->h_blob.data(), b->h_blob.size(), hipMemcpyHostToDevice, b->main));
    BSL_HIP(hipStreamSynchronize(b->main));
    return BSL_OK;
}

// one lockstep iteration of step parity `parity` on the batch stream
int batch_iteration(bsl_batch *b, int parity)
{
    int rc = BSL_OK;
    for (const Launch &l : b->launches[parity]) {
        if (l.kind == kBFinalize) {
            // members without a table potential kernel, after the kinetic
            // launches and before the finalize
            for (int i : b->direct) {
                bsl_ctx *c = b->ctx[i];
                double *X = parity == 0 ? c->d_X : c->d_R, *R = parity == 0 ? c->d_R : c->d_X;
                if (!rc && !kinetic_mfma_fits(c->op)) rc = launch_kinetic(c->op, X, R, c->d_scal, b->main);
                if (!rc) rc = launch_potential_part(c->op, X, R, c->d_scal, b->main);
            }
        }
        if (!rc)
            rc = launch_batch(l.kind, l.prec,
                              reinterpret_cast<const BatchEntry *>(b->d_tab + (size_t)l.offset * kBatchTableBytes),
                              l.nm, l.total, l.G, l.nz, b->main);
    }
    return rc;
}

struct Graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    ~Graph()
    {
        if (x) (void)hipGraphExecDestroy(x);
        if (g) (void)hipGraphDestroy(g);
    }
};

// captures two lockstep iterations (parity 0, then 1: every buffer ends where it began)
int capture_pair(bsl_batch *b, Graph &gr)
{
    BSL_HIP(hipStreamBeginCapture(b->main, hipStreamCaptureModeGlobal));
    int rc = batch_iteration(b, 0);
    if (!rc) rc = batch_iteration(b, 1);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->main, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    gr.g = g;
    BSL_HIP(hipGraphInstantiate(&gr.x, gr.g, nullptr, nullptr, 0));
    return BSL_OK;
}

// the members' X and R after `done` iterations from the tables' parity 0
void settle_buffers(bsl_batch *b, int done)
{
    if (done % 2)
        for (bsl_ctx *c : b->ctx) std::swap(c->d_X, c->d_R);
}

// niter lockstep iterations: graph replays of pairs, then an eager tail
int run_steps(bsl_batch *b, int niter, int use_graph)
{
    int rc = BSL_OK;
    int done = 0;
    if (use_graph && niter >= 2) {
        Graph gr;
        if ((rc = capture_pair(b, gr))) return rc;
        for (; done + 2 <= niter; done += 2) BSL_HIP(hipGraphLaunch(gr.x, b->main));
    }
    for (; done < niter && !rc; done++) rc = batch_iteration(b, done % 2);
    const hipError_t e = hipStreamSynchronize(b->main);
    if (!rc && e != hipSuccess) rc = hip_fail(e, "batch steps", __FILE__, __LINE__);
    settle_buffers(b, done);
    return rc;
}

int batch_start(bsl_batch *b, int niter, int start, uint64_t seed)
{
    int rc = BSL_OK;
    for (size_t i = 0; i < b->ctx.size() && !rc; i++) {
        rc = ctx_reserve_steps(b->ctx[i], niter);
        if (!rc) rc = ctx_start(b->ctx[i], nullptr, start, seed + i);
    }
    for (bsl_ctx *c : b->ctx)
        if (!rc) {
            hipError_t e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) rc = hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__);
        }
    if (!rc) rc = build_tables(b);
    return rc;
}

}  // namespace

extern "C" int bsl_batch_create(int32_t n, const bsl_params *params, bsl_batch **out)
{
    if (n <= 0 || !params || !out) return fail(BSL_E_ARG, "bsl_batch_create: bad arguments");
    *out = nullptr;
    std::unique_ptr<bsl_batch, void (*)(bsl_batch *)> b(new bsl_batch(), batch_free);
    BSL_HIP(hipGetDevice(&b->device));
    for (int i = 0; i < n; i++) {
        bsl_ctx *c = nullptr;
        int rc = bsl_create(&params[i], &c);
        if (rc) return rc;
        b->ctx.push_back(c);
    }
    BSL_HIP(hipStreamCreateWithFlags(&b->main, hipStreamNonBlocking));
    *out = b.release();
    return BSL_OK;
}

extern "C" int bsl_batch_destroy(bsl_batch *b)
{
    batch_free(b);
    return BSL_OK;
}

extern "C" int bsl_batch_size(const bsl_batch *b) { return b ? (int)b->ctx.size() : -1; }

extern "C" int bsl_batch_member(bsl_batch *b, int32_t i, bsl_ctx **out)
{
    if (!b || !out || i < 0 || i >= (int)b->ctx.size()) return fail(BSL_E_ARG, "bsl_batch_member: bad index");
    *out = b->ctx[i];
    return BSL_OK;
}

extern "C" int bsl_batch_lanczos(bsl_batch *b, int32_t niter, int32_t start, uint64_t seed, int32_t use_graph,
                                 double *alpha, double *beta, double *beta2)
{
    if (!b || niter <= 0 || !alpha || !beta || !beta2) return fail(BSL_E_ARG, "bsl_batch_lanczos: bad arguments");
    if (start != BSL_START_UNIFORM && start != BSL_START_RANDOM) return fail(BSL_E_ARG, "unknown start vector");
    BSL_HIP(hipSetDevice(b->device));
    int rc = batch_start(b, niter, start, seed);
    if (!rc) rc = run_steps(b, niter, use_graph);
    for (size_t i = 0; i < b->ctx.size() && !rc; i++) {
        bsl_ctx *c = b->ctx[i];
        double *a = alpha + i * (size_t)niter, *be = beta + i * (size_t)(niter + 1),
               *b2 = beta2 + i * (size_t)(niter + 1);
        hipError_t e = hipMemcpyAsync(a, c->d_alpha, sizeof(double) * niter, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(be + 1, c->d_beta + 1, sizeof(double) * niter, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(b2 + 1, c->d_beta2 + 1, sizeof(double) * niter, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(e, "download coefficients", __FILE__, __LINE__);
        be[0] = 0.0;
        b2[0] = 0.0;
        if (!rc) rc = check_breakdown(c->d_scal, c->stream);
    }
    return rc;
}

extern "C" int bsl_batch_time_steps(bsl_batch *b, int32_t warmup, int32_t steps, int32_t use_graph, double *ms)
{
    if (!b || warmup < 0 || steps <= 0 || !ms) return fail(BSL_E_ARG, "bsl_batch_time_steps: bad arguments");
    BSL_HIP(hipSetDevice(b->device));
    // an even warm-up keeps the tables' parity 0 current for the timed steps
    const int wu = warmup + (warmup % 2);
    int rc = batch_start(b, wu + steps + 2, BSL_START_RANDOM, 0);
    if (!rc) rc = run_steps(b, wu, use_graph);
    if (rc) return rc;
    hipEvent_t e0, e1;
    BSL_HIP(hipEventCreate(&e0));
    BSL_HIP(hipEventCreate(&e1));
    Graph gr;
    if (use_graph) rc = capture_pair(b, gr);
    int done = 0;
    if (!rc) {
        BSL_HIP(hipEventRecord(e0, b->main));
        if (use_graph)
            for (; done + 2 <= steps; done += 2) BSL_HIP(hipGraphLaunch(gr.x, b->main));
        for (; done < steps && !rc; done++) rc = batch_iteration(b, done % 2);
        BSL_HIP(hipEventRecord(e1, b->main));
        BSL_HIP(hipEventSynchronize(e1));
        float t = 0;
        BSL_HIP(hipEventElapsedTime(&t, e0, e1));
        ms[0] = t;
    }
    settle_buffers(b, done);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}
