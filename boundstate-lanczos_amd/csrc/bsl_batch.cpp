// bsl_batch.cpp -- many independent Lanczos recurrences (cages, para/ortho
// blocks, grid sizes) advanced in lockstep on one GPU.
//
// Each member is a full context (its own operator, Krylov vectors, stream
// and autotuned potential kernel).  One lockstep iteration enqueues every
// member's step on the member's stream; with graphs on, two iterations (so
// that every member's X/R buffers are back in place) are captured once into
// a hipGraph whose member chains are independent branches, and the graph is
// replayed: the coefficient outputs are indexed by each member's device step
// counter, so a replay needs no new arguments.  No reference counterpart:
// the reference runs one recurrence per process (lanczosEigenValuesVectors.cpp
// :158-242).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "bsl_hip.h"
#include "bsl_internal.h"

using namespace bsl;

struct bsl_batch {
    std::vector<bsl_ctx *> ctx;
    hipStream_t main = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    int device = 0;
};

namespace {

void batch_free(bsl_batch *b)
{
    if (!b) return;
    for (bsl_ctx *c : b->ctx) bsl_destroy(c);
    for (hipEvent_t e : b->join) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->main) (void)hipStreamDestroy(b->main);
    delete b;
}

// every member's stream waits for `main` (fork), or `main` waits for every member (join)
int batch_fork(bsl_batch *b)
{
    BSL_HIP(hipEventRecord(b->fork, b->main));
    for (bsl_ctx *c : b->ctx) BSL_HIP(hipStreamWaitEvent(c->stream, b->fork, 0));
    return BSL_OK;
}

int batch_join(bsl_batch *b)
{
    for (size_t i = 0; i < b->ctx.size(); i++) {
        BSL_HIP(hipEventRecord(b->join[i], b->ctx[i]->stream));
        BSL_HIP(hipStreamWaitEvent(b->main, b->join[i], 0));
    }
    return BSL_OK;
}

// one lockstep iteration on the member streams, between a fork and a join
int batch_iteration(bsl_batch *b)
{
    int rc = batch_fork(b);
    for (bsl_ctx *c : b->ctx)
        if (!rc) rc = ctx_step(c);
    if (!rc) rc = batch_join(b);
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

// captures two lockstep iterations (each member's buffers end where they began)
int capture_pair(bsl_batch *b, Graph &gr)
{
    BSL_HIP(hipStreamBeginCapture(b->main, hipStreamCaptureModeGlobal));
    int rc = batch_iteration(b);
    if (!rc) rc = batch_iteration(b);
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

// niter lockstep iterations: graph replays of pairs, then an eager tail
int run_steps(bsl_batch *b, int niter, int use_graph)
{
    int rc = BSL_OK;
    int done = 0;
    if (use_graph && niter >= 2) {
        Graph gr;
        if ((rc = capture_pair(b, gr))) return rc;
        for (; done + 2 <= niter; done += 2) BSL_HIP(hipGraphLaunch(gr.x, b->main));
        BSL_HIP(hipStreamSynchronize(b->main));
    }
    for (; done < niter && !rc; done++) rc = batch_iteration(b);
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
    BSL_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    b->join.resize(n, nullptr);
    for (auto &e : b->join) BSL_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
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
    int rc = batch_start(b, warmup + steps + 2, BSL_START_RANDOM, 0);
    if (!rc) rc = run_steps(b, warmup, use_graph);
    if (rc) return rc;
    BSL_HIP(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    BSL_HIP(hipEventCreate(&e0));
    BSL_HIP(hipEventCreate(&e1));
    Graph gr;
    if (use_graph) rc = capture_pair(b, gr);
    if (!rc) {
        BSL_HIP(hipEventRecord(e0, b->main));
        int done = 0;
        if (use_graph)
            for (; done + 2 <= steps; done += 2) BSL_HIP(hipGraphLaunch(gr.x, b->main));
        for (; done < steps && !rc; done++) rc = batch_iteration(b);
        BSL_HIP(hipEventRecord(e1, b->main));
        BSL_HIP(hipEventSynchronize(e1));
        float t = 0;
        BSL_HIP(hipEventElapsedTime(&t, e0, e1));
        ms[0] = t;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}
