// bsl_noalloc.hip -- bsl_lanczos_noalloc: Lanczos loop 1 on a prepared
// context without any device allocation in the call.
//
// bsl_lanczos keeps its per-run device state in buffers the library
// allocates: the recurrence scalars, the per-workgroup reduction partials of
// the potential kernel and the alpha/beta/beta^2 lists, which grow with the
// step count (hipMalloc on the first long run).  This variant takes that
// state from one file-scope device array per GPU instead, so a caller that
// may not allocate device memory (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc) can run the recurrence on a context
// it prepared beforehand.  The kernels are the same as bsl_lanczos's, launched
// in the same order on the same grids, so the coefficients are bitwise equal.
//
// Every call on one GPU shares that GPU's array and leaves its contents
// behind; the start of a run rewrites every word a run reads before reading
// it (scalars by the start-vector kernels, partials by each potential kernel,
// coefficients by each step's finalize).  Calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <utility>

#include "bsl_internal.h"

namespace bsl {
namespace {

constexpr int kNaMaxSteps = BSL_NOALLOC_MAX_STEPS;
constexpr int kNaMaxPartials = BSL_NOALLOC_MAX_PARTIALS;
// layout in doubles: the scalars (one 256-byte line), the three partial
// lists, then alpha, beta and beta^2 (kNaMaxSteps + 1 entries each)
constexpr size_t kNaScal = 0;
constexpr size_t kNaPart = 32;
constexpr size_t kNaCoef = kNaPart + 3 * (size_t)kNaMaxPartials;
constexpr size_t kNaDoubles = kNaCoef + 3 * (size_t)(kNaMaxSteps + 1);
static_assert(sizeof(Scal) <= 32 * sizeof(double), "scalars fit their line");
static_assert(kNaDoubles * sizeof(double) <= 2u << 20, "the scratch array stays below 2 MiB");

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mutex;
double *g_na_ptr[kNaMaxDevices] = {};

}  // namespace

// The per-GPU scratch of bsl_lanczos_noalloc.  In host code this name is a
// host-side stand-in: its device address comes from hipGetSymbolAddress, once
// per device (scratch_address).
static __device__ __align__(256) double scratch[kNaDoubles];

namespace {

// device address of `scratch` on the current device (which is `device`)
int scratch_address(int device, double **out)
{
    if (device < 0 || device >= kNaMaxDevices) return fail(BSL_E_ARG, "bsl_lanczos_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_ptr[device]) {
        void *p = nullptr;
        BSL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(BSL_E_HIP, "bsl_lanczos_noalloc: no device address for the scratch array");
        g_na_ptr[device] = static_cast<double *>(p);
    }
    *out = g_na_ptr[device];
    return BSL_OK;
}

}  // namespace
}  // namespace bsl

using namespace bsl;

extern "C" int bsl_lanczos_noalloc(bsl_ctx *c, int32_t niter, const double *v0, int32_t start, uint64_t seed,
                                   double *alpha, double *beta, double *beta2)
{
    if (!c || niter <= 0 || !alpha || !beta || !beta2) return fail(BSL_E_ARG, "bad arguments");
    if (niter > kNaMaxSteps)
        return fail(BSL_E_ARG, "bsl_lanczos_noalloc: at most " + std::to_string(kNaMaxSteps) +
                                   " steps (the size of its static scratch); use bsl_lanczos for longer runs");
    if (!v0 && start != BSL_START_UNIFORM && start != BSL_START_RANDOM) return fail(BSL_E_ARG, "unknown start vector");
    BSL_HIP(hipSetDevice(c->device));
    double *s = nullptr;
    int rc = scratch_address(c->device, &s);
    if (rc) return rc;
    // the context's operator with its partial lists redirected to the scratch
    DevOperator op = c->op;
    if (op.nalpha > kNaMaxPartials)
        return fail(BSL_E_ARG, "bsl_lanczos_noalloc: the potential kernel of this context uses " +
                                   std::to_string(op.nalpha) + " workgroups, above the scratch's " +
                                   std::to_string(kNaMaxPartials) + " partial sums");
    op.partial = s + kNaPart;
    op.partial_b = op.partial + kNaMaxPartials;
    op.partial_c = op.partial_b + kNaMaxPartials;
    op.npartial = kNaMaxPartials;
    Scal *sc = reinterpret_cast<Scal *>(s + kNaScal);
    double *d_alpha = s + kNaCoef, *d_beta = d_alpha + (kNaMaxSteps + 1), *d_beta2 = d_beta + (kNaMaxSteps + 1);
    // start vector into X (as bsl_lanczos), R = 0
    if (v0) {
        for (size_t k = 0; k < c->carried.size(); k++)
            BSL_HIP(hipMemcpyAsync(c->d_X + k * c->P, v0 + (size_t)c->carried[k] * c->P, sizeof(double) * c->P,
                                   hipMemcpyHostToDevice, c->stream));
        rc = launch_init_scal(sc, 1.0, 0.0, c->stream);
    } else {
        rc = launch_start_vector(c->d_X, c->n_kry, c->P, op.plane_ref, start, seed, sc, op.partial, op.npartial,
                                 c->stream);
    }
    if (!rc) rc = launch_zero(c->d_R, c->n_kry, c->stream);
    for (int j = 1; j <= niter && !rc; j++) {
        rc = launch_step(op, c->d_X, c->d_R, sc, d_alpha, d_beta, d_beta2, c->stream, nullptr);
        if (!rc) std::swap(c->d_X, c->d_R);
    }
    if (rc) {
        (void)hipStreamSynchronize(c->stream);
        return rc;
    }
    BSL_HIP(hipMemcpyAsync(alpha, d_alpha, sizeof(double) * niter, hipMemcpyDeviceToHost, c->stream));
    BSL_HIP(hipMemcpyAsync(beta + 1, d_beta + 1, sizeof(double) * niter, hipMemcpyDeviceToHost, c->stream));
    BSL_HIP(hipMemcpyAsync(beta2 + 1, d_beta2 + 1, sizeof(double) * niter, hipMemcpyDeviceToHost, c->stream));
    BSL_HIP(hipStreamSynchronize(c->stream));
    beta[0] = 0.0;
    beta2[0] = 0.0;
    return check_breakdown(sc, c->stream);
}
