// bsl_tridiag_dev.hip -- the Cullum-Willoughby tridiagonal eigensolve on the
// GPU (FP64): the Sturm bisections of BISEC (bisec.f:4-225) run one eigenvalue
// per thread, every index of (lo, hi] at once, each from the full interval
// instead of from the bounds refined by the previous indices; the spurious
// test counts of T and T-hat are taken at the same points as BISEC's.  The
// inverse-iteration error estimates of INVERR (inverr.f:4-167) run one good
// eigenvalue per thread.  The bookkeeping between them (clusters, ISOEV) is
// the host's, in bsl_tridiag.cpp.
#include <hip/hip_runtime.h>

#include "bsl_internal.h"

namespace bsl {

namespace {

constexpr double kEps2d = 2.0 * 1.1102230246251565e-16;
constexpr int kMaxBisect = 256;  // far above the ~60 halvings FP64 allows

// eigenvalues of T above x (negative pivots of x - T), top to bottom;
// be[i-1] = b(i), be[0] = 0, be[m] = 0
__device__ int sturm_forward_d(const double *__restrict__ al, const double *__restrict__ be, int m, double x)
{
    int count = 0;
    double d = 1.0;
    for (int i = 1; i <= m; i++) {
        const double b = be[i - 1];
        const double q = (d == 0.0) ? b / kEps2d : (b * b) / d;
        d = x - al[i - 1] - q;
        if (d < 0.0) count++;
    }
    return count;
}

// bottom to top; last_negative: the final pivot (row 1) was negative, so that
// count - last_negative is T-hat's count
__device__ void sturm_backward_d(const double *__restrict__ al, const double *__restrict__ be, int m, double x,
                                 int *count, int *last_negative)
{
    int c = 0, last = 0;
    double d = 1.0;
    for (int i = m; i >= 1; i--) {
        const double b = be[i];
        const double q = (d == 0.0) ? b / kEps2d : (b * b) / d;
        d = x - al[i - 1] - q;
        last = 0;
        if (d < 0.0) {
            c++;
            last = 1;
        }
    }
    *count = c;
    *last_negative = last;
}

// thread k (1-based, counted from the top of (lo, hi]): bisection, then the
// counts of the spurious test.  out[k-1][0..9] = x1, xl, xu, nu, kl, jl, ku,
// ju (T and T-hat counts at xl - ep0 and xu + ep0), klm, kum (T counts at
// xl - epm and xu + epm, the multiplicity width; equal to kl, ku when
// epm = ep0).
__global__ void k_tri_bisect(const double *__restrict__ al, const double *__restrict__ be, int m, double lo,
                             double hi, int above_hi, int mt, double ep0, double ep1, double epm,
                             double *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (k > mt) return;
    double xl = lo, xu = hi, x1 = 0.5 * (lo + hi);
    int nu = 0;
    for (int it = 0; it < kMaxBisect; it++) {
        x1 = (xl + xu) * 0.5;
        const double width = xu - xl;
        if (width <= kEps2d * (fabs(xl) + fabs(xu)) + ep1) break;
        const int nev = sturm_forward_d(al, be, m, x1) - above_hi;
        if (nev >= k) {
            xl = x1;
        } else {
            xu = x1;
            nu = nev;
        }
    }
    int kl, jl, ku, ju;
    sturm_backward_d(al, be, m, xl - ep0, &kl, &jl);
    kl -= above_hi;
    jl = kl - jl;
    sturm_backward_d(al, be, m, xu + ep0, &ku, &ju);
    ku -= above_hi;
    ju = ku - ju;
    int klm = kl, kum = ku;
    if (epm > ep0) {
        int l;
        sturm_backward_d(al, be, m, xl - epm, &klm, &l);
        klm -= above_hi;
        sturm_backward_d(al, be, m, xu + epm, &kum, &l);
        kum -= above_hi;
    }
    double *o = out + (size_t)(k - 1) * kTriOut;
    o[0] = x1;
    o[1] = xl;
    o[2] = xu;
    o[3] = nu;
    o[4] = kl;
    o[5] = jl;
    o[6] = ku;
    o[7] = ju;
    o[8] = klm;
    o[9] = kum;
}

// INVERR's inverse iteration for eigenvalue xs[j], thread j: (T - x) y = v
// with partial pivoting from the scaled start vector g, at most maxit sweeps
// until |y|_1 >= 1; err[j] = beta_m1 |y_m| of the normalised y, negative when
// unconverged.  Per-thread rows of V, M and the pivot flags live in scratch
// laid out [i][j].
__global__ void k_tri_inverr(const double *__restrict__ al, const double *__restrict__ be, int m,
                             const double *__restrict__ xs, int nx, const double *__restrict__ g, double eps3,
                             double eps4, double beta_m1, int maxit, double *__restrict__ sv, double *__restrict__ sm,
                             unsigned char *__restrict__ sf, double *__restrict__ err)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nx) return;
    const double x = xs[j];
    auto V = [&](int i) -> double & { return sv[(size_t)(i - 1) * nx + j]; };
    auto M = [&](int i) -> double & { return sm[(size_t)(i - 1) * nx + j]; };
    auto F = [&](int i) -> unsigned char & { return sf[(size_t)(i - 1) * nx + j]; };
    auto A = [&](int i) { return al[i - 1]; };
    auto B = [&](int i) { return i <= m ? be[i - 1] : 0.0; };  // b(m+1) = 0 here
    for (int i = 1; i <= m; i++) V(i) = g[i - 1];
    int it = 1;
    for (;;) {
        double u = A(1) - x;
        double z = B(2);
        for (int i = 2; i <= m; i++) {
            if (B(i) > fabs(u)) {
                const double ratio = u / B(i);
                F(i) = 1;
                M(i - 1) = A(i) - x;
                u = z - ratio * M(i - 1);
                z = -ratio * B(i + 1);
                const double keep = V(i - 1);
                V(i - 1) = V(i);
                V(i) = keep - ratio * V(i);
            } else {
                F(i) = 0;
                M(i - 1) = z / u;
                V(i - 1) = V(i - 1) / u;
                V(i) = V(i) - B(i) * V(i - 1);
                const double ratio = B(i) / u;
                u = A(i) - x - z * ratio;
                z = B(i + 1);
            }
        }
        if (u == 0.0) u = eps3;
        V(m) = V(m) / u;
        for (int i = m - 1; i >= 1; i--) {
            if (F(i + 1)) {
                const double far = (i + 2 <= m) ? B(i + 2) * V(i + 2) : 0.0;
                V(i) = (V(i) - M(i) * V(i + 1) - far) / B(i + 1);
            } else {
                V(i) = V(i) - M(i) * V(i + 1);
            }
        }
        double norm1 = fabs(V(m));
        for (int i = m - 1; i >= 1; i--) norm1 = norm1 + fabs(V(i));
        if (norm1 >= 1.0) break;
        if (++it > maxit) break;
        const double s = eps4 / norm1;
        for (int i = 1; i <= m; i++) V(i) = V(i) * s;
    }
    double ss = 0.0;
    for (int i = 1; i <= m; i++) ss = ss + V(i) * V(i);
    ss = 1.0 / sqrt(ss);
    double est = beta_m1 * fabs(ss * V(m));
    if (it > maxit) est = -est;
    err[j] = est;
}

}  // namespace

int launch_tri_bisect(const double *al, const double *be, int m, double lo, double hi, int above_hi, int mt,
                      double ep0, double ep1, double epm, double *out, hipStream_t s)
{
    if (mt <= 0) return BSL_OK;
    hipLaunchKernelGGL(k_tri_bisect, dim3((mt + 63) / 64), dim3(64), 0, s, al, be, m, lo, hi, above_hi, mt, ep0,
                       ep1, epm, out);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

int launch_tri_inverr(const double *al, const double *be, int m, const double *xs, int nx, const double *g,
                      double eps3, double eps4, double beta_m1, int maxit, double *sv, double *sm, unsigned char *sf,
                      double *err, hipStream_t s)
{
    if (nx <= 0) return BSL_OK;
    hipLaunchKernelGGL(k_tri_inverr, dim3((nx + 63) / 64), dim3(64), 0, s, al, be, m, xs, nx, g, eps3, eps4, beta_m1,
                       maxit, sv, sm, sf, err);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

}  // namespace bsl
