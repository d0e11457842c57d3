// bsl_tridiag.cpp -- Cullum-Willoughby post-processing of the Lanczos
// tridiagonal matrix T_m (host, FP64): Sturm bisection with the spurious
// (ghost) test, the isolation flag, inverse-iteration error estimates and the
// Tm eigenvectors used for Ritz vectors.
//
// Algorithms of the reference's F77 routines (bisec.f, isoev.f, inverr.f,
// trivec.f, inverm.f, sturmi.f, genran.f, scalar.f, driven by lancbis.cpp
// and lanczosvectors in lanczosEigenValuesVectors.cpp), rewritten as
// structured C++ with the same floating-point operations, so that eigenvalue
// lists are identical to the reference's for identical alpha/beta.
#include <cmath>
#include <cstdint>
#include <vector>

#include "bsl_internal.h"

namespace bsl {
namespace {

// twice the LAPACK relative machine precision dlamch('e') = 2^-53
constexpr double kEps2 = 2.0 * 1.1102230246251565e-16;

// GNU SRAND/RAND as used by GENRAN (genran.f:4-32): the Park-Miller
// "minimal standard" generator, a = 16807, m = 2^31 - 1, seed 0 -> 123459876;
// RAND returns (s - 1) with its low 9 bits cleared, over 2^31, a REAL*4 value
// (22 significant bits, so the quotient is exact in float and double).
void park_miller(int seed, double *r, int n)
{
    uint64_t s = seed ? (uint64_t)seed : 123459876ull;
    for (int i = 0; i < n; i++) {
        s = (16807ull * s) % 2147483647ull;
        r[i] = (double)((s - 1) & ~(uint64_t)511) * (1.0 / 2147483648.0);
    }
}

// Tridiagonal T with 1-based accessors a(i) = ALPHA(i), b(i) = BETA(i)
// (b(1) = 0, b(m+1) the residual norm) and b2(i) = BETA(i)^2.
struct Tri {
    int m;
    const double *al;
    std::vector<double> be, b2;
    double a(int i) const { return al[i - 1]; }
    double b(int i) const { return be[i - 1]; }
};

// number of eigenvalues of T above x (negative pivots of x - T), counted
// from the top (forward) or from the bottom (backward, which also reports
// whether the last pivot was negative: removing it gives T-hat's count).
int sturm_forward(const Tri &t, double x)
{
    int count = 0;
    double d = 1.0;
    for (int i = 1; i <= t.m; i++) {
        const double q = (d == 0.0) ? t.b(i) / kEps2 : t.b2[i - 1] / d;
        d = x - t.a(i) - q;
        if (d < 0.0) count++;
    }
    return count;
}

void sturm_backward(const Tri &t, double x, int *count, int *last_negative)
{
    int c = 0, last = 0;
    double d = 1.0;
    for (int i = t.m; i >= 1; i--) {
        const double q = (d == 0.0) ? t.b(i + 1) / kEps2 : t.b2[i] / d;
        d = x - t.a(i) - q;
        last = 0;
        if (d < 0.0) {
            c++;
            last = 1;
        }
    }
    *count = c;
    *last_negative = last;
}

// BISEC (bisec.f:4-225): distinct eigenvalues of T in (lo, hi], ascending,
// with multiplicities (0 = spurious).  Returns -1 when the Sturm budget is
// exceeded (the reference STOPs).  mtol > 0 (no reference counterpart)
// widens the interval [xl - ep0, xu + ep0] in which T's eigenvalues count as
// one multiple eigenvalue to at least mtol, while the spurious test of a
// simple eigenvalue keeps BISEC's ep0: with a reduced-precision operator the
// ghost copies of a converged level are spread by its rounding rather than
// by FP64's, and at BISEC's FP64-sized width each copy would be a simple
// eigenvalue of both T and T-hat, i.e. spurious.  mtol = 0 is BISEC itself.
int bisect(Tri &t, double lo, double hi, double mtol, std::vector<double> &vs, std::vector<int> &mp)
{
    const int m = t.m;
    double tkmax = 0.0;
    for (int i = 1; i <= m; i++) tkmax = std::fmax(std::fmax(std::fabs(t.a(i)), t.b(i)), tkmax);
    const double ttol = kEps2 * tkmax;
    const double scale = (double)(m + 1000);
    const double ep0 = scale * ttol, ep1 = std::sqrt(scale) * ttol;
    const double epm = std::fmax(ep0, mtol);  // multiplicity width
    t.be[m] = 0.0;  // BETA(MEV+1) is taken as zero during the search
    for (int i = 0; i <= m; i++) t.b2[i] = t.be[i] * t.be[i];

    const int above_hi = sturm_forward(t, hi);
    auto cnt = [&](double x) { return sturm_forward(t, x) - above_hi; };  // eigenvalues in (x, hi]
    const int mt = cnt(lo);
    vs.clear();
    mp.clear();
    if (30 * mt >= 100 * m) return -1;
    if (mt == 0) return 0;
    std::vector<double> lower(mt + 1, lo), upper(mt + 1, hi);  // bounds of the k-th from the top
    int k = mt;
    while (k > 0) {
        double xl = lower[k], xu = upper[k];
        int nu = cnt(xu);
        double x1;
        for (;;) {
            x1 = (xl + xu) * 0.5;
            const double width = xu - xl;
            if (width <= kEps2 * (std::fabs(xl) + std::fabs(xu)) + ep1) break;
            const int nev = cnt(x1);
            if (nev >= k) {
                xl = x1;
            } else {
                xu = x1;
                nu = nev;
                for (int i = 1; i <= nev; i++) lower[i] = std::fmax(x1, lower[i]);
                for (int i = nev + 1; i <= k; i++) upper[i] = std::fmin(x1, upper[i]);
            }
        }
        vs.push_back(x1);
        // spurious test: compare T and T-hat counts just outside the interval
        int kl, jl, ku, ju;
        sturm_backward(t, xl - epm, &kl, &jl);
        kl -= above_hi;
        jl = kl - jl;
        x1 = xu + epm;
        sturm_backward(t, x1, &ku, &ju);
        ku -= above_hi;
        ju = ku - ju;
        int mult, knew;
        if (kl - ku - 1 == 0) {
            if (epm > ep0) {
                // simple at the multiplicity width: T-hat counts at BISEC's own
                int c0, l0;
                sturm_backward(t, xl - ep0, &c0, &l0);
                jl = c0 - above_hi - l0;
                sturm_backward(t, xu + ep0, &c0, &l0);
                ju = c0 - above_hi - l0;
            }
            mult = (ju < jl) ? 0 : 1;
            knew = k - 1;
        } else {
            if (ku != nu) {
                int change;
                do {
                    x1 = x1 + epm;
                    const int nev = cnt(x1);
                    change = ku - nev;
                    ku = nev;
                } while (change != 0);
            }
            mult = kl - ku;
            knew = ku;
        }
        mp.push_back(mult);
        k = knew;
        if (k <= 0) break;
        for (int i = 1; i <= knew; i++) lower[i] = std::fmax(x1, lower[i]);
    }
    return (int)vs.size();
}

// ISOEV (isoev.f:4-55): a simple eigenvalue whose nearest neighbour is
// spurious and closer than max(1e-8, 1e-8 |lambda|) is flagged -1.
void isolate(const std::vector<double> &vs, std::vector<int> &mp)
{
    const int n = (int)vs.size();
    if (n < 2) return;
    for (int j = 0; j < n; j++) {
        if (mp[j] != 1) continue;
        double g;
        if (j > 0 && j < n - 1) {
            const double below = vs[j] - vs[j - 1], above = vs[j + 1] - vs[j];
            g = (below < above) ? -below : above;
        } else {
            g = (j == 0) ? vs[1] - vs[0] : vs[n - 2] - vs[n - 1];
        }
        const int nb = (g >= 0.0) ? j + 1 : j - 1;
        if (mp[nb] == 0 && std::fabs(g) <= std::fmax(1.e-8, 1.e-8 * std::fabs(vs[j]))) mp[j] = -mp[j];
    }
}

// One inverse-iteration solve of (T_order - x) y = v with partial pivoting,
// repeated until the 1-norm reaches 1 or `maxit` sweeps (the common core of
// INVERR inverr.f:173-215 and INVERM inverm.f:168-210).  v holds the start
// vector on entry and the normalised eigenvector on exit.  be is T's beta
// array (1-based via be[i-1]) with be[order] already zero; its sign is used
// as the pivot flag and restored.  Returns the number of sweeps.
int inverse_iteration(int order, const double *al, double *be, std::vector<double> &mult, double *v, double x,
                      double eps3, double eps4, int maxit)
{
    auto B = [&](int i) -> double & { return be[i - 1]; };
    auto A = [&](int i) { return al[i - 1]; };
    auto V = [&](int i) -> double & { return v[i - 1]; };
    auto M = [&](int i) -> double & { return mult[i - 1]; };
    int it = 1;
    for (;;) {
        double u = A(1) - x;
        double z = B(2);
        for (int i = 2; i <= order; i++) {
            if (B(i) > std::fabs(u)) {
                const double ratio = u / B(i);
                B(i) = -B(i);
                M(i - 1) = A(i) - x;
                u = z - ratio * M(i - 1);
                z = -ratio * B(i + 1);
                const double keep = V(i - 1);
                V(i - 1) = V(i);
                V(i) = keep - ratio * V(i);
            } else {
                M(i - 1) = z / u;
                V(i - 1) = V(i - 1) / u;
                V(i) = V(i) - B(i) * V(i - 1);
                const double ratio = B(i) / u;
                u = A(i) - x - z * ratio;
                z = B(i + 1);
            }
        }
        if (u == 0.0) u = eps3;
        V(order) = V(order) / u;
        for (int i = order - 1; i >= 1; i--) {
            if (B(i + 1) < 0.0) {
                B(i + 1) = -B(i + 1);
                const double far = (i + 2 <= order) ? B(i + 2) * V(i + 2) : 0.0;  // B(order+1) = 0
                V(i) = (V(i) - M(i) * V(i + 1) - far) / B(i + 1);
            } else {
                V(i) = V(i) - M(i) * V(i + 1);
            }
        }
        double norm1 = std::fabs(V(order));
        for (int i = order - 1; i >= 1; i--) norm1 = norm1 + std::fabs(V(i));
        if (norm1 >= 1.0) break;
        if (++it > maxit) break;
        const double s = eps4 / norm1;
        for (int i = 1; i <= order; i++) V(i) = V(i) * s;
    }
    double ss = 0.0;
    for (int i = 1; i <= order; i++) ss = ss + V(i) * V(i);
    ss = 1.0 / std::sqrt(ss);
    for (int i = 1; i <= order; i++) V(i) = ss * V(i);
    return it;
}

// INVERR (inverr.f:4-167): error estimate beta_{m+1} |y_m| of each good,
// isolated eigenvalue (0 for the others, negative if unconverged).
void error_estimates(Tri &t, double beta_m1, const std::vector<double> &vs, const std::vector<int> &mp,
                     std::vector<double> &err)
{
    const int m = t.m, maxit = 10;
    err.assign(vs.size(), 0.0);
    double tsum = std::fabs(t.a(1));
    for (int i = 2; i <= m; i++) tsum = tsum + std::fabs(t.a(i)) + t.b(i);
    const double eps3 = kEps2 * tsum, eps4 = (double)m * eps3;
    std::vector<double> g(m), v(m + 1), mult(m + 1);
    park_miller(1, g.data(), m);
    double gsum = 0.0;
    for (int i = 0; i < m; i++) gsum = gsum + std::fabs(g[i]);
    gsum = eps4 / gsum;
    for (int i = 0; i < m; i++) g[i] = gsum * g[i];
    for (size_t j = 0; j < vs.size(); j++) {
        if (mp[j] != 1) continue;
        for (int i = 0; i < m; i++) v[i] = g[i];
        const int it = inverse_iteration(m, t.al, t.be.data(), mult, v.data(), vs[j], eps3, eps4, maxit);
        double est = beta_m1 * std::fabs(v[m - 1]);
        if (it > maxit) est = -est;
        err[j] = est;
    }
}

}  // namespace

int cw_lancbis(int niter, const double *alpha, const double *beta, const double *beta2, double emin, double emax,
               double mtol, double *eval, double *err, int *ngood)
{
    (void)beta2;  // BISEC recomputes BETA**2 (bisec.f:91-92)
    if (niter <= 0) return fail(BSL_E_ARG, "lancbis: niter must be positive");
    Tri t{niter, alpha, std::vector<double>(niter + 2, 0.0), std::vector<double>(niter + 2, 0.0)};
    for (int j = 1; j <= niter; j++) t.be[j] = beta[j];
    const double beta_m1 = beta[niter];
    std::vector<double> vs, es;
    std::vector<int> mp;
    if (bisect(t, emin, emax, mtol, vs, mp) < 0)
        return fail(BSL_E_STURM, "BISEC: estimated number of Sturm sequences exceeds its limit");
    t.be[niter] = beta_m1;
    isolate(vs, mp);
    t.be[niter] = 0.0;
    error_estimates(t, beta_m1, vs, mp, es);
    int k = 0;
    for (size_t i = 0; i < vs.size(); i++)
        if (mp[i] != 0) {
            eval[k] = vs[i];
            err[k] = es[i];
            k++;
        }
    *ngood = k;
    return BSL_OK;
}

// The same post-processing with the Sturm bisections and the inverse
// iterations on the GPU (bsl_tridiag_dev.hip); the cluster bookkeeping of
// BISEC and ISOEV stay here.  Eigenvalues agree with cw_lancbis to the
// bisection tolerance (each index is bisected from the full interval rather
// than from the bounds the previous indices refined).
int cw_lancbis_device(int niter, const double *alpha, const double *beta, double emin, double emax, double mtol,
                      hipStream_t stream, double *eval, double *err, int *ngood)
{
    if (niter <= 0) return fail(BSL_E_ARG, "lancbis: niter must be positive");
    const int m = niter;
    Tri t{m, alpha, std::vector<double>(m + 2, 0.0), std::vector<double>(m + 2, 0.0)};
    for (int j = 1; j <= m; j++) t.be[j] = beta[j];
    const double beta_m1 = beta[m];
    double tkmax = 0.0;
    for (int i = 1; i <= m; i++) tkmax = std::fmax(std::fmax(std::fabs(t.a(i)), t.b(i)), tkmax);
    const double ttol = kEps2 * tkmax, scale = (double)(m + 1000);
    const double ep0 = scale * ttol, ep1 = std::sqrt(scale) * ttol;
    const double epm = std::fmax(ep0, mtol);  // multiplicity width, as in bisect()
    t.be[m] = 0.0;
    for (int i = 0; i <= m; i++) t.b2[i] = t.be[i] * t.be[i];
    const int above_hi = sturm_forward(t, emax);
    auto cnt = [&](double x) { return sturm_forward(t, x) - above_hi; };
    const int mt = cnt(emin);
    if (30 * mt >= 100 * m) return fail(BSL_E_STURM, "BISEC: estimated number of Sturm sequences exceeds its limit");
    *ngood = 0;
    if (mt == 0) return BSL_OK;
    // device copies: alpha, beta with b(1) = b(m+1) = 0
    std::vector<double> bev(t.be.begin(), t.be.begin() + m + 1);
    double *d_al = nullptr, *d_be = nullptr, *d_out = nullptr;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(stream);
        if (d_al) (void)hipFree(d_al);
        if (d_be) (void)hipFree(d_be);
        if (d_out) (void)hipFree(d_out);
    };
    hipError_t e = hipMalloc(&d_al, sizeof(double) * m);
    if (e == hipSuccess) e = hipMalloc(&d_be, sizeof(double) * (m + 1));
    if (e == hipSuccess) e = hipMalloc(&d_out, sizeof(double) * kTriOut * (size_t)mt);
    if (e == hipSuccess) e = hipMemcpyAsync(d_al, alpha, sizeof(double) * m, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_be, bev.data(), sizeof(double) * (m + 1), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) {
        cleanup();
        return hip_fail(e, "lancbis device buffers", __FILE__, __LINE__);
    }
    int rc = launch_tri_bisect(d_al, d_be, m, emin, emax, above_hi, mt, ep0, ep1, epm, d_out, stream);
    std::vector<double> res(kTriOut * (size_t)mt);
    if (!rc) {
        e = hipMemcpyAsync(res.data(), d_out, sizeof(double) * res.size(), hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) rc = hip_fail(e, "lancbis bisection", __FILE__, __LINE__);
    }
    if (rc) {
        cleanup();
        return rc;
    }
    // BISEC's walk from the top: one value per simple eigenvalue (spurious
    // when T-hat has one in the same window), one per cluster
    std::vector<double> vs;
    std::vector<int> mp;
    int k = mt;
    while (k > 0) {
        const double *r = res.data() + kTriOut * (size_t)(k - 1);
        double x1 = r[0];
        const double xu = r[2];
        const int nu = (int)r[3], jl = (int)r[5], ju = (int)r[7];  // T-hat counts at ep0
        const int kl = (int)r[8];                                   // T counts at epm
        int ku = (int)r[9];
        vs.push_back(x1);
        int mult, knew;
        if (kl - ku - 1 == 0) {
            mult = (ju < jl) ? 0 : 1;
            knew = k - 1;
        } else {
            x1 = xu + epm;
            if (ku != nu) {
                int change;
                do {
                    x1 = x1 + epm;
                    const int nev = cnt(x1);
                    change = ku - nev;
                    ku = nev;
                } while (change != 0);
            }
            mult = kl - ku;
            knew = ku;
        }
        mp.push_back(mult);
        k = knew;
    }
    isolate(vs, mp);
    // INVERR on the device for the good, isolated values
    std::vector<double> es(vs.size(), 0.0), xs;
    std::vector<int> idx;
    for (size_t j = 0; j < vs.size(); j++)
        if (mp[j] == 1) {
            xs.push_back(vs[j]);
            idx.push_back((int)j);
        }
    if (!xs.empty()) {
        const int nx = (int)xs.size(), maxit = 10;
        double tsum = std::fabs(t.a(1));
        for (int i = 2; i <= m; i++) tsum = tsum + std::fabs(t.a(i)) + t.b(i);
        const double eps3 = kEps2 * tsum, eps4 = (double)m * eps3;
        std::vector<double> g(m);
        park_miller(1, g.data(), m);
        double gsum = 0.0;
        for (int i = 0; i < m; i++) gsum = gsum + std::fabs(g[i]);
        gsum = eps4 / gsum;
        for (int i = 0; i < m; i++) g[i] = gsum * g[i];
        double *d_g = nullptr, *d_xs = nullptr, *d_sv = nullptr, *d_sm = nullptr, *d_err = nullptr;
        unsigned char *d_sf = nullptr;
        e = hipMalloc(&d_g, sizeof(double) * m);
        if (e == hipSuccess) e = hipMalloc(&d_xs, sizeof(double) * nx);
        if (e == hipSuccess) e = hipMalloc(&d_sv, sizeof(double) * (size_t)m * nx);
        if (e == hipSuccess) e = hipMalloc(&d_sm, sizeof(double) * (size_t)m * nx);
        if (e == hipSuccess) e = hipMalloc(&d_sf, (size_t)(m + 1) * nx);
        if (e == hipSuccess) e = hipMalloc(&d_err, sizeof(double) * nx);
        if (e == hipSuccess) e = hipMemcpyAsync(d_g, g.data(), sizeof(double) * m, hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(d_xs, xs.data(), sizeof(double) * nx, hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) rc = hip_fail(e, "lancbis error estimates", __FILE__, __LINE__);
        if (!rc)
            rc = launch_tri_inverr(d_al, d_be, m, d_xs, nx, d_g, eps3, eps4, beta_m1, maxit, d_sv, d_sm, d_sf, d_err,
                                   stream);
        std::vector<double> ev(nx);
        if (!rc) {
            e = hipMemcpyAsync(ev.data(), d_err, sizeof(double) * nx, hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            if (e != hipSuccess) rc = hip_fail(e, "lancbis error estimates", __FILE__, __LINE__);
        }
        (void)hipStreamSynchronize(stream);
        for (double *p : {d_g, d_xs, d_sv, d_sm, d_err})
            if (p) (void)hipFree(p);
        if (d_sf) (void)hipFree(d_sf);
        if (!rc)
            for (int q = 0; q < nx; q++) es[idx[q]] = ev[q];
    }
    cleanup();
    if (rc) return rc;
    // vs is ascending (the walk starts from the lowest index), as BISEC's
    int kk = 0;
    for (size_t i = 0; i < vs.size(); i++)
        if (mp[i] != 0) {
            eval[kk] = vs[i];
            err[kk] = es[i];
            kk++;
        }
    *ngood = kk;
    return BSL_OK;
}

namespace {
// STURMI (sturmi.f:4-116): smallest orders at which x first appears (mk1)
// and appears twice (mk2, or m).  Returns false when x is not an eigenvalue.
// mtol > 0 (no reference counterpart) widens the window x +- tol to at least
// mtol, the multiplicity width of bisect(): the value of a merged cluster of
// spread ghost copies is then found from the order at which its first copy
// converged, not only the order at which the copy bisect() returned did.
bool sturm_orders(const Tri &t, const double *b2, double x, double mtol, int *mk1, int *mk2)
{
    const int m = t.m;
    double tkmax = 0.0;
    for (int i = 1; i <= m; i++) tkmax = std::fmax(std::fmax(std::fabs(t.a(i)), t.b(i)), tkmax);
    const double stutol = 5.0 * ((double)(m + 1000) * (kEps2 * tkmax));
    const double tol = std::fmax(std::fmax(std::fabs(x) * 1e-10, stutol), mtol);
    const double lo = x - tol, hi = x + tol;
    double d1 = 1.0, d2 = 1.0;
    int c1 = 0, c2 = 0, found = 0, diff = 0, i;
    *mk1 = 0;
    bool twice = false;
    for (i = 1; i <= m; i++) {
        const double q1 = (d1 != 0.0) ? b2[i - 1] / d1 : t.b(i) / kEps2;
        d1 = lo - t.a(i) - q1;
        if (d1 < 0.0) c1++;
        const double q2 = (d2 != 0.0) ? b2[i - 1] / d2 : t.b(i) / kEps2;
        d2 = hi - t.a(i) - q2;
        if (d2 < 0.0) c2++;
        diff = c1 - c2;
        if (diff - found >= 1) {
            if (found == 0) *mk1 = i;
            if (++found > 1) {
                twice = true;
                break;
            }
        }
    }
    if (!twice) {
        i = m;
        if (found == 0) *mk1 = m;
    }
    *mk2 = i;
    return diff != 0;
}

// Eigenvalue of the leading order x order block of T nearest to x within
// [x - w, x + w] (Sturm bisection), or x when the window holds none.
double nearest_eigenvalue(const Tri &t, const double *be, int order, double x, double w)
{
    // eigenvalues of T_order below s
    auto below = [&](double s) {
        int c = 0;
        double d = 1.0;
        for (int i = 1; i <= order; i++) {
            const double b = i > 1 ? be[i - 1] : 0.0;
            const double q = (d != 0.0) ? (b * b) / d : std::fabs(b) / kEps2;
            d = t.a(i) - s - q;
            if (d < 0.0) c++;
        }
        return c;
    };
    // the k-th eigenvalue (0-based, ascending) inside (lo, hi]
    auto kth = [&](int k, double lo, double hi) {
        for (int it = 0; it < 200 && hi - lo > kEps2 * (std::fabs(lo) + std::fabs(hi)); it++) {
            const double mid = 0.5 * (lo + hi);
            if (below(mid) > k) hi = mid; else lo = mid;
        }
        return 0.5 * (lo + hi);
    };
    const int c_lo = below(x - w), c_x = below(x), c_hi = below(x + w);
    double best = x, dist = 2.0 * w + 1.0;
    if (c_x > c_lo) {  // the largest one below x
        const double e = kth(c_x - 1, x - w, x);
        if (x - e < dist) dist = x - e, best = e;
    }
    if (c_hi > c_x) {  // the smallest one above x
        const double e = kth(c_x, x, x + w);
        if (e - x < dist) dist = e - x, best = e;
    }
    return best;
}

// INVERM (inverm.f:4-142): eigenvector of T_order for x; returns |y_order|.
// shift_w > 0 (no reference counterpart, the multiplicity width of a
// mixed-precision eigensolve): the shift is T_order's own eigenvalue nearest
// to x within shift_w, because x, one copy of a merged cluster, can lie
// farther from the copy T_order holds than the inverse iteration's
// convergence test (a 1-norm of 1 from eps4) tolerates.
void tm_eigenvector(const Tri &t, double *be, int order, double x, double shift_w, const std::vector<double> &g,
                    double *v, std::vector<double> &mult, double *error, bool *converged)
{
    const int maxit = 10;
    if (shift_w > 0.0) x = nearest_eigenvalue(t, be, order, x, shift_w);
    const double keep = be[order];
    be[order] = 0.0;
    double tsum = std::fabs(t.a(1));
    for (int i = 2; i <= order; i++) tsum = tsum + std::fabs(t.a(i)) + be[i - 1];
    const double eps3 = kEps2 * tsum, eps4 = (double)order * eps3;
    double gsum = 0.0;
    for (int i = 0; i < order; i++) gsum = gsum + std::fabs(g[i]);
    gsum = eps4 / gsum;
    for (int i = 0; i < order; i++) v[i] = gsum * g[i];
    const int it = inverse_iteration(order, t.al, be, mult, v, x, eps3, eps4, maxit);
    *error = std::fabs(v[order - 1]);
    *converged = it <= maxit;
    be[order] = keep;
}
}  // namespace

// TRIVEC (trivec.f:4-95).  For each eigenvalue, orders between MK1 and MK2
// are tried and the one with the smallest |y_order| kept.  As in the
// reference, entries of a column beyond the chosen order keep the values of
// larger trial orders (INVERM writes only its own order).
int cw_trivec(int niter, const double *alpha, const double *beta, const double *beta2, const double *eval,
              int ngood, double mtol, double *evtr, int *mamin)
{
    if (niter <= 0 || ngood < 0) return fail(BSL_E_ARG, "trivec: bad sizes");
    Tri t{niter, alpha, std::vector<double>(niter + 2, 0.0), std::vector<double>()};
    for (int j = 1; j <= niter; j++) t.be[j] = beta[j];
    std::vector<double> b2(niter + 1, 0.0);
    for (int j = 1; j <= niter; j++) b2[j] = beta2[j];
    std::vector<double> g(niter), mult(niter + 1);
    park_miller(1, g.data(), niter);
    for (size_t i = 0; i < (size_t)niter * ngood; i++) evtr[i] = 0.0;
    for (int ig = 0; ig < ngood; ig++) {
        mamin[ig] = 0;
        int mk1, mk2;
        if (!sturm_orders(t, b2.data(), eval[ig], mtol, &mk1, &mk2))
            return fail(BSL_E_NOTFOUND, "STURMI: eigenvalue not found");
        const double shift_w = mtol > 0.0 ? std::fmax(mtol, std::fabs(eval[ig]) * 1e-10) : 0.0;
        int order, last, step;
        if (mk2 == niter) {
            order = std::min((5 * mk1) / 4 + 1, niter);
            last = niter;
            step = (std::min(niter, (13 * mk1) / 8 + 1) - order) / 10 + 1;
        } else {
            order = (3 * mk1 + mk2) / 4 + 1;
            last = (mk1 + 3 * mk2) / 4;
            step = ((3 * mk1 + 5 * mk2) / 8 + 1 - order) / 10 + 1;
        }
        double *col = evtr + (size_t)ig * niter;
        double best = 1.0;
        bool any = false;
        for (int tries = 1;; tries++) {
            double e;
            bool conv;
            tm_eigenvector(t, t.be.data(), order, eval[ig], shift_w, g, col, mult, &e, &conv);
            any = any || conv;
            if (e <= best) {
                best = e;
                mamin[ig] = order;
            }
            order += step;
            if (!(order < last && tries <= 10)) break;
        }
        if (!any) return fail(BSL_E_NOCONV, "TRIVEC: inverse iteration never converged");
        double e;
        bool conv;
        tm_eigenvector(t, t.be.data(), mamin[ig], eval[ig], shift_w, g, col, mult, &e, &conv);
    }
    return BSL_OK;
}

}  // namespace bsl
