// bsl_tables.cpp -- host-side basis tables of the calculator: lm index map,
// Gauss-Legendre / Gauss-Chebyshev quadrature, normalised associated Legendre
// functions, tesseral trig factors and the sinc-DVR kinetic row.  These are
// small (at most a few thousand entries) and are built once per context.
// The arithmetic follows the reference expression by expression so that the
// tables agree with it to the last bit.
#include <cfloat>
#include <cmath>
#include <algorithm>
#include <cstdlib>

#include "bsl_internal.h"

namespace bsl {

static constexpr double kPi = 3.141592653589793;  // lanczosUnits.h:6
static constexpr double kHbar = 0.0635077857002;  // lanczosUnits.h:9

// Newton iteration on P_n for the positive roots, mirrored
// (gauleg, linRotCartSphHvRoutines.cpp:219-255).
static void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w)
{
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    const double xm = 0.5 * (1.0 + -1.0), xl = 0.5 * (1.0 - -1.0);
    for (int i = 1; i <= (n + 1) / 2; i++) {
        double z = std::cos(M_PI * ((double)i - 0.25) / ((double)n + 0.5));
        double zprev, deriv;
        for (;;) {
            double p_cur = 1.0, p_prev = 0.0;
            for (int j = 1; j <= n; j++) {
                double p_prev2 = p_prev;
                p_prev = p_cur;
                p_cur = ((2.0 * (double)j - 1.0) * z * p_prev - ((double)j - 1.0) * p_prev2) / (double)j;
            }
            deriv = (double)n * (z * p_cur - p_prev) / (z * z - 1.0);
            zprev = z;
            z = zprev - p_cur / deriv;
            if (!(std::fabs(z - zprev) > DBL_EPSILON)) break;
        }
        x[i - 1] = xm - xl * z;
        x[n - i] = xm + xl * z;
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * deriv * deriv);
        w[n - i] = w[i - 1];
    }
}

static double fact(int k)
{
    double f = 1.0;
    for (int i = 2; i <= k; i++) f *= (double)i;
    return f;
}

// normalised P_l^{|m|}(x) for every reference lm index
// (normAssocLegendrePoly, linRotCartSphHvRoutines.cpp:293-403)
static void legendre_row(int lmax, const std::vector<int> &ls, const std::vector<int> &ms, double x,
                         double *out)
{
    const int L1 = lmax + 1;
    std::vector<double> p((size_t)L1 * L1, 0.0);
    auto at = [&](int l, int m) -> double & { return p[(size_t)l * L1 + m]; };
    const double s = std::sqrt(1.0 - x * x);
    at(0, 0) = 1.0;
    for (int l = 1; l <= lmax; l++) at(l, l) = (2.0 * (double)l - 1.0) * at(l - 1, l - 1) * s;
    for (int m = 0; m < lmax; m++)
        for (int l = m; l < lmax; l++) {
            const double below = (l - 1 < m) ? 0.0 : at(l - 1, m);
            at(l + 1, m) = (((2.0 * (double)l + 1.0) * x * at(l, m)) - (((double)l + (double)m) * below)) /
                           ((double)l - (double)m + 1.0);
        }
    for (int l = 0; l <= lmax; l++) {
        double sign = 1.0;
        for (int m = 0; m <= l; m++) {
            const double norm = fact(l - m) * (2.0 * (double)l + 1.0) / (2.0 * fact(l + m));
            at(l, m) *= sign * std::sqrt(norm);
            sign *= -1.0;
        }
    }
    for (size_t n = 0; n < ls.size(); n++) out[n] = at(ls[n], std::abs(ms[n]));
}

// tesseral trig factor (tesseralTrigTerm, linRotCartSphHvRoutines.cpp:405-434)
static double trig(int m, double phi)
{
    const double inv = 1.0 / std::sqrt(kPi);
    if (m == 0) return inv / std::sqrt(2.0);
    if (m > 0) return inv * std::cos((double)m * phi);
    return inv * std::sin((double)(-m) * phi);
}

void build_tables(int lmax, int na, int nb, Tables &t)
{
    t.lmax = lmax;
    t.na = na;
    t.nb = nb;
    t.l.clear();
    t.m.clear();
    // genIndices_lm (linRotCartSphHvRoutines.cpp:24-68): m outer, l inner
    for (int m = -lmax; m <= lmax; m++)
        for (int l = std::abs(m); l <= lmax; l++) {
            t.l.push_back(l);
            t.m.push_back(m);
        }
    t.nn = (int)t.l.size();
    // QuadraturePrep (linRotCartSphHvRoutines.cpp:1480-1532)
    gauss_legendre(na, t.ct, t.wt);
    t.th.resize(na);
    for (int a = 0; a < na; a++) t.th[a] = std::acos(t.ct[a]);
    t.cp.resize(nb);
    t.wp.resize(nb);
    t.ph1.resize(nb);
    t.ph2.resize(nb);
    for (int b = 0; b < nb; b++) {  // gaussChebyshev :450-463
        t.cp[b] = std::cos(kPi * (2.0 * (double)(b + 1) - 1.0) / (2.0 * (double)nb));
        t.wp[b] = kPi / (double)nb;
        t.ph1[b] = std::acos(t.cp[b]);
        t.ph2[b] = 2 * kPi - std::acos(t.cp[b]);
    }
    // TesseralPrep (:1534-1735)
    t.L.assign((size_t)t.nn * na, 0.0);
    std::vector<double> row(t.nn);
    for (int a = 0; a < na; a++) {
        legendre_row(lmax, t.l, t.m, t.ct[a], row.data());
        for (int n = 0; n < t.nn; n++) t.L[(size_t)n * na + a] = row[n];
    }
    const int nm = 2 * lmax + 1;
    t.S1.assign((size_t)nm * nb, 0.0);
    t.S2.assign((size_t)nm * nb, 0.0);
    for (int m = -lmax; m <= lmax; m++)
        for (int b = 0; b < nb; b++) {
            t.S1[(size_t)(m + lmax) * nb + b] = trig(m, std::acos(t.cp[b]));
            t.S2[(size_t)(m + lmax) * nb + b] = trig(m, 2.0 * kPi - std::acos(t.cp[b]));
        }
}

// Colbert-Miller sinc DVR (cartKinGrid, linRotCartSphHvRoutines.cpp:484-517).
// T[i][j] depends on i - j only, so its first row defines it.
void kinetic_row(double x_max, int n, double mass, std::vector<double> &row, std::vector<double> &grid)
{
    const double dx = x_max / (double)(n - 1);
    row.assign(n, 0.0);
    grid.assign(n, 0.0);
    for (int d = 0; d < n; d++) {
        double v = (kHbar * kHbar) / (2.0 * mass * dx * dx) * std::pow(-1.0, (double)(1 - (d + 1)));
        if (d == 0)
            v *= (kPi * kPi) / 3.0;
        else
            v *= 2.0 / (double)((1 - (d + 1)) * (1 - (d + 1)));
        row[d] = v;
    }
    for (int i = 0; i < n; i++) grid[i] = (double)i * dx - x_max / 2.0;
}

// harmonic columns of a block of lb: L = 0, 2, .., 2 lb (bsl_gaunt.h order)
static void harmonic_columns(int lb, std::vector<int> &Ls, std::vector<int> &Ms)
{
    Ls.clear();
    Ms.clear();
    for (int c = 0; c < gaunt::ncoef(lb); c++) {
        Ls.push_back(gaunt::coef_L(c));
        Ms.push_back(gaunt::coef_M(c));
    }
}

void harmonic_table(const Tables &t, int lb, std::vector<double> &hh)
{
    std::vector<int> Ls, Ms;
    harmonic_columns(lb, Ls, Ms);
    const int nc = (int)Ls.size(), na = t.na, nb = t.nb;
    std::vector<double> lrow(nc);
    hh.assign((size_t)2 * na * nb * nc, 0.0);
    for (int a = 0; a < na; a++) {
        legendre_row(2 * lb, Ls, Ms, t.ct[a], lrow.data());
        for (int half = 0; half < 2; half++)
            for (int b = 0; b < nb; b++) {
                // the nodes of the V-block table (S1 and S2 of build_tables)
                const double phi = half == 0 ? std::acos(t.cp[b]) : 2.0 * kPi - std::acos(t.cp[b]);
                const size_t q = (size_t)half * na * nb + (size_t)a * nb + b;
                for (int c = 0; c < nc; c++) hh[q * nc + c] = t.wt[a] * t.wp[b] * lrow[c] * trig(Ms[c], phi);
            }
    }
}

// G[i][j][c] = integral of Y_i Y_j Yh_c over the sphere with the functions of
// the tables above: Gauss-Legendre in cos(theta) with 2 lb + 2 nodes (exact
// to degree 4 lb + 3) and 4 lb + 4 equally spaced phi nodes (exact for
// frequencies below 4 lb + 4); the integrands have degree at most 4 lb.
void gaunt_values(int lb, int par, std::vector<double> &g, double *excluded)
{
    const int B = gaunt::block_size(lb, par), nc = gaunt::ncoef(lb);
    std::vector<int> ls, ms, Ls, Ms;
    for (int k = 0; k < B; k++) {
        ls.push_back(gaunt::block_l(lb, par, k));
        ms.push_back(gaunt::block_m(lb, par, k));
    }
    harmonic_columns(lb, Ls, Ms);
    const int nx = 2 * lb + 2, nphi = 4 * lb + 4;
    std::vector<double> x, w;
    gauss_legendre(nx, x, w);
    std::vector<double> full((size_t)B * B * nc, 0.0), yb(B), yc(nc), lb_row(B), lc_row(nc);
    for (int a = 0; a < nx; a++) {
        legendre_row(2 * lb, ls, ms, x[a], lb_row.data());
        legendre_row(2 * lb, Ls, Ms, x[a], lc_row.data());
        for (int f = 0; f < nphi; f++) {
            const double phi = kPi * (2.0 * f + 1.0) / (double)nphi, wq = w[a] * 2.0 * kPi / (double)nphi;
            for (int k = 0; k < B; k++) yb[k] = lb_row[k] * trig(ms[k], phi);
            for (int c = 0; c < nc; c++) yc[c] = lc_row[c] * trig(Ms[c], phi);
            for (int i = 0; i < B; i++)
                for (int j = 0; j < B; j++) {
                    const double yy = wq * yb[i] * yb[j];
                    double *o = &full[((size_t)i * B + j) * nc];
                    for (int c = 0; c < nc; c++) o[c] += yy * yc[c];
                }
        }
    }
    g.clear();
    double ex = 0.0;
    for (int i = 0; i < B; i++)
        for (int j = i; j < B; j++)
            for (int c = 0; c < nc; c++) {
                const double v = full[((size_t)i * B + j) * nc + c];
                if (gaunt::coupled(ls[i], ms[i], ls[j], ms[j], Ls[c], Ms[c]))
                    g.push_back(v);
                else
                    ex = std::max(ex, std::fabs(v));
            }
    *excluded = ex;
}

}  // namespace bsl
