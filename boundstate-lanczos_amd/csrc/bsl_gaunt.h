// bsl_gaunt.h -- the potential of a parity block in harmonic form.
//
// A NoQuad potential block is V_P[i][j] = sum_q w_q Y_i(q) Y_j(q) V(P, q)
// over the angular quadrature nodes q (calc_Vlmlpmp_NoQuad,
// linRotCartSphHvFullVPrepRoutines.cpp:116-131).  The product of two real
// spherical harmonics of one l parity is itself a finite sum of even-L real
// harmonics, Y_i Y_j = sum_c G[i][j][c] Yh_c with L_c <= l_i + l_j, an identity
// of functions that holds at every node of every grid.  Hence
//     V_P[i][j] = sum_c G[i][j][c] vh_c(P),  vh_c(P) = sum_q w_q Yh_c(q) V(P, q),
// exactly, whatever the quadrature: a block of B(B+1)/2 values per point is
// fixed by the (lb+1)(2 lb+1) harmonic coefficients vh_c of its point (66
// instead of 351 for the lmax 5 blocks).  G is sparse; which (i, j, c) can be
// non-zero follows from the selection rules of real harmonics, so the term
// list is known at compile time, and its values come from an exact
// quadrature at set-up.
//
// Everything here is constexpr so that the host (which fills G in this
// order) and the kernels (which unroll the same list) share one definition.
#pragma once

namespace bsl {
namespace gaunt {

constexpr int iabs(int v) { return v < 0 ? -v : v; }

// Planes of a parity block: every l = par (mod 2) up to lb, in the reference
// lm order (genIndices_lm, linRotCartSphHvRoutines.cpp:24-68: m outer from
// -lmax, l inner from |m|); the global lmax does not change the relative order.
constexpr int block_size(int lb, int par)
{
    int n = 0;
    for (int m = -lb; m <= lb; m++)
        for (int l = iabs(m); l <= lb; l++)
            if (l % 2 == par) n++;
    return n;
}
constexpr int block_lm(int lb, int par, int k, bool want_m)
{
    int n = 0;
    for (int m = -lb; m <= lb; m++)
        for (int l = iabs(m); l <= lb; l++)
            if (l % 2 == par) {
                if (n == k) return want_m ? m : l;
                n++;
            }
    return -1;
}
constexpr int block_l(int lb, int par, int k) { return block_lm(lb, par, k, false); }
constexpr int block_m(int lb, int par, int k) { return block_lm(lb, par, k, true); }

// harmonic columns: even L = 0, 2, .., 2 lb; M = -L .. L
constexpr int ncoef(int lb) { return (lb + 1) * (2 * lb + 1); }
constexpr int coef_index(int L, int M) { return (L / 2) * (L - 1) + M + L; }
constexpr int coef_L(int c)
{
    int L = 0;
    while (coef_index(L + 2, -(L + 2)) <= c) L += 2;
    return L;
}
constexpr int coef_M(int c) { return c - coef_index(coef_L(c), 0); }

// Can the integral of Y_{l1 m1} Y_{l2 m2} Y_{L M} over the sphere be
// non-zero?  Real harmonics: cos(m phi) for m > 0, sin(|m| phi) for m < 0,
// constant for m = 0.  The phi integral needs |M| = |m1| + |m2| or
// ||m1| - |m2|| and an even number of sines; the theta integral needs
// l1 + l2 + L even and the triangle |l1 - l2| <= L <= l1 + l2.  Accidental
// zeros of the remaining terms are kept (their coefficients are zero).
constexpr bool coupled(int l1, int m1, int l2, int m2, int L, int M)
{
    if ((l1 + l2 + L) % 2 != 0) return false;
    if (L < iabs(l1 - l2) || L > l1 + l2) return false;
    const int a = iabs(m1), b = iabs(m2), c = iabs(M);
    if (c != a + b && c != iabs(a - b)) return false;
    const int nsin = (m1 < 0 ? 1 : 0) + (m2 < 0 ? 1 : 0) + (M < 0 ? 1 : 0);
    return nsin % 2 == 0;
}

// One term of the block's expansion: V[i][j] += G * vh[c].  Terms are listed
// pair by pair (i <= j, i outer, j inner), columns ascending; `first` and
// `last` mark the ends of a pair's run; g is the term's place in the
// block's full list (where its value is stored).
struct Term {
    int i = 0, j = 0, c = 0, g = 0;
    bool first = false, last = false;
};

// Pair types of a sub-list: both planes cos-type (m >= 0), both sin-type
// (m < 0), or one of each.  Same-type pairs couple through the M >= 0
// harmonics only and cross pairs through M < 0 only (an even number of
// sines), so the sub-lists read disjoint coefficient columns.
enum { kAll = 0, kCC = 1, kSS = 2, kCross = 3 };
constexpr int pair_type(int mi, int mj) { return (mi < 0) != (mj < 0) ? kCross : (mi < 0 ? kSS : kCC); }
constexpr bool in_list(int type, int mi, int mj) { return type == kAll || pair_type(mi, mj) == type; }

// (l, m) of the block's planes and (L, M) of its harmonic columns
struct Shape {
    int B = 0, NC = 0;
    int l[64] = {}, m[64] = {}, L[128] = {}, M[128] = {};
    constexpr Shape(int lb, int par)
    {
        for (int mm = -lb; mm <= lb; mm++)
            for (int ll = iabs(mm); ll <= lb; ll++)
                if (ll % 2 == par) {
                    l[B] = ll;
                    m[B] = mm;
                    B++;
                }
        for (int LL = 0; LL <= 2 * lb; LL += 2)
            for (int MM = -LL; MM <= LL; MM++) {
                L[NC] = LL;
                M[NC] = MM;
                NC++;
            }
    }
};

constexpr int count_terms(int lb, int par, int type = kAll)
{
    const Shape s(lb, par);
    int n = 0;
    for (int i = 0; i < s.B; i++)
        for (int j = i; j < s.B; j++)
            if (in_list(type, s.m[i], s.m[j]))
                for (int c = 0; c < s.NC; c++)
                    if (coupled(s.l[i], s.m[i], s.l[j], s.m[j], s.L[c], s.M[c])) n++;
    return n;
}

// the block's term list (or the sub-list of one pair type) in order, with
// the rows and columns it touches
template <int NT>
struct TermList {
    Term t[NT > 0 ? NT : 1];
    bool row[64] = {}, col[128] = {};
    constexpr TermList(int lb, int par, int type) : t{}
    {
        const Shape s(lb, par);
        int n = 0, g = 0;
        for (int i = 0; i < s.B; i++)
            for (int j = i; j < s.B; j++) {
                const bool take = in_list(type, s.m[i], s.m[j]);
                const int f = n;
                for (int c = 0; c < s.NC; c++)
                    if (coupled(s.l[i], s.m[i], s.l[j], s.m[j], s.L[c], s.M[c])) {
                        if (take) {
                            t[n].i = i;
                            t[n].j = j;
                            t[n].c = c;
                            t[n].g = g;
                            t[n].first = n == f;
                            row[i] = row[j] = col[c] = true;
                            n++;
                        }
                        g++;
                    }
                if (n > f) t[n - 1].last = true;
            }
    }
};

template <int LB, int PAR, int TYPE = kAll>
struct Pattern {
    static constexpr int lb = LB;
    static constexpr int par = PAR;
    static constexpr int B = block_size(LB, PAR);
    static constexpr int NC = ncoef(LB);
    static constexpr int NT = count_terms(LB, PAR, TYPE);
    static constexpr TermList<NT> list{LB, PAR, TYPE};
};

// Split kernels take every row of a block from one same-type partial
// (cos-cos rows m >= 0, sin-sin rows m < 0) and one cross partial: each row
// must appear in the cross sub-list and in its same-type sub-list.
template <int LB, int PAR>
constexpr bool split_covers_rows()
{
    for (int r = 0; r < block_size(LB, PAR); r++) {
        if (!Pattern<LB, PAR, kCross>::list.row[r]) return false;
        if (!Pattern<LB, PAR, kCC>::list.row[r] && !Pattern<LB, PAR, kSS>::list.row[r]) return false;
    }
    return true;
}

// largest block held in harmonic form: its coefficients, vector entries and
// accumulators stay in registers (66 + 2 x 21 doubles at lb = 5)
constexpr int kMaxLb = 5;

}  // namespace gaunt
}  // namespace bsl
