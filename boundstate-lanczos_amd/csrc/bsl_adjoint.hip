// bsl_adjoint.hip -- Hellmann-Feynman gradients of selected eigenvalues.
//
// For an eigenpair (E, psi) of H = T + V with the potential acting on the
// quadrature grid (V_nn'(P) = sum_q Y_n(q) w_q V(P, q) Y_n'(q), the
// contraction of linRotCartSphHvContGridRoutines.cpp:7-152),
//   dE/dtheta = <psi| dV/dtheta |psi> = sum_{P,q} rho(P, q) dV(P, q)/dtheta,
//   rho(P, q) = w_q (sum_n Y_n(q) psi(P, n))^2,
// so rho is the gradient of E with respect to the potential grid.  Points
// where the ceiling clamp is active (V >= ceiling -> ceiling,
// Alavi_ContGrid_H2_Routines.cpp:101-139) do not depend on the parameters.
// With several eigenvectors and weights g_k (the output gradients of a
// caller), omega = sum_k g_k rho_k [V < ceiling] is accumulated once and
// contracted with the derivatives of the Alavi/water pair terms
// (Alavi_ContGrid_H2_Routines.cpp:6-147): site charges, site positions, the
// two Alavi charges and the LJ coefficients A, B.
#include <hip/hip_runtime.h>

#include "bsl_internal.h"

namespace bsl {

constexpr int kAWG = 256;
constexpr double kCoulombK = 138.9354846111096;

// Inversion pairs of nodes.  The Q = 2 na nb nodes are (half, a, b): theta
// node a (Gauss-Legendre, symmetric about pi/2) and phi = +-phi_b of the
// Chebyshev angles phi_b = pi (2b + 1) / (2 nb) (symmetric about pi / 2), at
// q = half na nb + a nb + b.  Inverting the molecular axis,
// (theta, phi) -> (pi - theta, phi + pi), maps node q < Q/2 onto node
// Q - 1 - q: the same H2 molecule with its two atoms exchanged, so both
// nodes have the same potential, the same pair derivatives, and one
// evaluation with weight omega(q) + omega(Q - 1 - q) serves both.
//
// The kept pairs of each group of npb consecutive points, in pair order:
// entry (pl << 24) | q for node q < Q/2 of the group's point pl, kept when
// q or Q - 1 - q is unclamped.  k_pair_count counts them per group (a ballot
// per wave, summed over the waves in LDS); k_pair_fill writes them from the
// group's offset, a ballot prefix per kAWG pairs keeping the pair order.  One
// workgroup per group.
__device__ inline bool hf_pair_kept(const uint64_t *__restrict__ unclamped, int64_t p, int Q, int q)
{
    const size_t i0 = (size_t)p * Q + q, i1 = (size_t)p * Q + (Q - 1 - q);
    return ((unclamped[i0 >> 6] >> (i0 & 63)) & 1) || ((unclamped[i1 >> 6] >> (i1 & 63)) & 1);
}

__global__ __launch_bounds__(kAWG) void k_pair_count(const uint64_t *__restrict__ unclamped, int64_t P, int Q,
                                                     int npb, int *__restrict__ count)
{
    __shared__ int wcnt[kAWG / 64];
    const int64_t p0 = (int64_t)blockIdx.x * npb;
    const int np = (int)(P - p0 < npb ? P - p0 : npb), Qh = Q / 2;
    const int npairs = np * Qh, lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int n = 0;  // the wave's count, the same in every lane
    int pl = (int)threadIdx.x / Qh, q = (int)threadIdx.x - pl * Qh;  // pair t0 + threadIdx.x
    for (int t0 = 0; t0 < npairs; t0 += kAWG) {
        const bool keep = t0 + (int)threadIdx.x < npairs && hf_pair_kept(unclamped, p0 + pl, Q, q);
        n += __popcll(__ballot(keep));
        for (q += kAWG; q >= Qh; q -= Qh) pl++;
    }
    if (lane == 0) wcnt[wid] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        int total = 0;
        for (int v = 0; v < kAWG / 64; v++) total += wcnt[v];
        count[blockIdx.x] = total;
    }
}

__global__ __launch_bounds__(kAWG) void k_pair_fill(const uint64_t *__restrict__ unclamped, int64_t P, int Q,
                                                    int npb, const int64_t *__restrict__ first,
                                                    uint32_t *__restrict__ pairs)
{
    __shared__ int wcnt[kAWG / 64];
    const int64_t p0 = (int64_t)blockIdx.x * npb;
    const int np = (int)(P - p0 < npb ? P - p0 : npb), Qh = Q / 2;
    const int npairs = np * Qh, lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t *out = pairs + first[blockIdx.x];
    int fill = 0;
    int pl = (int)threadIdx.x / Qh, q = (int)threadIdx.x - pl * Qh;  // pair t0 + threadIdx.x
    for (int t0 = 0; t0 < npairs; t0 += kAWG) {
        const bool keep = t0 + (int)threadIdx.x < npairs && hf_pair_kept(unclamped, p0 + pl, Q, q);
        const uint64_t bal = __ballot(keep);
        if (lane == 0) wcnt[wid] = __popcll(bal);
        __syncthreads();
        int woff = 0, total = 0;
#pragma unroll
        for (int v = 0; v < kAWG / 64; v++) {
            const int n = wcnt[v];
            woff += v < wid ? n : 0;
            total += n;
        }
        if (keep) out[fill + woff + __popcll(bal & ((1ull << lane) - 1ull))] = ((uint32_t)pl << 24) | (uint32_t)q;
        fill += total;
        for (q += kAWG; q >= Qh; q -= Qh) pl++;
        __syncthreads();
    }
}

// Up to kHfSets sets of eigenvectors at once (set m is vectors
// [first[m], first[m + 1]) of psi: [nvec][nk][P]).  For entry e of the kept
// node-pair list (point p, node q < Q/2, partner Q - 1 - q):
//   omega_m(p, n) = [n unclamped] w_n sum_{k in m} g_k (sum_r Y[r][n] psi_k[r][p])^2,
//   om[m][e] = omega_m(p, q) + omega_m(p, Q - 1 - q),
// and wcm[m][p] = the sum of om[m][e] over the entries of point p (= the
// sum of omega over every node of p: clamped nodes add zero), the weight of
// the H2 centre term.  One workgroup per group of kHfPoints points, their
// vectors' coefficients in LDS (zero rows pad nvec to a multiple of 4); a
// thread takes entries e = threadIdx.x + kAWG j and the vectors four at a
// time, so each Y[r][n] is read once per four vectors.  Every set's values
// take the operations of a launch with that set alone (sums over r in
// order, vectors in order; the point sums over a thread's entries in order,
// then over the threads in a fixed order).
constexpr int kHfSets = kHfDensitySets;  // most sets of eigenvectors per density launch
struct HfSets {
    int first[kHfSets + 1];
};

// omega of the node at grid index idx (node n of point p) into o[m]
__device__ inline void hf_node_omega(double *o, const double *__restrict__ c, int nvec, int nk, const HfSets &sets,
                                     int nset, const double *__restrict__ g, const double *__restrict__ Y,
                                     const double *__restrict__ wq, const uint64_t *__restrict__ unclamped,
                                     size_t idx, int n, int Q)
{
#pragma unroll
    for (int m = 0; m < kHfSets; m++) o[m] = 0.0;
    if (!((unclamped[idx >> 6] >> (idx & 63)) & 1)) return;
    for (int k0 = 0; k0 < nvec; k0 += 4) {
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        for (int r = 0; r < nk; r++) {
            const double y = Y[(size_t)r * Q + n];
#pragma unroll
            for (int j = 0; j < 4; j++) s[j] += y * c[(size_t)(k0 + j) * nk + r];
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = k0 + j;
#pragma unroll
            for (int m = 0; m < kHfSets; m++)
                if (m < nset && k >= sets.first[m] && k < sets.first[m + 1]) o[m] += g[k] * s[j] * s[j];
        }
    }
#pragma unroll
    for (int m = 0; m < kHfSets; m++) o[m] *= wq[n];
}

__global__ __launch_bounds__(kAWG) void k_hf_density(const double *__restrict__ psi, int nvec, HfSets sets, int nset,
                                                     const double *__restrict__ g, const double *__restrict__ Y,
                                                     const double *__restrict__ wq,
                                                     const uint64_t *__restrict__ unclamped,
                                                     const uint32_t *__restrict__ pairs,
                                                     const int64_t *__restrict__ first, int64_t ne,
                                                     double *__restrict__ om, double *__restrict__ wcm, int64_t P,
                                                     int nk, int Q)
{
    extern __shared__ double cpsi[];  // [kHfPoints][nvec rounded up to 4][nk]
    __shared__ double red[kAWG / 64][kHfPoints * kHfSets];
    const int64_t p0 = (int64_t)blockIdx.x * kHfPoints;
    const int np = (int)(P - p0 < kHfPoints ? P - p0 : kHfPoints);
    const int nv4 = (nvec + 3) & ~3, per = nv4 * nk;
    for (int t = threadIdx.x; t < np * per; t += blockDim.x) {
        const int pl = t / per, kr = t - pl * per, k = kr / nk, r = kr - k * nk;
        cpsi[t] = k < nvec ? psi[((size_t)k * nk + r) * P + p0 + pl] : 0.0;
    }
    __syncthreads();
    const int64_t beg = first[blockIdx.x];
    const int nkeep = (int)(first[blockIdx.x + 1] - beg);
    double cw[kHfPoints][kHfSets];
#pragma unroll
    for (int j = 0; j < kHfPoints; j++)
#pragma unroll
        for (int m = 0; m < kHfSets; m++) cw[j][m] = 0.0;
    for (int u = threadIdx.x; u < nkeep; u += kAWG) {
        const uint32_t e = pairs[beg + u];
        const int q = (int)(e & 0xffffffu), pl = (int)(e >> 24);
        const size_t base = (size_t)(p0 + pl) * Q;
        double o0[kHfSets], o1[kHfSets];
        hf_node_omega(o0, cpsi + (size_t)pl * per, nvec, nk, sets, nset, g, Y, wq, unclamped, base + q, q, Q);
        hf_node_omega(o1, cpsi + (size_t)pl * per, nvec, nk, sets, nset, g, Y, wq, unclamped, base + (Q - 1 - q),
                      Q - 1 - q, Q);
#pragma unroll
        for (int m = 0; m < kHfSets; m++) {
            const double w = o0[m] + o1[m];
            if (m < nset) om[m * ne + beg + u] = w;
#pragma unroll
            for (int j = 0; j < kHfPoints; j++)
                if (j == pl) cw[j][m] += w;
        }
    }
    // point sums: within each wave by a fixed shuffle tree, then the waves in order
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < kHfPoints; j++)
#pragma unroll
        for (int m = 0; m < kHfSets; m++) {
            double v = cw[j][m];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
            if (lane == 0) red[wid][j * kHfSets + m] = v;
        }
    __syncthreads();
    if ((int)threadIdx.x < kHfPoints * kHfSets) {
        const int j = threadIdx.x / kHfSets, m = threadIdx.x % kHfSets;
        double v = 0.0;
        for (int w = 0; w < kAWG / 64; w++) v += red[w][threadIdx.x];
        if (j < np && m < nset) wcm[m * P + p0 + j] = v;
    }
}

__device__ inline double a_block_sum(double v, double *lds)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += lds[w];
    return s;  // valid in thread 0
}

// 1/sqrt(x) for x > 0: the hardware estimate (about 23 correct bits) and two
// Newton steps y <- y (3/2 - x y^2 / 2), each of which squares the relative
// error, so the result is within a few ulp of the rounded value.
__device__ inline double a_rsqrt(double x)
{
    const double h = 0.5 * x;
    double y = __builtin_amdgcn_rsq(x);
    y = fma(y, fma(-h * y, y, 0.5), y);
    y = fma(y, fma(-h * y, y, 0.5), y);
    return y;
}

// Per workgroup (npb consecutive points) and set m < NS of eigenvectors
// (omega of set m at om + m P Q):
// partial[wg][m][4*s + {0,1,2,3}] = d/d(q_s, x_s, y_s, z_s), then
// partial[wg][m][4*ns + {0,1,2,3}] = d/d(q_h2_h, q_h2_cm, A, B).
// Only unclamped pairs (point, node) contribute (omega = 0 at clamped nodes,
// most of a box that reaches into the cage walls), and the two nodes of an
// inversion pair share one evaluation (weight omega(q) + omega(Q - 1 - q),
// the geometry of q < Q/2; k_hf_density's om and wcm): the workgroup walks
// its list of kept node pairs (k_pair_fill, pair order) in chunks of kAWG,
// each thread
// staging one pair (the sets' omegas and the two H positions) into one half
// of a double-buffered LDS stage; the loads of the next chunk are issued
// before the current one is evaluated, and one barrier per chunk suffices
// (a half is rewritten two chunks later, after every wave has passed the
// barrier that follows its evaluation).  Thread t accumulates site
// sb + t % 64 over a quarter (t / 64) of each chunk.  The chunks depend on
// the clamp mask alone, not on the sets.
// Per site and set only two sums are carried over the H atoms h of the pairs,
//   S = sum omega / |h - s|,   G = sum omega (h - s) / |h - s|^3,
// since d/dq_s = k q_h S, its share of d/dq_h = k q_s S and d/ds = k q_h q_s G;
// the geometry (|h - s|^-1 of both atoms, and their sum and cubed-weighted
// offset) is shared by the sets, so a set adds four multiply-adds per pair,
// and each set's sums take the same operations in the same order as with
// that set alone.  The quarters combine in a fixed order.  The H2 centre term only
// depends on the point: its weight is omega_cm(p) = sum_q omega(p, q).
constexpr int kHfSites = 64;  // sites per pass

template <int NS>
__global__ __launch_bounds__(kAWG) void k_hf_grad(const double *__restrict__ om, int64_t ne,
                                                  const double *__restrict__ wcm_g, int64_t P,
                                                  const double *__restrict__ gx, const double *__restrict__ gy,
                                                  const double *__restrict__ gz, int nx, int ny,
                                                  const double *__restrict__ off, const double *__restrict__ site_g,
                                                  const int *__restrict__ lj_g, int ns, double q_h, double q_cm,
                                                  double lj_a, double lj_b, int npb,
                                                  const uint32_t *__restrict__ pairs,
                                                  const int64_t *__restrict__ first, double *__restrict__ partial)
{
    extern __shared__ double site[];  // [ns][4], omega_cm [NS][npb], point coordinates [npb][3], int lj [ns]
    double *wcm = site + 4 * ns;
    double *pc = wcm + NS * npb;
    int *lj = (int *)(pc + 3 * npb);
    __shared__ double red[kAWG / 64];
    __shared__ double stw[NS][2 * kAWG];  // staged pairs: omega of each set
    __shared__ double2 stg[3][2 * kAWG];        // and the H positions (x1, y1), (z1, x2), (y2, z2)
    __shared__ double acc[4][4][kHfSites];      // per quarter: S, Gx, Gy, Gz of a site (one set at a time)
    for (int t = threadIdx.x; t < 4 * ns; t += blockDim.x) site[t] = site_g[t];
    for (int t = threadIdx.x; t < ns; t += blockDim.x) lj[t] = lj_g[t];
    const int64_t p0 = (int64_t)blockIdx.x * npb;
    const int np = (int)(P - p0 < npb ? P - p0 : npb);
    for (int pl = threadIdx.x; pl < np; pl += blockDim.x) {
        const int64_t p = p0 + pl;
        pc[3 * pl] = gx[p % nx];
        pc[3 * pl + 1] = gy[(p / nx) % ny];
        pc[3 * pl + 2] = gz[p / ((int64_t)nx * ny)];
    }
    for (int t = threadIdx.x; t < NS * np; t += blockDim.x) {
        const int m = t / np, pl = t - m * np;
        wcm[m * npb + pl] = wcm_g[m * P + p0 + pl];
    }
    __syncthreads();
    const int sl = threadIdx.x % kHfSites, quarter = threadIdx.x / kHfSites;
    const int nent = 4 * ns + 4;
    double d_qh[NS], d_qcm[NS], d_a[NS], d_b[NS];
#pragma unroll
    for (int m = 0; m < NS; m++) d_qh[m] = d_qcm[m] = d_a[m] = d_b[m] = 0.0;
    for (int sb = 0; sb < ns; sb += kHfSites) {
        const int s = sb + sl;
        const bool has = s < ns;
        const double sx = has ? site[4 * s] : 0.0, sy = has ? site[4 * s + 1] : 0.0;
        const double sz = has ? site[4 * s + 2] : 0.0, qs = has ? site[4 * s + 3] : 0.0;
        double S[NS], Gx[NS], Gy[NS], Gz[NS];
#pragma unroll
        for (int m = 0; m < NS; m++) S[m] = Gx[m] = Gy[m] = Gz[m] = 0.0;
        // staged pairs [u0, u1) into this site's sums
        auto eval = [&](int u0, int u1) {
            for (int u = u0; u < u1; u++) {
                const double2 a = stg[0][u], b = stg[1][u], c = stg[2][u];
                const double e1x = a.x - sx, e1y = a.y - sy, e1z = b.x - sz;
                const double e2x = b.y - sx, e2y = c.x - sy, e2z = c.y - sz;
                const double i1 = a_rsqrt(e1x * e1x + e1y * e1y + e1z * e1z);
                const double i2 = a_rsqrt(e2x * e2x + e2y * e2y + e2z * e2z);
                // the pair's geometry, shared by the sets: |h1 - s|^-1 + |h2 - s|^-1
                // and sum_a (h_a - s) |h_a - s|^-3; each set adds its weight times them
                const double si = i1 + i2, b1 = i1 * i1 * i1, b2 = i2 * i2 * i2;
                const double tx = b1 * e1x + b2 * e2x, ty = b1 * e1y + b2 * e2y, tz = b1 * e1z + b2 * e2z;
#pragma unroll
                for (int m = 0; m < NS; m++) {
                    const double w = stw[m][u];
                    S[m] = fma(w, si, S[m]);
                    Gx[m] = fma(w, tx, Gx[m]);
                    Gy[m] = fma(w, ty, Gy[m]);
                    Gz[m] = fma(w, tz, Gz[m]);
                }
            }
        };
        const int64_t beg = first[blockIdx.x];
        const int nkeep = (int)(first[blockIdx.x + 1] - beg);
        // registers of the thread's pair in the next chunk
        int npl = 0;
        double nw[NS], nox = 0.0, noy = 0.0, noz = 0.0;
#pragma unroll
        for (int m = 0; m < NS; m++) nw[m] = 0.0;
        auto fetch = [&](int c) {
            const int u = c + (int)threadIdx.x;
            if (u < nkeep) {
                const uint32_t e = pairs[beg + u];
                const int q = (int)(e & 0xffffffu);
                npl = (int)(e >> 24);
#pragma unroll
                for (int m = 0; m < NS; m++) nw[m] = om[m * ne + beg + u];
                nox = off[3 * q];
                noy = off[3 * q + 1];
                noz = off[3 * q + 2];
            }
        };
        fetch(0);
        for (int c = 0, h = 0; c < nkeep; c += kAWG, h ^= kAWG) {
            const int n = nkeep - c < kAWG ? nkeep - c : kAWG;
            if ((int)threadIdx.x < n) {
                const int slot = h + (int)threadIdx.x;
                const double cx = pc[3 * npl], cy = pc[3 * npl + 1], cz = pc[3 * npl + 2];
#pragma unroll
                for (int m = 0; m < NS; m++) stw[m][slot] = nw[m];
                stg[0][slot] = make_double2(cx + nox, cy + noy);
                stg[1][slot] = make_double2(cz + noz, cx - nox);
                stg[2][slot] = make_double2(cy - noy, cz - noz);
            }
            fetch(c + kAWG);
            __syncthreads();
            if (has) {
                // a full chunk by quarters of kAWG / 4, the last one by quarters of ceil(n / 4)
                const int part = (n + 3) / 4;
                const int u0 = quarter * part, u1 = u0 + part < n ? u0 + part : n;
                eval(h + u0, h + u1);
            }
        }
        __syncthreads();
        for (int m = 0; m < NS; m++) {
            acc[quarter][0][sl] = S[m];
            acc[quarter][1][sl] = Gx[m];
            acc[quarter][2][sl] = Gy[m];
            acc[quarter][3][sl] = Gz[m];
            __syncthreads();
            // the quarters in order, then the centre term, by thread sl of quarter 0
            if (quarter == 0 && has) {
                double r[4];
#pragma unroll
                for (int e = 0; e < 4; e++) r[e] = ((acc[0][e][sl] + acc[1][e][sl]) + acc[2][e][sl]) + acc[3][e][sl];
                double g0 = kCoulombK * q_h * r[0];
                double g1 = kCoulombK * q_h * qs * r[1], g2 = kCoulombK * q_h * qs * r[2];
                double g3 = kCoulombK * q_h * qs * r[3];
                double dqh = kCoulombK * qs * r[0], dqcm = 0.0, da = 0.0, db = 0.0;
                for (int pl = 0; pl < np; pl++) {
                    const double w = wcm[m * npb + pl];
                    if (w == 0.0) continue;
                    const double ex = pc[3 * pl] - sx, ey = pc[3 * pl + 1] - sy, ez = pc[3 * pl + 2] - sz;
                    const double r2 = ex * ex + ey * ey + ez * ez;
                    const double ir = 1.0 / sqrt(r2);
                    const double ir3 = ir * ir * ir;
                    g0 += w * kCoulombK * q_cm * ir;
                    dqcm += w * kCoulombK * qs * ir;
                    double c = w * kCoulombK * q_cm * qs * ir3;
                    if (lj[s]) {
                        const double i2 = ir * ir, i6 = i2 * i2 * i2, i12 = i6 * i6;
                        da += w * i12;
                        db -= w * i6;
                        // d/ds of A r^-12 - B r^-6 = (12 A r^-14 - 6 B r^-8) (x - s)
                        c += w * (12.0 * lj_a * i12 * i2 - 6.0 * lj_b * i6 * i2);
                    }
                    g1 += c * ex;
                    g2 += c * ey;
                    g3 += c * ez;
                }
                double *out = partial + ((size_t)blockIdx.x * NS + m) * nent;
                out[4 * s + 0] = g0;
                out[4 * s + 1] = g1;
                out[4 * s + 2] = g2;
                out[4 * s + 3] = g3;
#pragma unroll
                for (int mm = 0; mm < NS; mm++)
                    if (mm == m) {
                        d_qh[mm] += dqh;
                        d_qcm[mm] += dqcm;
                        d_a[mm] += da;
                        d_b[mm] += db;
                    }
            }
            __syncthreads();
        }
    }
    for (int m = 0; m < NS; m++) {
        double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int mm = 0; mm < NS; mm++)
            if (mm == m) {
                v[0] = d_qh[mm];
                v[1] = d_qcm[mm];
                v[2] = d_a[mm];
                v[3] = d_b[mm];
            }
        double *out = partial + ((size_t)blockIdx.x * NS + m) * nent;
        for (int e = 0; e < 4; e++) {
            const double t = a_block_sum(v[e], red);
            if (threadIdx.x == 0) out[4 * ns + e] = t;
        }
    }
}

// out[e] = sum over workgroups of partial[wg][e]: one workgroup per entry,
// thread t sums rows t, t + kAWG, ... in order, then a fixed tree
__global__ __launch_bounds__(kAWG) void k_hf_reduce(const double *__restrict__ partial, int nwg, int nent,
                                                    double *__restrict__ out)
{
    __shared__ double red[kAWG];
    const int e = blockIdx.x;
    double s = 0.0;
    for (int w = threadIdx.x; w < nwg; w += kAWG) s += partial[(size_t)w * nent + e];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = kAWG / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[e] = red[0];
}

int launch_hf_density(const double *psi, int nset, const int *counts, const double *g, const double *Y,
                      const double *wq, const uint64_t *unclamped, const uint32_t *pairs, const int64_t *first,
                      int64_t ne, double *om, double *wcm, int64_t P, int nk, int Q, hipStream_t s)
{
    if (nset < 1 || nset > kHfSets) return fail(BSL_E_ARG, "adjoint: 1 to 4 sets of eigenvectors per pass");
    HfSets sets{};
    for (int m = 0; m < nset; m++) sets.first[m + 1] = sets.first[m] + counts[m];
    const int nvec = sets.first[nset];
    const size_t lds = sizeof(double) * kHfPoints * (size_t)((nvec + 3) & ~3) * nk;
    if (lds > 65536) return fail(BSL_E_ARG, "too many eigenvectors for one density pass");
    hipLaunchKernelGGL(k_hf_density, dim3(hf_grad_groups(P, kHfPoints)), dim3(kAWG), lds, s, psi, nvec, sets, nset, g,
                       Y, wq, unclamped, pairs, first, ne, om, wcm, P, nk, Q);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

int hf_grad_groups(int64_t P, int npb) { return (int)((P + npb - 1) / npb); }

int launch_pair_count(const uint64_t *unclamped, int64_t P, int Q, int npb, int *count, hipStream_t s)
{
    if (npb < 1 || npb > 256 || Q < 2 || Q % 2 || Q > (1 << 24)) return fail(BSL_E_ARG, "adjoint: pair list limits");
    hipLaunchKernelGGL(k_pair_count, dim3(hf_grad_groups(P, npb)), dim3(kAWG), 0, s, unclamped, P, Q, npb, count);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

int launch_pair_fill(const uint64_t *unclamped, int64_t P, int Q, int npb, const int64_t *first, uint32_t *pairs,
                     hipStream_t s)
{
    if (npb < 1 || npb > 256 || Q < 2 || Q % 2 || Q > (1 << 24)) return fail(BSL_E_ARG, "adjoint: pair list limits");
    hipLaunchKernelGGL(k_pair_fill, dim3(hf_grad_groups(P, npb)), dim3(kAWG), 0, s, unclamped, P, Q, npb, first,
                       pairs);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

int launch_hf_grad(const double *om, int64_t ne, const double *wcm, int nset, int64_t P, const double *gx,
                   const double *gy, const double *gz, int nx, int ny, const double *off, const double *site4,
                   const int *lj, int ns, double q_h, double q_cm, double lj_a, double lj_b, int npb,
                   const uint32_t *pairs, const int64_t *first, double *partial, double *out, hipStream_t s)
{
    if (nset < 1 || nset > kHfGradSets) return fail(BSL_E_ARG, "adjoint: 1 to 8 sets of eigenvectors per pass");
    const int nwg = hf_grad_groups(P, npb);
    const size_t lds = sizeof(double) * (4 * (size_t)ns + (size_t)(nset + 3) * npb) + sizeof(int) * ns;
    // dynamic LDS beside the static staging arrays of k_hf_grad
    const size_t lds_static = sizeof(double) * (kAWG / 64 + (size_t)nset * 2 * kAWG + 3 * 2 * 2 * kAWG + 4 * 4 * kHfSites);
    if (npb > kAWG || lds + lds_static > 96 * 1024) return fail(BSL_E_ARG, "adjoint: too many sites");
    // one instance per set count: a pass of fewer sets carries fewer sums
    // (the one-set instance keeps the occupancy of a single-set kernel)
    static_assert(kHfGradSets == 8, "one kernel instance per set count");
    void (*const kerns[8])(const double *, int64_t, const double *, int64_t, const double *, const double *,
                           const double *, int, int, const double *, const double *, const int *, int, double,
                           double, double, double, int, const uint32_t *, const int64_t *, double *) = {
        k_hf_grad<1>, k_hf_grad<2>, k_hf_grad<3>, k_hf_grad<4>,
        k_hf_grad<5>, k_hf_grad<6>, k_hf_grad<7>, k_hf_grad<8>};
    auto kern = kerns[nset - 1];
    hipLaunchKernelGGL(kern, dim3(nwg), dim3(kAWG), lds, s, om, ne, wcm, P, gx, gy, gz, nx, ny, off, site4, lj, ns, q_h,
                       q_cm, lj_a, lj_b, npb, pairs, first, partial);
    BSL_HIP(hipGetLastError());
    const int nent = (4 * ns + 4) * nset;
    hipLaunchKernelGGL(k_hf_reduce, dim3(nent), dim3(kAWG), 0, s, partial, nwg, nent, out);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

}  // namespace bsl
