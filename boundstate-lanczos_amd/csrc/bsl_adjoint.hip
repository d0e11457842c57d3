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

// mask bit n = [vg[n] < ceiling], 64 nodes per word: one wave per word, the
// wave's lanes read 64 consecutive entries (coalesced) and lane 0 stores the
// ballot.  The loop bound is uniform over a wave, so every ballot is whole.
__global__ __launch_bounds__(kAWG) void k_unclamped_mask(const double *__restrict__ vg, int64_t n, double ceiling,
                                                         uint64_t *__restrict__ mask)
{
    const int lane = threadIdx.x & 63;
    const int64_t nw = (n + 63) / 64;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t stride = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t w = wave; w < nw; w += stride) {
        const int64_t idx = w * 64 + lane;
        const bool keep = idx < n && vg[idx] < ceiling;
        const uint64_t bits = __ballot(keep);
        if (lane == 0) mask[w] = bits;
    }
}

// om[p][q] = [unclamped bit p*Q + q] w_q sum_k g_k (sum_r Y[r][q] psi_k[r][p])^2.
// psi: [nvec][nk][P]; Y: [nk][Q]; one workgroup per point.
__global__ __launch_bounds__(kAWG) void k_hf_density(const double *__restrict__ psi, int nvec,
                                                     const double *__restrict__ g, const double *__restrict__ Y,
                                                     const double *__restrict__ wq,
                                                     const uint64_t *__restrict__ unclamped, double *__restrict__ om,
                                                     int64_t P, int nk, int Q)
{
    extern __shared__ double cpsi[];  // [nvec][nk]
    const int64_t p = blockIdx.x;
    for (int t = threadIdx.x; t < nvec * nk; t += blockDim.x) {
        const int k = t / nk, r = t - k * nk;
        cpsi[t] = psi[((size_t)k * nk + r) * P + p];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < Q; q += blockDim.x) {
        const size_t idx = (size_t)p * Q + q;
        double o = 0.0;
        if ((unclamped[idx >> 6] >> (idx & 63)) & 1) {
            for (int k = 0; k < nvec; k++) {
                double s = 0.0;
                for (int r = 0; r < nk; r++) s += Y[(size_t)r * Q + q] * cpsi[k * nk + r];
                o += g[k] * s * s;
            }
            o *= wq[q];
        }
        om[idx] = o;
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

// Per workgroup (npb consecutive points) and set m < nset of eigenvectors
// (omega of set m at om + m P Q):
// partial[wg][m][4*s + {0,1,2,3}] = d/d(q_s, x_s, y_s, z_s), then
// partial[wg][m][4*ns + {0,1,2,3}] = d/d(q_h2_h, q_h2_cm, A, B).
// Only unclamped pairs (point, node) contribute (omega = 0 at clamped nodes,
// most of a box that reaches into the cage walls), so the workgroup compacts
// its pairs: each pass over kAWG pairs appends the unclamped ones (in pair
// order, by a ballot prefix) to an LDS stage of the sets' omegas and the two
// H positions; whenever kAWG are staged, thread t accumulates site
// sb + t % 64 over a quarter (t / 64) of them.  The stage depends on the
// clamp mask alone, not on the sets.
// Per site and set only two sums are carried over the H atoms h of the pairs,
//   S = sum omega / |h - s|,   G = sum omega (h - s) / |h - s|^3,
// since d/dq_s = k q_h S, its share of d/dq_h = k q_s S and d/ds = k q_h q_s G;
// the geometry (|h - s|^-1 of both atoms) is shared by the sets, and each
// set's sums take the same operations in the same order as with that set
// alone.  The quarters combine in a fixed order.  The H2 centre term only depends on the point: its weight
// is omega_cm(p) = sum_q omega(p, q).
constexpr int kHfSites = 64;  // sites per pass
constexpr int kHfSets = 4;    // sets of eigenvectors per launch

__global__ __launch_bounds__(kAWG) void k_hf_grad(const double *__restrict__ om, int64_t pq, int nset, int64_t P,
                                                  int Q, const double *__restrict__ gx, const double *__restrict__ gy,
                                                  const double *__restrict__ gz, int nx, int ny,
                                                  const double *__restrict__ off, const double *__restrict__ site_g,
                                                  const int *__restrict__ lj_g, int ns, double q_h, double q_cm,
                                                  double lj_a, double lj_b, int npb,
                                                  const uint64_t *__restrict__ unclamped, double *__restrict__ partial)
{
    extern __shared__ double site[];  // [ns][4], then omega_cm [kHfSets][npb], then int lj [ns]
    double *wcm = site + 4 * ns;
    int *lj = (int *)(wcm + kHfSets * npb);
    __shared__ double red[kAWG / 64];
    __shared__ int wcnt[kAWG / 64];
    __shared__ double stw[kHfSets][2 * kAWG];  // staged pairs: omega of each set
    __shared__ double2 stg[3][2 * kAWG];        // and the H positions (x1, y1), (z1, x2), (y2, z2)
    __shared__ double acc[4][4][kHfSites];      // per quarter: S, Gx, Gy, Gz of a site (one set at a time)
    for (int t = threadIdx.x; t < 4 * ns; t += blockDim.x) site[t] = site_g[t];
    for (int t = threadIdx.x; t < ns; t += blockDim.x) lj[t] = lj_g[t];
    const int64_t p0 = (int64_t)blockIdx.x * npb;
    const int np = (int)(P - p0 < npb ? P - p0 : npb);
    for (int m = 0; m < nset; m++)
        for (int pl = 0; pl < np; pl++) {
            double s = 0.0;
            for (int q = threadIdx.x; q < Q; q += blockDim.x) s += om[m * pq + (size_t)(p0 + pl) * Q + q];
            s = a_block_sum(s, red);
            if (threadIdx.x == 0) wcm[m * npb + pl] = s;
        }
    __syncthreads();
    const int npairs = np * Q;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int sl = threadIdx.x % kHfSites, quarter = threadIdx.x / kHfSites;
    const int nent = 4 * ns + 4;
    double d_qh[kHfSets], d_qcm[kHfSets], d_a[kHfSets], d_b[kHfSets];
#pragma unroll
    for (int m = 0; m < kHfSets; m++) d_qh[m] = d_qcm[m] = d_a[m] = d_b[m] = 0.0;
    for (int sb = 0; sb < ns; sb += kHfSites) {
        const int s = sb + sl;
        const bool has = s < ns;
        const double sx = has ? site[4 * s] : 0.0, sy = has ? site[4 * s + 1] : 0.0;
        const double sz = has ? site[4 * s + 2] : 0.0, qs = has ? site[4 * s + 3] : 0.0;
        double S[kHfSets], Gx[kHfSets], Gy[kHfSets], Gz[kHfSets];
#pragma unroll
        for (int m = 0; m < kHfSets; m++) S[m] = Gx[m] = Gy[m] = Gz[m] = 0.0;
        // staged pairs [u0, u1) into this site's sums
        auto eval = [&](int u0, int u1) {
            for (int u = u0; u < u1; u++) {
                const double2 a = stg[0][u], b = stg[1][u], c = stg[2][u];
                const double e1x = a.x - sx, e1y = a.y - sy, e1z = b.x - sz;
                const double e2x = b.y - sx, e2y = c.x - sy, e2z = c.y - sz;
                const double i1 = a_rsqrt(e1x * e1x + e1y * e1y + e1z * e1z);
                const double i2 = a_rsqrt(e2x * e2x + e2y * e2y + e2z * e2z);
                const double i1s = i1 * i1, i2s = i2 * i2;
#pragma unroll
                for (int m = 0; m < kHfSets; m++) {
                    if (m >= nset) break;
                    const double w = stw[m][u];
                    const double w1 = w * i1, w2 = w * i2;
                    S[m] += w1 + w2;
                    const double c1 = w1 * i1s, c2 = w2 * i2s;
                    Gx[m] += c1 * e1x + c2 * e2x;
                    Gy[m] += c1 * e1y + c2 * e2y;
                    Gz[m] += c1 * e1z + c2 * e2z;
                }
            }
        };
        int fill = 0;  // staged pairs, the same in every thread
        for (int t0 = 0; t0 < npairs; t0 += kAWG) {
            const int t = t0 + (int)threadIdx.x;
            int pl = 0, q = 0;
            double w[kHfSets];
            bool keep = false;
#pragma unroll
            for (int m = 0; m < kHfSets; m++) w[m] = 0.0;
            if (t < npairs) {
                pl = t / Q;
                q = t - pl * Q;
                const size_t idx = (size_t)(p0 + pl) * Q + q;
                keep = (unclamped[idx >> 6] >> (idx & 63)) & 1;
#pragma unroll
                for (int m = 0; m < kHfSets; m++)
                    if (m < nset && keep) w[m] = om[m * pq + idx];
            }
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
            if (keep) {
                // omega != 0 only at unclamped nodes, where no H atom sits on a site
                const int slot = fill + woff + __popcll(bal & ((1ull << lane) - 1ull));
                const int64_t p = p0 + pl;
                const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((int64_t)nx * ny));
                const double ox = off[3 * q], oy = off[3 * q + 1], oz = off[3 * q + 2];
#pragma unroll
                for (int m = 0; m < kHfSets; m++)
                    if (m < nset) stw[m][slot] = w[m];
                stg[0][slot] = make_double2(gx[i] + ox, gy[j] + oy);
                stg[1][slot] = make_double2(gz[k] + oz, gx[i] - ox);
                stg[2][slot] = make_double2(gy[j] - oy, gz[k] - oz);
            }
            fill += total;
            __syncthreads();
            if (fill >= kAWG) {
                if (has) eval(quarter * (kAWG / 4), (quarter + 1) * (kAWG / 4));
                __syncthreads();
                const int rest = fill - kAWG;  // < kAWG: moved from [kAWG, fill) to [0, rest)
                if ((int)threadIdx.x < rest) {
#pragma unroll
                    for (int m = 0; m < kHfSets; m++)
                        if (m < nset) stw[m][threadIdx.x] = stw[m][kAWG + threadIdx.x];
#pragma unroll
                    for (int f = 0; f < 3; f++) stg[f][threadIdx.x] = stg[f][kAWG + threadIdx.x];
                }
                fill = rest;
                __syncthreads();
            }
        }
        if (has) {
            const int part = (fill + 3) / 4;
            const int u0 = quarter * part, u1 = u0 + part < fill ? u0 + part : fill;
            eval(u0, u1);
        }
        for (int m = 0; m < nset; m++) {
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
                    const int64_t p = p0 + pl;
                    const double w = wcm[m * npb + pl];
                    if (w == 0.0) continue;
                    const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((int64_t)nx * ny));
                    const double ex = gx[i] - sx, ey = gy[j] - sy, ez = gz[k] - sz;
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
                double *out = partial + ((size_t)blockIdx.x * nset + m) * nent;
                out[4 * s + 0] = g0;
                out[4 * s + 1] = g1;
                out[4 * s + 2] = g2;
                out[4 * s + 3] = g3;
#pragma unroll
                for (int mm = 0; mm < kHfSets; mm++)
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
    for (int m = 0; m < nset; m++) {
        double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int mm = 0; mm < kHfSets; mm++)
            if (mm == m) {
                v[0] = d_qh[mm];
                v[1] = d_qcm[mm];
                v[2] = d_a[mm];
                v[3] = d_b[mm];
            }
        double *out = partial + ((size_t)blockIdx.x * nset + m) * nent;
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

int launch_unclamped_mask(const double *vg, int64_t n, double ceiling, uint64_t *mask, hipStream_t s)
{
    const int64_t nw = (n + 63) / 64;
    const int64_t groups = (nw + kAWG / 64 - 1) / (kAWG / 64);
    hipLaunchKernelGGL(k_unclamped_mask, dim3((unsigned)(groups < 65536 ? groups : 65536)), dim3(kAWG), 0, s, vg, n,
                       ceiling, mask);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

int launch_hf_density(const double *psi, int nvec, const double *g, const double *Y, const double *wq,
                      const uint64_t *unclamped, double *om, int64_t P, int nk, int Q, hipStream_t s)
{
    const size_t lds = sizeof(double) * (size_t)nvec * nk;
    if (lds > 65536) return fail(BSL_E_ARG, "too many eigenvectors for one density pass");
    hipLaunchKernelGGL(k_hf_density, dim3((unsigned)P), dim3(kAWG), lds, s, psi, nvec, g, Y, wq, unclamped, om, P,
                       nk, Q);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

int hf_grad_groups(int64_t P, int npb) { return (int)((P + npb - 1) / npb); }

int launch_hf_grad(const double *om, int nset, int64_t P, int Q, const double *gx, const double *gy, const double *gz,
                   int nx, int ny, const double *off, const double *site4, const int *lj, int ns, double q_h,
                   double q_cm, double lj_a, double lj_b, int npb, const uint64_t *unclamped, double *partial,
                   double *out, hipStream_t s)
{
    if (nset < 1 || nset > kHfSets) return fail(BSL_E_ARG, "adjoint: 1 to 4 sets of eigenvectors per pass");
    const int nwg = hf_grad_groups(P, npb);
    const size_t lds = sizeof(double) * (4 * (size_t)ns + (size_t)kHfSets * npb) + sizeof(int) * ns;
    // dynamic LDS beside the static staging arrays of k_hf_grad
    const size_t lds_static = sizeof(double) * (kAWG / 64 + kHfSets * 2 * kAWG + 3 * 2 * 2 * kAWG + 4 * 4 * kHfSites) +
                              sizeof(int) * (kAWG / 64);
    if (npb > kAWG || lds + lds_static > 96 * 1024) return fail(BSL_E_ARG, "adjoint: too many sites");
    hipLaunchKernelGGL(k_hf_grad, dim3(nwg), dim3(kAWG), lds, s, om, (int64_t)P * Q, nset, P, Q, gx, gy, gz, nx, ny,
                       off, site4, lj, ns, q_h, q_cm, lj_a, lj_b, npb, unclamped, partial);
    BSL_HIP(hipGetLastError());
    const int nent = (4 * ns + 4) * nset;
    hipLaunchKernelGGL(k_hf_reduce, dim3(nent), dim3(kAWG), 0, s, partial, nwg, nent, out);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

}  // namespace bsl
