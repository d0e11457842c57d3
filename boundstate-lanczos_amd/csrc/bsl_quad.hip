// bsl_quad.hip -- potential term of the quadrature ("on-the-fly")
// calculator: per grid point P,
//   f(a, m)   = sum_l  L_lm(theta_a) c_lm                (Legendre, forward)
//   g(a, phi) = sum_m  f(a, m) S_m(phi)                  (phi, forward)
//   h(a, phi) = w_a w_phi V(P, a, phi) g(a, phi)         (pointwise potential)
//   f'(a, m)  = sum_phi h(a, phi) S_m(phi)               (phi, backward)
//   u_lm      = sum_a  L_lm(theta_a) f'(a, m)            (Legendre, backward)
// which is <lm|V|l'm'>(P) c_l'm' evaluated through the quadrature grid
// instead of precomputed blocks (the algorithm the reference's calc_ulm,
// linRotCartSphHvRoutines.cpp:990-1179, intends; its loop-1 and loop-5 sums
// are corrected here).  The weighted potential grid W = w_a w_phi V stays
// resident in HBM and is streamed once per Hv.
//
// FP64: one workgroup walks a run of grid points, with the transform tables
// in LDS.  FP32: the two phi transforms are batched GEMMs over the points of
// a tile on MFMA (v_mfma_f32_32x32x2f32), with the pointwise multiply fused
// between them.
#include <hip/hip_runtime.h>

#include "bsl_internal.h"

namespace bsl {

constexpr int kQWG = 256;

__device__ inline double q_wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__device__ inline double q_block_sum(double v, double *lds)
{
    v = q_wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += lds[w];
    __syncthreads();
    return s;
}

// W: [P][na][nphi] (weights folded in), L: [nk][na], S: [nm][nphi],
// mplane[r]: m + lmax of carried plane r.
template <typename WT>
__global__ __launch_bounds__(kQWG) void k_quad_fp64(const double *__restrict__ X, double *__restrict__ R,
                                                    const WT *__restrict__ W, const double *__restrict__ Lg,
                                                    const double *__restrict__ Sg, const int *__restrict__ mplane,
                                                    const double *__restrict__ mask, const Scal *__restrict__ sc,
                                                    int64_t P, int nk, int na, int nphi, int nm, int ppw,
                                                    double *__restrict__ partial)
{
    extern __shared__ double lds[];
    double *Ls = lds;                      // nk * na
    double *Ss = Ls + nk * na;             // nm * nphi
    double *c = Ss + nm * nphi;            // nk
    double *f = c + nk;                    // na * nm  (forward, then backward)
    double *h = f + na * nm;               // na * nphi
    __shared__ double red[kQWG / 64];
    __shared__ int ms[64];
    for (int t = threadIdx.x; t < nk * na; t += blockDim.x) Ls[t] = Lg[t];
    for (int t = threadIdx.x; t < nm * nphi; t += blockDim.x) Ss[t] = Sg[t];
    for (int t = threadIdx.x; t < nk; t += blockDim.x) ms[t] = mplane[t];
    __syncthreads();
    const double sx = sc->sx;
    double dot = 0.0;
    const int64_t p0 = (int64_t)blockIdx.x * ppw;
    for (int64_t p = p0; p < p0 + ppw && p < P; p++) {
        for (int r = threadIdx.x; r < nk; r += blockDim.x) c[r] = sx * X[(size_t)r * P + p];
        __syncthreads();
        for (int t = threadIdx.x; t < na * nm; t += blockDim.x) {
            const int a = t / nm, mi = t - a * nm;
            double s = 0.0;
            for (int r = 0; r < nk; r++)
                if (ms[r] == mi) s += Ls[r * na + a] * c[r];
            f[t] = s;
        }
        __syncthreads();
        const WT *Wp = W + (size_t)p * na * nphi;
        for (int t = threadIdx.x; t < na * nphi; t += blockDim.x) {
            const int a = t / nphi, ph = t - a * nphi;
            double g = 0.0;
            for (int mi = 0; mi < nm; mi++) g += f[a * nm + mi] * Ss[mi * nphi + ph];
            h[t] = (double)Wp[t] * g;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < na * nm; t += blockDim.x) {
            const int a = t / nm, mi = t - a * nm;
            double s = 0.0;
            for (int ph = 0; ph < nphi; ph++) s += h[a * nphi + ph] * Ss[mi * nphi + ph];
            f[t] = s;
        }
        __syncthreads();
        for (int r = threadIdx.x; r < nk; r += blockDim.x) {
            const int mi = ms[r];
            double u = 0.0;
            for (int a = 0; a < na; a++) u += Ls[r * na + a] * f[a * nm + mi];
            const size_t o = (size_t)r * P + p;
            const double nr = R[o] + mask[r] * u;
            R[o] = nr;
            dot += c[r] * nr;
        }
        __syncthreads();
    }
    const double s = q_block_sum(dot, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// W[p][a][half*nb + b] = w_a w_b vg[p][half*na*nb + a*nb + b]
template <typename WT>
__global__ void k_weight_grid(const double *__restrict__ vg, WT *__restrict__ W, const double *__restrict__ wt,
                              const double *__restrict__ wp, int64_t P, int na, int nb)
{
    const int64_t Q = (int64_t)2 * na * nb;
    const int64_t total = P * Q;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = idx / Q;
        const int q = (int)(idx - p * Q);
        const int a = q / (2 * nb), rem = q - a * 2 * nb, half = rem / nb, b = rem - half * nb;
        W[idx] = (WT)(wt[a] * wp[b] * vg[(size_t)p * Q + (size_t)half * na * nb + (size_t)a * nb + b]);
    }
}

int launch_weight_grid(const double *vg, void *W, int wprec, const double *wt, const double *wp, int64_t P,
                       int na, int nb, hipStream_t s)
{
    const int64_t total = P * 2 * na * nb;
    int64_t g = (total + 255) / 256;
    if (g > 65536) g = 65536;
    if (wprec == BSL_PREC_FP64)
        hipLaunchKernelGGL(k_weight_grid<double>, dim3((unsigned)g), dim3(256), 0, s, vg, (double *)W, wt, wp, P,
                           na, nb);
    else
        hipLaunchKernelGGL(k_weight_grid<float>, dim3((unsigned)g), dim3(256), 0, s, vg, (float *)W, wt, wp, P,
                           na, nb);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

int launch_quad_potential(const DevOperator &op, const double *X, double *R, Scal *sc, hipStream_t s)
{
    const int ppw = 8;
    const unsigned grid = (unsigned)((op.P + ppw - 1) / ppw);
    if ((int64_t)grid > op.npartial || (int64_t)grid != op.nalpha)
        return fail(BSL_E_STATE, "partial buffer does not match the quadrature kernel grid");
    const size_t lds = sizeof(double) * ((size_t)op.nk * op.na + (size_t)op.nm * op.nphi + op.nk +
                                         (size_t)op.na * op.nm + (size_t)op.na * op.nphi);
    if (op.wprec == BSL_PREC_FP64)
        hipLaunchKernelGGL(k_quad_fp64<double>, dim3(grid), dim3(kQWG), lds, s, X, R, (const double *)op.wgrid,
                           op.qL, op.qS, op.mplane, op.mask, sc, op.P, op.nk, op.na, op.nphi, op.nm, ppw,
                           op.partial);
    else
        hipLaunchKernelGGL(k_quad_fp64<float>, dim3(grid), dim3(kQWG), lds, s, X, R, (const float *)op.wgrid,
                           op.qL, op.qS, op.mplane, op.mask, sc, op.P, op.nk, op.na, op.nphi, op.nm, ppw,
                           op.partial);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

}  // namespace bsl
