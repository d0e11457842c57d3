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

// omega[p][q] = [vg[p][q] < ceiling] w_q sum_k g_k (sum_r Y[r][q] psi_k[r][p])^2,
// written over vg.  psi: [nvec][nk][P]; Y: [nk][Q]; one workgroup per point.
__global__ __launch_bounds__(kAWG) void k_hf_density(const double *__restrict__ psi, int nvec,
                                                     const double *__restrict__ g, const double *__restrict__ Y,
                                                     const double *__restrict__ wq, double *vg, int64_t P, int nk,
                                                     int Q, double ceiling)
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
        const double v = vg[idx];
        double om = 0.0;
        if (v < ceiling) {
            for (int k = 0; k < nvec; k++) {
                double s = 0.0;
                for (int r = 0; r < nk; r++) s += Y[(size_t)r * Q + q] * cpsi[k * nk + r];
                om += g[k] * s * s;
            }
            om *= wq[q];
        }
        vg[idx] = om;
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

// Per workgroup (npb consecutive points, threads over (point, q) pairs):
// partial[wg][4*s + {0,1,2,3}] = d/d(q_s, x_s, y_s, z_s), then
// partial[wg][4*ns + {0,1,2,3}] = d/d(q_h2_h, q_h2_cm, A, B).
// The H2 centre term only depends on the point: its weight is
// omega_cm(p) = sum_q omega(p, q) (unclamped nodes only, as omega).
__global__ __launch_bounds__(kAWG) void k_hf_grad(const double *__restrict__ om, int64_t P, int Q,
                                                  const double *__restrict__ gx, const double *__restrict__ gy,
                                                  const double *__restrict__ gz, int nx, int ny,
                                                  const double *__restrict__ off, const double *__restrict__ site_g,
                                                  const int *__restrict__ lj_g, int ns, double q_h, double q_cm,
                                                  double lj_a, double lj_b, int npb, double *__restrict__ partial)
{
    extern __shared__ double site[];  // [ns][4], then omega_cm [npb], then int lj [ns]
    double *wcm = site + 4 * ns;
    int *lj = (int *)(wcm + npb);
    __shared__ double red[kAWG / 64];
    for (int t = threadIdx.x; t < 4 * ns; t += blockDim.x) site[t] = site_g[t];
    for (int t = threadIdx.x; t < ns; t += blockDim.x) lj[t] = lj_g[t];
    const int64_t p0 = (int64_t)blockIdx.x * npb;
    const int np = (int)(P - p0 < npb ? P - p0 : npb);
    for (int pl = 0; pl < np; pl++) {
        double s = 0.0;
        for (int q = threadIdx.x; q < Q; q += blockDim.x) s += om[(size_t)(p0 + pl) * Q + q];
        s = a_block_sum(s, red);
        if (threadIdx.x == 0) wcm[pl] = s;
    }
    __syncthreads();
    const int npairs = np * Q;
    double d_qh = 0.0, d_qcm = 0.0, d_a = 0.0, d_b = 0.0;
    double *out = partial + (size_t)blockIdx.x * (4 * ns + 4);
    for (int s = 0; s < ns; s++) {
        const double sx = site[4 * s], sy = site[4 * s + 1], sz = site[4 * s + 2], qs = site[4 * s + 3];
        double dq = 0.0, dx = 0.0, dy = 0.0, dz = 0.0;
        for (int t = threadIdx.x; t < npairs; t += blockDim.x) {
            const int pl = t / Q, q = t - pl * Q;
            const int64_t p = p0 + pl;
            const double w = om[(size_t)p * Q + q];
            if (w == 0.0) continue;
            const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((int64_t)nx * ny));
            const double x = gx[i], y = gy[j], z = gz[k];
            const double ox = off[3 * q], oy = off[3 * q + 1], oz = off[3 * q + 2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const double sg = h ? -1.0 : 1.0;
                const double ex = x + sg * ox - sx, ey = y + sg * oy - sy, ez = z + sg * oz - sz;
                const double ir = 1.0 / sqrt(ex * ex + ey * ey + ez * ez);
                const double ir3 = ir * ir * ir;
                dq += w * kCoulombK * q_h * ir;
                d_qh += w * kCoulombK * qs * ir;
                const double c = w * kCoulombK * q_h * qs * ir3;
                dx += c * ex;
                dy += c * ey;
                dz += c * ez;
            }
        }
        // the centre term, by thread pl for point pl
        if (threadIdx.x < np) {
            const int pl = threadIdx.x;
            const int64_t p = p0 + pl;
            const double w = wcm[pl];
            if (w != 0.0) {
                const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((int64_t)nx * ny));
                const double ex = gx[i] - sx, ey = gy[j] - sy, ez = gz[k] - sz;
                const double r2 = ex * ex + ey * ey + ez * ez;
                const double ir = 1.0 / sqrt(r2);
                const double ir3 = ir * ir * ir;
                dq += w * kCoulombK * q_cm * ir;
                d_qcm += w * kCoulombK * qs * ir;
                double c = w * kCoulombK * q_cm * qs * ir3;
                if (lj[s]) {
                    const double ir2 = ir * ir, ir6 = ir2 * ir2 * ir2, ir12 = ir6 * ir6;
                    d_a += w * ir12;
                    d_b -= w * ir6;
                    // d/ds of A r^-12 - B r^-6 = (12 A r^-14 - 6 B r^-8) (x - s)
                    c += w * (12.0 * lj_a * ir12 * ir2 - 6.0 * lj_b * ir6 * ir2);
                }
                dx += c * ex;
                dy += c * ey;
                dz += c * ez;
            }
        }
        dq = a_block_sum(dq, red);
        dx = a_block_sum(dx, red);
        dy = a_block_sum(dy, red);
        dz = a_block_sum(dz, red);
        if (threadIdx.x == 0) {
            out[4 * s + 0] = dq;
            out[4 * s + 1] = dx;
            out[4 * s + 2] = dy;
            out[4 * s + 3] = dz;
        }
    }
    d_qh = a_block_sum(d_qh, red);
    d_qcm = a_block_sum(d_qcm, red);
    d_a = a_block_sum(d_a, red);
    d_b = a_block_sum(d_b, red);
    if (threadIdx.x == 0) {
        out[4 * ns + 0] = d_qh;
        out[4 * ns + 1] = d_qcm;
        out[4 * ns + 2] = d_a;
        out[4 * ns + 3] = d_b;
    }
}

// out[e] = sum over workgroups of partial[wg][e], in workgroup order
__global__ void k_hf_reduce(const double *__restrict__ partial, int nwg, int nent, double *__restrict__ out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nent) return;
    double s = 0.0;
    for (int w = 0; w < nwg; w++) s += partial[(size_t)w * nent + e];
    out[e] = s;
}

int launch_hf_density(const double *psi, int nvec, const double *g, const double *Y, const double *wq, double *vg,
                      int64_t P, int nk, int Q, double ceiling, hipStream_t s)
{
    const size_t lds = sizeof(double) * (size_t)nvec * nk;
    if (lds > 65536) return fail(BSL_E_ARG, "too many eigenvectors for one density pass");
    hipLaunchKernelGGL(k_hf_density, dim3((unsigned)P), dim3(kAWG), lds, s, psi, nvec, g, Y, wq, vg, P, nk, Q,
                       ceiling);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

int hf_grad_groups(int64_t P, int npb) { return (int)((P + npb - 1) / npb); }

int launch_hf_grad(const double *om, int64_t P, int Q, const double *gx, const double *gy, const double *gz, int nx,
                   int ny, const double *off, const double *site4, const int *lj, int ns, double q_h, double q_cm,
                   double lj_a, double lj_b, int npb, double *partial, double *out, hipStream_t s)
{
    const int nwg = hf_grad_groups(P, npb);
    const size_t lds = sizeof(double) * (4 * (size_t)ns + npb) + sizeof(int) * ns;
    if (npb > kAWG || lds > 65536) return fail(BSL_E_ARG, "adjoint: too many sites");
    hipLaunchKernelGGL(k_hf_grad, dim3(nwg), dim3(kAWG), lds, s, om, P, Q, gx, gy, gz, nx, ny, off, site4, lj, ns,
                       q_h, q_cm, lj_a, lj_b, npb, partial);
    BSL_HIP(hipGetLastError());
    const int nent = 4 * ns + 4;
    hipLaunchKernelGGL(k_hf_reduce, dim3((nent + 255) / 256), dim3(256), 0, s, partial, nwg, nent, out);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

}  // namespace bsl
