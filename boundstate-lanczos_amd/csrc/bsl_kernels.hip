// bsl_kernels.hip -- gfx950 kernels of the NoQuad Hv, the Lanczos recurrence
// and the device-side preparation (potential grid, V-block contraction).
//
// Vector layout in HBM: [n][k][j][i] (i fastest), n over the carried lm
// planes.  The per-point potential blocks are symmetric and, for a
// homonuclear rotor on the parity-symmetric quadrature, block diagonal in
// the parity of l: they are stored as packed upper triangles of each parity
// block, structure-of-arrays [pair][P], so that one thread per grid point
// streams them with fully coalesced loads.
//
// All reductions are deterministic: fixed per-workgroup partial sums and a
// fixed-order final sum done by the last workgroup to finish.
#include <hip/hip_runtime.h>

#include <cmath>

#include "bsl_internal.h"

namespace bsl {

constexpr int kWG = 256;

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// sum over the workgroup; result valid in thread 0
__device__ inline double block_sum(double v, double *lds)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += lds[w];
    return s;
}

// three sums over the workgroup at once (one barrier); each is the sum
// block_sum would give, bitwise; valid in thread 0.  lds: 3 per wave.
__device__ inline void block_sum3(double a, double b, double c, double *lds, double &sa, double &sb, double &sc)
{
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        lds[3 * wid] = a;
        lds[3 * wid + 1] = b;
        lds[3 * wid + 2] = c;
    }
    __syncthreads();
    sa = sb = sc = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
            sa += lds[3 * w];
            sb += lds[3 * w + 1];
            sc += lds[3 * w + 2];
        }
}

// Fixed-order total of partial[0..n) computed by one workgroup; valid in
// every thread.  Partials are written with plain stores by the kernel before
// (kernel boundaries make them visible), so no device-scope fence or atomic
// is needed in the producers.
__device__ inline double total_of_partials(const double *__restrict__ partial, unsigned int n, double *lds)
{
    __shared__ double tot;
    // four independent chains per thread (the loads of a long list overlap),
    // combined in a fixed order
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    const unsigned int st = blockDim.x;
    unsigned int i = threadIdx.x;
    for (; i + 3 * st < n; i += 4 * st) {
        s0 += partial[i];
        s1 += partial[i + st];
        s2 += partial[i + 2 * st];
        s3 += partial[i + 3 * st];
    }
    for (; i < n; i += st) s0 += partial[i];
    double s = (s0 + s1) + (s2 + s3);
    s = block_sum(s, lds);
    if (threadIdx.x == 0) tot = s;
    __syncthreads();
    return tot;
}

// End of a Lanczos step: alpha = v.R, beta^2 = |R|^2 - 2 alpha v.R +
// alpha^2 |v|^2 = |R - alpha v|^2 from the potential kernel's three partial
// lists (fixed order); outputs at the step counter; the next step forms
// v' = (R - alpha v) / beta and R' = -beta v.  Called by the whole workgroup.
__device__ inline void finalize_step(const double *__restrict__ pa, const double *__restrict__ pb,
                                     const double *__restrict__ pc, unsigned int n, Scal *sc, double *alpha_out,
                                     double *beta_out, double *beta2_out, double *red)
{
    // the three lists are read together (one memory round trip), each summed
    // in the order of total_of_partials
    double sa[4] = {0.0, 0.0, 0.0, 0.0}, sb[4] = {0.0, 0.0, 0.0, 0.0}, sv[4] = {0.0, 0.0, 0.0, 0.0};
    const unsigned int st = blockDim.x;
    unsigned int i = threadIdx.x;
    for (; i + 3 * st < n; i += 4 * st)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            sa[e] += pa[i + e * st];
            sb[e] += pb[i + e * st];
            sv[e] += pc[i + e * st];
        }
    for (; i < n; i += st) {
        sa[0] += pa[i];
        sb[0] += pb[i];
        sv[0] += pc[i];
    }
    double t, rr, vv;
    block_sum3((sa[0] + sa[1]) + (sa[2] + sa[3]), (sb[0] + sb[1]) + (sb[2] + sb[3]),
               (sv[0] + sv[1]) + (sv[2] + sv[3]), red, t, rr, vv);
    if (threadIdx.x != 0) return;
    const double alpha = t;
    const double b2 = rr - 2.0 * alpha * t + alpha * alpha * vv;
    const double beta = sqrt(b2);
    const long long j = sc->step;
    // no correct digit left in b2 (a vanishing or cancelled residual): record
    // the first such step; the host reports it as an error after the run
    if (!(b2 > 64.0 * 2.220446049250313e-16 * rr) && sc->breakdown == 0) sc->breakdown = j + 1;
    if (alpha_out) alpha_out[j] = alpha;
    if (beta_out) beta_out[j + 1] = beta;
    if (beta2_out) beta2_out[j + 1] = b2;
    sc->alpha = alpha;
    sc->beta = beta;
    sc->beta2 = b2;
    sc->step = j + 1;
    sc->ca = 1.0 / beta;
    sc->cb = -alpha / beta;
    sc->cr = -beta;
}

// ---------------------------------------------------------------------------
// kinetic terms on MFMA (v_mfma_f64_16x16x4f64): the contractions are small
// GEMMs with the slab (or column block) staged in LDS, zero padded to 64,
// and the Toeplitz operand read from a reflected first row, t[64 + d] =
// t(|d|).  FP64 accumulator layout: lane l holds rows (l/16) + 4*q, q = 0..3,
// of column l%16; A operand lane l: row l%16, k = l/16; B operand lane l:
// k = l/16, column l%16.
// ---------------------------------------------------------------------------
typedef double kd4 __attribute__((ext_vector_type(4)));
typedef float kf4 __attribute__((ext_vector_type(4)));

__device__ inline kd4 kmfma(double a, double b, kd4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ inline kf4 kmfma(float a, float b, kf4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

template <typename T>
struct KinAcc {
    typedef kd4 type;
};
template <>
struct KinAcc<float> {
    typedef kf4 type;
};

// accumulator row of item q in lane group kq: FP64 instruction kq + 4q,
// FP32 instruction 4kq + q
template <typename T>
__device__ inline int kin_row(int kq, int q)
{
    return sizeof(T) == 8 ? kq + 4 * q : 4 * kq + q;
}

// BF16x3 kinetic contractions (KM = 1, FP32 operands in LDS): each FP32
// operand x = hi + lo is split into two bfloat16 values (round to nearest
// even) and a product is hi*hi + hi*lo + lo*hi, on gfx950's K = 32 BF16 MFMA
// (v_mfma_f32_16x16x32_bf16: A lane l row l%16, k = 8(l/16) + e; B lane l
// k = 8(l/16) + e, column l%16; the FP32 accumulator as 16x16x4) with FP32
// accumulation.  One instruction covers the whole half-size contraction
// (at most 32 terms).
typedef __bf16 kbf8 __attribute__((ext_vector_type(8)));
__device__ inline __bf16 bf16_from_bits(unsigned short b) { return __builtin_bit_cast(__bf16, b); }
__device__ inline unsigned short bf16_rne_bits(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
// element e of the hi and lo vectors from x
__device__ inline void bf16x3_split(float x, kbf8 &hi, kbf8 &lo, int e)
{
    const unsigned short h = bf16_rne_bits(x);
    hi[e] = bf16_from_bits(h);
    lo[e] = bf16_from_bits(bf16_rne_bits(x - __uint_as_float((unsigned)h << 16)));
}
__device__ inline kf4 kmfma_bf16x3(const kbf8 &ah, const kbf8 &al, const kbf8 &bh, const kbf8 &bl, kf4 c)
{
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
}

constexpr int kSlabLd = 65;

// XCD-aware order of a (tile, plane) grid: workgroups are dealt to the 8 XCDs
// round-robin by linear id, so neighbouring tiles would run on different XCDs
// and fetch the cache lines they share twice; each XCD gets a contiguous range
// of the (plane-major) tile sequence instead.  A bijection of the grid (the
// identity when its size is not a multiple of 8).
constexpr int kXcds = 8;
// the same for a one-dimensional grid of any size: XCD x (ids = x mod 8)
// takes the contiguous range starting at x * (total / 8) + min(x, total % 8)
__device__ inline int xcd_linear(int id, int total)
{
    const int q = total / kXcds, r = total % kXcds, x = id % kXcds;
    return x * q + (x < r ? x : r) + id / kXcds;
}
__device__ inline void xcd_tile(int &tile, int &plane)
{
    const int nt = (int)gridDim.x, total = nt * (int)gridDim.y;
    const int id = (int)(blockIdx.x + gridDim.x * blockIdx.y);
    const int lin = (total % kXcds == 0) ? (id % kXcds) * (total / kXcds) + id / kXcds : id;
    tile = lin % nt;
    plane = lin / nt;
}

// T is symmetric Toeplitz and centrosymmetric, so with mirror index
// i' = n-1-i the sums and differences v(i) +- v(i'), i < m = ceil(n/2), are
// mapped separately by the m x m matrices t(|i - i2|) +- t(n-1-i-i2): half
// the multiply-adds of the full contraction.  For odd n the middle point
// h = (n-1)/2 is its own mirror: its sum 2 v(h) enters the sum block with
// half weight (the operand t(|i - h|) instead of 2 t(|i - h|)), its
// difference is zero, and the outputs written to i and i' at the middle
// are the same value.
constexpr int kCompLd = 33;

// x and y axes on the four x/y parity components of the slab,
// v_pq(j, i) = v(j, i) + p v(j, i') + q v(j', i) + pq v(j', i'), j, i < n/2;
// wave w owns the 16 x 16 tile (w/2, w%2) of all four components, so each lane
// holds the four transformed components of its elements and recombines them
// into the four mirror points.
template <typename T, int KM>
__device__ __forceinline__ void kin_xy_half_body(int k, int n, double *__restrict__ X, double *__restrict__ R,
                                                 const double *__restrict__ tx, const double *__restrict__ ty,
                                                 const double *__restrict__ rot, const double *__restrict__ mask,
                                                 const Scal *__restrict__ sc, int nx, int ny, int nz)
{
    typedef typename KinAcc<T>::type acc_t;
    extern __shared__ __align__(16) unsigned char kin_lds[];
    T *comp = (T *)kin_lds;                // [4][32][kCompLd]
    T *rx = comp + 4 * 32 * kCompLd;       // [129]
    T *ry = rx + 129;                      // [129]
    const double ca = sc->ca, cb = sc->cb, cr = sc->cr;
    const size_t base = ((size_t)n * nz + k) * (size_t)ny * nx;
    const int mx = (nx + 1) >> 1, my = (ny + 1) >> 1;
    const int hx = (nx & 1) ? nx >> 1 : -1, hy = (ny & 1) ? ny >> 1 : -1;  // middle points of odd axes
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int l16 = lane & 15, kq = lane >> 4;
    const int ta = wid >> 1, tb = wid & 1;
    const int i = 16 * tb + l16;  // this lane's component column
    // this lane's 4 items x 4 mirror points: R (kept for the epilogue) and
    // v = ca X + cb R (written back to X), then the four parity components
    // of v, zero padded to 32 x 32; the lanes cover the 32 x 32 component
    // grid once
    double rp[4][4];
    {
        double xp[4][4], vp[4][4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int j = 16 * ta + kin_row<T>(kq, q);
            const bool ok = j < my && i < mx;
            const int jj[2] = {j, ny - 1 - j}, ii[2] = {i, nx - 1 - i};
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const size_t o = base + (size_t)jj[m >> 1] * nx + ii[m & 1];
                rp[q][m] = ok ? R[o] : 0.0;
                xp[q][m] = ok ? X[o] : 0.0;
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int j = 16 * ta + kin_row<T>(kq, q);
            const bool ok = j < my && i < mx;
            const int jj[2] = {j, ny - 1 - j}, ii[2] = {i, nx - 1 - i};
#pragma unroll
            for (int m = 0; m < 4; m++) {
                vp[q][m] = ca * xp[q][m] + cb * rp[q][m];
                if (ok) X[base + (size_t)jj[m >> 1] * nx + ii[m & 1]] = vp[q][m];
            }
            const double a = vp[q][0] + vp[q][1], b = vp[q][0] - vp[q][1];
            const double cc = vp[q][2] + vp[q][3], d = vp[q][2] - vp[q][3];
            comp[(0 * 32 + j) * kCompLd + i] = (T)(a + cc);  // p = +, q = +
            comp[(1 * 32 + j) * kCompLd + i] = (T)(b + d);   // p = -, q = +
            comp[(2 * 32 + j) * kCompLd + i] = (T)(a - cc);  // p = +, q = -
            comp[(3 * 32 + j) * kCompLd + i] = (T)(b - d);   // p = -, q = -
        }
    }
    for (int t = threadIdx.x; t < 129; t += blockDim.x) {
        const int d = abs(t - 64);
        rx[t] = d < nx ? (T)tx[d] : (T)0;
        ry[t] = d < ny ? (T)ty[d] : (T)0;
    }
    __syncthreads();
    acc_t acc[4];
#pragma unroll
    for (int c = 0; c < 4; c++) acc[c] = acc_t{0, 0, 0, 0};
    const int jr = 16 * ta + l16;  // A-operand row of this lane
    if constexpr (KM == 0) {
        const int kx = (mx + 3) >> 2, ky = (my + 3) >> 2;
        for (int s = 0; s < kx; s++) {
            const int i2 = 4 * s + kq;
            const T t1 = rx[64 + i2 - i], t2 = rx[64 + (nx - 1 - i2 - i)];
            const T bp = i2 == hx ? t1 : t1 + t2, bm = t1 - t2;
#pragma unroll
            for (int c = 0; c < 4; c++)
                acc[c] = kmfma(comp[(c * 32 + jr) * kCompLd + i2], (c & 1) ? bm : bp, acc[c]);
        }
        for (int s = 0; s < ky; s++) {
            const int j2 = 4 * s + kq;
            const T t1 = ry[64 + jr - j2], t2 = ry[64 + (ny - 1 - jr - j2)];
            const T ap = j2 == hy ? t1 : t1 + t2, am = t1 - t2;
#pragma unroll
            for (int c = 0; c < 4; c++)
                acc[c] = kmfma((c & 2) ? am : ap, comp[(c * 32 + j2) * kCompLd + i], acc[c]);
        }
    } else {
        // x: K = i2 over [0, 32) (zero components beyond mx), lane k = 8 kq + e
        kbf8 ph, pl, mh, ml;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int i2 = 8 * kq + e;
            const float t1 = rx[64 + i2 - i], t2 = rx[64 + (nx - 1 - i2 - i)];
            bf16x3_split(i2 == hx ? t1 : t1 + t2, ph, pl, e);
            bf16x3_split(t1 - t2, mh, ml, e);
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            kbf8 ah, al;
#pragma unroll
            for (int e = 0; e < 8; e++) bf16x3_split(comp[(c * 32 + jr) * kCompLd + 8 * kq + e], ah, al, e);
            acc[c] = (c & 1) ? kmfma_bf16x3(ah, al, mh, ml, acc[c]) : kmfma_bf16x3(ah, al, ph, pl, acc[c]);
        }
        // y: A the Toeplitz rows jr, K = j2
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int j2 = 8 * kq + e;
            const float t1 = ry[64 + jr - j2], t2 = ry[64 + (ny - 1 - jr - j2)];
            bf16x3_split(j2 == hy ? t1 : t1 + t2, ph, pl, e);
            bf16x3_split(t1 - t2, mh, ml, e);
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            kbf8 bh, bl;
#pragma unroll
            for (int e = 0; e < 8; e++) bf16x3_split(comp[(c * 32 + 8 * kq + e) * kCompLd + i], bh, bl, e);
            acc[c] = (c & 2) ? kmfma_bf16x3(mh, ml, bh, bl, acc[c]) : kmfma_bf16x3(ph, pl, bh, bl, acc[c]);
        }
    }
    const double rn = rot[n], mn = mask[n];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int j = 16 * ta + kin_row<T>(kq, q);
        if (j >= my || i >= mx) continue;
        const double d0 = (double)acc[0][q], d1 = (double)acc[1][q], d2 = (double)acc[2][q], d3 = (double)acc[3][q];
        // mirror point m = (y mirrored) * 2 + (x mirrored): signs p^mx q^my
        const double o[4] = {0.25 * ((d0 + d1) + (d2 + d3)), 0.25 * ((d0 - d1) + (d2 - d3)),
                             0.25 * ((d0 + d1) - (d2 + d3)), 0.25 * ((d0 - d1) - (d2 - d3))};
        const double c0 = (double)comp[(0 * 32 + j) * kCompLd + i], c1 = (double)comp[(1 * 32 + j) * kCompLd + i];
        const double c2 = (double)comp[(2 * 32 + j) * kCompLd + i], c3 = (double)comp[(3 * 32 + j) * kCompLd + i];
        const double v[4] = {0.25 * ((c0 + c1) + (c2 + c3)), 0.25 * ((c0 - c1) + (c2 - c3)),
                             0.25 * ((c0 + c1) - (c2 + c3)), 0.25 * ((c0 - c1) - (c2 - c3))};
        const int jj[2] = {j, ny - 1 - j}, ii[2] = {i, nx - 1 - i};
#pragma unroll
        for (int m = 0; m < 4; m++)
            R[base + (size_t)jj[m >> 1] * nx + ii[m & 1]] = cr * rp[q][m] + mn * (o[m] + rn * v[m]);
    }
}

template <typename T, int KM>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_kin_xy_half(
    double *__restrict__ X, double *__restrict__ R, const double *__restrict__ tx, const double *__restrict__ ty,
    const double *__restrict__ rot, const double *__restrict__ mask, const Scal *__restrict__ sc, int nx, int ny,
    int nz)
{
    kin_xy_half_body<T, KM>(blockIdx.x, blockIdx.y, X, R, tx, ty, rot, mask, sc, nx, ny, nz);
}

// z axis of 64 columns with the z parity split: s = v(k) + v(k'), d = v(k) - v(k'),
// k < nz/2; wave w owns columns 16w.. of both components, every k tile.
template <typename T, int KM>
__device__ __forceinline__ void kin_z_half_body(int tile, int n, const double *__restrict__ X, double *__restrict__ R,
                                                const double *__restrict__ tz, const double *__restrict__ mask,
                                                int nxy, int nz)
{
    typedef typename KinAcc<T>::type acc_t;
    extern __shared__ __align__(16) unsigned char kin_lds[];
    T *cs = (T *)kin_lds;                  // [32][kSlabLd]: sums
    T *cd = cs + 32 * kSlabLd;             // [32][kSlabLd]: differences
    T *rz = cd + 32 * kSlabLd;             // [129]
    const int c0 = tile * 64;
    const size_t base = (size_t)n * nz * nxy;
    const int mz = (nz + 1) >> 1;
    const int hz = (nz & 1) ? nz >> 1 : -1;  // middle point of an odd axis
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int l16 = lane & 15, kq = lane >> 4;
    const int cw = 16 * wid + l16, c = c0 + cw;
    double rp[2][4][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int kk = 16 * a + kin_row<T>(kq, q);
            const bool ok = kk < mz && c < nxy;
            rp[a][q][0] = ok ? R[base + (size_t)kk * nxy + c] : 0.0;
            rp[a][q][1] = ok ? R[base + (size_t)(nz - 1 - kk) * nxy + c] : 0.0;
        }
    {
        // the 32 x 64 mirror pairs of the tile, all loads issued before the
        // LDS stores (8 pairs per thread at 256 threads)
        static_assert(32 * 64 == 8 * kWG, "z tile staging assumes 256 threads");
        double v0[8], v1[8];
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int t = threadIdx.x + e * kWG, kk = t >> 6, cc = t & 63;
            const bool ok = kk < mz && c0 + cc < nxy;
            v0[e] = ok ? X[base + (size_t)kk * nxy + c0 + cc] : 0.0;
            v1[e] = ok ? X[base + (size_t)(nz - 1 - kk) * nxy + c0 + cc] : 0.0;
        }
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int t = threadIdx.x + e * kWG, kk = t >> 6, cc = t & 63;
            cs[kk * kSlabLd + cc] = (T)(v0[e] + v1[e]);
            cd[kk * kSlabLd + cc] = (T)(v0[e] - v1[e]);
        }
    }
    for (int t = threadIdx.x; t < 129; t += blockDim.x) {
        const int d = abs(t - 64);
        rz[t] = d < nz ? (T)tz[d] : (T)0;
    }
    __syncthreads();
    acc_t as[2], ad[2];
#pragma unroll
    for (int a = 0; a < 2; a++) as[a] = ad[a] = acc_t{0, 0, 0, 0};
    if constexpr (KM == 0) {
        const int kz = (mz + 3) >> 2;
        for (int s = 0; s < kz; s++) {
            const int kp = 4 * s + kq;
            const T bs = cs[kp * kSlabLd + cw], bd = cd[kp * kSlabLd + cw];
#pragma unroll
            for (int a = 0; a < 2; a++) {
                const int kr = 16 * a + l16;
                const T t1 = rz[64 + kr - kp], t2 = rz[64 + (nz - 1 - kr - kp)];
                as[a] = kmfma(kp == hz ? t1 : t1 + t2, bs, as[a]);
                ad[a] = kmfma(t1 - t2, bd, ad[a]);
            }
        }
    } else {
        // K = kp over [0, 32) (zero rows beyond mz), lane k = 8 kq + e
        kbf8 sh, sl, dh, dl;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            bf16x3_split(cs[(8 * kq + e) * kSlabLd + cw], sh, sl, e);
            bf16x3_split(cd[(8 * kq + e) * kSlabLd + cw], dh, dl, e);
        }
#pragma unroll
        for (int a = 0; a < 2; a++) {
            const int kr = 16 * a + l16;
            kbf8 ph, pl, mh, ml;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int kp = 8 * kq + e;
                const float t1 = rz[64 + kr - kp], t2 = rz[64 + (nz - 1 - kr - kp)];
                bf16x3_split(kp == hz ? t1 : t1 + t2, ph, pl, e);
                bf16x3_split(t1 - t2, mh, ml, e);
            }
            as[a] = kmfma_bf16x3(ph, pl, sh, sl, as[a]);
            ad[a] = kmfma_bf16x3(mh, ml, dh, dl, ad[a]);
        }
    }
    const double mn = mask[n];
    if (c >= nxy) return;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int kk = 16 * a + kin_row<T>(kq, q);
            if (kk >= mz) continue;
            const double S = (double)as[a][q], D = (double)ad[a][q];
            R[base + (size_t)kk * nxy + c] = rp[a][q][0] + mn * (0.5 * (S + D));
            R[base + (size_t)(nz - 1 - kk) * nxy + c] = rp[a][q][1] + mn * (0.5 * (S - D));
        }
}

template <typename T, int KM>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_kin_z_half(
    const double *__restrict__ X, double *__restrict__ R, const double *__restrict__ tz, const double *__restrict__ mask,
    const Scal *__restrict__ sc, int nxy, int nz)
{
    (void)sc;  // X already holds v
    int tile, n;
    xcd_tile(tile, n);
    kin_z_half_body<T, KM>(tile, n, X, R, tz, mask, nxy, nz);
}

// grid-stride launches: workgroups of kWG threads for n items, at most cap
static int grid_for(int64_t n, int cap = 4096)
{
    int64_t g = (n + kWG - 1) / kWG;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

// ---------------------------------------------------------------------------
// kinetic term for axes of more than 64 points (FP64, any length): the
// recurrence prologue of the x-y kernel as a pass of its own, then one
// thread per (point, plane) summing the three axis contractions along its
// lines, read through the caches (Tv_5D, linRotCartSphHvRoutines.cpp:1872-1969:
// R += T_x v + T_y v + T_z v with T_a(i, i') = t_a(|i - i'|)).
// ---------------------------------------------------------------------------
// v = ca X + cb R into X; R = cr R + mask B l(l+1) v
__global__ __launch_bounds__(kWG) void k_kin_form(double *__restrict__ X, double *__restrict__ R,
                                                  const double *__restrict__ rot, const double *__restrict__ mask,
                                                  const Scal *__restrict__ sc, int64_t P, int64_t n)
{
    const double ca = sc->ca, cb = sc->cb, cr = sc->cr;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / P);
        const double v = ca * X[i] + cb * R[i];
        X[i] = v;
        R[i] = cr * R[i] + mask[r] * (rot[r] * v);
    }
}

// R += mask (T_x + T_y + T_z) v; tx, ty, tz: first rows t(d), d < n
__global__ __launch_bounds__(kWG) void k_kin_lines(const double *__restrict__ X, double *__restrict__ R,
                                                   const double *__restrict__ tx, const double *__restrict__ ty,
                                                   const double *__restrict__ tz, const double *__restrict__ mask,
                                                   int nx, int ny, int nz, int64_t P, int64_t n)
{
    const int64_t nxy = (int64_t)nx * ny;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(idx / P);
        const int64_t p = idx - (int64_t)r * P;
        const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / nxy);
        const double *v = X + (size_t)r * P;
        const double *lx = v + (size_t)k * nxy + (size_t)j * nx;
        const double *ly = v + (size_t)k * nxy + i;
        const double *lz = v + (size_t)j * nx + i;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int a = 0; a < nx; a++) sx = fma(tx[a > i ? a - i : i - a], lx[a], sx);
        for (int a = 0; a < ny; a++) sy = fma(ty[a > j ? a - j : j - a], ly[(size_t)a * nx], sy);
        for (int a = 0; a < nz; a++) sz = fma(tz[a > k ? a - k : k - a], lz[(size_t)a * nxy], sz);
        R[idx] += mask[r] * ((sx + sy) + sz);
    }
}

static int launch_kinetic_lines(const DevOperator &op, double *X, double *R, Scal *sc, hipStream_t s)
{
    const int64_t n = op.P * op.nk;
    const int g = grid_for(n, 16384);
    timed_launch(k_kin_form, dim3(g), dim3(kWG), 0, s, X, R, (const double *)op.rot, (const double *)op.mask,
                 (const Scal *)sc, op.P, n);
    BSL_HIP(hipGetLastError());
    timed_launch(k_kin_lines, dim3(g), dim3(kWG), 0, s, (const double *)X, R, (const double *)op.tx,
                 (const double *)op.ty, (const double *)op.tz, (const double *)op.mask, op.nx, op.ny, op.nz, op.P,
                 n);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

bool kinetic_mfma_fits(const DevOperator &op) { return op.nx <= 64 && op.ny <= 64 && op.nz <= 64; }

// with_z false: the x-y kernel only (the fused kernel applies T_z)
template <typename T, int KM>
static int launch_kinetic_t(const DevOperator &op, double *X, double *R, Scal *sc, hipStream_t s, bool with_z = true)
{
    // the half-size (centrosymmetric) contractions, for every axis length
    const size_t lds_xy = sizeof(T) * ((size_t)4 * 32 * kCompLd + 2 * 129);
    timed_launch((k_kin_xy_half<T, KM>), dim3(op.nz, op.nk), dim3(kWG), lds_xy, s, X, R, op.tx, op.ty, op.rot,
                       op.mask, sc, op.nx, op.ny, op.nz);
    BSL_HIP(hipGetLastError());
    if (!with_z) return BSL_OK;
    const int nxy = op.nx * op.ny;
    const size_t lds_z = sizeof(T) * ((size_t)2 * 32 * kSlabLd + 129);
    timed_launch((k_kin_z_half<T, KM>), dim3((nxy + 63) / 64, op.nk), dim3(kWG), lds_z, s, X, R, op.tz, op.mask,
                       sc, nxy, op.nz);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

// ---------------------------------------------------------------------------
// potential term: one thread per (grid point, parity block)
// ---------------------------------------------------------------------------
// R[n_r][P] += mask * sum_c V_P[r][c] X[n_c][P] over the packed upper
// triangle; then the contributions of these planes to v.R, |R|^2, |v|^2.
// One thread per grid point.  The packed upper triangle of a B x B block is
// stored row pair by row pair: row r (columns r..B-1) followed by row
// B-1-r (columns B-1-r..B-1), for r = 0 .. (B-1)/2, so that each row pair
// is B+1 consecutive streams.  The B+1 loads of a row pair are issued back
// to back through a running pointer before they are consumed, which keeps
// B+1 loads per wave in flight.
template <int B, typename VT>
__device__ inline void vblock_point(const double *__restrict__ X, double *__restrict__ R, const VT *__restrict__ V,
                                    int plane0, double mask, int64_t P, int p, double &dot, double &rr2, double &vv2)
{
    double x[B], acc[B];
    const double *xp = X + (size_t)plane0 * P + p;
#pragma unroll
    for (int r = 0; r < B; r++) {
        x[r] = *xp;
        xp += P;
        acc[r] = 0.0;
    }
    // this point's tile of the block: B(B+1)/2 pairs of kVTile points
    const VT *vp = V + (size_t)(p / kVTile) * (B * (B + 1) / 2) * kVTile + p % kVTile;
#pragma unroll
    for (int r = 0; r <= (B - 1) / 2; r++) {
        const int r2 = B - 1 - r;
        VT vv[B + 1];
        int k = 0;
#pragma unroll
        for (int c = r; c < B; c++) {
            vv[k++] = __builtin_nontemporal_load(vp);
            vp += kVTile;
        }
        if (r2 != r) {
#pragma unroll
            for (int c = r2; c < B; c++) {
                vv[k++] = __builtin_nontemporal_load(vp);
                vp += kVTile;
            }
        }
        k = 0;
#pragma unroll
        for (int c = r; c < B; c++) {
            const double v = (double)vv[k++];
            acc[r] += v * x[c];
            if (c != r) acc[c] += v * x[r];
        }
        if (r2 != r) {
#pragma unroll
            for (int c = r2; c < B; c++) {
                const double v = (double)vv[k++];
                acc[r2] += v * x[c];
                if (c != r2) acc[c] += v * x[r2];
            }
        }
    }
    double *rp = R + (size_t)plane0 * P + p;
    double rr[B];
#pragma unroll
    for (int r = 0; r < B; r++) rr[r] = rp[(size_t)r * P];
#pragma unroll
    for (int r = 0; r < B; r++) {
        const double nr = rr[r] + mask * acc[r];
        rp[(size_t)r * P] = nr;
        dot += x[r] * nr;
        rr2 += nr * nr;
        vv2 += x[r] * x[r];
    }
}

// The carried planes are ordered block by block: block 0 occupies planes
// [0, B0), block 1 planes [B0, B0 + B1).  The symmetry mask is uniform
// within a parity block.
template <int B0, int B1, typename VT, int W>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(W, 8))) void k_vblock(
    const double *__restrict__ X, double *__restrict__ R, const VT *__restrict__ V0, const VT *__restrict__ V1,
    double mask0, double mask1, int64_t P, double *__restrict__ partial, double *__restrict__ partial_b,
    double *__restrict__ partial_c)
{
    __shared__ double red[3 * kWG / 64];
    double dot = 0.0, rr2 = 0.0, vv2 = 0.0;
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    // the larger block's workgroups are dispatched first (blockIdx.y = 0), so
    // that the shorter ones fill the tail of the launch
    const int blk = (B1 > B0) ? (int)(gridDim.y - 1 - blockIdx.y) : (int)blockIdx.y;
    if (p < P) {
        if (blk == 0)
            vblock_point<B0, VT>(X, R, V0, 0, mask0, P, p, dot, rr2, vv2);
        else if (B1 > 0)
            vblock_point<(B1 > 0 ? B1 : 1), VT>(X, R, V1, B0, mask1, P, p, dot, rr2, vv2);
    }
    const int w = blockIdx.y * gridDim.x + blockIdx.x;
    double s0, s1, s2;
    block_sum3(dot, rr2, vv2, red, s0, s1, s2);
    if (threadIdx.x == 0) {
        partial[w] = s0;
        partial_b[w] = s1;
        partial_c[w] = s2;
    }
}

// Packed position of element (a, c), a <= c, of a B x B block in the row
// pair order above: row pair m = min(a, B-1-a) starts at m (B+1); row m
// comes first (B - m columns), then row B-1-m.
__device__ inline int vblock_pair(int a, int c, int B)
{
    const int m = a < B - 1 - a ? a : B - 1 - a;
    return m * (B + 1) + (a == m ? 0 : B - m) + (c - a);
}

// Blocks of any size (those without a register kernel): 64 points per
// workgroup, one per lane, the block's X rows staged in LDS ([row][64]);
// wave w computes rows w, w + 4, ..., each packed element read for both of
// its rows (twice the V bytes of the register kernels, no register arrays).
// Grid (ceil(P/64), blocks); partial index blockIdx.y * gridDim.x + blockIdx.x.
template <typename VT>
__global__ __launch_bounds__(kWG) void k_vblock_rows(const double *__restrict__ X, double *__restrict__ R,
                                                     const VT *__restrict__ V0, const VT *__restrict__ V1, int B0,
                                                     int B1, double mask0, double mask1, int64_t P,
                                                     double *__restrict__ partial, double *__restrict__ partial_b,
                                                     double *__restrict__ partial_c)
{
    extern __shared__ double xrow[];  // [B][64]
    __shared__ double red[3 * kWG / 64];
    const int blk = (int)blockIdx.y;
    const int B = blk == 0 ? B0 : B1, plane0 = blk == 0 ? 0 : B0;
    const double mask = blk == 0 ? mask0 : mask1;
    const VT *V = blk == 0 ? V0 : V1;
    const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
    const int64_t g0 = (int64_t)blockIdx.x * 64;
    for (int t = threadIdx.x; t < B * 64; t += blockDim.x) {
        const int r = t >> 6, l = t & 63;
        const int64_t q = g0 + l < P ? g0 + l : P - 1;
        xrow[t] = X[(size_t)(plane0 + r) * P + q];
    }
    __syncthreads();
    const int64_t p = g0 + lane;
    const bool live = p < P;
    const int64_t pl = live ? p : P - 1;
    const VT *vp = V + (size_t)(pl / kVTile) * ((size_t)B * (B + 1) / 2) * kVTile + pl % kVTile;
    double dot = 0.0, rr2 = 0.0, vv2 = 0.0;
    for (int i = wave; i < B; i += kWG / 64) {
        double acc = 0.0;
        for (int c = 0; c < B; c++) {
            const int off = c >= i ? vblock_pair(i, c, B) : vblock_pair(c, i, B);
            acc = fma((double)vp[(size_t)off * kVTile], xrow[c * 64 + lane], acc);
        }
        if (live) {
            const size_t o = (size_t)(plane0 + i) * P + p;
            const double x = xrow[i * 64 + lane];
            const double nr = R[o] + mask * acc;
            R[o] = nr;
            dot += x * nr;
            rr2 += nr * nr;
            vv2 += x * x;
        }
    }
    const int w = blockIdx.y * gridDim.x + blockIdx.x;
    double s0, s1, s2;
    block_sum3(dot, rr2, vv2, red, s0, s1, s2);
    if (threadIdx.x == 0) {
        partial[w] = s0;
        partial_b[w] = s1;
        partial_c[w] = s2;
    }
}

// ---------------------------------------------------------------------------
// potential term, harmonic form: one thread per (grid point, parity block)
// ---------------------------------------------------------------------------
// The block of a point is rebuilt from its harmonic coefficients h[c] as the
// sparse sum V[i][j] = sum_c G[t] h[c] over the block's term list (one run
// of terms per packed pair), each element applied to both of its entries
// at once.  The list is a compile-time constant, so h, x and acc stay in
// registers; the coefficients G[t] are uniform across the wave (scalar
// loads at constant offsets).
template <class Pat, int T>
__device__ __forceinline__ void gaunt_term(const double *__restrict__ g, const double (&h)[Pat::NC],
                                           const double (&x)[Pat::B], double (&acc)[Pat::B], double &e)
{
    constexpr gaunt::Term tm = Pat::list.t[T];
    if constexpr (tm.first)
        e = g[tm.g] * h[tm.c];
    else
        e = fma(g[tm.g], h[tm.c], e);
    if constexpr (tm.last) {
        acc[tm.i] = fma(e, x[tm.j], acc[tm.i]);
        if constexpr (tm.i != tm.j) acc[tm.j] = fma(e, x[tm.i], acc[tm.j]);
    }
}

template <class Pat, int... T>
__device__ __forceinline__ void gaunt_terms(const double *__restrict__ g, const double (&h)[Pat::NC],
                                            const double (&x)[Pat::B], double (&acc)[Pat::B],
                                            std::integer_sequence<int, T...>)
{
    double e = 0.0;
    (gaunt_term<Pat, T>(g, h, x, acc, e), ...);
}

// H is [tile][column][kVTile] with nc columns per point; a block of lb < the
// context's largest reads the leading ncoef(lb) columns.
template <int LB, int PAR, typename VT>
__device__ inline void gaunt_point(const double *__restrict__ X, double *__restrict__ R, const VT *__restrict__ H,
                                   const double *__restrict__ G, int nc, int plane0, double mask, int64_t P, int p,
                                   double &dot, double &rr2, double &vv2)
{
    using Pat = gaunt::Pattern<LB, PAR>;
    constexpr int B = Pat::B, NC = Pat::NC;
    double x[B], acc[B], h[NC];
    const double *xp = X + (size_t)plane0 * P + p;
#pragma unroll
    for (int r = 0; r < B; r++) {
        x[r] = xp[(size_t)r * P];
        acc[r] = 0.0;
    }
    const VT *hp = H + (size_t)(p / kVTile) * nc * kVTile + p % kVTile;
#pragma unroll
    for (int c = 0; c < NC; c++) h[c] = (double)hp[(size_t)c * kVTile];
    gaunt_terms<Pat>(G, h, x, acc, std::make_integer_sequence<int, Pat::NT>{});
    double *rp = R + (size_t)plane0 * P + p;
    double rr[B];
#pragma unroll
    for (int r = 0; r < B; r++) rr[r] = rp[(size_t)r * P];
#pragma unroll
    for (int r = 0; r < B; r++) {
        const double nr = rr[r] + mask * acc[r];
        rp[(size_t)r * P] = nr;
        dot += x[r] * nr;
        rr2 += nr * nr;
        vv2 += x[r] * x[r];
    }
}

// NB parity blocks per workgroup: threads [0, kWG) take block 0 of the
// workgroup's kVTile points, threads [kWG, 2 kWG) block 1 of the same points,
// so that the coefficient columns both blocks read are fetched once.
// Block 0 occupies planes [0, plane1), block 1 planes [plane1, ...).
template <int LB0, int PAR0, int LB1, int PAR1, typename VT, int W>
__global__ __launch_bounds__((LB1 >= 0 ? 2 : 1) * kWG) __attribute__((amdgpu_waves_per_eu(W, 8))) void k_vgaunt(
    const double *__restrict__ X, double *__restrict__ R, const VT *__restrict__ H, const double *__restrict__ G0,
    const double *__restrict__ G1, int nc, int plane1, double mask0, double mask1, int64_t P,
    double *__restrict__ partial, double *__restrict__ partial_b, double *__restrict__ partial_c)
{
    __shared__ double red[3 * 2 * kWG / 64];
    double dot = 0.0, rr2 = 0.0, vv2 = 0.0;
    const int half = (int)threadIdx.x / kWG;
    const int p = (int)blockIdx.x * kWG + ((int)threadIdx.x & (kWG - 1));
    if (p < P) {
        if constexpr (LB1 >= 0) {
            if (half == 0)
                gaunt_point<LB0, PAR0, VT>(X, R, H, G0, nc, 0, mask0, P, p, dot, rr2, vv2);
            else
                gaunt_point<LB1, PAR1, VT>(X, R, H, G1, nc, plane1, mask1, P, p, dot, rr2, vv2);
        } else {
            gaunt_point<LB0, PAR0, VT>(X, R, H, G0, nc, 0, mask0, P, p, dot, rr2, vv2);
        }
    }
    double s0, s1, s2;
    block_sum3(dot, rr2, vv2, red, s0, s1, s2);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s0;
        partial_b[blockIdx.x] = s1;
        partial_c[blockIdx.x] = s2;
    }
}

// One sub-list of a block for the wave's 64 points: the x entries (from the
// workgroup's copy of its X rows in LDS) and coefficient columns it touches
// are loaded, its terms accumulated, and the rows it touches added to
// part[row][lane] (LDS).  Every index is a
// compile-time constant (index-sequence expansions), so x, acc and h stay in
// registers.  Lanes past the last point load the last point (in bounds,
// unconditional loads issue back to back) and are never written back.
template <class Pat, int R_>
__device__ __forceinline__ void sub_load_x(double (&x)[Pat::B], double (&acc)[Pat::B], const double *__restrict__ xp,
                                           int64_t P)
{
    acc[R_] = 0.0;
    if constexpr (Pat::list.row[R_])
        x[R_] = xp[(size_t)R_ * P];
    else
        x[R_] = 0.0;
}
template <class Pat, typename VT, int C>
__device__ __forceinline__ void sub_load_h(double (&h)[Pat::NC], const VT *__restrict__ hp)
{
    if constexpr (Pat::list.col[C])
        h[C] = (double)hp[(size_t)C * kVTile];
    else
        h[C] = 0.0;
}
// Every row takes exactly one same-type and one cross partial, added onto
// zero: two adds give a + b whatever their order, so the LDS atomic keeps
// the result deterministic.
template <class Pat, int R_>
__device__ __forceinline__ void sub_store(const double (&acc)[Pat::B], double (*part)[64], int lane)
{
    if constexpr (Pat::list.row[R_])
        __hip_atomic_fetch_add(&part[R_][lane], acc[R_], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class Pat, typename VT, int... Rs, int... Cs>
__device__ __forceinline__ void sub_run(const double *__restrict__ xp, const VT *__restrict__ hp,
                                        const double *__restrict__ G, int64_t P, double (*part)[64], int lane,
                                        std::integer_sequence<int, Rs...>, std::integer_sequence<int, Cs...>)
{
    double x[Pat::B], acc[Pat::B], h[Pat::NC];
    (sub_load_x<Pat, Rs>(x, acc, xp, P), ...);
    (sub_load_h<Pat, VT, Cs>(h, hp), ...);
    gaunt_terms<Pat>(G, h, x, acc, std::make_integer_sequence<int, Pat::NT>{});
    (sub_store<Pat, Rs>(acc, part, lane), ...);
}
// Two sub-lists in one wave sharing the coefficient columns: B's columns
// must be among A's (B of a smaller or equal lb: its columns are a prefix).
template <class PA, class PB, typename VT, int... RAs, int... RBs, int... Cs>
__device__ __forceinline__ void sub_run2(const double *__restrict__ xpa, const double *__restrict__ xpb,
                                         const VT *__restrict__ hp, const double *__restrict__ GA,
                                         const double *__restrict__ GB, int64_t P, double (*parta)[64],
                                         double (*partb)[64], int lane, std::integer_sequence<int, RAs...>,
                                         std::integer_sequence<int, RBs...>, std::integer_sequence<int, Cs...>)
{
    static_assert(PB::NC <= PA::NC, "the second sub-list reads a prefix of the first one's columns");
    double xa[PA::B], acca[PA::B], xb[PB::B], accb[PB::B], h[PA::NC];
    (sub_load_x<PA, RAs>(xa, acca, xpa, P), ...);
    (sub_load_x<PB, RBs>(xb, accb, xpb, P), ...);
    (sub_load_h<PA, VT, Cs>(h, hp), ...);
    gaunt_terms<PA>(GA, h, xa, acca, std::make_integer_sequence<int, PA::NT>{});
    gaunt_terms<PB>(GB, reinterpret_cast<const double(&)[PB::NC]>(h), xb, accb,
                    std::make_integer_sequence<int, PB::NT>{});
    (sub_store<PA, RAs>(acca, parta, lane), ...);
    (sub_store<PB, RBs>(accb, partb, lane), ...);
}

// x0: this lane's entry of plane 0 (the workgroup's LDS rows, stride 64, or
// X itself, stride P)
template <class Pat, typename VT>
__device__ __forceinline__ void gaunt_sublist(const double *x0, int64_t xst, const VT *__restrict__ hp,
                                              const double *__restrict__ G, int plane0, double (*part)[64], int lane)
{
    sub_run<Pat, VT>(x0 + (size_t)plane0 * xst, hp, G, xst, part, lane, std::make_integer_sequence<int, Pat::B>{},
                     std::make_integer_sequence<int, Pat::NC>{});
}

// Where the split kernels' sub-lists read X: the FP64-coefficient kernels
// stage the group's rows in LDS once (coalesced, each row fetched once); the
// FP32-coefficient kernels, whose coefficient loads are half the size, read
// X directly (one memory round trip fewer: 42.9 against 45.1 us at 54^3).
template <typename VT>
constexpr bool split_reads_x_directly()
{
    return sizeof(VT) == sizeof(float);
}

// part zeroed by the NT threads
template <int NR, int NT>
__device__ __forceinline__ void zero_rows(double (*part)[64])
{
    for (int t = (int)threadIdx.x; t < NR * 64; t += NT) (&part[0][0])[t] = 0.0;
}

// The workgroup's X rows [0, NR) at points group*64 + [0, 64) into xs (the
// tail clamped to the last point, never written back) and part zeroed; NT
// threads, NR*64 a multiple of NT so that every load issues before the stores.
template <int NR, int NT>
__device__ __forceinline__ void stage_rows(const double *__restrict__ X, int64_t P, int group, double (*xs)[64],
                                           double (*part)[64])
{
    static_assert((NR * 64) % NT == 0, "whole rounds of loads");
    constexpr int K = NR * 64 / NT;
    double v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int t = (int)threadIdx.x + k * NT, r = t >> 6, l = t & 63;
        const int64_t q = (int64_t)group * 64 + l < P ? (int64_t)group * 64 + l : P - 1;
        v[k] = X[(size_t)r * P + q];
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int t = (int)threadIdx.x + k * NT, r = t >> 6, l = t & 63;
        xs[r][l] = v[k];
        part[r][l] = 0.0;
    }
}

// The lmax-5 pair of blocks (l = 0, 2, 4 and l = 1, 3, 5; 15 + 21 planes,
// 1597 terms) split over four waves per 64 points by pair type, so that a
// wave holds only the coefficient columns and rows of its sub-lists:
//   wave 0: odd cross (695 multiply-adds), wave 1: odd cos-cos (512),
//   wave 2: odd sin-sin + even sin-sin (505), wave 3: even cos-cos + even cross (551).
// Each row of a block gets one same-type partial (cos-cos for m >= 0 rows,
// sin-sin for m < 0 rows) and one cross partial, exchanged through LDS; after
// the barrier each wave finishes 9 of the 36 rows as R += mask (same + cross).
template <typename VT>
__device__ __forceinline__ void vgaunt_split_body(int group, const double *__restrict__ X, double *__restrict__ R,
                                                  const VT *__restrict__ H, const double *__restrict__ G0,
                                                  const double *__restrict__ G1, int nc, int plane1, double mask0,
                                                  double mask1, int64_t P, double *__restrict__ partial,
                                                  double *__restrict__ partial_b, double *__restrict__ partial_c)
{
    constexpr int B0 = gaunt::block_size(4, 0), B1 = gaunt::block_size(5, 1), NR = B0 + B1;
    static_assert(NR % 4 == 0, "rows are shared out evenly over the four waves");
    static_assert(gaunt::split_covers_rows<4, 0>() && gaunt::split_covers_rows<5, 1>(),
                  "every row needs a same-type and a cross partial");
    constexpr bool kDirect = split_reads_x_directly<VT>();
    __shared__ double xs[kDirect ? 1 : NR][64];  // X rows of the 64 points (even rows, then odd rows)
    __shared__ double part[NR][64];              // same-type + cross partials [row][lane]
    __shared__ double red[3 * 4];
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int lane = (int)threadIdx.x & 63;
    const int p = group * 64 + lane;
    const bool live = p < P;
    const int pl = live ? p : (int)(P - 1);
    const VT *hp = H + (size_t)(pl / kVTile) * nc * kVTile + pl % kVTile;
    if constexpr (kDirect)
        zero_rows<NR, 256>(part);
    else
        stage_rows<NR, 256>(X, P, group, xs, part);
    __syncthreads();
    const double *x0 = kDirect ? X + pl : &xs[0][lane];
    const int64_t xst = kDirect ? P : 64;
    if (wave == 0) {
        gaunt_sublist<gaunt::Pattern<5, 1, gaunt::kCross>, VT>(x0, xst, hp, G1, plane1, &part[B0], lane);
    } else if (wave == 1) {
        gaunt_sublist<gaunt::Pattern<5, 1, gaunt::kCC>, VT>(x0, xst, hp, G1, plane1, &part[B0], lane);
    } else if (wave == 2) {
        // the even sin-sin columns (M >= 0, L <= 8) are among the odd ones
        using PA = gaunt::Pattern<5, 1, gaunt::kSS>;
        using PB = gaunt::Pattern<4, 0, gaunt::kSS>;
        static_assert(gaunt::ncoef(4) <= gaunt::ncoef(5), "column prefix");
        sub_run2<PA, PB, VT>(x0 + (size_t)plane1 * xst, x0, hp, G1, G0, xst, &part[B0], &part[0], lane,
                             std::make_integer_sequence<int, PA::B>{}, std::make_integer_sequence<int, PB::B>{},
                             std::make_integer_sequence<int, PA::NC>{});
    } else {
        gaunt_sublist<gaunt::Pattern<4, 0, gaunt::kCC>, VT>(x0, xst, hp, G0, 0, &part[0], lane);
        gaunt_sublist<gaunt::Pattern<4, 0, gaunt::kCross>, VT>(x0, xst, hp, G0, 0, &part[0], lane);
    }
    __syncthreads();
    // rows [9 w, 9 w + 9) of (even rows | odd rows); planes are contiguous
    double dot = 0.0, rr2 = 0.0, vv2 = 0.0;
    if (live) {
        constexpr int NW = NR / 4;
        const int r0 = wave * NW;
        double rr[NW], xx[NW];
#pragma unroll
        for (int k = 0; k < NW; k++) {
            rr[k] = R[(size_t)(r0 + k) * P + p];
            xx[k] = kDirect ? X[(size_t)(r0 + k) * P + p] : xs[r0 + k][lane];
        }
#pragma unroll
        for (int k = 0; k < NW; k++) {
            const int r = r0 + k;
            const double mk = r < B0 ? mask0 : mask1;
            const double nr = rr[k] + mk * part[r][lane];
            R[(size_t)r * P + p] = nr;
            dot += xx[k] * nr;
            rr2 += nr * nr;
            vv2 += xx[k] * xx[k];
        }
    }
    double s0, s1, s2;
    block_sum3(dot, rr2, vv2, red, s0, s1, s2);
    if (threadIdx.x == 0) {
        partial[group] = s0;
        partial_b[group] = s1;
        partial_c[group] = s2;
    }
}

template <typename VT, int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, 8))) void k_vgaunt_split(
    const double *__restrict__ X, double *__restrict__ R, const VT *__restrict__ H, const double *__restrict__ G0,
    const double *__restrict__ G1, int nc, int plane1, double mask0, double mask1, int64_t P,
    double *__restrict__ partial, double *__restrict__ partial_b, double *__restrict__ partial_c)
{
    vgaunt_split_body<VT>(xcd_linear((int)blockIdx.x, (int)gridDim.x), X, R, H, G0, G1, nc, plane1, mask0, mask1,
                          P, partial, partial_b, partial_c);
}

// One parity block of the reduced basis (para l = 0, 2, 4 or ortho
// l = 1, 3, 5) split the same way over three waves per 64 points: wave 0 the
// cross pairs, wave 1 cos-cos, wave 2 sin-sin; then each wave finishes
// ceil(B/3) rows as R += mask (same + cross).
template <int LB, int PAR, typename VT>
__device__ __forceinline__ void vgaunt_split1_body(int group, const double *__restrict__ X, double *__restrict__ R,
                                                   const VT *__restrict__ H, const double *__restrict__ G0, int nc,
                                                   double mask0, int64_t P, double *__restrict__ partial,
                                                   double *__restrict__ partial_b, double *__restrict__ partial_c)
{
    constexpr int B = gaunt::block_size(LB, PAR), NW = (B + 2) / 3;
    static_assert(gaunt::split_covers_rows<LB, PAR>(), "every row needs a same-type and a cross partial");
    constexpr bool kDirect = split_reads_x_directly<VT>();
    __shared__ double xs[kDirect ? 1 : B][64];  // X rows of the 64 points
    __shared__ double part[B][64];              // same-type + cross partials [row][lane]
    __shared__ double red[3 * 3];
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int lane = (int)threadIdx.x & 63;
    const int p = group * 64 + lane;
    const bool live = p < P;
    const int pl = live ? p : (int)(P - 1);
    const VT *hp = H + (size_t)(pl / kVTile) * nc * kVTile + pl % kVTile;
    if constexpr (kDirect)
        zero_rows<B, 192>(part);
    else
        stage_rows<B, 192>(X, P, group, xs, part);
    __syncthreads();
    const double *x0 = kDirect ? X + pl : &xs[0][lane];
    const int64_t xst = kDirect ? P : 64;
    if (wave == 0)
        gaunt_sublist<gaunt::Pattern<LB, PAR, gaunt::kCross>, VT>(x0, xst, hp, G0, 0, &part[0], lane);
    else if (wave == 1)
        gaunt_sublist<gaunt::Pattern<LB, PAR, gaunt::kCC>, VT>(x0, xst, hp, G0, 0, &part[0], lane);
    else
        gaunt_sublist<gaunt::Pattern<LB, PAR, gaunt::kSS>, VT>(x0, xst, hp, G0, 0, &part[0], lane);
    __syncthreads();
    double dot = 0.0, rr2 = 0.0, vv2 = 0.0;
    if (live) {
        const int r0 = wave * NW;
        double rr[NW], xx[NW];
#pragma unroll
        for (int k = 0; k < NW; k++) {
            const int r = r0 + k < B ? r0 + k : B - 1;
            rr[k] = R[(size_t)r * P + p];
            xx[k] = kDirect ? X[(size_t)r * P + p] : xs[r][lane];
        }
#pragma unroll
        for (int k = 0; k < NW; k++) {
            const int r = r0 + k;
            if (r < B) {
                const double nr = rr[k] + mask0 * part[r][lane];
                R[(size_t)r * P + p] = nr;
                dot += xx[k] * nr;
                rr2 += nr * nr;
                vv2 += xx[k] * xx[k];
            }
        }
    }
    double s0, s1, s2;
    block_sum3(dot, rr2, vv2, red, s0, s1, s2);
    if (threadIdx.x == 0) {
        partial[group] = s0;
        partial_b[group] = s1;
        partial_c[group] = s2;
    }
}

template <int LB, int PAR, typename VT, int W>
__global__ __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(W, 8))) void k_vgaunt_split1(
    const double *__restrict__ X, double *__restrict__ R, const VT *__restrict__ H, const double *__restrict__ G0,
    int nc, double mask0, int64_t P, double *__restrict__ partial, double *__restrict__ partial_b,
    double *__restrict__ partial_c)
{
    vgaunt_split1_body<LB, PAR, VT>(xcd_linear((int)blockIdx.x, (int)gridDim.x), X, R, H, G0, nc, mask0, P, partial,
                                    partial_b, partial_c);
}

// ---------------------------------------------------------------------------
// z kinetic term fused with the harmonic-form potential of one parity block
// ---------------------------------------------------------------------------
// The z contraction couples the nz points of one column (c = j nx + i) within
// a plane, the potential the planes of one point.  A workgroup that owns every
// point of a few columns for every plane of a parity block applies both with
// one read of X and one read-modify-write of R: three vector passes where the
// z kernel and the potential kernel make six.  (Tv_5D's z term,
// linRotCartSphHvRoutines.cpp:1872-1969, and Vv_5D, linRotCartSphHvFullVPrepRoutines.cpp:675-734.)
//
// Tile: kZpCols consecutive columns x nz x the block's B planes, staged once
// in LDS as xs[row][k][col].  Its points are taken by half index h
// (k = h or nz-1-h, h < ceil(nz/2)) in M-tiles of 16 (at most two, nz <= 64):
// M-tile mt holds h in [16 mt, 16 mt + 16), i.e. two mirror groups of 64
// points with lane (h - 16 mt) * kZpCols + col.  First the six waves run the
// block's three sub-lists (cross, cos-cos, sin-sin) on every group (waves
// 0-2 the direct points, waves 3-5 the mirrored ones, M-tile after M-tile)
// into part[2 mt + mirror][row][lane] (LDS atomics onto zero, as in the
// split kernels).  Then wave w takes N-tile w of the z contraction (its 16
// (row, col) columns; the sums and differences of mirror points on MFMA, as
// in kin_z_half_body, for every M-tile) and finishes its elements,
// R = R + mask (z + part), with their contributions to v.R, |R|^2 and |v|^2.
// One barrier separates the phases, so a workgroup makes few dependent
// memory round trips (staging, coefficients, R).
constexpr int kZpCols = 4;
constexpr int kZpThreads = 384;

__host__ __device__ inline int zpot_mtiles(int nz) { return (((nz + 1) >> 1) + 15) >> 4; }

template <int LB, int PAR, typename VT, typename T>
__device__ __forceinline__ void zpot_body(int tile, const double *__restrict__ X, double *__restrict__ R,
                                          const VT *__restrict__ H, const double *__restrict__ G, int nc, int plane0,
                                          double mask, const double *__restrict__ tz, int nxy, int nz, int64_t P,
                                          double &dot, double &rr2, double &vv2)
{
    using PX = gaunt::Pattern<LB, PAR, gaunt::kCross>;
    using PC = gaunt::Pattern<LB, PAR, gaunt::kCC>;
    using PS = gaunt::Pattern<LB, PAR, gaunt::kSS>;
    typedef typename KinAcc<T>::type acc_t;
    constexpr int B = gaunt::block_size(LB, PAR);
    constexpr int NTN = (B * kZpCols + 15) / 16;  // N-tiles of the z contraction
    static_assert(gaunt::split_covers_rows<LB, PAR>(), "every row needs a same-type and a cross partial");
    static_assert(NTN <= kZpThreads / 64, "one N-tile per wave");
    static_assert(kZpCols == 4, "lane (h - 16 mt) * 4 + col covers a mirror group of 64 points");
    extern __shared__ __align__(16) unsigned char zp_lds[];
    const int RS = nz * kZpCols;  // row stride of xs
    const int mz = (nz + 1) >> 1, hz = (nz & 1) ? nz >> 1 : -1;
    const int nmt = (mz + 15) >> 4, nks = (mz + 3) >> 2;
    double *xs = (double *)zp_lds;
    double(*part)[B][64] = reinterpret_cast<double(*)[B][64]>(xs + (size_t)B * RS);  // [2 nmt][B][64]
    T *rz = reinterpret_cast<T *>(xs + (size_t)B * RS + (size_t)nmt * 2 * B * 64);   // [129]: t[64 + d] = t(|d|)
    const int c0 = tile * kZpCols;
    const int tid = (int)threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    {
        // the tile's B x nz x 4 values (columns past nxy read as the last
        // column, never written back); all loads issued before the stores
        constexpr int K = (B * 64 * kZpCols + kZpThreads - 1) / kZpThreads;
        const int n = B * RS;
        double v[K];
#pragma unroll
        for (int e = 0; e < K; e++) {
            const int t = min(tid + e * kZpThreads, n - 1);
            const int row = t / RS, r = t - row * RS, cc = min(c0 + (r & 3), nxy - 1);
            v[e] = X[(size_t)(plane0 + row) * P + (size_t)(r >> 2) * nxy + cc];
        }
#pragma unroll
        for (int e = 0; e < K; e++)
            if (tid + e * kZpThreads < n) xs[tid + e * kZpThreads] = v[e];
    }
    for (int t = tid; t < 129; t += kZpThreads) {
        const int d = abs(t - 64);
        rz[t] = d < nz ? (T)tz[d] : (T)0;
    }
    for (int t = tid; t < nmt * 2 * B * 64; t += kZpThreads) (&part[0][0][0])[t] = 0.0;
    __syncthreads();
    {
        // potential: this wave's sub-list on its mirror group of every
        // M-tile (lanes past the last half index take the last one:
        // computed, never read)
        const int mirror = wave >= 3 ? 1 : 0, sub = wave - 3 * mirror;
        const int hl = lane >> 2, cl = lane & 3;
        const int ccl = min(c0 + cl, nxy - 1);
        for (int mt = 0; mt < nmt; mt++) {
            const int h = min(16 * mt + hl, mz - 1);
            const int k = mirror ? nz - 1 - h : h;
            const int64_t p = (int64_t)k * nxy + ccl;
            const VT *hp = H + (size_t)(p / kVTile) * nc * kVTile + p % kVTile;
            const double *xp = xs + k * kZpCols + cl;
            double(*pt)[64] = part[2 * mt + mirror];
            if (sub == 0)
                sub_run<PX, VT>(xp, hp, G, RS, pt, lane, std::make_integer_sequence<int, B>{},
                                std::make_integer_sequence<int, PX::NC>{});
            else if (sub == 1)
                sub_run<PC, VT>(xp, hp, G, RS, pt, lane, std::make_integer_sequence<int, B>{},
                                std::make_integer_sequence<int, PC::NC>{});
            else
                sub_run<PS, VT>(xp, hp, G, RS, pt, lane, std::make_integer_sequence<int, B>{},
                                std::make_integer_sequence<int, PS::NC>{});
        }
    }
    __syncthreads();
    if (wave >= NTN) return;
    const int l16 = lane & 15, kq = lane >> 4;
    const int zcol = 16 * wave + l16, zrow = min(zcol >> 2, B - 1), zc = zcol & 3;
    const bool zok = (zcol >> 2) < B && c0 + zc < nxy;
    const int zcc = min(c0 + zc, nxy - 1);
    const double *xr = xs + zrow * RS + zc;
    const size_t rbase = (size_t)(plane0 + zrow) * P + zcc;
    // R of this lane's elements (rows kk and nz-1-kk of its column, every
    // M-tile), loaded ahead of the contractions (clamped: never written back)
    double r1[2][4], r2[2][4];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int kk = min(16 * m + kin_row<T>(kq, q), mz - 1);
            r1[m][q] = R[rbase + (size_t)kk * nxy];
            r2[m][q] = R[rbase + (size_t)(nz - 1 - kk) * nxy];
        }
#pragma unroll
    for (int m = 0; m < 2; m++) {
        if (m >= nmt) break;
        const int kr = 16 * m + l16;  // A-operand row (output half index)
        acc_t as = acc_t{0, 0, 0, 0}, ad = acc_t{0, 0, 0, 0};
        for (int s = 0; s < nks; s++) {
            const int kp = 4 * s + kq;
            const bool in = kp < mz;
            const int kpc = in ? kp : 0;
            const double x1 = xr[kpc * kZpCols], x2 = xr[(nz - 1 - kpc) * kZpCols];
            const T bs = in ? (T)(x1 + x2) : (T)0, bd = in ? (T)(x1 - x2) : (T)0;
            const T t1 = rz[64 + kr - kp], t2 = rz[64 + (nz - 1 - kr - kp)];
            as = kmfma(kp == hz ? t1 : t1 + t2, bs, as);
            ad = kmfma(t1 - t2, bd, ad);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int hq = kin_row<T>(kq, q), kk = 16 * m + hq;
            if (!zok || kk >= mz) continue;
            const int lp = hq * kZpCols + zc;
            const double S = (double)as[q], D = (double)ad[q];
            const double n1 = r1[m][q] + mask * (0.5 * (S + D) + part[2 * m][zrow][lp]);
            const double x1 = xr[kk * kZpCols];
            R[rbase + (size_t)kk * nxy] = n1;
            dot += x1 * n1;
            rr2 += n1 * n1;
            vv2 += x1 * x1;
            if (kk != hz) {
                const int k2 = nz - 1 - kk;
                const double n2 = r2[m][q] + mask * (0.5 * (S - D) + part[2 * m + 1][zrow][lp]);
                const double x2 = xr[k2 * kZpCols];
                R[rbase + (size_t)k2 * nxy] = n2;
                dot += x2 * n2;
                rr2 += n2 * n2;
                vv2 += x2 * x2;
            }
        }
    }
}

// the workgroup's three partial sums at index idx (fixed order)
__device__ inline void zpot_partials(int idx, double dot, double rr2, double vv2, double *__restrict__ pa,
                                     double *__restrict__ pb, double *__restrict__ pc)
{
    __shared__ double red[3 * kZpThreads / 64];
    double s0, s1, s2;
    block_sum3(dot, rr2, vv2, red, s0, s1, s2);
    if (threadIdx.x == 0) {
        pa[idx] = s0;
        pb[idx] = s1;
        pc[idx] = s2;
    }
}

// dynamic LDS of a tile of a block of B planes on an axis of nz points
// (gfx950: up to 160 KiB per workgroup; B = 21 at nz = 54 takes 80328 B,
// two workgroups per CU)
inline size_t zpot_lds(int B, int nz, size_t tsize)
{
    return sizeof(double) * ((size_t)B * nz * kZpCols + (size_t)zpot_mtiles(nz) * 2 * B * 64) + tsize * 129;
}

// the lmax-5 pair of blocks of the full basis (l = 0, 2, 4 from plane 0;
// l = 1, 3, 5 from plane1): workgroup 2 t takes block 0 of tile t, 2 t + 1
// block 1, so that the two share the tile's coefficient lines in one XCD's L2
template <typename VT, typename T>
__global__ __launch_bounds__(kZpThreads) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_zpot2(
    const double *__restrict__ X, double *__restrict__ R, const VT *__restrict__ H, const double *__restrict__ G0,
    const double *__restrict__ G1, int nc, int plane1, double mask0, double mask1, const double *__restrict__ tz,
    int nxy, int nz, int64_t P, double *__restrict__ pa, double *__restrict__ pb, double *__restrict__ pc)
{
    const int lin = xcd_linear((int)blockIdx.x, (int)gridDim.x);
    double dot = 0.0, rr2 = 0.0, vv2 = 0.0;
    if (lin & 1)
        zpot_body<5, 1, VT, T>(lin >> 1, X, R, H, G1, nc, plane1, mask1, tz, nxy, nz, P, dot, rr2, vv2);
    else
        zpot_body<4, 0, VT, T>(lin >> 1, X, R, H, G0, nc, 0, mask0, tz, nxy, nz, P, dot, rr2, vv2);
    zpot_partials(lin, dot, rr2, vv2, pa, pb, pc);
}

// one lmax-5 block of the reduced basis (l = 0, 2, 4 or l = 1, 3, 5)
template <int LB, int PAR, typename VT, typename T>
__global__ __launch_bounds__(kZpThreads) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_zpot1(
    const double *__restrict__ X, double *__restrict__ R, const VT *__restrict__ H, const double *__restrict__ G0,
    int nc, double mask0, const double *__restrict__ tz, int nxy, int nz, int64_t P, double *__restrict__ pa,
    double *__restrict__ pb, double *__restrict__ pc)
{
    const int lin = xcd_linear((int)blockIdx.x, (int)gridDim.x);
    double dot = 0.0, rr2 = 0.0, vv2 = 0.0;
    zpot_body<LB, PAR, VT, T>(lin, X, R, H, G0, nc, 0, mask0, tz, nxy, nz, P, dot, rr2, vv2);
    zpot_partials(lin, dot, rr2, vv2, pa, pb, pc);
}

// ---------------------------------------------------------------------------
// recurrence scalars
// ---------------------------------------------------------------------------
// mode 0 (end of a step): finalize_step.
// mode 1 (start vector): |X|^2 in `pa`; v_1 = X / |X|, R' = 0.
__global__ __launch_bounds__(kWG) void k_finalize(const double *__restrict__ pa, const double *__restrict__ pb,
                                                  const double *__restrict__ pc, unsigned int n,
                                                  Scal *__restrict__ sc, int mode, double *alpha_out,
                                                  double *beta_out, double *beta2_out)
{
    __shared__ double red[3 * kWG / 64];
    if (mode == 0) {
        finalize_step(pa, pb, pc, n, sc, alpha_out, beta_out, beta2_out, red);
        return;
    }
    const double t = total_of_partials(pa, n, red);
    if (threadIdx.x == 0) {
        sc->ca = 1.0 / sqrt(t);
        sc->cb = 0.0;
        sc->cr = 0.0;
        sc->step = 0;
        sc->breakdown = 0;
    }
}

__global__ void k_set_scal(Scal *sc, double ca, double cr)
{
    sc->ca = ca;
    sc->cb = 0.0;
    sc->cr = cr;
    sc->alpha = 0.0;
    sc->beta = 0.0;
    sc->beta2 = 0.0;
    sc->step = 0;
    sc->breakdown = 0;
}

__global__ void k_fill(double *p, int64_t n, double v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// counter-based U[-1,1) start vector (SplitMix64 finaliser of seed and index)
__device__ inline double start_value(uint64_t seed, uint64_t idx)
{
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + (idx + 1) * 0xBF58476D1CE4E5B9ull + 0x94D049BB133111EBull;
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    return 2.0 * u - 1.0;
}

// value of device element i is that of reference element plane_ref[i / P] * P + i % P
__global__ __launch_bounds__(kWG) void k_start_random(double *X, int64_t n, int64_t P,
                                                      const int *__restrict__ plane_ref, uint64_t seed,
                                                      double *partial)
{
    __shared__ double red[3 * kWG / 64];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t plane = i / P;
        const double v = start_value(seed, (uint64_t)(plane_ref[plane] * P + (i - plane * P)));
        X[i] = v;
        s += v * v;
    }
    const double b = block_sum(s, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

// Rayleigh quotient pieces of y and hy = H y, per workgroup (grid-stride
// shares, summed in a fixed order): mode 0 writes sum y^2 and sum y.hy to
// partial[2 b], partial[2 b + 1]; mode 1 writes sum (hy - e y)^2 to
// partial[2 b] (the residual in a pass of its own: |hy|^2 - 2 e y.hy + e^2 |y|^2
// would cancel to nothing for a converged vector)
__global__ __launch_bounds__(kWG) void k_rayleigh_dots(const double *__restrict__ y, const double *__restrict__ hy,
                                                       int64_t n, double e, int mode, double *partial)
{
    __shared__ double red[3 * kWG / 64];
    double s0 = 0.0, s1 = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double a = y[i], b = hy[i];
        if (mode == 0) {
            s0 += a * a;
            s1 += a * b;
        } else {
            const double d = b - e * a;
            s0 += d * d;
        }
    }
    const double t0 = block_sum(s0, red);
    __syncthreads();
    const double t1 = block_sum(s1, red);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = t0;
        partial[2 * blockIdx.x + 1] = t1;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

template <typename VT>
static int launch_vblock_t(const DevOperator &op, const double *X, double *R, hipStream_t s)
{
    const VT *V = (const VT *)op.vpk;
    static_assert(kVTile == kWG, "one potential-kernel workgroup per V tile");
    const int64_t ptile = vpk_ntile(op.P) * kVTile;
    const VT *V0 = V + op.blk[0].pair_offset * ptile;
    const VT *V1 = op.nblk > 1 ? V + op.blk[1].pair_offset * ptile : V0;
    const double mk0 = op.blk[0].mask, mk1 = op.nblk > 1 ? op.blk[1].mask : 0.0;
    dim3 grid((unsigned)((op.P + kWG - 1) / kWG), (unsigned)op.nblk);
    const int b0 = op.blk[0].size, b1 = op.nblk > 1 ? op.blk[1].size : 0;
    if (vblock_has_kernel(b0, b1) && ((int64_t)grid.x * grid.y > op.npartial || (int64_t)grid.x * grid.y != op.nalpha))
        return fail(BSL_E_STATE, "partial buffer does not match the potential kernel grid");
#define BSL_VB(A, Bb)                                                                                  \
    if (b0 == A && b1 == Bb) {                                                                         \
        switch (op.vb_variant) {                                                                       \
        case 0:                                                                                        \
            timed_launch((k_vblock<A, Bb, VT, 1>), grid, dim3(kWG), 0, s, X, R, V0, V1, mk0, mk1,     \
                               op.P, op.partial, op.partial_b, op.partial_c);                          \
            break;                                                                                     \
        case 1:                                                                                        \
            timed_launch((k_vblock<A, Bb, VT, 2>), grid, dim3(kWG), 0, s, X, R, V0, V1, mk0, mk1,     \
                               op.P, op.partial, op.partial_b, op.partial_c);                          \
            break;                                                                                     \
        default:                                                                                       \
            timed_launch((k_vblock<A, Bb, VT, 3>), grid, dim3(kWG), 0, s, X, R, V0, V1, mk0, mk1,     \
                               op.P, op.partial, op.partial_b, op.partial_c);                          \
            break;                                                                                     \
        }                                                                                              \
        return hipGetLastError() == hipSuccess ? BSL_OK : fail(BSL_E_HIP, "k_vblock launch");           \
    }
    BSL_VB(1, 3) BSL_VB(6, 10) BSL_VB(6, 3) BSL_VB(15, 21) BSL_VB(15, 10) BSL_VB(28, 21)
    BSL_VB(1, 0) BSL_VB(3, 0) BSL_VB(6, 0) BSL_VB(10, 0) BSL_VB(15, 0) BSL_VB(21, 0) BSL_VB(28, 0)
#undef BSL_VB
    // any other sizes: the row-parallel kernel
    const int bmax = b0 > b1 ? b0 : b1;
    if ((size_t)bmax * 64 * sizeof(double) > 65536) return fail(BSL_E_ARG, "potential block too large");
    dim3 rgrid((unsigned)((op.P + 63) / 64), (unsigned)op.nblk);
    if ((int64_t)rgrid.x * rgrid.y != op.nalpha) return fail(BSL_E_STATE, "partial buffer does not match the potential kernel grid");
    timed_launch((k_vblock_rows<VT>), rgrid, dim3(kWG), (uint32_t)(bmax * 64 * sizeof(double)), s, X, R, V0, V1, b0,
                 b1, mk0, mk1, op.P, op.partial, op.partial_b, op.partial_c);
    return hipGetLastError() == hipSuccess ? BSL_OK : fail(BSL_E_HIP, "k_vblock_rows launch");
}

bool vblock_has_kernel(int b0, int b1)
{
    static const int sizes[][2] = {{1, 3},  {6, 10}, {6, 3},  {15, 21}, {15, 10}, {28, 21}, {1, 0},
                                   {3, 0},  {6, 0},  {10, 0}, {15, 0},  {21, 0},  {28, 0}};
    for (const auto &z : sizes)
        if (z[0] == b0 && z[1] == b1) return true;
    return false;
}

template <typename VT>
static int launch_vgaunt_t(const DevOperator &op, const double *X, double *R, hipStream_t s)
{
    const VT *H = (const VT *)op.vpk;
    const Block &b0 = op.blk[0], &b1 = op.nblk > 1 ? op.blk[1] : op.blk[0];
    const double mk1 = op.nblk > 1 ? b1.mask : 0.0;
    const int l0 = b0.lb, p0 = b0.par, l1 = op.nblk > 1 ? b1.lb : -1, p1 = op.nblk > 1 ? b1.par : 0;
    const int split = harmonic_split(op.nblk, l0, p0, l1, p1, b1.plane0, b0.size);
    dim3 grid((unsigned)harmonic_workgroups(op.P, split));
    if ((int64_t)grid.x > op.npartial || (int64_t)grid.x != op.nalpha)
        return fail(BSL_E_STATE, "partial buffer does not match the potential kernel grid");
    if (split == 2) {
        if (op.ncoef != gaunt::ncoef(l0)) return fail(BSL_E_STATE, "coefficient columns do not match the block");
#define BSL_VG1(A, PA)                                                                                      \
    if (l0 == A && p0 == PA) {                                                                              \
        if (op.vb_variant == 0)                                                                             \
            timed_launch((k_vgaunt_split1<A, PA, VT, 3>), grid, dim3(192), 0, s, X, R, H, op.gc[0],       \
                               op.ncoef, b0.mask, op.P, op.partial, op.partial_b, op.partial_c);            \
        else                                                                                                \
            timed_launch((k_vgaunt_split1<A, PA, VT, 4>), grid, dim3(192), 0, s, X, R, H, op.gc[0],       \
                               op.ncoef, b0.mask, op.P, op.partial, op.partial_b, op.partial_c);            \
        return hipGetLastError() == hipSuccess ? BSL_OK : fail(BSL_E_HIP, "k_vgaunt_split1 launch");       \
    }
        BSL_VG1(5, 1) BSL_VG1(4, 0)
#undef BSL_VG1
        return fail(BSL_E_STATE, "no single-block split kernel for this block");
    }
    if (split == 1) {
        if (op.ncoef != gaunt::ncoef(5)) return fail(BSL_E_STATE, "coefficient columns do not match the blocks");
        if (op.vb_variant == 0)
            timed_launch((k_vgaunt_split<VT, 3>), grid, dim3(256), 0, s, X, R, H, op.gc[0], op.gc[1],
                               op.ncoef, b1.plane0, b0.mask, mk1, op.P, op.partial, op.partial_b, op.partial_c);
        else
            timed_launch((k_vgaunt_split<VT, 4>), grid, dim3(256), 0, s, X, R, H, op.gc[0], op.gc[1],
                               op.ncoef, b1.plane0, b0.mask, mk1, op.P, op.partial, op.partial_b, op.partial_c);
        return hipGetLastError() == hipSuccess ? BSL_OK : fail(BSL_E_HIP, "k_vgaunt_split launch");
    }
    const dim3 threads((unsigned)(op.nblk * kWG));
#define BSL_VG(A, PA, Bb, PB)                                                                              \
    if (l0 == A && p0 == PA && l1 == Bb && (Bb < 0 || p1 == PB)) {                                         \
        if (op.vb_variant == 0)                                                                            \
            timed_launch((k_vgaunt<A, PA, Bb, PB, VT, 1>), grid, threads, 0, s, X, R, H, op.gc[0],       \
                               op.gc[1], op.ncoef, b1.plane0, b0.mask, mk1, op.P, op.partial, op.partial_b,    \
                               op.partial_c);                                                              \
        else                                                                                               \
            timed_launch((k_vgaunt<A, PA, Bb, PB, VT, 2>), grid, threads, 0, s, X, R, H, op.gc[0],       \
                               op.gc[1], op.ncoef, b1.plane0, b0.mask, mk1, op.P, op.partial, op.partial_b,    \
                               op.partial_c);                                                              \
        return hipGetLastError() == hipSuccess ? BSL_OK : fail(BSL_E_HIP, "k_vgaunt launch");             \
    }
    BSL_VG(4, 0, 5, 1) BSL_VG(4, 0, 3, 1) BSL_VG(2, 0, 3, 1) BSL_VG(2, 0, 1, 1) BSL_VG(0, 0, 1, 1)
    BSL_VG(0, 0, -1, 0) BSL_VG(1, 1, -1, 0) BSL_VG(2, 0, -1, 0) BSL_VG(3, 1, -1, 0) BSL_VG(4, 0, -1, 0)
    BSL_VG(5, 1, -1, 0)
#undef BSL_VG
    return fail(BSL_E_ARG, "no harmonic-form potential kernel for blocks of l <= " + std::to_string(l0) + "/" +
                               std::to_string(l1));
}

// ---------------------------------------------------------------------------
// fused z kinetic + potential launches
// ---------------------------------------------------------------------------
bool zpot_fits(const DevOperator &op)
{
    if (op.calc == BSL_CALC_QUAD || op.vform != BSL_VFORM_HARMONICS || !kinetic_mfma_fits(op)) return false;
    if (!((op.kprec == BSL_PREC_FP64 && op.vprec == BSL_PREC_FP64) ||
          (op.kprec == BSL_PREC_FP32 && op.vprec == BSL_PREC_FP32)))
        return false;
    const Block &b0 = op.blk[0], &b1 = op.nblk > 1 ? op.blk[1] : op.blk[0];
    const int split = harmonic_split(op.nblk, b0.lb, b0.par, op.nblk > 1 ? b1.lb : -1, b1.par, b1.plane0, b0.size);
    if (split == 1) return op.ncoef == gaunt::ncoef(5);
    if (split == 2) return op.ncoef == gaunt::ncoef(b0.lb);
    return false;
}

int64_t zpot_workgroups(const DevOperator &op)
{
    const int64_t nt = (op.nx * op.ny + kZpCols - 1) / kZpCols;
    return op.nblk > 1 ? 2 * nt : nt;
}

template <typename VT, typename T>
static int launch_zpot_t(const DevOperator &op, const double *X, double *R, hipStream_t s)
{
    const VT *H = (const VT *)op.vpk;
    const int nxy = op.nx * op.ny;
    const dim3 grid((unsigned)zpot_workgroups(op));
    if ((int64_t)grid.x > op.npartial || (int64_t)grid.x != op.nalpha)
        return fail(BSL_E_STATE, "partial buffer does not match the fused kernel grid");
    const Block &b0 = op.blk[0];
    if (op.nblk > 1) {
        const Block &b1 = op.blk[1];
        timed_launch((k_zpot2<VT, T>), grid, dim3(kZpThreads), (uint32_t)zpot_lds(b1.size, op.nz, sizeof(T)), s, X,
                     R, H, op.gc[0], op.gc[1], op.ncoef, b1.plane0, b0.mask, b1.mask, op.tz, nxy, op.nz, op.P,
                     op.partial, op.partial_b, op.partial_c);
    } else if (b0.par == 1) {
        timed_launch((k_zpot1<5, 1, VT, T>), grid, dim3(kZpThreads), (uint32_t)zpot_lds(b0.size, op.nz, sizeof(T)),
                     s, X, R, H, op.gc[0], op.ncoef, b0.mask, op.tz, nxy, op.nz, op.P, op.partial, op.partial_b,
                     op.partial_c);
    } else {
        timed_launch((k_zpot1<4, 0, VT, T>), grid, dim3(kZpThreads), (uint32_t)zpot_lds(b0.size, op.nz, sizeof(T)),
                     s, X, R, H, op.gc[0], op.ncoef, b0.mask, op.tz, nxy, op.nz, op.P, op.partial, op.partial_b,
                     op.partial_c);
    }
    return hipGetLastError() == hipSuccess ? BSL_OK : fail(BSL_E_HIP, "k_zpot launch");
}

int launch_zpot(const DevOperator &op, const double *X, double *R, hipStream_t s)
{
    if (!zpot_fits(op)) return fail(BSL_E_STATE, "no fused z-kinetic and potential kernel for this operator");
    if (!op.vpk || !op.gc[0] || (op.nblk > 1 && !op.gc[1]))
        return fail(BSL_E_STATE, "the context has no potential to apply");
    return op.vprec == BSL_PREC_FP64 ? launch_zpot_t<double, double>(op, X, R, s)
                                     : launch_zpot_t<float, float>(op, X, R, s);
}

// ---------------------------------------------------------------------------
// batch kernels: the bodies above over a table of members, one launch per
// variant.  Workgroup vb belongs to the member whose [first, first + count)
// holds it (at most 64 members per table).  The member's workgroups run exactly the
// single-context kernel's code with the same indices, so the results are
// bitwise those of the member run alone.
// ---------------------------------------------------------------------------
// The table is followed by the launch's first[] list (kBatchMaxTable ints,
// ascending, unused slots INT_MAX): uniform scalar loads, then a count.
__device__ inline int batch_find(const BatchEntry *__restrict__ tab, int nm, int vb)
{
    (void)nm;
    const int *__restrict__ first = reinterpret_cast<const int *>(tab + kBatchMaxTable);
    int m = 0;
#pragma unroll
    for (int i = 1; i < kBatchMaxTable; i++) m += first[i] <= vb ? 1 : 0;
    return m;
}

// the XCD-aware bijection of xcd_tile over a one-dimensional grid
__device__ inline int batch_xcd(int id, int total)
{
    return (total % kXcds == 0) ? (id % kXcds) * (total / kXcds) + id / kXcds : id;
}

template <typename T, int KM>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(4, 8))) void kb_xy_half(
    const BatchEntry *__restrict__ tab, int nm)
{
    const int vb = (int)blockIdx.x, m = batch_find(tab, nm, vb);
    const BatchEntry &e = tab[m];
    const int loc = vb - e.first;
    kin_xy_half_body<T, KM>(loc % e.nz, loc / e.nz, e.X, e.R, e.tx, e.ty, e.rot, e.mask, e.sc, e.nx, e.ny, e.nz);
}

template <typename T, int KM>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(4, 8))) void kb_z_half(
    const BatchEntry *__restrict__ tab, int nm)
{
    const int vb = batch_xcd((int)blockIdx.x, (int)gridDim.x), m = batch_find(tab, nm, vb);
    const BatchEntry &e = tab[m];
    const int loc = vb - e.first, nxy = e.nx * e.ny, nt = (nxy + 63) / 64;
    kin_z_half_body<T, KM>(loc % nt, loc / nt, e.X, e.R, e.tz, e.mask, nxy, e.nz);
}

// The Gaunt values of a block depend on its shape only, so every member of
// the launch shares them; they come in as a kernel argument (read-only, so
// that the unrolled term list takes them through scalar loads).
// Occupancy: the ortho block with FP64 coefficients at 3 waves/SIMD (at 4
// the table lookup's registers spill), with FP32 ones (X read directly) at
// 4; the para block at 4, or 5 with FP32 coefficients.
template <int LB, int PAR, typename VT>
constexpr int kb_split1_waves()
{
    return LB == 5 ? (sizeof(VT) == 8 ? 3 : 4) : (sizeof(VT) == 8 ? 4 : 5);
}
template <int LB, int PAR, typename VT>
__global__ __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(kb_split1_waves<LB, PAR, VT>(), 8))) void kb_split1(
    const BatchEntry *__restrict__ tab, int nm, const double *__restrict__ G)
{
    const int vb = (int)blockIdx.x, m = batch_find(tab, nm, vb);
    const BatchEntry &e = tab[m];
    vgaunt_split1_body<LB, PAR, VT>(vb - e.first, e.X, e.R, (const VT *)e.H, G, e.nc, e.mask0, e.P, e.pa, e.pb,
                                    e.pc);
}

// fused z kinetic + potential of one reduced-basis lmax-5 block (kinetic
// and potential precision VT); workgroup = the member's tile, as in k_zpot1
template <int LB, int PAR, typename VT>
__global__ __launch_bounds__(kZpThreads) __attribute__((amdgpu_waves_per_eu(3, 8))) void kb_zpot(
    const BatchEntry *__restrict__ tab, int nm, const double *__restrict__ G)
{
    const int vb = xcd_linear((int)blockIdx.x, (int)gridDim.x), m = batch_find(tab, nm, vb);
    const BatchEntry &e = tab[m];
    const int tile = vb - e.first;
    double dot = 0.0, rr2 = 0.0, vv2 = 0.0;
    zpot_body<LB, PAR, VT, VT>(tile, e.X, e.R, (const VT *)e.H, G, e.nc, 0, e.mask0, e.tz, e.nx * e.ny, e.nz, e.P,
                               dot, rr2, vv2);
    zpot_partials(tile, dot, rr2, vv2, e.pa, e.pb, e.pc);
}

__global__ __launch_bounds__(kWG) void kb_finalize(const BatchEntry *__restrict__ tab, int nm)
{
    __shared__ double red[3 * kWG / 64];
    if ((int)blockIdx.x >= nm) return;
    const BatchEntry &e = tab[blockIdx.x];
    finalize_step(e.pa, e.pb, e.pc, (unsigned)e.nalpha, e.sc, e.alpha, e.beta, e.beta2, red);
}

template <typename T, int KM>
static int launch_batch_t(int kind, const BatchEntry *tab, int nm, int total, const double *G, int lds_nz,
                          hipStream_t s)
{
    const dim3 grid((unsigned)total);
    switch (kind) {
    case kBXyHalf:
        hipLaunchKernelGGL((kb_xy_half<T, KM>), grid, dim3(kWG), sizeof(T) * ((size_t)4 * 32 * kCompLd + 2 * 129), s,
                           tab, nm);
        break;
    case kBZHalf:
        hipLaunchKernelGGL((kb_z_half<T, KM>), grid, dim3(kWG), sizeof(T) * ((size_t)2 * 32 * kSlabLd + 129), s, tab,
                           nm);
        break;
    case kBPotSplitPara:
        hipLaunchKernelGGL((kb_split1<4, 0, T>), grid, dim3(192), 0, s, tab, nm, G);
        break;
    case kBPotSplitOrtho:
        hipLaunchKernelGGL((kb_split1<5, 1, T>), grid, dim3(192), 0, s, tab, nm, G);
        break;
    case kBZPotPara:  // LDS for the longest z axis among the members
        hipLaunchKernelGGL((kb_zpot<4, 0, T>), grid, dim3(kZpThreads),
                           zpot_lds(gaunt::block_size(4, 0), lds_nz, sizeof(T)), s, tab, nm, G);
        break;
    case kBZPotOrtho:
        hipLaunchKernelGGL((kb_zpot<5, 1, T>), grid, dim3(kZpThreads),
                           zpot_lds(gaunt::block_size(5, 1), lds_nz, sizeof(T)), s, tab, nm, G);
        break;
    default:
        return fail(BSL_E_ARG, "unknown batch kernel");
    }
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

int launch_batch(int kind, int prec, const BatchEntry *d_tab, int nm, int total, const double *G, int lds_nz,
                 hipStream_t s)
{
    if ((kind == kBZPotPara || kind == kBZPotOrtho) && (lds_nz < 1 || lds_nz > 64))
        return fail(BSL_E_ARG, "fused batch kernel: z axis out of range");
    if (nm <= 0 || nm > kBatchMaxTable || total <= 0) return fail(BSL_E_ARG, "batch table size out of range");
    if (kind == kBFinalize) {
        hipLaunchKernelGGL(kb_finalize, dim3((unsigned)nm), dim3(kWG), 0, s, d_tab, nm);
        BSL_HIP(hipGetLastError());
        return BSL_OK;
    }
    if (prec == BSL_PREC_BF16X3) {
        if (kind != kBXyHalf && kind != kBZHalf) return fail(BSL_E_ARG, "BF16x3 batch tables are kinetic only");
        return launch_batch_t<float, 1>(kind, d_tab, nm, total, G, lds_nz, s);
    }
    return prec == BSL_PREC_FP64 ? launch_batch_t<double, 0>(kind, d_tab, nm, total, G, lds_nz, s)
                                 : launch_batch_t<float, 0>(kind, d_tab, nm, total, G, lds_nz, s);
}

int launch_potential(const DevOperator &op, const double *X, double *R, Scal *sc, hipStream_t s)
{
    if (op.calc == BSL_CALC_QUAD) return launch_quad_potential(op, X, R, sc, s);
    // a context whose potential was never built (or not rebuilt) launches nothing
    if (!op.vpk || (op.vform == BSL_VFORM_HARMONICS && (!op.gc[0] || (op.nblk > 1 && !op.gc[1]))))
        return fail(BSL_E_STATE, "the context has no potential to apply");
    if (op.vform == BSL_VFORM_HARMONICS)
        return op.vprec == BSL_PREC_FP64 ? launch_vgaunt_t<double>(op, X, R, s) : launch_vgaunt_t<float>(op, X, R, s);
    return op.vprec == BSL_PREC_FP64 ? launch_vblock_t<double>(op, X, R, s) : launch_vblock_t<float>(op, X, R, s);
}

int launch_kinetic(const DevOperator &op, double *X, double *R, Scal *sc, hipStream_t s)
{
    if (!kinetic_mfma_fits(op)) return launch_kinetic_lines(op, X, R, sc, s);
    return op.kprec == BSL_PREC_FP64   ? launch_kinetic_t<double, 0>(op, X, R, sc, s)
           : op.kprec == BSL_PREC_FP32 ? launch_kinetic_t<float, 0>(op, X, R, sc, s)
                                       : launch_kinetic_t<float, 1>(op, X, R, sc, s);
}

LaunchTimer &launch_timer()
{
    static thread_local LaunchTimer t;
    return t;
}

int launch_hv(const DevOperator &op, double *X, double *R, Scal *sc, hipStream_t s, hipEvent_t *ev)
{
    LaunchTimer &t = launch_timer();
    if (ev) t = {ev[0], ev[1]};
    int krc;
    if (op.fused)
        krc = op.kprec == BSL_PREC_FP64 ? launch_kinetic_t<double, 0>(op, X, R, sc, s, false)
                                        : launch_kinetic_t<float, 0>(op, X, R, sc, s, false);
    else
        krc = launch_kinetic(op, X, R, sc, s);
    if (ev) t = {ev[2], ev[3]};
    int rc = krc ? krc : launch_potential_part(op, X, R, sc, s);
    t = {};
    return rc;
}

int launch_potential_part(const DevOperator &op, const double *X, double *R, Scal *sc, hipStream_t s)
{
    return op.fused ? launch_zpot(op, X, R, s) : launch_potential(op, X, R, sc, s);
}

int launch_finalize(const DevOperator &op, Scal *sc, double *alpha_out, double *beta_out, double *beta2_out,
                    hipStream_t s)
{
    timed_launch(k_finalize, dim3(1), dim3(kWG), 0, s, op.partial, op.partial_b, op.partial_c,
                 (unsigned)op.nalpha, sc, 0, alpha_out, beta_out, beta2_out);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

int launch_step(const DevOperator &op, double *X, double *R, Scal *sc, double *alpha_out, double *beta_out,
                double *beta2_out, hipStream_t s, hipEvent_t *ev)
{
    int rc = launch_hv(op, X, R, sc, s, ev);
    if (rc) return rc;
    if (ev) launch_timer() = {ev[4], ev[5]};
    rc = launch_finalize(op, sc, alpha_out, beta_out, beta2_out, s);
    launch_timer() = {};
    return rc;
}

int check_breakdown(const Scal *sc, hipStream_t s)
{
    Scal h;
    BSL_HIP(hipMemcpyAsync(&h, sc, sizeof(Scal), hipMemcpyDeviceToHost, s));
    BSL_HIP(hipStreamSynchronize(s));
    if (h.breakdown != 0)
        return fail(BSL_E_STATE, "Lanczos breakdown at step " + std::to_string(h.breakdown) +
                                     ": beta^2 = |r|^2 - alpha^2 kept no correct digit (the residual vanished: "
                                     "the Krylov space of the start vector is invariant); restart from another "
                                     "start vector");
    return BSL_OK;
}

int launch_init_scal(Scal *sc, double ca, double cr, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_scal, dim3(1), dim3(1), 0, s, sc, ca, cr);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

int launch_zero(double *p, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kWG), 0, s, p, n, 0.0);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

int launch_start_vector(double *X, int64_t n, int64_t P, const int *plane_ref, int kind, uint64_t seed, Scal *sc,
                        double *partial, int npartial, hipStream_t s)
{
    if (kind == BSL_START_UNIFORM) {
        // X = 1, ca = 1/sqrt(N): ca*X is the reference's 1./sqrt((double)N) exactly
        hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kWG), 0, s, X, n, 1.0);
        BSL_HIP(hipGetLastError());
        return launch_init_scal(sc, 1. / std::sqrt((double)n), 0.0, s);
    }
    int g = grid_for(n, 2048);
    if (g > npartial) g = npartial;
    hipLaunchKernelGGL(k_start_random, dim3(g), dim3(kWG), 0, s, X, n, P, plane_ref, seed, partial);
    BSL_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kWG), 0, s, partial, partial, partial, (unsigned)g, sc, 1, nullptr,
                       nullptr, nullptr);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

int launch_rayleigh_dots(const double *y, const double *hy, int64_t n, double e, int mode, double *partial,
                         int npartial, int *groups, hipStream_t s)
{
    int g = grid_for(n, 1024);
    if (2 * g > npartial) g = npartial / 2;
    if (g < 1) return fail(BSL_E_STATE, "no room for the Rayleigh partial sums");
    hipLaunchKernelGGL(k_rayleigh_dots, dim3(g), dim3(kWG), 0, s, y, hy, n, e, mode, partial);
    BSL_HIP(hipGetLastError());
    *groups = g;
    return BSL_OK;
}

// ritz[g][i] += coef[k][g] w[k][i] for k = 0 .. nw-1 in order (w: the
// Lanczos vectors of nw consecutive steps), skipping zero coefficients, with
// each Ritz vector read and written once per window
__global__ void k_ritz_flush(double *ritz, const double *__restrict__ w, const double *__restrict__ coef, int nw,
                             int64_t n, int ngood)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double wv[kRitzWindow];  // the staged entries of this element, loaded once
#pragma unroll
        for (int k = 0; k < kRitzWindow; k++) wv[k] = k < nw ? w[(size_t)k * n + i] : 0.0;
        for (int g = 0; g < ngood; g++) {
            bool any = false;
            for (int k = 0; k < nw; k++) any = any || coef[(size_t)k * ngood + g] != 0.0;
            if (!any) continue;
            double r = ritz[(size_t)g * n + i];
#pragma unroll
            for (int k = 0; k < kRitzWindow; k++)
                if (k < nw) {
                    const double c = coef[(size_t)k * ngood + g];
                    if (c != 0.0) r += c * wv[k];
                }
            ritz[(size_t)g * n + i] = r;
        }
    }
}

int launch_ritz_flush(double *ritz, const double *w, const double *coef, int nw, int64_t n, int ngood, hipStream_t s)
{
    hipLaunchKernelGGL(k_ritz_flush, dim3(grid_for(n)), dim3(kWG), 0, s, ritz, w, coef, nw, n, ngood);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

// ---------------------------------------------------------------------------
// preparation: potential grid
// ---------------------------------------------------------------------------
// Coulomb sum over the sites in site order, K q q_s / |s - pos|, and the LJ
// term A/r^12 - B/r^6 on the flagged sites; FP contraction off so that the
// grid matches the FP64 CPU evaluation of the same expression.
__device__ inline double coulomb_sum(const double *site, int ns, double px, double py, double pz, double q)
{
#pragma clang fp contract(off)
    double e = 0;
    for (int s = 0; s < ns; s++) {
        const double qs = site[4 * s + 3];
        if (qs == 0.0) continue;
        const double d0 = site[4 * s + 0] - px, d1 = site[4 * s + 1] - py, d2 = site[4 * s + 2] - pz;
        double r2 = 0.;
        r2 += d0 * d0;
        r2 += d1 * d1;
        r2 += d2 * d2;
        e += 138.9354846111096 * q * qs / sqrt(r2);
    }
    return e;
}

__global__ __launch_bounds__(kWG) void k_cm_potential(double *cm, int64_t P, const double *__restrict__ gx,
                                                      const double *__restrict__ gy, const double *__restrict__ gz,
                                                      int nx, int ny, const double *__restrict__ site_g,
                                                      const int *__restrict__ lj_g, int ns, double q_cm,
                                                      double lj_a, double lj_b)
{
#pragma clang fp contract(off)
    extern __shared__ double site[];
    int *lj = (int *)(site + 4 * ns);
    for (int t = threadIdx.x; t < 4 * ns; t += blockDim.x) site[t] = site_g[t];
    for (int t = threadIdx.x; t < ns; t += blockDim.x) lj[t] = lj_g[t];
    __syncthreads();
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((int64_t)nx * ny));
    const double x = gx[i], y = gy[j], z = gz[k];
    double e = coulomb_sum(site, ns, x, y, z, q_cm);
    double l = 0;
    for (int s = 0; s < ns; s++) {
        if (!lj[s]) continue;
        const double d0 = site[4 * s + 0] - x, d1 = site[4 * s + 1] - y, d2 = site[4 * s + 2] - z;
        double r2 = 0.;
        r2 += d0 * d0;
        r2 += d1 * d1;
        r2 += d2 * d2;
        const double r = sqrt(r2);
        l += (lj_a / pow(r, 12.0)) - (lj_b / pow(r, 6.0));
    }
    cm[p] = e + l;
}

// 1/sqrt(x): the hardware estimate and two Newton steps y <- y (3/2 - x y^2 / 2),
// each of which squares the relative error (a few ulp of the rounded value)
__device__ inline double rsqrt_nr(double x)
{
    const double h = 0.5 * x;
    double y = __builtin_amdgcn_rsq(x);
    y = fma(y, fma(-h * y, y, 0.5), y);
    y = fma(y, fma(-h * y, y, 0.5), y);
    return y;
}

// The charged sites of the cage with their H-site weights K q_H q_s, in site
// order: hs[n][4] = x, y, z, K q_H q_s, n < *nh (uncharged sites, such as
// TIP4P's O, give no term, as in the reference's sum).  One thread: a few
// hundred sites at most.
__global__ void k_cage_hsites(const double *__restrict__ site_g, int ns, double q_h, double *__restrict__ hs,
                              int *__restrict__ nh)
{
    int n = 0;
    for (int t = 0; t < ns; t++) {
        const double w = 138.9354846111096 * q_h * site_g[4 * t + 3];
        if (w == 0.0) continue;
        hs[4 * n + 0] = site_g[4 * t + 0];
        hs[4 * n + 1] = site_g[4 * t + 1];
        hs[4 * n + 2] = site_g[4 * t + 2];
        hs[4 * n + 3] = w;
        n++;
    }
    *nh = n;
}

// vg[p][q], q = (half, a, b): H2 centre at grid point p, H atoms at +-off[q].
// The H-site terms K q_H q_s / |h - s| are summed in site order over the
// charged sites hs (wave-uniform reads, no branch) with 1/|h - s| from
// rsqrt_nr (a few ulp from the quotient of a correctly rounded square root,
// per term).  MASK: instead
// of the grid, one bit per (point, node), [vg < ceiling] (the clamp mask of
// the gradients, bit idx & 63 of word idx >> 6): each wave takes 64
// consecutive idx (the loop runs in whole waves) and stores the ballot.
template <bool MASK>
__global__ __launch_bounds__(kWG) void k_vgrid_cage(double *vg, uint64_t *mask, int64_t P, int Q,
                                                    const double *__restrict__ gx, const double *__restrict__ gy,
                                                    const double *__restrict__ gz, int nx, int ny,
                                                    const double *__restrict__ off, const double *__restrict__ hs,
                                                    const int *__restrict__ nh_g, const double *__restrict__ cm,
                                                    double ceiling)
{
    const int nh = __builtin_amdgcn_readfirstlane(*nh_g);
    const int64_t total = P * (int64_t)Q;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // whole waves: blockDim and the stride are multiples of 64
    const int64_t end = MASK ? (total + 63) / 64 * 64 : total;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < end; idx += stride) {
        const bool in = idx < total;
        const int64_t ic = in ? idx : total - 1;
        const int64_t p = ic / Q;
        const int q = (int)(ic - p * Q);
        const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((int64_t)nx * ny));
        const double x = gx[i], y = gy[j], z = gz[k];
        const double ox = off[3 * q], oy = off[3 * q + 1], oz = off[3 * q + 2];
        const double ax = x + ox, ay = y + oy, az = z + oz, bx = x - ox, by = y - oy, bz = z - oz;
        double va = 0.0, vb = 0.0;
        for (int t = 0; t < nh; t++) {
            const double sx = hs[4 * t], sy = hs[4 * t + 1], sz = hs[4 * t + 2], w = hs[4 * t + 3];
            const double a0 = sx - ax, a1 = sy - ay, a2 = sz - az;
            const double b0 = sx - bx, b1 = sy - by, b2 = sz - bz;
            va = fma(w, rsqrt_nr(fma(a0, a0, fma(a1, a1, a2 * a2))), va);
            vb = fma(w, rsqrt_nr(fma(b0, b0, fma(b1, b1, b2 * b2))), vb);
        }
        double v = va + vb + cm[p];
        if (v >= ceiling) v = ceiling;
        if constexpr (MASK) {
            const uint64_t bits = __ballot(in && v < ceiling);
            if ((threadIdx.x & 63) == 0) mask[idx >> 6] = bits;
        } else {
            vg[idx] = v;
        }
    }
}

__global__ __launch_bounds__(kWG) void k_vgrid_iho(double *vg, int64_t P, int Q, const double *__restrict__ gx,
                                                   const double *__restrict__ gy, const double *__restrict__ gz,
                                                   int nx, int ny, double factor, double iho_ceil, double ceiling)
{
#pragma clang fp contract(off)
    const int64_t total = P * (int64_t)Q;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = idx / Q;
        const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((int64_t)nx * ny));
        const double x = gx[i], y = gy[j], z = gz[k];
        double d2 = 0.;
        d2 += x * x;
        d2 += y * y;
        d2 += z * z;
        double v = factor * d2;
        if (fabs(v) > fabs(iho_ceil)) v = iho_ceil;
        if (v >= ceiling) v = ceiling;
        vg[idx] = v;
    }
}

int launch_vgrid_cage(double *vg, uint64_t *mask, int64_t P, int Q, const double *gx, const double *gy,
                      const double *gz, int nx, int ny, const double *off, const CageDev &cage, double q_h,
                      double q_cm, double lj_a, double lj_b, double ceiling, hipStream_t s)
{
    double *cm = nullptr;
    BSL_HIP(hipMallocAsync((void **)&cm, sizeof(double) * (size_t)P, s));
    const size_t lds_cm = sizeof(double) * 4 * cage.nsites + sizeof(int) * cage.nsites;
    hipLaunchKernelGGL(k_cm_potential, dim3((unsigned)((P + kWG - 1) / kWG)), dim3(kWG), lds_cm, s, cm, P, gx, gy,
                       gz, nx, ny, cage.site, cage.lj, cage.nsites, q_cm, lj_a, lj_b);
    BSL_HIP(hipGetLastError());
    // charged sites [nsites][4] and their count after them
    double *hs = nullptr;
    BSL_HIP(hipMallocAsync((void **)&hs, sizeof(double) * (4 * (size_t)cage.nsites + 1), s));
    int *nh = (int *)(hs + 4 * (size_t)cage.nsites);
    hipLaunchKernelGGL(k_cage_hsites, dim3(1), dim3(1), 0, s, cage.site, cage.nsites, q_h, hs, nh);
    BSL_HIP(hipGetLastError());
    if (mask)
        hipLaunchKernelGGL(k_vgrid_cage<true>, dim3(grid_for(P * Q, 65536)), dim3(kWG), 0, s, nullptr, mask, P, Q, gx,
                           gy, gz, nx, ny, off, hs, nh, cm, ceiling);
    else
        hipLaunchKernelGGL(k_vgrid_cage<false>, dim3(grid_for(P * Q, 65536)), dim3(kWG), 0, s, vg, nullptr, P, Q, gx,
                           gy, gz, nx, ny, off, hs, nh, cm, ceiling);
    BSL_HIP(hipGetLastError());
    BSL_HIP(hipFreeAsync(hs, s));
    BSL_HIP(hipFreeAsync(cm, s));
    return BSL_OK;
}

int launch_vgrid_iho(double *vg, int64_t P, int Q, const double *gx, const double *gy, const double *gz, int nx,
                     int ny, double factor, double iho_ceil, double ceiling, hipStream_t s)
{
    hipLaunchKernelGGL(k_vgrid_iho, dim3(grid_for(P * Q, 65536)), dim3(kWG), 0, s, vg, P, Q, gx, gy, gz, nx, ny,
                       factor, iho_ceil, ceiling);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

// Potential universe (Alavi_H2_Routines.cpp:189-317, gridFcns.cpp:6-58):
// point u = i0 + n0 (i1 + n1 i2) at c_d = i_d dx_d - max_d / 2, dx_d =
// max_d / (n_d - 1); cm[u] = centre Coulomb + LJ, hp[u] = H-atom Coulomb.
__global__ __launch_bounds__(kWG) void k_universe_cage(double *cm, double *hp, Universe u,
                                                       const double *__restrict__ site_g,
                                                       const int *__restrict__ lj_g, int ns, double q_cm, double q_h,
                                                       double lj_a, double lj_b)
{
#pragma clang fp contract(off)
    extern __shared__ double site[];
    int *lj = (int *)(site + 4 * ns);
    for (int t = threadIdx.x; t < 4 * ns; t += blockDim.x) site[t] = site_g[t];
    for (int t = threadIdx.x; t < ns; t += blockDim.x) lj[t] = lj_g[t];
    __syncthreads();
    const int64_t total = (int64_t)u.n[0] * u.n[1] * u.n[2];
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int i0 = (int)(idx % u.n[0]), i1 = (int)((idx / u.n[0]) % u.n[1]),
                  i2 = (int)(idx / ((int64_t)u.n[0] * u.n[1]));
        const double x = ((double)i0 * u.d[0]) - (u.mx[0] / 2.0), y = ((double)i1 * u.d[1]) - (u.mx[1] / 2.0),
                     z = ((double)i2 * u.d[2]) - (u.mx[2] / 2.0);
        double e = coulomb_sum(site, ns, x, y, z, q_cm);
        double l = 0;
        for (int s = 0; s < ns; s++) {
            if (!lj[s]) continue;
            const double d0 = site[4 * s + 0] - x, d1 = site[4 * s + 1] - y, d2 = site[4 * s + 2] - z;
            double r2 = 0.;
            r2 += d0 * d0;
            r2 += d1 * d1;
            r2 += d2 * d2;
            const double r = sqrt(r2);
            l += (lj_a / pow(r, 12.0)) - (lj_b / pow(r, 6.0));
        }
        cm[idx] = e + l;
        hp[idx] = coulomb_sum(site, ns, x, y, z, q_h);
    }
}

// nearest universe index along axis d of coordinate c (AlaviMapH2GridToHGrid
// Alavi_H2_Routines.cpp:5-54: round half away from zero, kept inside)
__device__ inline int universe_index(const Universe &u, int d, double c)
{
#pragma clang fp contract(off)
    const double cmin = (-0.5) * u.mx[d];
    int i = (int)round((c - cmin) / u.d[d]);
    if (i <= 0) i = 0;
    else if (i >= u.n[d]) i = u.n[d] - 1;
    return i;
}

// vg[p][q] = cm[nearest(c)] + hp[nearest(c + dist e_q)] + hp[nearest(c - dist e_q)],
// V >= ceiling -> ceiling (Alavi_H2_Eng_Point Alavi_H2_Routines.cpp:133-187;
// calc_Vlmlpmp_NoQuad linRotCartSphHvFullVPrepRoutines.cpp:134-184)
__global__ __launch_bounds__(kWG) void k_vgrid_mapped(double *vg, int64_t P, int Q, const double *__restrict__ gx,
                                                      const double *__restrict__ gy, const double *__restrict__ gz,
                                                      int nx, int ny, const double *__restrict__ dir,
                                                      const double *__restrict__ cm, const double *__restrict__ hp,
                                                      Universe u, double dist, double ceiling)
{
#pragma clang fp contract(off)
    const int64_t total = P * (int64_t)Q;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = idx / Q;
        const int q = (int)(idx - p * Q);
        const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((int64_t)nx * ny));
        const double c[3] = {gx[i], gy[j], gz[k]};
        int ic[3], i1[3], i2[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            ic[d] = universe_index(u, d, c[d]);
            i1[d] = universe_index(u, d, c[d] + dir[3 * q + d] * dist);
            i2[d] = universe_index(u, d, c[d] + dir[3 * q + d] * (-dist));
        }
        const int64_t n01 = (int64_t)u.n[0] * u.n[1];
        double v = 0.0;
        v += cm[ic[0] + (int64_t)u.n[0] * ic[1] + n01 * ic[2]];
        v += hp[i1[0] + (int64_t)u.n[0] * i1[1] + n01 * i1[2]];
        v += hp[i2[0] + (int64_t)u.n[0] * i2[1] + n01 * i2[2]];
        if (v >= ceiling) v = ceiling;
        vg[idx] = v;
    }
}

// V = K q_centre q_mol / |c| (a charge at the origin), |V| > |v_ceil| -> v_ceil,
// then V >= ceiling -> ceiling (CoulombPotential Coulomb_Routines.cpp:6-31)
__global__ __launch_bounds__(kWG) void k_vgrid_coulomb(double *vg, int64_t P, int Q, const double *__restrict__ gx,
                                                       const double *__restrict__ gy, const double *__restrict__ gz,
                                                       int nx, int ny, double q1, double q2, double v_ceil,
                                                       double ceiling)
{
#pragma clang fp contract(off)
    const int64_t total = P * (int64_t)Q;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = idx / Q;
        const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((int64_t)nx * ny));
        const double x = gx[i], y = gy[j], z = gz[k];
        double r2 = 0.;
        r2 += x * x;
        r2 += y * y;
        r2 += z * z;
        double v = 138.9354846111096 * q1 * q2 / sqrt(r2);
        if (fabs(v) > fabs(v_ceil)) v = v_ceil;
        if (v >= ceiling) v = ceiling;
        vg[idx] = v;
    }
}

int launch_universe_cage(double *cm, double *hp, const Universe &u, const CageDev &cage, double q_cm, double q_h,
                         double lj_a, double lj_b, hipStream_t s)
{
    const int64_t total = (int64_t)u.n[0] * u.n[1] * u.n[2];
    const size_t lds = sizeof(double) * 4 * cage.nsites + sizeof(int) * cage.nsites;
    hipLaunchKernelGGL(k_universe_cage, dim3(grid_for(total, 65536)), dim3(kWG), lds, s, cm, hp, u, cage.site,
                       cage.lj, cage.nsites, q_cm, q_h, lj_a, lj_b);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

int launch_vgrid_mapped(double *vg, int64_t P, int Q, const double *gx, const double *gy, const double *gz, int nx,
                        int ny, const double *dir, const double *cm, const double *hp, const Universe &u, double dist,
                        double ceiling, hipStream_t s)
{
    hipLaunchKernelGGL(k_vgrid_mapped, dim3(grid_for(P * Q, 65536)), dim3(kWG), 0, s, vg, P, Q, gx, gy, gz, nx, ny,
                       dir, cm, hp, u, dist, ceiling);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

int launch_vgrid_coulomb(double *vg, int64_t P, int Q, const double *gx, const double *gy, const double *gz, int nx,
                         int ny, double q1, double q2, double v_ceil, double ceiling, hipStream_t s)
{
    hipLaunchKernelGGL(k_vgrid_coulomb, dim3(grid_for(P * Q, 65536)), dim3(kWG), 0, s, vg, P, Q, gx, gy, gz, nx, ny,
                       q1, q2, v_ceil, ceiling);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

// ---------------------------------------------------------------------------
// preparation: V blocks, Vpk[pair][P] = sum_q Vg[P][q] H[q][pair]  (FP64 GEMM)
// ---------------------------------------------------------------------------
// FP64 MFMA (16x16x4) GEMM.  A workgroup owns 64 points and up to
// kGemmCols columns (the 66 harmonic columns of lmax 5 in one workgroup, so
// the P x Q grid, 7 GB at 54^3 x 5600 nodes, is read once); wave w owns
// points 16w .. 16w+15 and every column tile.  The grid and the angular table
// are staged through LDS in K blocks of kGemmK nodes, double buffered: the
// loads of block b+1 are issued before the MFMAs of block b.  Accumulator
// layout of mfma_f64_16x16x4: lane l holds rows (l/16) + 4q of column l%16.
constexpr int kGemmK = 16;
constexpr int kGemmColTiles = 5;
constexpr int kGemmCols = 16 * kGemmColTiles;

template <typename VT>
__global__ __launch_bounds__(kWG) void k_vblock_gemm(const double *__restrict__ vg, const double *__restrict__ h,
                                                     VT *__restrict__ vpk, int64_t P, int Q, int64_t npair0,
                                                     int64_t npair)
{
    __shared__ double As[2][64][kGemmK + 1];         // [buffer][point][node]
    __shared__ double Bs[2][kGemmK][kGemmCols + 1];  // [buffer][node][column]
    const int64_t ntile = vpk_ntile(P);
    const int64_t p0 = (int64_t)blockIdx.x * 64;
    const int64_t c0 = (int64_t)blockIdx.y * kGemmCols;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int l16 = lane & 15, kq = lane >> 4;
    // staging map: A 64 x 16 = 4 nodes per thread (one 32-byte run),
    // B 16 x 80 = 5 columns per thread
    const int ap = threadIdx.x >> 2, aq = (threadIdx.x & 3) * 4;
    const int bq = threadIdx.x >> 4, bc = (threadIdx.x & 15) * kGemmColTiles;
    const bool a_live = p0 + ap < P;
    const double *arow = vg + (size_t)(a_live ? p0 + ap : 0) * Q;
    double ra[4], rb[kGemmColTiles];
    // every load address is clamped into its array; the values outside are
    // replaced by zeros
    auto load = [&](int q0) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int q = q0 + aq + e;
            const double v = arow[q < Q ? q : Q - 1];
            ra[e] = (a_live && q < Q) ? v : 0.0;
        }
        const int qb = q0 + bq;
        const int qc = qb < Q ? qb : Q - 1;
#pragma unroll
        for (int e = 0; e < kGemmColTiles; e++) {
            const int64_t c = c0 + bc + e;
            const double v = h[(size_t)qc * npair + (c < npair ? c : npair - 1)];
            rb[e] = (qb < Q && c < npair) ? v : 0.0;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int e = 0; e < 4; e++) As[buf][ap][aq + e] = ra[e];
#pragma unroll
        for (int e = 0; e < kGemmColTiles; e++) Bs[buf][bq][bc + e] = rb[e];
    };
    kd4 acc[kGemmColTiles];
#pragma unroll
    for (int t = 0; t < kGemmColTiles; t++) acc[t] = kd4{0, 0, 0, 0};
    const int nblk = (Q + kGemmK - 1) / kGemmK;
    load(0);
    store(0);
    __syncthreads();
    for (int b = 0; b < nblk; b++) {
        const int cur = b & 1;
        if (b + 1 < nblk) load((b + 1) * kGemmK);
#pragma unroll
        for (int ks = 0; ks < kGemmK / 4; ks++) {
            const double a = As[cur][16 * wid + l16][4 * ks + kq];
#pragma unroll
            for (int t = 0; t < kGemmColTiles; t++) acc[t] = kmfma(a, Bs[cur][4 * ks + kq][16 * t + l16], acc[t]);
        }
        if (b + 1 < nblk) store(cur ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < kGemmColTiles; t++) {
        const int64_t col = c0 + 16 * t + l16;
        if (col >= npair) continue;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int64_t p = p0 + 16 * wid + kq + 4 * q;
            if (p < P) vpk[vpk_index(col, p, npair0, npair, ntile)] = (VT)acc[t][q];
        }
    }
}

int launch_vblock_gemm(const double *vg, const double *h, void *vpk, int vprec, int64_t P, int Q, int64_t npair0,
                       int64_t npair, hipStream_t s)
{
    dim3 grid((unsigned)((P + 63) / 64), (unsigned)((npair + kGemmCols - 1) / kGemmCols));
    if (vprec == BSL_PREC_FP64)
        hipLaunchKernelGGL(k_vblock_gemm<double>, grid, dim3(kWG), 0, s, vg, h, (double *)vpk, P, Q, npair0, npair);
    else
        hipLaunchKernelGGL(k_vblock_gemm<float>, grid, dim3(kWG), 0, s, vg, h, (float *)vpk, P, Q, npair0, npair);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

}  // namespace bsl
