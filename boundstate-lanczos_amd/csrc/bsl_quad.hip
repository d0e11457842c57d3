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
// Hemisphere fold.  The H2 potential is unchanged by inverting the molecular
// axis, V(pi - theta, phi + pi) = V(theta, phi), both quadrature grids map
// onto themselves under that inversion, and Y_lm(pi - theta, phi + pi) =
// (-1)^l Y_lm(theta, phi).  So over the whole grid the products of two
// harmonics of one l parity count twice what they count over one
// hemisphere and those of opposite parity cancel: u = 2 [T(c_even)]_even +
// 2 [T(c_odd)]_odd with T the transform chain above over the hemisphere's
// nodes a < ceil(na/2) (the equator row of an odd na maps onto itself
// through phi + pi and keeps weight 1).  The kernel carries the two parities
// through separate phi transforms (the forward Legendre sum split by the
// plane's parity, the backward one reading the plane's own) over half the
// columns, and W holds the hemisphere only, with the factor 2 folded in.
//
// The two phi transforms are GEMMs batched over columns (P, a): with
// columns as N and phi as M (forward) or K (backward), on MFMA 16x16x4
// (FP64 or FP32, FP32 accumulation for FP32) or gfx950's 16x16x32 BF16
// (three split products hi*hi + hi*lo + lo*hi of FP32 operands, FP32
// accumulation).  The forward accumulator tile (16 phi x 16 columns) is
// multiplied by W in registers and fed straight back as the B operand of the
// backward MFMA, so the (a, phi) intermediate never leaves the registers.  W
// is stored in that accumulator layout: [column tile][phi tile][lane][4], one
// contiguous 16-byte (FP32) or 32-byte (FP64) load per lane.  Accumulator
// rows of lane j (item i): 4*(j/16) + i for the FP32 and BF16 instructions,
// j/16 + 4*i for the FP64 one.
//
// BF16x3 on the K = 32 instruction (A lane l: row l%16, k = 8(l/16) + e; B
// lane l: k = 8(l/16) + e, column l%16).  Forward: K packs the split terms,
// k < 16 taking S_hi x F_hi and S_hi x F_lo (two instructions) and k >= 16
// S_lo x F_hi, over the 16 m rows.  Backward: K runs over 32 phi, so that the
// phi tiles are taken in pairs (2u, 2u + 1) whose rows are permuted: row
// 4 kq + i of tile 2u + s holds phi 32u + 8 kq + 4 s + i, and lane group kq
// of the pair then holds exactly the 8 phi values its backward B operand
// needs; an odd last tile keeps the natural rows and fills half of its
// backward K.  The forward B operand needs m = 8 (kq & 1) + e, e < 8: the
// forward Legendre transform, itself on the K = 32 instruction (A: exact 0/1
// selections of the carried planes by m and parity, B: hi + lo of the FP32
// products L c over 32 planes), fills two accumulators whose selection rows
// are permuted the same way (m = 8 (kq & 1) + i and 8 (kq & 1) + 4 + i).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "bsl_internal.h"

namespace bsl {

constexpr int kQWG = 256;  // 4 waves

typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d4v __attribute__((ext_vector_type(4)));

template <typename T>
struct AccT;
template <>
struct AccT<float> {
    typedef f4v type;
};
template <>
struct AccT<double> {
    typedef d4v type;
};

__device__ inline f4v mfma16(float a, float b, f4v c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ inline d4v mfma16(double a, double b, d4v c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// bfloat16 split of a float: x ~ hi + lo, both round-to-nearest-even
__device__ inline short bf16_rne(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ inline float bf16_to_f(short h) { return __uint_as_float(((unsigned)(unsigned short)h) << 16); }
__device__ inline void bf16_split(float x, short &hi, short &lo)
{
    hi = bf16_rne(x);
    lo = bf16_rne(x - bf16_to_f(hi));
}
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));
// the same split of two floats through the packed conversion instruction
// (round to nearest even): hi = bf16(x), lo = bf16(x - hi), element 0 in the
// low half of each word
__device__ inline void split_pk(float x0, float x1, unsigned &hi, unsigned &lo)
{
    const f2v x = {x0, x1};
    const unsigned h = __builtin_bit_cast(unsigned, __builtin_convertvector(x, bf2v));
    const f2v r = {x0 - __uint_as_float(h << 16), x1 - __uint_as_float(h & 0xffff0000u)};
    hi = h;
    lo = __builtin_bit_cast(unsigned, __builtin_convertvector(r, bf2v));
}
// hi and lo operands of eight floats: elements 0-3 from a, 4-7 from b
__device__ inline void split8(const f4v &a, const f4v &b, bf8v &hi, bf8v &lo)
{
    unsigned h0, l0, h1, l1, h2, l2, h3, l3;
    split_pk(a[0], a[1], h0, l0);
    split_pk(a[2], a[3], h1, l1);
    split_pk(b[0], b[1], h2, l2);
    split_pk(b[2], b[3], h3, l3);
    const u4v hv = {h0, h1, h2, h3}, lv = {l0, l1, l2, l3};
    hi = __builtin_bit_cast(bf8v, hv);
    lo = __builtin_bit_cast(bf8v, lv);
}
__device__ inline bf8v bf8_zero()
{
    const u4v z = {0u, 0u, 0u, 0u};
    return __builtin_bit_cast(bf8v, z);
}
__device__ inline f4v mfma_bf16_32(const bf8v &a, const bf8v &b, f4v c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline double q_block_sum(double v, double *lds)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += lds[w];
    __syncthreads();
    return s;
}

// accumulator row of item i in lane group kq = lane/16
template <bool F64>
__host__ __device__ inline int acc_row(int kq, int i)
{
    return F64 ? kq + 4 * i : 4 * kq + i;
}

// Virtual plane layout of the BF16 Legendre transform: blocks of 8 planes
// of one l parity, the even-l planes (r < ne) in the first ceil(ne/8)
// blocks and the odd-l ones after them when both parities are carried,
// otherwise the planes in order; padding zero; K steps of 4 blocks.
__host__ __device__ inline bool quad_two(int nk, int ne) { return ne > 0 && ne < nk; }
__host__ __device__ inline int quad_even_blocks(int nk, int ne) { return quad_two(nk, ne) ? (ne + 7) / 8 : (nk + 7) / 8; }
__host__ __device__ inline int quad_vsteps(int nk, int ne)
{
    const int nb = quad_even_blocks(nk, ne) + (quad_two(nk, ne) ? (nk - ne + 7) / 8 : 0);
    return (nb + 3) / 4;
}
// virtual index of plane r, and the plane of virtual index k (-1: padding)
__host__ __device__ inline int quad_vindex(int r, int nk, int ne)
{
    return quad_two(nk, ne) && r >= ne ? 8 * quad_even_blocks(nk, ne) + r - ne : r;
}
__host__ __device__ inline int quad_vplane(int k, int nk, int ne)
{
    if (!quad_two(nk, ne)) return k < nk ? k : -1;
    const int e8 = 8 * quad_even_blocks(nk, ne);
    if (k < e8) return k < ne ? k : -1;
    return k - e8 < nk - ne ? ne + k - e8 : -1;
}

// Shared-memory plan of one workgroup: kQW waves, each owning ppw grid
// points (ppw*na a multiple of 16: whole column tiles per wave).
struct QuadLds {
    int sp, fm, bp, kps;          // row strides (odd or padded, against bank conflicts)
    size_t off_s, off_b, off_x, off_u, off_l, off_c, off_sel, bytes;
};

constexpr int kQW = kQWG / 64;
// sched_barrier mask: VALU, SALU and LDS instructions may cross; global
// loads and MFMAs keep their order around it
constexpr int kKeepLoads = 0x2 | 0x4 | 0x100 | 0x200;

__host__ __device__ inline size_t align16(size_t o) { return (o + 15) & ~(size_t)15; }

__host__ __device__ inline QuadLds quad_lds(int nk, int ne, int na, int npts, int kf, int mode, int nphp)
{
    QuadLds q;
    const int es = mode == 0 ? 8 : 4;
    q.sp = 17;
    q.fm = kf + 1;
    q.bp = nphp + 4;
    q.kps = 0;
    q.off_c = 0;
    // c [npts][nk] in FP64 (FP64 and FP32 modes; the BF16 mode keeps an FP32
    // copy in its virtual layout instead)
    size_t o = mode == 2 ? 0 : sizeof(double) * (size_t)npts * nk;
    o = align16(o);
    q.off_s = o;
    if (mode == 2) {
        // bf16 splits: forward A [phi][16] hi, lo; backward A [16][phi] hi,
        // lo (rows of 16-byte aligned runs of 8 phi)
        q.bp = nphp + 8;
        o += 2 * sizeof(short) * (size_t)nphp * 16;
        q.off_b = o;
        o += 2 * sizeof(short) * (size_t)16 * q.bp;
    } else {
        o += (size_t)es * nphp * q.sp;             // S [phi][m] (m padded to 16)
        q.off_b = o;
    }
    o = align16(o);
    q.off_x = o;
    o += (size_t)es * 2 * kQW * 16 * q.fm;         // per-wave F' of one column tile [16][m], both parities
    o = align16(o);
    q.off_u = o;
    o += sizeof(double) * (size_t)npts * nk;       // u [npts][nk]
    o = align16(o);
    q.off_l = o;
    if (mode == 2) {
        const int vs = quad_vsteps(nk, ne);
        q.kps = 32 * vs + 4;
        o += sizeof(float) * (size_t)na * q.kps;   // L [na][virtual plane]
        o = align16(o);
        q.off_c = o;
        o += sizeof(float) * (size_t)npts * q.kps; // c [npts][virtual plane]
        o = align16(o);
        q.off_sel = o;
        o += (size_t)16 * vs * 4 * 64;             // m selections, bf16 A operand [K step][table][lane][8]
    } else {
        o += (size_t)es * nk * na;                 // L [nk][na]
        o = align16(o);
        q.off_sel = o;
        o += (size_t)es * ((nk + 3) / 4) * 64;     // m selection, A operand [K step][lane]
    }
    q.bytes = o;
    return q;
}

template <int MODE>
struct QuadTypes;
template <>
struct QuadTypes<0> {
    typedef double T;
    typedef d4v acc;
};
template <>
struct QuadTypes<1> {
    typedef float T;
    typedef f4v acc;
};
template <>
struct QuadTypes<2> {
    typedef float T;
    typedef f4v acc;
};

template <int MODE>
__device__ inline typename QuadTypes<MODE>::acc load_w(const typename QuadTypes<MODE>::T *p)
{
    if constexpr (MODE == 0) {
        const double2 w0 = *(const double2 *)p, w1 = *(const double2 *)(p + 2);
        d4v w = {w0.x, w0.y, w1.x, w1.y};
        return w;
    } else {
        return *(const f4v *)p;
    }
}

// W: [tile][NT][64][4] (tile = global 16-column tile, NT phi tiles, the
// ones beyond nphi zero; see k_weight_grid); Lg: [nk][na]; Sg: [nm][nphi];
// mplane[r]: m + lmax of carried plane r.  MODE 0: FP64 MFMA; 1: FP32 MFMA;
// 2: BF16x3 MFMA (W in FP32).  KS: forward K steps of 4 m (m padded to
// 4*KS with zeros).  W of the next column tile is prefetched into NT register
// tiles while the current one is transformed.
template <int MODE, int NT, int KS>
__global__ __launch_bounds__(kQWG) void k_quad_mfma(const double *__restrict__ X, double *__restrict__ R,
                                                    const void *__restrict__ Wv, const double *__restrict__ Lg,
                                                    const double *__restrict__ Sg, const int *__restrict__ mplane,
                                                    const double *__restrict__ mask,
                                                    const Scal *__restrict__ sc, int64_t P, int nk, int na,
                                                    int ne, int nphi, int nm, int ppw, int64_t ngroups,
                                                    double *__restrict__ partial, double *__restrict__ partial_b,
                                                    double *__restrict__ partial_c)
{
    typedef typename QuadTypes<MODE>::T T;
    typedef typename QuadTypes<MODE>::acc acc_t;
    // Legendre backward accumulation: FP64, FP32 for the BF16 mode
    typedef typename std::conditional<MODE == 2, float, double>::type U;
    constexpr bool F64 = MODE == 0;
    constexpr int KF = 4 * KS;  // m rows carried through the transforms
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double red[kQW];
    const int npts = kQW * ppw;
    const QuadLds q = quad_lds(nk, ne, na, npts, KF, MODE, 16 * NT);
    double *cs = (double *)smem;
    double *us = (double *)(smem + q.off_u);
    T *Ls = (T *)(smem + q.off_l);
    (void)sc;  // X already holds v (formed by the kinetic kernel)
    // both l parities carried (full basis, uniform): two transform chains;
    // otherwise the one parity takes the first chain
    const bool two = quad_two(nk, ne);
    // K steps, and the blocks of 8 virtual planes of the first chain
    const int vs = quad_vsteps(nk, ne), nbe = quad_even_blocks(nk, ne);

    // ---- tables and inputs ----
    if constexpr (MODE == 2) {
        // L and c in the virtual plane layout (FP32), and the A operands of
        // the forward Legendre MFMA: for K step s, table tb and lane (row
        // l16, lane group kq) element e is 1 where virtual plane
        // k = 32 s + 8 kq + e has m = 8 ((row / 4) & 1) + 4 (tb & 1) +
        // row % 4 and the parity of table tb (tb < 2: the first chain, even
        // l or the one parity carried; tb >= 2: odd l), so that accumulator
        // item i of lane group kq holds m = 8 (kq & 1) + 4 (tb & 1) + i, the
        // m of the forward phi operand of that lane
        float *Lt = (float *)(smem + q.off_l), *cf = (float *)(smem + q.off_c);
        for (int t = threadIdx.x; t < na * q.kps; t += blockDim.x) {
            const int a = t / q.kps, r = quad_vplane(t - a * q.kps, nk, ne);
            Lt[t] = r >= 0 ? (float)Lg[(size_t)r * na + a] : 0.f;
        }
        for (int t = threadIdx.x; t < npts * q.kps; t += blockDim.x) cf[t] = 0.f;
        unsigned short *selb = (unsigned short *)(smem + q.off_sel);
        for (int t = threadIdx.x; t < vs * 4 * 64 * 8; t += blockDim.x) {
            const int e = t & 7, ln = (t >> 3) & 63, tb = (t >> 9) & 3, st = t >> 11;
            const int row = ln & 15, r = quad_vplane(32 * st + 8 * (ln >> 4) + e, nk, ne);
            const int msel = 8 * ((row >> 2) & 1) + 4 * (tb & 1) + (row & 3);
            const bool par = r >= 0 && (tb < 2 ? !two || r < ne : two && r >= ne);
            selb[t] = (par && mplane[r] == msel) ? (unsigned short)0x3f80 : (unsigned short)0;
        }
        short *fh = (short *)(smem + q.off_s), *fl = fh + 16 * NT * 16;
        short *bh = (short *)(smem + q.off_b), *bl = bh + 16 * q.bp;
        for (int t = threadIdx.x; t < NT * 16 * 16; t += blockDim.x) {
            const int ph = t >> 4, mi = t & 15;
            const float v = (ph < nphi && mi < nm) ? (float)Sg[(size_t)mi * nphi + ph] : 0.f;
            short h, l;
            bf16_split(v, h, l);
            fh[ph * 16 + mi] = h;
            fl[ph * 16 + mi] = l;
            bh[mi * q.bp + ph] = h;
            bl[mi * q.bp + ph] = l;
        }
    } else {
        for (int t = threadIdx.x; t < nk * na; t += blockDim.x) Ls[t] = (T)Lg[t];
        // A operand of the forward Legendre MFMA: Sel[row][r] = 1 where
        // carried plane r has m = msel(row), for K step s4 (r = 4 s4 +
        // lane/16) and row lane%16.  msel orders the rows so that
        // accumulator item i of lane group kq holds the m of the forward phi
        // operand of that lane (m = 4i + kq for the FP32 instruction, whose
        // D rows are 4kq + i; the FP64 one's rows kq + 4i are already m).
        const int nks = (nk + 3) >> 2;
        T *Sel = (T *)(smem + q.off_sel);
        for (int t = threadIdx.x; t < nks * 64; t += blockDim.x) {
            const int r = 4 * (t >> 6) + ((t & 63) >> 4), row = t & 15;
            const int msel = (MODE == 1) ? 4 * (row & 3) + (row >> 2) : row;
            Sel[t] = (r < nk && mplane[r] == msel) ? (T)1 : (T)0;
        }
        T *Ss = (T *)(smem + q.off_s);
        for (int t = threadIdx.x; t < NT * 16 * 16; t += blockDim.x) {
            const int ph = t >> 4, mi = t & 15;
            Ss[ph * q.sp + mi] = (ph < nphi && mi < nm) ? (T)Sg[(size_t)mi * nphi + ph] : (T)0;
        }
    }

    // ---- per wave: ppw points of each group, tw column tiles ----
    const int lane = threadIdx.x & 63, wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int l16 = lane & 15, kq = lane >> 4;
    const int tw = ppw * na / 16;
    const int pw0 = wid * ppw;  // first point of this wave within the group
    const T *W = (const T *)Wv + (size_t)lane * 4;
    T *xs = (T *)(smem + q.off_x) + wid * 16 * q.fm;          // even-l F'
    T *xo = (T *)(smem + q.off_x) + (kQW + wid) * 16 * q.fm;  // odd-l F'
    const int mr = lane < nk ? mplane[lane] : 0;
    // the lane's plane in the backward Legendre transform: its L row (the
    // BF16 mode reads the virtual layout, stride q.kps over the nodes)
    const T *lrow = MODE == 2 ? Ls + (lane < nk ? quad_vindex(lane, nk, ne) : 0) : Ls + lane * na;
    const int lstride = MODE == 2 ? q.kps : 1;
    U uacc = 0;

    // One column tile.  Each phi tile's W register tile is refilled with the
    // W of the next tile (Wn: the next tile of this wave, or its first tile in
    // the workgroup's next group) right after its use (prefetch:
    // std::true_type); the scheduling barrier keeps the load there (the
    // compiler would otherwise sink it to the end of the tile), so its latency
    // overlaps the rest of this tile's transforms.
    // this lane's column 16 n + l16 of the wave: point pc, node ac (advanced
    // by one tile at the end of each tile, set at each group's start)
    int pc = 0, ac = 0;
    auto tile = [&](int n, acc_t(&wc)[NT], const T *Wn, auto prefetch) {
        acc_t ae = {0, 0, 0, 0}, ao = {0, 0, 0, 0};
        if constexpr (MODE == 2) {
            // Legendre forward on the BF16 instruction: f(m, col) = sum_k
            // Sel[m][k] L_k(a) c_k(p) over 32 virtual planes per K step, the
            // B operand (k = 8 kq + e, column l16) this lane's own products
            // split into hi + lo (Sel is exact).  Two accumulators per
            // parity: item i of lane group kq holds m = 8 (kq & 1) + i (A)
            // and 8 (kq & 1) + 4 + i (B); a step without blocks of a parity
            // skips that parity's instructions.
            const float *Lt = (const float *)(smem + q.off_l), *cf = (const float *)(smem + q.off_c);
            const bf8v *selb = (const bf8v *)(smem + q.off_sel);
            f4v feA = {0, 0, 0, 0}, feB = {0, 0, 0, 0}, foA = {0, 0, 0, 0}, foB = {0, 0, 0, 0};
            for (int st = 0; st < vs; st++) {
                const float *cp = cf + pc * q.kps + 32 * st + 8 * kq, *lp = Lt + ac * q.kps + 32 * st + 8 * kq;
                const f4v p0 = *(const f4v *)cp * *(const f4v *)lp;
                const f4v p1 = *(const f4v *)(cp + 4) * *(const f4v *)(lp + 4);
                bf8v ph, pl;
                split8(p0, p1, ph, pl);
                const bf8v *sl = selb + 4 * st * 64 + lane;
                if (4 * st < nbe) {
                    feA = mfma_bf16_32(sl[0], ph, feA);
                    feA = mfma_bf16_32(sl[0], pl, feA);
                    feB = mfma_bf16_32(sl[64], ph, feB);
                    feB = mfma_bf16_32(sl[64], pl, feB);
                }
                if (two && 4 * st + 4 > nbe) {
                    foA = mfma_bf16_32(sl[128], ph, foA);
                    foA = mfma_bf16_32(sl[128], pl, foA);
                    foB = mfma_bf16_32(sl[192], ph, foB);
                    foB = mfma_bf16_32(sl[192], pl, foB);
                }
            }
            // forward phi operands, k = 8 kq + e <-> m = 8 (kq & 1) + e.  A:
            // S_hi for k < 16 and S_lo for k >= 16 (rows: phi of the tile);
            // B1 = F_hi, B2 = F_lo for k < 16 and zero for k >= 16, so that
            // A B1 + A B2 = S_hi F_hi + S_lo F_hi + S_hi F_lo.
            const bool lo_half = kq < 2;
            const bf8v zero8 = bf8_zero();
            bf8v b1e, b2e, b1o, b2o;
            split8(feA, feB, b1e, b2e);
            b2e = lo_half ? b2e : zero8;
            if (two) {
                split8(foA, foB, b1o, b2o);
                b2o = lo_half ? b2o : zero8;
            }
            const short *fa = lo_half ? (const short *)(smem + q.off_s) + 8 * kq
                                      : (const short *)(smem + q.off_s) + 16 * NT * 16 + 8 * (kq - 2);
            // pairs of tiles: row l16 of tile 2u + sx is phi 32u + 8 (l16 / 4) + 4 sx + l16 % 4
            const short *fap = fa + (8 * (l16 >> 2) + (l16 & 3)) * 16;
            const short *bhr = (const short *)(smem + q.off_b) + l16 * q.bp + 8 * kq;
            const short *blr = bhr + 16 * q.bp;
#pragma unroll
            for (int u = 0; u < NT / 2; u++) {
                f4v he[2], ho[2];
#pragma unroll
                for (int sx = 0; sx < 2; sx++) {
                    const int t = 2 * u + sx;
                    const bf8v a1 = *(const bf8v *)(fap + (32 * u + 4 * sx) * 16);
                    f4v ge = {0.f, 0.f, 0.f, 0.f}, go = {0.f, 0.f, 0.f, 0.f};
                    ge = mfma_bf16_32(a1, b1e, ge);
                    ge = mfma_bf16_32(a1, b2e, ge);
                    if (two) {
                        go = mfma_bf16_32(a1, b1o, go);
                        go = mfma_bf16_32(a1, b2o, go);
                    }
                    he[sx] = ge * wc[t];
                    ho[sx] = go * wc[t];
                    if constexpr (decltype(prefetch)::value) {
                        wc[t] = load_w<MODE>(Wn + (size_t)t * 256);
                        __builtin_amdgcn_sched_barrier(kKeepLoads);
                    }
                }
                // backward over the pair's 32 phi: A [m = l16][phi 32u + 8 kq + e]
                const bf8v a3h = *(const bf8v *)(bhr + 32 * u);
                const bf8v a3l = *(const bf8v *)(blr + 32 * u);
                bf8v hh, hl;
                split8(he[0], he[1], hh, hl);
                ae = mfma_bf16_32(a3h, hh, ae);
                ae = mfma_bf16_32(a3h, hl, ae);
                ae = mfma_bf16_32(a3l, hh, ae);
                if (two) {
                    split8(ho[0], ho[1], hh, hl);
                    ao = mfma_bf16_32(a3h, hh, ao);
                    ao = mfma_bf16_32(a3h, hl, ao);
                    ao = mfma_bf16_32(a3l, hh, ao);
                }
            }
            if constexpr (NT & 1) {
                // odd last tile, rows in natural order (phi 16t + l16 and
                // accumulator item i of lane group kq at phi 16t + 4kq + i);
                // its backward fills half of K (k = 8kq + e, e < 4, at phi
                // 16t + 4kq + e), the other half multiplying zeros
                constexpr int t = NT - 1;
                const bf8v a1 = *(const bf8v *)(fa + (16 * t + l16) * 16);
                f4v ge = {0.f, 0.f, 0.f, 0.f}, go = {0.f, 0.f, 0.f, 0.f};
                ge = mfma_bf16_32(a1, b1e, ge);
                ge = mfma_bf16_32(a1, b2e, ge);
                if (two) {
                    go = mfma_bf16_32(a1, b1o, go);
                    go = mfma_bf16_32(a1, b2o, go);
                }
                const f4v he = ge * wc[t], ho = go * wc[t];
                if constexpr (decltype(prefetch)::value) {
                    wc[t] = load_w<MODE>(Wn + (size_t)t * 256);
                    __builtin_amdgcn_sched_barrier(kKeepLoads);
                }
                const f4v z4 = {0.f, 0.f, 0.f, 0.f};
                const unsigned *bq = (const unsigned *)(smem + q.off_b);
                // S rows m = l16 at phi 16t + 4kq .. + 3 (8-byte aligned)
                const int wo = (l16 * q.bp + 16 * t + 4 * kq) >> 1;
                const u4v h4 = {bq[wo], bq[wo + 1], 0u, 0u};
                const u4v l4 = {bq[wo + 8 * q.bp], bq[wo + 8 * q.bp + 1], 0u, 0u};
                const bf8v a3h = __builtin_bit_cast(bf8v, h4), a3l = __builtin_bit_cast(bf8v, l4);
                bf8v hh, hl;
                split8(he, z4, hh, hl);
                ae = mfma_bf16_32(a3h, hh, ae);
                ae = mfma_bf16_32(a3h, hl, ae);
                ae = mfma_bf16_32(a3l, hh, ae);
                if (two) {
                    split8(ho, z4, hh, hl);
                    ao = mfma_bf16_32(a3h, hh, ao);
                    ao = mfma_bf16_32(a3h, hl, ao);
                    ao = mfma_bf16_32(a3l, hh, ao);
                }
            }
        } else {
            // Legendre forward on MFMA: f(m, col) = sum_r Sel[m][r] L_r(a) c_r(p)
            // over the carried planes, the B operand (k = r, column l16) being
            // this lane's own products; item i of lane group kq ends up
            // holding the m of the forward phi operand (see Sel).  Even-l
            // planes (r < ne) feed fe, odd-l planes fo.
            const int nks = (nk + 3) >> 2;
            const T *Sel = (const T *)(smem + q.off_sel);
            acc_t fe = {0, 0, 0, 0}, fo = {0, 0, 0, 0};
            for (int s4 = 0; s4 < nks; s4++) {
                const int r = 4 * s4 + kq;
                const double prod = r < nk ? (double)Ls[r * na + ac] * cs[pc * nk + r] : 0.0;
                const T sel = Sel[s4 * 64 + lane];
                fe = mfma16(sel, (T)(two && r >= ne ? 0.0 : prod), fe);
                if (two) fo = mfma16(sel, (T)(r >= ne ? prod : 0.0), fo);
            }
            const T *Ss = (const T *)(smem + q.off_s);
            T bFe[KS], bFo[KS];  // m = 4s + kq of this lane's column
#pragma unroll
            for (int s = 0; s < KS; s++) {
                bFe[s] = fe[s];
                bFo[s] = fo[s];
            }
#pragma unroll
            for (int t = 0; t < NT; t++) {
                acc_t ge = {0, 0, 0, 0}, go = {0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < KS; s++) ge = mfma16(Ss[(16 * t + l16) * q.sp + 4 * s + kq], bFe[s], ge);
                if (two) {
#pragma unroll
                    for (int s = 0; s < KS; s++) go = mfma16(Ss[(16 * t + l16) * q.sp + 4 * s + kq], bFo[s], go);
                }
                const acc_t he = ge * wc[t], ho = go * wc[t];
                if constexpr (decltype(prefetch)::value) {
                    wc[t] = load_w<MODE>(Wn + (size_t)t * 256);
                    __builtin_amdgcn_sched_barrier(kKeepLoads);
                }
#pragma unroll
                for (int i = 0; i < 4; i++) ae = mfma16(Ss[(16 * t + acc_row<F64>(kq, i)) * q.sp + l16], he[i], ae);
                if (two) {
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        ao = mfma16(Ss[(16 * t + acc_row<F64>(kq, i)) * q.sp + l16], ho[i], ao);
                }
            }
        }
        // F' of the tile (both parities) -> per-wave LDS, then the Legendre
        // backward of its 16 columns by lane r (one carried plane per lane,
        // reading its own parity's F')
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int m = acc_row<F64>(kq, i);
            if (m < KF) {
                xs[l16 * q.fm + m] = ae[i];
                xo[l16 * q.fm + m] = ao[i];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < nk) {
            const T *xf = (two && lane >= ne) ? xo : xs;
            // independent products first, then the sums over each point's
            // nodes in order; column c of the tile is node (16n + c) % na of
            // point pw0 + (16n + c) / na (wave-uniform), and a point's sum is
            // stored at its last node (it may begin in an earlier tile)
            const int a0 = (16 * n) % na;
            U v[16];
            int aa = a0;
#pragma unroll
            for (int c = 0; c < 16; c++) {
                v[c] = (U)lrow[aa * lstride] * (U)xf[c * q.fm + mr];
                if (++aa == na) aa = 0;
            }
            int pp = pw0 + (16 * n) / na;
            aa = a0;
#pragma unroll
            for (int c = 0; c < 16; c++) {
                uacc += v[c];
                if (++aa == na) {
                    us[pp * nk + lane] = (double)uacc;
                    uacc = 0;
                    aa = 0;
                    pp++;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        ac += 16;
        while (ac >= na) {
            ac -= na;
            pc++;
        }
    };
    // The workgroup is persistent: it takes the groups blockIdx.x,
    // blockIdx.x + gridDim.x, ... (npts points each); the tables above are
    // loaded once, and the W prefetch runs on across group boundaries.
    double dot = 0.0, rr2 = 0.0, vv2 = 0.0;
    acc_t wb[NT];
    int64_t g = blockIdx.x;
    if (g < ngroups) {
        const int64_t t0 = g * kQW * tw + (int64_t)wid * tw;
#pragma unroll
        for (int t = 0; t < NT; t++) wb[t] = load_w<MODE>(W + (size_t)(t0 * NT + t) * 256);
    }
    for (; g < ngroups; g += gridDim.x) {
        const int64_t p0 = g * npts;
        const int64_t tile0 = g * kQW * tw + (int64_t)wid * tw;
        const bool more = g + gridDim.x < ngroups;  // uniform over the workgroup
        for (int t = threadIdx.x; t < nk * npts; t += blockDim.x) {
            const int r = t / npts, p = t - r * npts;
            const double x = p0 + p < P ? X[(size_t)r * P + p0 + p] : 0.0;
            if constexpr (MODE == 2)
                ((float *)(smem + q.off_c))[p * q.kps + quad_vindex(r, nk, ne)] = (float)x;
            else
                cs[p * nk + r] = x;
        }
        __syncthreads();
        uacc = 0;
        pc = pw0 + l16 / na;
        ac = l16 % na;
        // one call site: the successor of the group's last tile is the wave's
        // first tile of the next group, or (none left) the tile itself
        const int64_t next0 = more ? (g + gridDim.x) * kQW * tw + (int64_t)wid * tw : tile0 + tw - 1;
        for (int n = 0; n < tw; n++) {
            const int64_t nt = n + 1 < tw ? tile0 + n + 1 : next0;
            tile(n, wb, W + (size_t)(nt * NT) * 256, std::true_type());
        }
        __syncthreads();

        // ---- R update and the partials of v.R, |R|^2, |v|^2 ----
        for (int t = threadIdx.x; t < nk * npts; t += blockDim.x) {
            const int r = t / npts, p = t - r * npts;
            if (p0 + p >= P) continue;
            const size_t o = (size_t)r * P + p0 + p;
            const double nr = R[o] + mask[r] * us[p * nk + r];
            R[o] = nr;
            const double v = MODE == 2 ? X[o] : cs[p * nk + r];
            dot += v * nr;
            rr2 += nr * nr;
            vv2 += v * v;
        }
        __syncthreads();  // c and us are rewritten by the next group
    }
    const double s0 = q_block_sum(dot, red);
    const double s1 = q_block_sum(rr2, red);
    const double s2 = q_block_sum(vv2, red);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s0;
        partial_b[blockIdx.x] = s1;
        partial_c[blockIdx.x] = s2;
    }
}

// W[tile][t][lane][i] = f_a w_a w_b vg[p][half*na*nb + a*nb + b] for column
// col = 16*tile + lane%16 = p*nah + a (the hemisphere's nodes a < nah =
// ceil(na/2)) and phi = 16*t + acc_row(lane/16, i) = half*nb + b, or for the
// BF16 pairs of tiles (PAIR, all tiles but the last of an odd count)
// phi = 32 (t/2) + 8 (lane/16) + 4 (t%2) + i;
// f_a = 2, or 1 on the equator row of an odd na (hemisphere fold); zero
// outside the grid (padding columns and phi).  FP64 grids feed the FP64
// instruction, FP32 grids the FP32 and BF16 ones.
template <typename WT, bool PAIR>
__global__ void k_weight_grid(const double *__restrict__ vg, WT *__restrict__ W, const double *__restrict__ wt,
                              const double *__restrict__ wp, int64_t P, int na, int nb, int64_t ntiles, int ntp)
{
    const int nah = (na + 1) / 2;
    const int nphi = 2 * nb;
    const int64_t total = ntiles * ntp * 256;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(idx & 3), lane = (int)((idx >> 2) & 63);
        const int64_t rest = idx >> 8;
        const int t = (int)(rest % ntp);
        const int64_t tile = rest / ntp;
        const int64_t col = 16 * tile + (lane & 15);
        const int ph = PAIR && t < (ntp & ~1) ? 32 * (t >> 1) + 8 * (lane >> 4) + 4 * (t & 1) + i
                                              : 16 * t + acc_row<sizeof(WT) == 8>(lane >> 4, i);
        const int64_t p = col / nah;
        const int a = (int)(col - p * nah);
        double v = 0.0;
        if (p < P && ph < nphi) {
            const int half = ph / nb, b = ph - half * nb;
            const double fold = (na % 2 == 1 && a == nah - 1) ? 1.0 : 2.0;
            v = fold * wt[a] * wp[b] * vg[(size_t)p * 2 * na * nb + (size_t)half * na * nb + (size_t)a * nb + b];
        }
        W[idx] = (WT)v;
    }
}

// phi tiles of the kernel instantiations (zero padded beyond nphi)
int quad_phi_tiles(int nphi)
{
    const int nt = (nphi + 15) / 16;
    return nt <= 1 ? 1 : nt <= 2 ? 2 : nt <= 4 ? 4 : nt <= 9 ? 9 : 16;
}

int quad_points_per_group(int na, int nk)
{
    // each wave owns ppw points with ppw*na a multiple of 16 (whole column
    // tiles), at least 4 tiles and 4 points per wave (16 points per group:
    // 128-byte segments of each plane in the group's X and R accesses)
    int g = 16;
    while (na % g) g >>= 1;
    const int unit = 16 / g, tiles_per_unit = na / g;
    int k = (4 + tiles_per_unit - 1) / tiles_per_unit;
    while (unit * k < 4) k++;
    while (k > 1 && quad_lds(nk, 0, na, kQW * unit * k, 16, 0, 256).bytes > 65536) k--;
    return kQW * unit * k;
}

int quad_kf(int nm, int mode);

std::vector<int> quad_group_candidates(const DevOperator &op)
{
    // 1x, 2x and 4x the smallest group, as long as the shared memory fits
    std::vector<int> out{op.q_npts};
    const int kf = quad_kf(op.nm, op.qmode), ntp = quad_phi_tiles(op.nphi);
    for (int f = 2; f <= 4; f *= 2) {
        const int npts = op.q_npts * f;
        if ((int64_t)npts > op.P) break;
        if (quad_lds(op.nk, op.q_ne, op.na, npts, kf, op.qmode, 16 * ntp).bytes > 65536) break;
        out.push_back(npts);
    }
    return out;
}

int64_t quad_tiles(int64_t P, int na, int npts)
{
    const int64_t groups = (P + npts - 1) / npts;
    return groups * (int64_t)npts * na / 16;
}

int launch_weight_grid(const double *vg, void *W, int wprec, int mode, const double *wt, const double *wp,
                       int64_t P, int na, int nb, int npts, hipStream_t s)
{
    const int64_t ntiles = quad_tiles(P, (na + 1) / 2, npts);
    const int ntp = quad_phi_tiles(2 * nb);
    const int64_t total = ntiles * ntp * 256;
    int64_t g = (total + 255) / 256;
    if (g > 65536) g = 65536;
    if (wprec == BSL_PREC_FP64)
        hipLaunchKernelGGL((k_weight_grid<double, false>), dim3((unsigned)g), dim3(256), 0, s, vg, (double *)W, wt,
                           wp, P, na, nb, ntiles, ntp);
    else if (mode == 2)
        hipLaunchKernelGGL((k_weight_grid<float, true>), dim3((unsigned)g), dim3(256), 0, s, vg, (float *)W, wt, wp,
                           P, na, nb, ntiles, ntp);
    else
        hipLaunchKernelGGL((k_weight_grid<float, false>), dim3((unsigned)g), dim3(256), 0, s, vg, (float *)W, wt,
                           wp, P, na, nb, ntiles, ntp);
    BSL_HIP(hipGetLastError());
    return BSL_OK;
}

#define BSL_QUAD_LAUNCH(M, NTV, KSV)                                                                       \
    do {                                                                                                   \
        if (occ) {                                                                                         \
            *occ = 0;                                                                                      \
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, (const void *)k_quad_mfma<M, NTV, KSV>,   \
                                                             kQWG, lds) != hipSuccess)                     \
                *occ = 0;                                                                                  \
        } else {                                                                                           \
            timed_launch((k_quad_mfma<M, NTV, KSV>), dim3(grid), dim3(kQWG), (uint32_t)lds, s, X, R, op.wgrid, \
                               op.qL, op.qS, op.mplane, op.mask, sc, op.P, op.nk, op.na, op.q_ne, op.nphi, \
                               op.nm, ppw, ngroups, op.partial, op.partial_b, op.partial_c);               \
        }                                                                                                  \
    } while (0)

// with occ != null: only the resident workgroups per CU of the instance
template <int M, int NTV>
static void launch_quad_ks(const DevOperator &op, const double *X, double *R, Scal *sc, unsigned grid, size_t lds,
                           int ppw, int64_t ngroups, int ks, hipStream_t s, int *occ)
{
    if constexpr (M == 2) {
        BSL_QUAD_LAUNCH(M, NTV, 4);
    } else {
        if (ks >= 4)
            BSL_QUAD_LAUNCH(M, NTV, 4);
        else if (ks <= 2)
            BSL_QUAD_LAUNCH(M, NTV, 2);
        else
            BSL_QUAD_LAUNCH(M, NTV, 3);
    }
}
#undef BSL_QUAD_LAUNCH

template <int M>
static void launch_quad_mode(const DevOperator &op, const double *X, double *R, Scal *sc, unsigned grid, size_t lds,
                             int ppw, int64_t ngroups, int ks, hipStream_t s, int *occ)
{
    switch (quad_phi_tiles(op.nphi)) {
    case 1: launch_quad_ks<M, 1>(op, X, R, sc, grid, lds, ppw, ngroups, ks, s, occ); break;
    case 2: launch_quad_ks<M, 2>(op, X, R, sc, grid, lds, ppw, ngroups, ks, s, occ); break;
    case 4: launch_quad_ks<M, 4>(op, X, R, sc, grid, lds, ppw, ngroups, ks, s, occ); break;
    case 9: launch_quad_ks<M, 9>(op, X, R, sc, grid, lds, ppw, ngroups, ks, s, occ); break;
    default: launch_quad_ks<M, 16>(op, X, R, sc, grid, lds, ppw, ngroups, ks, s, occ); break;
    }
}

int quad_kf(int nm, int mode)
{
    // m rows carried: 4*KS with KS >= 2 (16 for the BF16 instruction's K)
    const int ks = (nm + 3) / 4;
    return mode == 2 ? 16 : 4 * (ks < 2 ? 2 : ks);
}

// launch (occ == null) or report the resident workgroups per CU (occ)
static int quad_dispatch(const DevOperator &op, const double *X, double *R, Scal *sc, hipStream_t s, int *occ)
{
    const int npts = op.q_npts, ppw = npts / kQW;
    const int64_t ngroups = (op.P + npts - 1) / npts;
    const unsigned grid = (unsigned)op.q_grid;
    if (!occ && (grid == 0 || (int64_t)grid > op.npartial || (int64_t)grid != op.nalpha))
        return fail(BSL_E_STATE, "partial buffer does not match the quadrature kernel grid");
    const int kf = quad_kf(op.nm, op.qmode), ntp = quad_phi_tiles(op.nphi);
    const size_t lds = quad_lds(op.nk, op.q_ne, op.na, npts, kf, op.qmode, 16 * ntp).bytes;
    if (op.nm > 16 || op.nk > 64 || op.nphi > 256 || npts % kQW || (ppw * op.na) % 16 || lds > 65536)
        return fail(BSL_E_ARG, "quadrature kernel: unsupported sizes");
    if (op.qmode == 0)
        launch_quad_mode<0>(op, X, R, sc, grid, lds, ppw, ngroups, kf / 4, s, occ);
    else if (op.qmode == 1)
        launch_quad_mode<1>(op, X, R, sc, grid, lds, ppw, ngroups, kf / 4, s, occ);
    else
        launch_quad_mode<2>(op, X, R, sc, grid, lds, ppw, ngroups, kf / 4, s, occ);
    if (!occ) BSL_HIP(hipGetLastError());
    return BSL_OK;
}

int launch_quad_potential(const DevOperator &op, const double *X, double *R, Scal *sc, hipStream_t s)
{
    return quad_dispatch(op, X, R, sc, s, nullptr);
}

int quad_grid(const DevOperator &op, int device, int *grid)
{
    int per_cu = 0, cus = 0;
    int rc = quad_dispatch(op, nullptr, nullptr, nullptr, nullptr, &per_cu);
    if (rc) return rc;
    BSL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    if (per_cu < 1) per_cu = 1;
    const int64_t ngroups = (op.P + op.q_npts - 1) / op.q_npts;
    const int64_t g = (int64_t)per_cu * (cus > 0 ? cus : 1);
    *grid = (int)(g < ngroups ? g : ngroups);
    return BSL_OK;
}

}  // namespace bsl
