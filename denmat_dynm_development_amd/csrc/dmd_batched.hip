// dmd_batched.hip -- B trajectories sharing one pass over P (nb = 4, complex64 or BF16 P):
// the per-pair GEMVs of electronphonon::evolve (src_v4.5.7/evolve_eph.cpp:99-106)
// become GEMMs over the trajectory index and run on the matrix cores.
//
// For a destination row k and each of its visits v (a stored pair (k, q) in
// the forward direction, or (q, k) in the reverse direction):
//   F1_k[b] += A1_v vec(X_q[b])          A1 = P1      (forward) | conj(P2)   (reverse)
//   G_k[b]  += A2_v^T vec(X_q[b])        A2 = P2      (forward) | conj(P1)   (reverse)
// with X = drho_S = rho_S - diag(f) the FP32-class part of the state; the
// equilibrium part F(diag f, 1 - diag f) is the FP64 F_eq shared with the
// unbatched reduced forms, and dm1P = F2_eq - G (1 - rho = (1 - f) - drho).
//
// Matrix-core mapping (v_mfma_f32_32x32x16_bf16, FP32 accumulation).  X is
// Hermitian (rho_S is, f is real and diagonal), so its 16 complex entries are
// 16 reals x = (Re X_ii; Re X_ij, Im X_ij for i < j).  A vec(X) is then the
// real 32 x 16 product M x with M = [A_ii | A_ij + A_ji | i (A_ij - A_ji)]
// (rows: Re and Im of the 16 outputs), one K = 16 MFMA per 32 trajectories.
// M and x are carried as BF16 hi + lo (16-bit significands); the three
// leading cross terms make 3 MFMAs per product, 6 per visit per 32
// trajectories.
//
// FP16 records (BF16 store).  For P already rounded to BF16 (8-bit
// significands) a second 16-bit word of M is wasted precision: M is kept as
// one FP16 word (11-bit significand: a BF16 word, or the sum or difference of
// two whose exponents differ by at most 2, is exact; otherwise the rounding is
// <= 2^-11 relative, 8x below the BF16 store's own 2^-8) with
// v_mfma_f32_32x32x16_f16, 2 KiB per visit, and x
// as FP16 hi + lo (22-bit significand).  FP16's range is met by two powers of
// two: M x G with G = 2^14 / max per-pair scale of the store (every entry
// under 2^15), and x x s_b per trajectory with s_b = 2^14 / max |drho_b|
// (every operand under 2^14); the row sums are divided by G s_b at assembly.
// Entries below 2^-14 of those maxima lose relative precision (FP16
// subnormals), with absolute error under 2^-25 of the maximum.  4 MFMAs per
// visit per 32 trajectories (M xh, M xl for each product).
//
// Layouts in HBM.
//  * Visit records, CSR by destination row: M1 (from A1) and M2 (from A2^T),
//    32 x 16 BF16 row-major each, hi then lo: [M1 hi | M2 hi | M1 lo | M2 lo],
//    4 KiB per visit (8 KiB per stored pair), conj and the per-pair scale of
//    the BF16 store applied; FP16 records [M1 | M2] x G, 2 KiB per visit.
//  * X: per partner row q, [term][k-half][trajectory][8] BF16 (FP16 with FP16
//    records; term 0 hi, 1 lo), so a
//    wavefront's 32 trajectories of one 8-element operand are 512 contiguous
//    bytes (Bp x 64 bytes per row q).
//  * Row sums: FP32 [k][64][B] (F1 re, F1 im, G re, G im).
// Packed pass (B <= 16 with visit records, Bp = 16): the 32 columns of one
// MFMA carry x_hi of the 16 trajectories and x_lo, so one MFMA per product and
// record word forms both terms (2 MFMAs per visit with FP16 records instead of
// 4 at 32 padded trajectories, 4 instead of 6 with hi + lo records) and the
// partner operand is half as large; the two column halves meet in a lane
// exchange before the row-sum atomics.
// Uneven rows: every row's visits are cut into segments of at most seg_len
// visits, so a 5000-partner row is as many independent work items; segments
// of one row meet in FP32 atomics.  A persistent pass of resident wavefronts
// takes the segments from a queue counter, longest first within each block
// of partner rows.  seg_len, the partner-block size and the trajectory blocks
// per wavefront are picked by timing at the first batched call for a padded B
// (autotune).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dmd_internal.h"

namespace dmd {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) _Float16 halfx8;
typedef __attribute__((ext_vector_type(16))) float floatx16;

__device__ __forceinline__ uint32_t f2bf(float f)  // round to nearest even (finite)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ float bflo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ uint16_t f2h(float f) { return __builtin_bit_cast(uint16_t, (_Float16)f); }  // nearest even
__device__ __forceinline__ double h2d(uint16_t w) { return (double)__builtin_bit_cast(_Float16, w); }

// FP16 operand scale of a trajectory from the bits of its max |drho|: a power
// of two s with max |drho| s < 2^14 (1 for an all-zero or non-finite max).
__device__ __forceinline__ double xscale_of(uint32_t bits)
{
    const float m = __uint_as_float(bits);
    if (!(m > 0.0f) || !(m <= 3.0e38f)) return 1.0;
    int ex;
    (void)frexpf(m, &ex);  // m < 2^ex
    return ldexp(1.0, min(14 - ex, 900));
}
__device__ __forceinline__ float bfhi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }

// Complex element (row, col) of A1 = P1 | conj(P2) or A2^T = P2^T | conj(P1)^T
// of a stored pair, from the blocked store (BF16 with scale, or complex64).
template <bool C64>
__device__ __forceinline__ float2 a_elem(const void *rec, float sc, bool rev, int mat, int r, int c)
{
    const int src_mat = mat == 0 ? (rev ? 1 : 0) : (rev ? 0 : 1);
    const int row = mat == 0 ? r : c, col = mat == 0 ? c : r;
    float2 w;
    if (C64) {
        w = ((const float2 *)rec)[blk_off(src_mat, row, col, 2)];
    } else {
        const uint32_t u = ((const uint32_t *)rec)[blk_off(src_mat, row, col, 4)];
        w = make_float2(bflo(u) * sc, bfhi(u) * sc);
    }
    if (rev) w.y = -w.y;
    return w;
}

// Visit records: M = [A_ii | A_ij + A_ji | i (A_ij - A_ji)] real-ified (32 x 16)
// for A1 and A2^T, BF16 hi and lo (HALF: M x gsc as one FP16 word).  One
// 64-lane group per visit; lane t covers rows (t / 16) of the 4 row blocks and
// K column t % 16.
template <bool C64, bool HALF>
__global__ __launch_bounds__(64) void k_build_visits(int64_t nvis, const uint32_t *__restrict__ vsd,
                                                     const void *__restrict__ Pq, const float *__restrict__ pscale,
                                                     double gsc, uint16_t *__restrict__ vrec)
{
    static_assert(!(C64 && HALF), "FP16 records are built from the BF16 store");
    const int64_t v = blockIdx.x;
    if (v >= nvis) return;
    const int64_t s = vsd[v] >> 1;
    const bool rev = (vsd[v] & 1u) != 0;
    const void *rec = C64 ? (const void *)((const float2 *)Pq + s * 512) : (const void *)((const uint32_t *)Pq + s * 512);
    const float sc = C64 ? 1.0f : (float)((double)pscale[s] * gsc);
    uint16_t *out = vrec + v * (HALF ? 1024 : 2048);
    // K column p -> (i, j, kind): 0..3 diagonal, 4..9 Re upper, 10..15 Im upper
    const int pi[16] = { 0, 1, 2, 3, 0, 0, 0, 1, 1, 2, 0, 0, 0, 1, 1, 2 };
    const int pj[16] = { 0, 1, 2, 3, 1, 2, 3, 2, 3, 3, 1, 2, 3, 2, 3, 3 };
    for (int t = threadIdx.x; t < 2 * 16 * 16; t += 64) {
        const int mat = t >> 8, r = (t >> 4) & 15, p = t & 15;
        const int i = pi[p], j = pj[p];
        const float2 aij = a_elem<C64>(rec, sc, rev, mat, r, 4 * i + j);
        float2 m;
        if (p < 4) {
            m = aij;
        } else {
            const float2 aji = a_elem<C64>(rec, sc, rev, mat, r, 4 * j + i);
            m = p < 10 ? make_float2(aij.x + aji.x, aij.y + aji.y)      // A_ij + A_ji
                       : make_float2(-(aij.y - aji.y), aij.x - aji.x);  // i (A_ij - A_ji)
        }
        // real-ified rows: r (Re of output r) and 16 + r (Im of output r)
        const float vals[2] = { m.x, m.y };
        for (int part = 0; part < 2; part++) {
            const float x = vals[part];
            const int o = mat * 512 + (part * 16 + r) * 16 + p;
            if (HALF) {
                out[o] = f2h(x);
            } else {
                const uint32_t hi = f2bf(x), lo = f2bf(x - __uint_as_float(hi << 16));
                out[o] = (uint16_t)hi;
                out[1024 + o] = (uint16_t)lo;
            }
        }
    }
}

__global__ void k_zero_u32(int n, uint32_t *__restrict__ a)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) a[i] = 0u;
}

// Bits of an upper bound of max |x| of every trajectory (|rho_ij| for the
// off-diagonal terms, which the rotation keeps; |rho_ii - f_i| on the
// diagonal), for the FP16 operand scale.  Grid (k blocks of 256, B); one
// atomic max per block on the FP32 bits (non-negative floats order as their
// bits).
__global__ __launch_bounds__(256) void k_xmax(int64_t nk, const double2 *__restrict__ rho,
                                              const double *__restrict__ f, uint32_t *__restrict__ xmax)
{
    __shared__ float part[4];
    const int b = blockIdx.y;
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float m = 0.0f;
    if (k < nk) {
        const double2 *r = rho + ((int64_t)b * nk + k) * 16;
        double mx = 0.0;
#pragma unroll
        for (int l = 0; l < 16; l++) {
            const int i = l >> 2, j = l & 3;
            if (i > j) continue;
            const double2 w = r[l];
            const double v = i == j ? fabs(w.x - (f ? f[k * 4 + i] : 0.0)) : sqrt(w.x * w.x + w.y * w.y);
            mx = fmax(mx, v);
        }
        m = (float)(mx * (1.0 + 1.0 / 1048576));  // FP32 rounding must not fall below the bound
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
        if (m > 0.0f) atomicMax(xmax + b, __float_as_uint(m));
    }
}

// x = (Re drho_ii; Re drho_ij, Im drho_ij, i < j) of every trajectory as the
// hi/lo BF16 operand (with xmax: x s_b as FP16 hi/lo), and rho_S in FP64 for
// the assembly.  rho: B x nk x 16 (interaction picture).  One thread per
// (b, k, K column p).
__global__ void k_prepare_x(int64_t nk, int B, int Bp, const double2 *__restrict__ rho, const double *__restrict__ e,
                            const DClock *__restrict__ clk, double a, int expt, const double *__restrict__ f,
                            const uint32_t *__restrict__ xmax, double2 *__restrict__ rhoS, uint16_t *__restrict__ xb)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)Bp * nk * 16) return;
    const int64_t bk = idx >> 4;
    const int b = (int)(bk / nk);
    const int64_t k = bk - (int64_t)b * nk;
    const int l = (int)(idx & 15), i = l >> 2, j = l & 3;
    double x = 0.0;
    if (b < B) {
        const double t = clock_t_at(clk, a);
        double2 v = rho[idx];
        if (expt && i != j) v = cmul(v, ccis(-(e[k * 4 + i] - e[k * 4 + j]) * t));
        rhoS[idx] = v;
        // this thread's K column p = l: its (i, j, kind)
        const int pi[16] = { 0, 1, 2, 3, 0, 0, 0, 1, 1, 2, 0, 0, 0, 1, 1, 2 };
        const int pj[16] = { 0, 1, 2, 3, 1, 2, 3, 2, 3, 3, 1, 2, 3, 2, 3, 3 };
        const int ii = pi[l], jj = pj[l];
        double2 w = rho[bk * 16 + 4 * ii + jj];
        if (expt && ii != jj) w = cmul(w, ccis(-(e[k * 4 + ii] - e[k * 4 + jj]) * t));
        if (l < 4) x = w.x - (f ? f[k * 4 + ii] : 0.0);
        else x = l < 10 ? w.x : w.y;
    }
    uint32_t hi, lo;
    if (xmax) {
        const double xs = x * xscale_of(xmax[b]);
        hi = f2h((float)xs);
        lo = f2h((float)(xs - h2d((uint16_t)hi)));
    } else {
        hi = f2bf((float)x);
        lo = f2bf((float)(x - (double)__uint_as_float(hi << 16)));
    }
    const int half = l >> 3, el = l & 7;
    // [q][term][k-half][b][8]
    xb[(((k * 2 + 0) * 2 + half) * Bp + b) * 8 + el] = (uint16_t)hi;
    xb[(((k * 2 + 1) * 2 + half) * Bp + b) * 8 + el] = (uint16_t)lo;
}

__device__ __forceinline__ bf16x8 as_bf8(uint4 u)
{
    bf16x8 r;
    __builtin_memcpy(&r, &u, 16);
    return r;
}
__device__ __forceinline__ halfx8 as_h8(uint4 u)
{
    halfx8 r;
    __builtin_memcpy(&r, &u, 16);
    return r;
}

// Visits per segment at most (the autotuner's segment lengths are 64 to
// 1024; longer segments mean fewer row-sum flushes, whose FP32 atomics share
// the memory system with the record stream).
constexpr int kSegMax = 1024;

// One segment of the record pass (a wavefront): the row sums of its visits
// over NBLK blocks of 32 trajectories from b0, added into F at its end.  BF16
// hi + lo records (HALF = false): M_hi x_hi + M_hi x_lo + M_lo x_hi (the
// M_lo x_lo term is below the 16-bit significands), 6 MFMAs per visit per
// block; FP16 records (HALF): M x_hi + M x_lo, 4 MFMAs.  Timing variants that
// staged the segment's partner rows in LDS and kept two visits' records and
// operands in flight per wavefront ran no faster (the extra registers cost a
// wave per SIMD), so a visit's loads are issued when it starts.
// PK (B <= 16, Bp = 16, NBLK = 1): the 32 MFMA columns carry x_hi of the 16
// trajectories (columns 0-15) and x_lo (16-31), so one MFMA per product and
// record word forms both terms; the two column halves are added at the flush.
template <bool HALF, int NBLK, bool PK = false>
__device__ __forceinline__ void batched_segment(int4 S, int b0, int Bp, const uint16_t *__restrict__ vrec,
                                                const int32_t *__restrict__ vq, const uint16_t *__restrict__ xb,
                                                float *__restrict__ F)
{
    constexpr int NR = HALF ? 2 : 4;            // 16-byte record words per lane
    constexpr int64_t RS = HALF ? 1024 : 2048;  // record stride (16-bit words)
    const int lane = threadIdx.x & 63;
    const int64_t k = S.x;  // (k, v0 lo, v0 hi, count)
    const int64_t v0 = (int64_t)(uint32_t)S.y | ((int64_t)S.z << 32);
    const int nv = S.w;
    const int i = lane & 31, h = lane >> 5;  // A row (real-ified), K half
    floatx16 f1[NBLK], g2[NBLK];
#pragma unroll
    for (int nb = 0; nb < NBLK; nb++)
#pragma unroll
        for (int t = 0; t < 16; t++) { f1[nb][t] = 0.0f; g2[nb][t] = 0.0f; }
    // the lane's 8 K columns of row i in each 512-word matrix of a record
    const uint16_t *rb = vrec + v0 * RS + i * 16 + 8 * h;
    for (int u = 0; u < nv; u++) {
        uint4 m[NR];
#pragma unroll
        for (int w = 0; w < NR; w++) m[w] = ld_stream_u4(rb + u * RS + w * 512);
        const uint16_t *xq = xb + (int64_t)vq[v0 + u] * 4 * (int64_t)Bp * 8;
        if (PK) {
            const uint4 xv = *(const uint4 *)(xq + (((i >> 4) * 2 + h) * (int64_t)Bp + (i & 15)) * 8);
            if (HALF) {
                const halfx8 M1 = as_h8(m[0]), M2 = as_h8(m[1]), a = as_h8(xv);
                f1[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(M1, a, f1[0], 0, 0, 0);
                g2[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(M2, a, g2[0], 0, 0, 0);
            } else {  // M_hi (x_hi | x_lo) + M_lo (x_hi | 0)
                const bf16x8 M1h = as_bf8(m[0]), M2h = as_bf8(m[1]), M1l = as_bf8(m[2]), M2l = as_bf8(m[3]);
                const bf16x8 a = as_bf8(xv), az = as_bf8(i < 16 ? xv : make_uint4(0u, 0u, 0u, 0u));
                f1[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M1h, a, f1[0], 0, 0, 0);
                g2[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M2h, a, g2[0], 0, 0, 0);
                f1[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M1l, az, f1[0], 0, 0, 0);
                g2[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M2l, az, g2[0], 0, 0, 0);
            }
            continue;
        }
#pragma unroll
        for (int nb = 0; nb < NBLK; nb++) {
            const int b = b0 + nb * 32 + i;
            const uint4 xh = *(const uint4 *)(xq + ((0 * 2 + h) * (int64_t)Bp + b) * 8);
            const uint4 xl = *(const uint4 *)(xq + ((1 * 2 + h) * (int64_t)Bp + b) * 8);
            if (HALF) {
                const halfx8 M1 = as_h8(m[0]), M2 = as_h8(m[1]), ah = as_h8(xh), al = as_h8(xl);
                f1[nb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(M1, ah, f1[nb], 0, 0, 0);
                g2[nb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(M2, ah, g2[nb], 0, 0, 0);
                f1[nb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(M1, al, f1[nb], 0, 0, 0);
                g2[nb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(M2, al, g2[nb], 0, 0, 0);
            } else {
                const bf16x8 M1h = as_bf8(m[0]), M2h = as_bf8(m[1]), M1l = as_bf8(m[2]), M2l = as_bf8(m[3]);
                const bf16x8 ah = as_bf8(xh), al = as_bf8(xl);
                f1[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M1h, ah, f1[nb], 0, 0, 0);
                g2[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M2h, ah, g2[nb], 0, 0, 0);
                f1[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M1h, al, f1[nb], 0, 0, 0);
                g2[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M2h, al, g2[nb], 0, 0, 0);
                f1[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M1l, ah, f1[nb], 0, 0, 0);
                g2[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M2l, ah, g2[nb], 0, 0, 0);
            }
        }
    }
    // D layout: element t of lane l is row 8 (t / 4) + 4 h + t % 4, column (trajectory) l % 32
    if (PK) {  // column i + 16 (x_lo) of the same row sits in lane l ^ 16
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const int row = 8 * (t >> 2) + 4 * h + (t & 3);
            const float a = f1[0][t] + __shfl_xor(f1[0][t], 16), c = g2[0][t] + __shfl_xor(g2[0][t], 16);
            if (i < 16) {
                unsafeAtomicAdd(F + (k * 64 + row) * Bp + i, a);
                unsafeAtomicAdd(F + (k * 64 + 32 + row) * Bp + i, c);
            }
        }
        return;
    }
#pragma unroll
    for (int nb = 0; nb < NBLK; nb++) {
        const int b = b0 + nb * 32 + i;
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const int row = 8 * (t >> 2) + 4 * h + (t & 3);
            unsafeAtomicAdd(F + (k * 64 + row) * Bp + b, f1[nb][t]);
            unsafeAtomicAdd(F + (k * 64 + 32 + row) * Bp + b, g2[nb][t]);
        }
    }
}

// One wavefront per segment (the grid's y index selects the trajectory pass).
template <bool HALF, int NBLK, bool PK = false>
__global__ __launch_bounds__(256) void k_batched(int64_t nseg, int Bp, const int4 *__restrict__ seg,
                                                 const uint16_t *__restrict__ vrec, const int32_t *__restrict__ vq,
                                                 const uint16_t *__restrict__ xb, float *__restrict__ F)
{
    const int64_t sg = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (sg >= nseg) return;
    batched_segment<HALF, NBLK, PK>(seg[sg], blockIdx.y * NBLK * 32, Bp, vrec, vq, xb, F);
}

// Persistent form: as many wavefronts as the GPU holds at once, each taking
// the next segment from a queue counter (one per trajectory pass, zero at
// launch) until the segments run out, so long and short rows balance over
// the wavefronts that are resident.  Every wavefront leaves once the counter
// passes nseg.
template <bool HALF, int NBLK, bool PK = false>
__global__ __launch_bounds__(256) void k_batched_q(int64_t nseg, int Bp, const int4 *__restrict__ seg,
                                                   const uint16_t *__restrict__ vrec, const int32_t *__restrict__ vq,
                                                   const uint16_t *__restrict__ xb, float *__restrict__ F,
                                                   uint32_t *__restrict__ queue)
{
    const int lane = threadIdx.x & 63;
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(queue + blockIdx.y, 1u);
        const int64_t sg = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane(t);
        if (sg >= nseg) return;
        batched_segment<HALF, NBLK, PK>(seg[sg], blockIdx.y * NBLK * 32, Bp, vrec, vq, xb, F);
    }
}

// The same pass without visit records, for operators whose records do not
// fit in device memory: the folded M of each visit is formed from the pair's
// stored P (the blocked record of the tile store: BF16 with its scale, or
// complex64) staged through LDS, so a visit streams the stored P (2 KiB per
// visit for BF16, 4 KiB for complex64) and no 4 KiB-per-visit record array is
// kept.  The forward visit reads rows of P1 and columns of P2, the reverse one
// rows of P2 and columns of P1 (conjugated); LDS serves both.  M (scale,
// sums, hi/lo split) is formed as k_build_visits forms it, so every visit's
// products equal the record pass's; only the order of the row-sum atomics
// differs.  vsd: slot << 1 | reverse.
template <bool C64, int NBLK>
__global__ __launch_bounds__(256) void k_batched_p(int64_t nseg, int Bp, const int4 *__restrict__ seg,
                                                   const void *__restrict__ Pq, const float *__restrict__ pscale,
                                                   const uint32_t *__restrict__ vsd, const int32_t *__restrict__ vq,
                                                   const uint16_t *__restrict__ xb, float *__restrict__ F)
{
    constexpr int NCH = C64 ? 4 : 2;  // 16-byte chunks of a record per lane
    __shared__ uint4 stage[4][64 * NCH];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t sg = (int64_t)blockIdx.x * 4 + wv;
    if (sg >= nseg) return;
    const int4 S = seg[sg];
    const int64_t k = S.x;
    const int64_t v0 = (int64_t)(uint32_t)S.y | ((int64_t)S.z << 32);
    const int nv = S.w;
    const int i = lane & 31, h = lane >> 5;
    const int r = i & 15, part = i >> 4;
    const int b0 = blockIdx.y * NBLK * 32;
    floatx16 f1[NBLK], g2[NBLK];
#pragma unroll
    for (int nb = 0; nb < NBLK; nb++)
#pragma unroll
        for (int t = 0; t < 16; t++) { f1[nb][t] = 0.0f; g2[nb][t] = 0.0f; }
    uint4 *buf = stage[wv];
    auto rec_of = [&](uint32_t sd) {
        const int64_t s = sd >> 1;
        return C64 ? (const uint4 *)((const float2 *)Pq + s * 512) : (const uint4 *)((const uint32_t *)Pq + s * 512);
    };
    uint32_t sd_next = vsd[v0];
    uint4 nxt[NCH];
    {
        const uint4 *rec = rec_of(sd_next);
#pragma unroll
        for (int q = 0; q < NCH; q++) nxt[q] = ld_stream_u4(rec + q * 64 + lane);
    }
    for (int u = 0; u < nv; u++) {
        const int64_t v = v0 + u;
        const uint32_t sd = sd_next;
        uint4 cur[NCH];
#pragma unroll
        for (int q = 0; q < NCH; q++) cur[q] = nxt[q];
        if (u + 1 < nv) {
            sd_next = vsd[v + 1];
            const uint4 *rec = rec_of(sd_next);
#pragma unroll
            for (int q = 0; q < NCH; q++) nxt[q] = ld_stream_u4(rec + q * 64 + lane);
        }
        __builtin_amdgcn_wave_barrier();  // the previous visit's reads of the stage are done
#pragma unroll
        for (int q = 0; q < NCH; q++) buf[q * 64 + lane] = cur[q];
        __builtin_amdgcn_wave_barrier();
        const bool rev = sd & 1u;
        const float sc = C64 ? 1.0f : pscale[sd >> 1];
        // complex element (row, col) of A1 = P1 | conj(P2) (mat 0) or A2^T = P2^T | conj(P1)^T (mat 1)
        auto a_el = [&](int mat, int row, int col) {
            const int src = mat == 0 ? (rev ? 1 : 0) : (rev ? 0 : 1);
            const int rr = mat == 0 ? row : col, cc = mat == 0 ? col : row;
            float2 w;
            if (C64) {
                w = ((const float2 *)buf)[blk_off(src, rr, cc, 2)];
            } else {
                const uint32_t x = ((const uint32_t *)buf)[blk_off(src, rr, cc, 4)];
                w = make_float2(bflo(x) * sc, bfhi(x) * sc);
            }
            if (rev) w.y = -w.y;
            return w;
        };
        uint32_t mh[2][4], ml[2][4];  // the lane's 8 K columns of row i, BF16 pairs
#pragma unroll
        for (int mat = 0; mat < 2; mat++)
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const int p = 8 * h + t;
                // K column p -> (ii, jj): 0..3 diagonal, 4..9 Re upper, 10..15 Im upper
                const int ii = p < 4 ? p : (p < 10 ? (p < 7 ? 0 : (p < 9 ? 1 : 2)) : (p < 13 ? 0 : (p < 15 ? 1 : 2)));
                const int jj = p < 4 ? p : (p < 10 ? (p < 7 ? p - 3 : (p < 9 ? p - 5 : 3)) : (p < 13 ? p - 9 : (p < 15 ? p - 11 : 3)));
                const float2 aij = a_el(mat, r, 4 * ii + jj);
                float2 m;
                if (p < 4) {
                    m = aij;
                } else {
                    const float2 aji = a_el(mat, r, 4 * jj + ii);
                    m = p < 10 ? make_float2(aij.x + aji.x, aij.y + aji.y)
                               : make_float2(-(aij.y - aji.y), aij.x - aji.x);
                }
                const float x = part ? m.y : m.x;
                const uint32_t hi = f2bf(x), lo = f2bf(x - __uint_as_float(hi << 16));
                if (t & 1) {
                    mh[mat][t >> 1] |= hi << 16;
                    ml[mat][t >> 1] |= lo << 16;
                } else {
                    mh[mat][t >> 1] = hi;
                    ml[mat][t >> 1] = lo;
                }
            }
        const bf16x8 M1h = as_bf8(make_uint4(mh[0][0], mh[0][1], mh[0][2], mh[0][3]));
        const bf16x8 M2h = as_bf8(make_uint4(mh[1][0], mh[1][1], mh[1][2], mh[1][3]));
        const bf16x8 M1l = as_bf8(make_uint4(ml[0][0], ml[0][1], ml[0][2], ml[0][3]));
        const bf16x8 M2l = as_bf8(make_uint4(ml[1][0], ml[1][1], ml[1][2], ml[1][3]));
        const uint16_t *xq = xb + (int64_t)vq[v] * 4 * (int64_t)Bp * 8;
#pragma unroll
        for (int nb = 0; nb < NBLK; nb++) {
            const int b = b0 + nb * 32 + i;
            const bf16x8 xh = as_bf8(*(const uint4 *)(xq + ((0 * 2 + h) * (int64_t)Bp + b) * 8));
            const bf16x8 xl = as_bf8(*(const uint4 *)(xq + ((1 * 2 + h) * (int64_t)Bp + b) * 8));
            f1[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M1h, xh, f1[nb], 0, 0, 0);
            g2[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M2h, xh, g2[nb], 0, 0, 0);
            f1[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M1h, xl, f1[nb], 0, 0, 0);
            g2[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M2h, xl, g2[nb], 0, 0, 0);
            f1[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M1l, xh, f1[nb], 0, 0, 0);
            g2[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(M2l, xh, g2[nb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int nb = 0; nb < NBLK; nb++) {
        const int b = b0 + nb * 32 + i;
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const int row = 8 * (t >> 2) + 4 * h + (t & 3);
            unsafeAtomicAdd(F + (k * 64 + row) * Bp + b, f1[nb][t]);
            unsafeAtomicAdd(F + (k * 64 + 32 + row) * Bp + b, g2[nb][t]);
        }
    }
}

// ddm of every trajectory: C = (1-rho)(F1_eq + F1) - (F2_eq - G) rho, ddm =
// prefac/2 (C + C^+), rotated back, minus ddmdt_eq.  Clears F.  One thread
// per (k, b) with b fastest, so each of the 64 row sums of k is one coalesced
// row across the trajectories; the thread owns its 64 sums, reads and clears
// them, and writes its 16 outputs.
__global__ __launch_bounds__(256) void k_assemble_b(int64_t nk, int B, int Bp, float *__restrict__ F,
                                                    const uint32_t *__restrict__ xmax, double gsc,
                                                    const double *__restrict__ feq, const double2 *__restrict__ rhoS,
                                                    const double *__restrict__ e, const DClock *__restrict__ clk,
                                                    double a, int expt, int schrodinger, double prefac,
                                                    const double2 *__restrict__ eq, double2 *__restrict__ out)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t k = idx / Bp;
    const int b = (int)(idx - k * Bp);
    if (k >= nk) return;
    float *Fk = F + k * 64 * Bp + b;
    if (b < B) {
        const int64_t bk = (int64_t)b * nk + k;
        const double *fe = feq + k * 64;
        const double fs = xmax ? 1.0 / (gsc * xscale_of(xmax[b])) : 1.0;  // FP16 operands: undo G s_b
        double2 rk[16];
#pragma unroll
        for (int q = 0; q < 16; q++) rk[q] = rhoS[bk * 16 + q];
        // C = (1 - rho) F1 - F2 rho; 1 - rho and rho read as Hermitian from
        // their upper triangles, as zhemm does (evolve_eph.cpp:142-143)
        double2 C[16];
#pragma unroll
        for (int q = 0; q < 16; q++) C[q] = make_double2(0, 0);
#pragma unroll
        for (int m = 0; m < 4; m++) {
#pragma unroll
            for (int jj = 0; jj < 4; jj++) {
                const int r = 4 * m + jj;  // F1[m][jj] = F1_eq + F
                const double2 f1 = make_double2(fe[2 * r] + fs * (double)Fk[r * Bp],
                                                fe[2 * r + 1] + fs * (double)Fk[(16 + r) * Bp]);
#pragma unroll
                for (int ii = 0; ii < 4; ii++) {
                    double2 r1;
                    if (ii < m) r1 = make_double2(-rk[ii * 4 + m].x, -rk[ii * 4 + m].y);
                    else if (ii > m) r1 = make_double2(-rk[m * 4 + ii].x, rk[m * 4 + ii].y);
                    else r1 = make_double2(1.0 - rk[ii * 5].x, 0.0);
                    cfma(C[ii * 4 + jj], r1, f1);
                }
            }
        }
#pragma unroll
        for (int ii = 0; ii < 4; ii++) {
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const int r = 4 * ii + m;  // F2[ii][m] = F2_eq - G
                const double2 g = make_double2(fe[32 + 2 * r] - fs * (double)Fk[(32 + r) * Bp],
                                               fe[32 + 2 * r + 1] - fs * (double)Fk[(48 + r) * Bp]);
                const double2 ng = make_double2(-g.x, -g.y);
#pragma unroll
                for (int jj = 0; jj < 4; jj++) {
                    double2 rr;
                    if (m < jj) rr = rk[m * 4 + jj];
                    else if (m > jj) rr = cconj(rk[jj * 4 + m]);
                    else rr = make_double2(rk[m * 5].x, 0.0);
                    cfma(C[ii * 4 + jj], ng, rr);
                }
            }
        }
        const double t = clock_t_at(clk, a);
#pragma unroll
        for (int l = 0; l < 16; l++) {
            const int i = l >> 2, j = l & 3;
            double2 d = cadd(C[l], cconj(C[j * 4 + i]));
            d = make_double2(d.x * (prefac * 0.5), d.y * (prefac * 0.5));
            const double arg = (e[k * 4 + i] - e[k * 4 + j]) * t;
            if (expt && i != j) d = cmul(d, ccis(arg));
            if (eq) {
                double2 q = eq[k * 16 + l];
                if (i != j && !schrodinger) q = cmul(q, ccis(arg));
                d = csub(d, q);
            }
            out[bk * 16 + l] = d;
        }
    }
    // this thread's 64 row sums are read; clear them for the next pass
#pragma unroll
    for (int r = 0; r < 64; r++) Fk[r * Bp] = 0.0f;
}

// max of non-negative floats as bits (one atomic per block)
__global__ __launch_bounds__(256) void k_umax_f(int64_t n, const float *__restrict__ a, uint32_t *__restrict__ out)
{
    __shared__ float part[4];
    float m = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) m = fmaxf(m, a[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
        if (m > 0.0f) atomicMax(out, __float_as_uint(m));
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

// G = 2^14 / (power of two above the largest per-pair scale of the BF16
// store), so that every FP16 record entry |M| <= 2 sc stays under 2^15.
static int half_record_factor(dmd_scatt *h, double *g)
{
    uint32_t *d = nullptr, bits = 0;
    int rc;
    if ((rc = dalloc_t(&d, 1, "scale max"))) return rc;
    hipError_t e = hipMemsetAsync(d, 0, 4, h->st);
    if (e == hipSuccess && h->nslots > 0) {
        hipLaunchKernelGGL(k_umax_f, dim3(1024), dim3(256), 0, h->st, h->nslots, (const float *)h->d_pscale, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&bits, d, 4, hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    dfree(d);
    if (e != hipSuccess) return fail(DMD_ERR_HIP, std::string("record scale: ") + hipGetErrorString(e));
    float m;
    std::memcpy(&m, &bits, 4);
    int ex = 0;
    if (m > 0.0f && std::isfinite(m)) (void)std::frexp((double)m, &ex);  // m < 2^ex
    *g = std::ldexp(1.0, 14 - ex);
    return DMD_OK;
}

// The visits' partner rows and slots on the device, and their records built
// there from the stored P (when records are kept).
static int upload_visits(dmd_scatt *h, const std::vector<int32_t> &vq, const std::vector<uint32_t> &vsd)
{
    const int64_t nvis = (int64_t)vq.size();
    if (!nvis) return DMD_OK;
    DMD_HIP(hipStreamSynchronize(h->st));  // no pass may still read the arrays
    DMD_HIP(hipMemcpyAsync(h->d_vq, vq.data(), nvis * 4, hipMemcpyHostToDevice, h->st));
    DMD_HIP(hipMemcpyAsync(h->d_vsd, vsd.data(), nvis * 4, hipMemcpyHostToDevice, h->st));
    if (!h->bdirect) {
        const dim3 g((unsigned)nvis);
        if (h->precision == DMD_PREC_C64)
            hipLaunchKernelGGL((k_build_visits<true, false>), g, dim3(64), 0, h->st, nvis, h->d_vsd,
                               (const void *)h->d_P, h->d_pscale, 1.0, h->d_vrec);
        else if (h->bhalf)
            hipLaunchKernelGGL((k_build_visits<false, true>), g, dim3(64), 0, h->st, nvis, h->d_vsd,
                               (const void *)h->d_P, h->d_pscale, h->bhalf_g, h->d_vrec);
        else
            hipLaunchKernelGGL((k_build_visits<false, false>), g, dim3(64), 0, h->st, nvis, h->d_vsd,
                               (const void *)h->d_P, h->d_pscale, 1.0, h->d_vrec);
        DMD_HIP(hipGetLastError());
    }
    DMD_HIP(hipStreamSynchronize(h->st));  // the host arrays may go now
    return DMD_OK;
}

// visit records, segments (built once per P; rebuilt when P changes)
static int build_visits(dmd_scatt *h)
{
    if (h->bvis_ready && !h->bvis_dirty) return DMD_OK;
    const int64_t nk = h->nk;
    std::vector<int64_t> rowptr(nk + 1), revptr(nk + 1);
    DMD_HIP(hipMemcpy(rowptr.data(), h->d_rowptr, (nk + 1) * 8, hipMemcpyDeviceToHost));
    DMD_HIP(hipMemcpy(revptr.data(), h->d_revptr, (nk + 1) * 8, hipMemcpyDeviceToHost));
    std::vector<int64_t> fwdpos(h->np), revpos(h->nrev);
    std::vector<int32_t> col(h->np), revcol(h->nrev);
    if (h->np) {
        DMD_HIP(hipMemcpy(fwdpos.data(), h->d_fwdpos, h->np * 8, hipMemcpyDeviceToHost));
        DMD_HIP(hipMemcpy(col.data(), h->d_col, h->np * 4, hipMemcpyDeviceToHost));
    }
    if (h->nrev) {
        DMD_HIP(hipMemcpy(revpos.data(), h->d_revpos, h->nrev * 8, hipMemcpyDeviceToHost));
        DMD_HIP(hipMemcpy(revcol.data(), h->d_revcol, h->nrev * 4, hipMemcpyDeviceToHost));
    }
    if (h->nslots >= (1ll << 31)) return fail(DMD_ERR_UNSUPPORTED, "batched: 2^31 or more stored slots");
    const int64_t nvis = h->np + h->nrev;
    std::vector<uint32_t> vsd(nvis);  // slot << 1 | reverse
    std::vector<int32_t> vq(nvis);
    int64_t v = 0;
    h->h_vptr.assign(nk + 1, 0);
    std::vector<int64_t> perm;
    std::vector<uint32_t> s2;
    std::vector<int32_t> q2;
    for (int64_t k = 0; k < nk; k++) {
        h->h_vptr[k] = v;
        const int64_t v0 = v;
        for (int64_t p = rowptr[k]; p < rowptr[k + 1]; p++, v++) { vsd[v] = (uint32_t)(fwdpos[p] << 1); vq[v] = col[p]; }
        for (int64_t p = revptr[k]; p < revptr[k + 1]; p++, v++) { vsd[v] = (uint32_t)(revpos[p] << 1) | 1u; vq[v] = revcol[p]; }
        // a row's visits by partner, so that a partner-block segment order is contiguous
        perm.resize(v - v0);
        for (int64_t q = 0; q < v - v0; q++) perm[q] = v0 + q;
        std::stable_sort(perm.begin(), perm.end(), [&](int64_t x, int64_t y) { return vq[x] < vq[y]; });
        s2.resize(perm.size());
        q2.resize(perm.size());
        for (size_t q = 0; q < perm.size(); q++) { s2[q] = vsd[perm[q]]; q2[q] = vq[perm[q]]; }
        for (size_t q = 0; q < perm.size(); q++) { vsd[v0 + q] = s2[q]; vq[v0 + q] = q2[q]; }
    }
    h->h_vptr[nk] = v;
    int rc;
    dfree(h->d_vrec); dfree(h->d_vq); dfree(h->d_vsd); dfree(h->d_seg);
    h->d_vrec = nullptr; h->d_vq = nullptr; h->d_vsd = nullptr; h->d_seg = nullptr;
    h->nseg = 0;
    const bool c64 = h->precision == DMD_PREC_C64;
    // Visit records (4 KiB per visit; 2 KiB as FP16) when they fit in device
    // memory: forming M per visit from the stored P reads half the bytes for
    // BF16 P but costs more instructions than it saves (BF16, preset 4: 35
    // against 25 ms per evaluation at B = 16 with 4 KiB records); without room
    // for the records the pass forms M from P.  DMD_BATCH_RECORDS=0 forces
    // the record-free pass (tests).
    const char *renv = std::getenv("DMD_BATCH_RECORDS");
    h->bdirect = renv && renv[0] == '0';
    // The BF16 store's records are one FP16 word per entry (2 KiB per visit);
    // DMD_BATCH_HALF=0 keeps BF16 hi + lo records for it (tests, timing).
    const char *henv = std::getenv("DMD_BATCH_HALF");
    const bool want_half = !c64 && !(henv && henv[0] == '0');
    h->bhalf = false;
    h->bhalf_g = 1.0;
    if (!h->bdirect && (rc = dalloc_t(&h->d_vrec, (size_t)nvis * (want_half ? 1024 : 2048), "visit records"))) {
        if (rc != DMD_ERR_NOMEM) return rc;
        (void)hipGetLastError();  // the failed allocation: continue without records
        h->bdirect = true;
    }
    if (!h->bdirect && want_half) {
        if ((rc = half_record_factor(h, &h->bhalf_g))) return rc;
        h->bhalf = true;
    }
    if ((rc = dalloc_t(&h->d_vq, (size_t)std::max<int64_t>(nvis, 1), "visit partners"))) return rc;
    if ((rc = dalloc_t(&h->d_vsd, (size_t)std::max<int64_t>(nvis, 1), "visit slots"))) return rc;
    if ((rc = upload_visits(h, vq, vsd))) return rc;
    h->h_vq = std::move(vq);
    h->bvis_ready = true;
    h->bvis_dirty = false;
    h->seg_len = 0;
    h->seg_pblock = 0;
    h->tuned_Bp = 0;
    return DMD_OK;
}

// Segments of at most seg_len visits of one row, longest first.  With
// pblock > 0 a segment also keeps to one block of pblock partner rows, and
// the segments run block after block, so the partner operands of the running
// wavefronts share the L2 (a block of X at B trajectories is 64 B x B per row).
static int build_segments(dmd_scatt *h, int seg_len, int pblock)
{
    if (seg_len < 1 || seg_len > kSegMax) return fail(DMD_ERR_ARG, "batched: segment length outside 1 .. 1024");
    if (h->seg_len == seg_len && h->seg_pblock == pblock && h->d_seg) return DMD_OK;
    std::vector<int4> seg;
    std::vector<int32_t> blk;
    for (int64_t k = 0; k < h->nk; k++) {
        const int64_t e = h->h_vptr[k + 1];
        for (int64_t s = h->h_vptr[k]; s < e;) {
            int64_t t = s + 1;
            const int32_t b = pblock > 0 ? h->h_vq[s] / pblock : 0;
            while (t < e && t - s < seg_len && (pblock <= 0 || h->h_vq[t] / pblock == b)) t++;
            seg.push_back(make_int4((int)k, (int)(uint32_t)(s & 0xFFFFFFFF), (int)(s >> 32), (int)(t - s)));
            blk.push_back(b);
            s = t;
        }
    }
    std::vector<int64_t> order(seg.size());
    for (size_t q = 0; q < order.size(); q++) order[q] = (int64_t)q;
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        return blk[a] != blk[b] ? blk[a] < blk[b] : seg[a].w > seg[b].w;
    });
    {
        std::vector<int4> sorted(seg.size());
        for (size_t q = 0; q < order.size(); q++) sorted[q] = seg[order[q]];
        seg.swap(sorted);
    }
    dfree(h->d_seg);
    h->d_seg = nullptr;
    int rc;
    if ((rc = dalloc_t(&h->d_seg, std::max<size_t>(seg.size(), 1), "segments"))) return rc;
    if (!seg.empty()) DMD_HIP(hipMemcpy(h->d_seg, seg.data(), seg.size() * sizeof(int4), hipMemcpyHostToDevice));
    h->nseg = (int64_t)seg.size();
    h->seg_len = seg_len;
    h->seg_pblock = pblock;
    return DMD_OK;
}

static void launch_batched_core(dmd_scatt *h, int Bp, int nblk_req);
static void persistent_residency(dmd_scatt *h);

// Autotune: segment length x trajectory blocks per wavefront, by timing the
// matrix-core pass on the current operands; the row sums are cleared after.
static int autotune(dmd_scatt *h, int Bp)
{
    const char *env = std::getenv("DMD_BATCH_TUNE");
    const bool tune = !(env && env[0] == '0');
    // DMD_BATCH_PBLOCK=n fixes the partner-block size of the segment order (0: row order)
    const char *penv = std::getenv("DMD_BATCH_PBLOCK");
    const int forced_pb = penv && *penv ? std::max(0, std::atoi(penv)) : -1;
    int best_seg = 128, best_nb = (Bp % 64 == 0) ? 2 : 1, best_pb = forced_pb > 0 ? forced_pb : 0;
    auto trial = [&](int nb, float *ms) -> int {
        launch_batched_core(h, Bp, nb);  // warm
        DMD_HIP(hipEventRecord(h->ev0, h->st));
        for (int r = 0; r < 2; r++) launch_batched_core(h, Bp, nb);
        DMD_HIP(hipEventRecord(h->ev1, h->st));
        DMD_HIP(hipEventSynchronize(h->ev1));
        DMD_HIP(hipEventElapsedTime(ms, h->ev0, h->ev1));
        return DMD_OK;
    };
    if (tune && h->np > 0) {
        float best = 1e30f, ms = 0;
        int rc;
        for (int sl : { 64, 128, 256, 512, 1024 }) {
            if ((rc = build_segments(h, sl, best_pb))) return rc;
            for (int nb : { 1, 2 }) {
                if (nb == 2 && Bp % 64) continue;
                if ((rc = trial(nb, &ms))) return rc;
                if (ms < best) { best = ms; best_seg = sl; best_nb = nb; }
            }
        }
        // partner blocks that fit the L2 of one XCD (4 MiB), and 2x .. 16x that
        if (forced_pb < 0) {
            const int p0 = std::max(256, 65536 / Bp);
            for (int pb : { p0, 2 * p0, 4 * p0, 8 * p0, 16 * p0 }) {
                if ((rc = build_segments(h, best_seg, pb))) return rc;
                if ((rc = trial(best_nb, &ms))) return rc;
                if (ms < best) { best = ms; best_pb = pb; }
            }
        }
        // the segment length again within the chosen partner blocks (blocks cut
        // segments, so the best length can differ from the row order's)
        if (best_pb > 0) {
            const int seg0 = best_seg;
            for (int sl : { 64, 128, 256, 512, 1024 }) {
                if (sl == seg0) continue;
                if ((rc = build_segments(h, sl, best_pb))) return rc;
                if ((rc = trial(best_nb, &ms))) return rc;
                if (ms < best) { best = ms; best_seg = sl; }
            }
        }
    }
    int rc = build_segments(h, best_seg, best_pb);
    if (rc) return rc;
    h->batch_nblk = best_nb;
    h->tuned_Bp = Bp;
    DMD_HIP(hipMemsetAsync(h->d_Fb, 0, (size_t)h->nk * 64 * h->b_cap * sizeof(float), h->st));
    DMD_HIP(hipStreamSynchronize(h->st));
    return DMD_OK;
}

// The padded trajectory count: 16 (the packed pass) for B <= 16 with visit
// records, else B rounded up to 32 (DMD_BATCH_PACK=0: always 32-wide).
static int padded_B(const dmd_scatt *h, int B)
{
    const char *env = std::getenv("DMD_BATCH_PACK");
    if (B <= 16 && !h->bdirect && !(env && env[0] == '0')) return 16;
    return (B + 31) / 32 * 32;
}

int batched_prepare(dmd_scatt *h, int B)
{
    int rc;
    if ((rc = tiles4_prepare(h))) return rc;
    if ((rc = build_visits(h))) return rc;
    if (!h->bq_resident[0]) persistent_residency(h);
    const int Bp = padded_B(h, B);
    if (Bp > h->b_cap) {
        dfree(h->d_xb); dfree(h->d_Fb); dfree(h->d_rhoSB); dfree(h->d_bxmax); dfree(h->d_bqueue);
        h->d_xb = nullptr; h->d_Fb = nullptr; h->d_rhoSB = nullptr; h->d_bxmax = nullptr; h->d_bqueue = nullptr;
        h->b_cap = 0;
        if ((rc = dalloc_t(&h->d_bxmax, (size_t)Bp, "batched operand scales"))) return rc;
        if ((rc = dalloc_t(&h->d_bqueue, (size_t)std::max(1, Bp / 32), "segment queues"))) return rc;
        if ((rc = dalloc_t(&h->d_xb, (size_t)h->nk * 32 * Bp, "batched X"))) return rc;
        if ((rc = dalloc_t(&h->d_Fb, (size_t)h->nk * 64 * Bp, "batched row sums"))) return rc;
        if ((rc = dalloc_t(&h->d_rhoSB, (size_t)h->nk * 16 * Bp, "batched rho_S"))) return rc;
        DMD_HIP(hipMemset(h->d_Fb, 0, (size_t)h->nk * 64 * Bp * sizeof(float)));
        DMD_HIP(hipMemset(h->d_xb, 0, (size_t)h->nk * 32 * Bp * sizeof(uint16_t)));
        h->b_cap = Bp;
    }
    if (h->tuned_Bp != Bp && (rc = autotune(h, Bp))) return rc;
    return DMD_OK;
}

// the matrix-core pass alone (X prepared, F accumulated)
static void launch_batched_core(dmd_scatt *h, int Bp, int nblk_req)
{
    const bool pk = Bp == 16;
    const int nb2 = !pk && (Bp / 32) % 2 == 0 && nblk_req == 2 ? 2 : 1;
    dim3 g(nblk(h->nseg, 4), std::max(1, Bp / (32 * nb2)));
    if (h->bdirect) {
        const bool c64 = h->precision == DMD_PREC_C64;
#define DMD_BP(C, N)                                                                                              \
    hipLaunchKernelGGL((k_batched_p<C, N>), g, dim3(256), 0, h->st, h->nseg, Bp, h->d_seg, (const void *)h->d_P,  \
                       h->d_pscale, h->d_vsd, h->d_vq, (const uint16_t *)h->d_xb, h->d_Fb)
        if (c64 && nb2 == 2) DMD_BP(true, 2);
        else if (c64) DMD_BP(true, 1);
        else if (nb2 == 2) DMD_BP(false, 2);
        else DMD_BP(false, 1);
#undef DMD_BP
        return;
    }
    // DMD_BATCH_PERSIST=0: one wavefront per segment instead of the persistent pass
    const char *penv = std::getenv("DMD_BATCH_PERSIST");
    const bool persist = !(penv && penv[0] == '0') && h->nseg < (1ll << 31);
    if (persist) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, h->st, (int)g.y, h->d_bqueue);
    auto launch = [&](auto kq, auto kg, int resident) {
        if (!persist || resident <= 0) {
            hipLaunchKernelGGL(kg, g, dim3(256), 0, h->st, h->nseg, Bp, h->d_seg, h->d_vrec, h->d_vq,
                               (const uint16_t *)h->d_xb, h->d_Fb);
            return;
        }
        const dim3 gq((unsigned)std::min<int64_t>(g.x, resident), g.y);
        hipLaunchKernelGGL(kq, gq, dim3(256), 0, h->st, h->nseg, Bp, h->d_seg, h->d_vrec, h->d_vq,
                           (const uint16_t *)h->d_xb, h->d_Fb, h->d_bqueue);
    };
    const int *res = h->bq_resident;
    if (pk && h->bhalf) launch(k_batched_q<true, 1, true>, k_batched<true, 1, true>, res[5]);
    else if (pk) launch(k_batched_q<false, 1, true>, k_batched<false, 1, true>, res[4]);
    else if (h->bhalf && nb2 == 2) launch(k_batched_q<true, 2>, k_batched<true, 2>, res[3]);
    else if (h->bhalf) launch(k_batched_q<true, 1>, k_batched<true, 1>, res[2]);
    else if (nb2 == 2) launch(k_batched_q<false, 2>, k_batched<false, 2>, res[1]);
    else launch(k_batched_q<false, 1>, k_batched<false, 1>, res[0]);
}

// Workgroups of each persistent pass instance the GPU holds at once (set up
// outside any stream capture; 0 leaves that instance on the per-segment grid).
static void persistent_residency(dmd_scatt *h)
{
    int cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 0;
    auto one = [&](auto kq) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kq, 256, 0) != hipSuccess) per_cu = 0;
        return cus > 0 && per_cu > 0 ? per_cu * cus : 0;
    };
    h->bq_resident[0] = one(k_batched_q<false, 1>);
    h->bq_resident[1] = one(k_batched_q<false, 2>);
    h->bq_resident[2] = one(k_batched_q<true, 1>);
    h->bq_resident[3] = one(k_batched_q<true, 2>);
    h->bq_resident[4] = one(k_batched_q<false, 1, true>);
    h->bq_resident[5] = one(k_batched_q<true, 1, true>);
    (void)hipGetLastError();
}

// ddm of B trajectories (rho: B x nk x 16 device, interaction picture at the
// clock time a) into out (B x nk x 16).  The X and row-sum buffers are sized
// for the padded B the last batched_prepare saw.
int launch_batched(dmd_scatt *h, int B, const double2 *rho, double a, double2 *out)
{
    const int Bp = padded_B(h, B);
    if (Bp > h->b_cap) return fail(DMD_ERR_STATE, "batched buffers too small");
    if (h->has_dP)  // built from one state at a rebuild: not a term the trajectories share
        return fail(DMD_ERR_UNSUPPORTED, "the batched forms do not carry the linearize_dPee term");
    const int64_t nx = (int64_t)Bp * h->nk * 16;
    const uint32_t *xmax = h->bhalf ? h->d_bxmax : nullptr;
    if (h->bhalf) {  // the FP16 operand scale of every trajectory
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, h->st, Bp, h->d_bxmax);
        hipLaunchKernelGGL(k_xmax, dim3(nblk(h->nk, 256), B), dim3(256), 0, h->st, h->nk, rho, h->d_fsplit,
                           h->d_bxmax);
    }
    hipLaunchKernelGGL(k_prepare_x, dim3(nblk(nx, 256)), dim3(256), 0, h->st, h->nk, B, Bp, rho, h->d_e, h->d_clk, a,
                       h->expt, h->d_fsplit, xmax, h->d_rhoSB, (uint16_t *)h->d_xb);
    if (h->nseg > 0) launch_batched_core(h, Bp, h->batch_nblk);
    const int64_t nout = (int64_t)Bp * h->nk;
    hipLaunchKernelGGL(k_assemble_b, dim3(nblk(nout, 256)), dim3(256), 0, h->st, h->nk, B, Bp, h->d_Fb, xmax,
                       h->bhalf_g, h->d_feq, h->d_rhoSB, h->d_e, h->d_clk, a, h->expt, h->schrodinger, h->prefac,
                       h->has_eq ? h->d_eq : nullptr, out);
    DMD_HIP(hipGetLastError());
    return DMD_OK;
}

}  // namespace dmd
