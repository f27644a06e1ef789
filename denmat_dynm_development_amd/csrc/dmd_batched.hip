// dmd_batched.hip -- B trajectories sharing one pass over P (nb = 4, BF16 P):
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
// Layouts in HBM.
//  * Visit records, CSR by destination row: M1 (from A1) and M2 (from A2^T),
//    32 x 16 BF16 row-major each, hi then lo: [M1 hi | M2 hi | M1 lo | M2 lo],
//    4 KiB per visit (8 KiB per stored pair), conj and the per-pair scale of
//    the BF16 store applied.
//  * X: per partner row q, [term][k-half][trajectory][8] BF16, so a
//    wavefront's 32 trajectories of one 8-element operand are 512 contiguous
//    bytes (Bp x 64 bytes per row q).
//  * Row sums: FP32 [k][64][B] (F1 re, F1 im, G re, G im).
// Uneven rows: every row's visits are cut into segments of at most seg_len
// visits, one wavefront each, so a 5000-partner row is as many independent
// work items; segments of one row meet in FP32 atomics.  seg_len and the
// trajectory blocks per wavefront are picked by timing at the first batched
// call for a padded B (autotune).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dmd_internal.h"

namespace dmd {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float floatx16;

__device__ __forceinline__ uint32_t f2bf(float f)  // round to nearest even (finite)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ float bflo(uint32_t w) { return __uint_as_float(w << 16); }
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
// for A1 and A2^T, BF16 hi and lo.  One 64-lane group per visit; lane t
// covers rows (t / 16) of the 4 row blocks and K column t % 16.
template <bool C64>
__global__ __launch_bounds__(64) void k_build_visits(int64_t nvis, const int64_t *__restrict__ vslot,
                                                     const uint8_t *__restrict__ vdir, const void *__restrict__ Pq,
                                                     const float *__restrict__ pscale, uint16_t *__restrict__ vrec)
{
    const int64_t v = blockIdx.x;
    if (v >= nvis) return;
    const int64_t s = vslot[v];
    const bool rev = vdir[v] != 0;
    const void *rec = C64 ? (const void *)((const float2 *)Pq + s * 512) : (const void *)((const uint32_t *)Pq + s * 512);
    const float sc = C64 ? 1.0f : pscale[s];
    uint16_t *out = vrec + v * 2048;
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
            const uint32_t hi = f2bf(x), lo = f2bf(x - __uint_as_float(hi << 16));
            const int o = mat * 512 + (part * 16 + r) * 16 + p;
            out[o] = (uint16_t)hi;
            out[1024 + o] = (uint16_t)lo;
        }
    }
}

// x = (Re drho_ii; Re drho_ij, Im drho_ij, i < j) of every trajectory as the
// hi/lo BF16 operand, and rho_S in FP64 for the assembly.  rho: B x nk x 16
// (interaction picture).  One thread per (b, k, K column p).
__global__ void k_prepare_x(int64_t nk, int B, int Bp, const double2 *__restrict__ rho, const double *__restrict__ e,
                            const DClock *__restrict__ clk, double a, int expt, const double *__restrict__ f,
                            double2 *__restrict__ rhoS, uint16_t *__restrict__ xb)
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
    const uint32_t hi = f2bf((float)x);
    const uint32_t lo = f2bf((float)(x - (double)__uint_as_float(hi << 16)));
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

// One wavefront per segment; NBLK blocks of 32 trajectories per pass (the
// grid's y index selects the pass).  Products M_hi x_hi + M_hi x_lo + M_lo
// x_hi (the M_lo x_lo term is below the 16-bit significands).
template <int NBLK>
__global__ __launch_bounds__(256) void k_batched(int64_t nseg, int Bp, const int4 *__restrict__ seg,
                                                 const uint16_t *__restrict__ vrec, const int32_t *__restrict__ vq,
                                                 const uint16_t *__restrict__ xb, float *__restrict__ F)
{
    const int lane = threadIdx.x & 63;
    const int64_t sg = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (sg >= nseg) return;
    const int4 S = seg[sg];  // (k, v0 lo, v0 hi, count)
    const int64_t k = S.x;
    const int64_t v0 = (int64_t)(uint32_t)S.y | ((int64_t)S.z << 32);
    const int nv = S.w;
    const int i = lane & 31, h = lane >> 5;  // A row (real-ified), K half
    const int b0 = blockIdx.y * NBLK * 32;
    floatx16 f1[NBLK], g2[NBLK];
#pragma unroll
    for (int nb = 0; nb < NBLK; nb++)
#pragma unroll
        for (int t = 0; t < 16; t++) { f1[nb][t] = 0.0f; g2[nb][t] = 0.0f; }
    const int oa = i * 16 + 8 * h;  // the lane's 8 K columns of row i
    for (int u = 0; u < nv; u++) {
        const int64_t v = v0 + u;
        const uint16_t *rec = vrec + v * 2048;
        const bf16x8 M1h = as_bf8(ld_stream_u4(rec + oa)), M2h = as_bf8(ld_stream_u4(rec + 512 + oa));
        const bf16x8 M1l = as_bf8(ld_stream_u4(rec + 1024 + oa)), M2l = as_bf8(ld_stream_u4(rec + 1536 + oa));
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
    // D layout: element t of lane l is row 8 (t / 4) + 4 h + t % 4, column (trajectory) l % 32
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
                const double2 f1 = make_double2(fe[2 * r] + (double)Fk[r * Bp], fe[2 * r + 1] + (double)Fk[(16 + r) * Bp]);
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
                const double2 g = make_double2(fe[32 + 2 * r] - (double)Fk[(32 + r) * Bp],
                                               fe[32 + 2 * r + 1] - (double)Fk[(48 + r) * Bp]);
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

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

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
    const int64_t nvis = h->np + h->nrev;
    std::vector<int64_t> vslot(nvis);
    std::vector<uint8_t> vdir(nvis);
    std::vector<int32_t> vq(nvis);
    std::vector<int4> seg;
    int64_t v = 0;
    h->h_vptr.assign(nk + 1, 0);
    std::vector<int64_t> perm;
    for (int64_t k = 0; k < nk; k++) {
        h->h_vptr[k] = v;
        const int64_t v0 = v;
        for (int64_t p = rowptr[k]; p < rowptr[k + 1]; p++, v++) { vslot[v] = fwdpos[p]; vdir[v] = 0; vq[v] = col[p]; }
        for (int64_t p = revptr[k]; p < revptr[k + 1]; p++, v++) { vslot[v] = revpos[p]; vdir[v] = 1; vq[v] = revcol[p]; }
        // a row's visits by partner, so that a partner-block segment order is contiguous
        perm.resize(v - v0);
        for (int64_t q = 0; q < v - v0; q++) perm[q] = v0 + q;
        std::stable_sort(perm.begin(), perm.end(), [&](int64_t x, int64_t y) { return vq[x] < vq[y]; });
        std::vector<int64_t> s2(perm.size());
        std::vector<uint8_t> d2(perm.size());
        std::vector<int32_t> q2(perm.size());
        for (size_t q = 0; q < perm.size(); q++) { s2[q] = vslot[perm[q]]; d2[q] = vdir[perm[q]]; q2[q] = vq[perm[q]]; }
        for (size_t q = 0; q < perm.size(); q++) { vslot[v0 + q] = s2[q]; vdir[v0 + q] = d2[q]; vq[v0 + q] = q2[q]; }
    }
    h->h_vptr[nk] = v;
    int rc;
    dfree(h->d_vrec); dfree(h->d_vq); dfree(h->d_vsd);
    h->d_vrec = nullptr; h->d_vq = nullptr; h->d_vsd = nullptr;
    const bool c64 = h->precision == DMD_PREC_C64;
    if (h->nslots >= (1ll << 31)) return fail(DMD_ERR_UNSUPPORTED, "batched: 2^31 or more stored slots");
    // Visit records (4 KiB per visit) when they fit in device memory: forming
    // M per visit from the stored P reads half the bytes for BF16 P but costs
    // more instructions than it saves (BF16, preset 4: 35 against 25 ms per
    // evaluation at B = 16); without room for the records the pass forms M from
    // P.  DMD_BATCH_RECORDS=0 forces the record-free pass (tests).
    const char *renv = std::getenv("DMD_BATCH_RECORDS");
    h->bdirect = renv && renv[0] == '0';
    std::vector<uint32_t> vsd(nvis);
    for (int64_t q = 0; q < nvis; q++) vsd[q] = (uint32_t)(vslot[q] << 1) | (vdir[q] ? 1u : 0u);
    if (!h->bdirect && (rc = dalloc_t(&h->d_vrec, (size_t)nvis * 2048, "visit records"))) {
        if (rc != DMD_ERR_NOMEM) return rc;
        (void)hipGetLastError();  // the failed allocation: continue without records
        h->bdirect = true;
    }
    if ((rc = dalloc_t(&h->d_vq, (size_t)nvis, "visit partners"))) return rc;
    if ((rc = dalloc_t(&h->d_vsd, (size_t)std::max<int64_t>(nvis, 1), "visit slots"))) return rc;
    int64_t *d_vslot = nullptr;
    uint8_t *d_vdir = nullptr;
    if (!h->bdirect) {
        if ((rc = dalloc_t(&d_vslot, (size_t)nvis, "visit slots"))) return rc;
        if ((rc = dalloc_t(&d_vdir, (size_t)nvis, "visit directions"))) { dfree(d_vslot); return rc; }
    }
    hipError_t e = hipSuccess;
    if (nvis) {
        e = hipMemcpy(h->d_vq, vq.data(), nvis * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h->d_vsd, vsd.data(), nvis * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess && !h->bdirect) e = hipMemcpy(d_vslot, vslot.data(), nvis * 8, hipMemcpyHostToDevice);
        if (e == hipSuccess && !h->bdirect) e = hipMemcpy(d_vdir, vdir.data(), nvis, hipMemcpyHostToDevice);
    }

    if (e == hipSuccess && nvis && !h->bdirect) {
        if (c64)
            hipLaunchKernelGGL(k_build_visits<true>, dim3((unsigned)nvis), dim3(64), 0, h->st, nvis, d_vslot, d_vdir,
                               (const void *)h->d_P, h->d_pscale, h->d_vrec);
        else
            hipLaunchKernelGGL(k_build_visits<false>, dim3((unsigned)nvis), dim3(64), 0, h->st, nvis, d_vslot, d_vdir,
                               (const void *)h->d_P, h->d_pscale, h->d_vrec);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    }
    dfree(d_vslot); dfree(d_vdir);
    if (e != hipSuccess) return fail(DMD_ERR_HIP, std::string("building visit records: ") + hipGetErrorString(e));
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
    if ((rc = dalloc_t(&h->d_seg, seg.size(), "segments"))) return rc;
    if (!seg.empty()) DMD_HIP(hipMemcpy(h->d_seg, seg.data(), seg.size() * sizeof(int4), hipMemcpyHostToDevice));
    h->nseg = (int64_t)seg.size();
    h->seg_len = seg_len;
    h->seg_pblock = pblock;
    return DMD_OK;
}

static void launch_batched_core(dmd_scatt *h, int Bp, int nblk_req);

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
        for (int sl : { 64, 128, 256 }) {
            if ((rc = build_segments(h, sl, best_pb))) return rc;
            for (int nb : { 1, 2 }) {
                if (nb == 2 && Bp % 64) continue;
                if ((rc = trial(nb, &ms))) return rc;
                if (ms < best) { best = ms; best_seg = sl; best_nb = nb; }
            }
        }
        // partner blocks that fit the L2 of one XCD (4 MiB), and 4x and 16x that
        if (forced_pb < 0) {
            const int p0 = std::max(256, 65536 / Bp);
            for (int pb : { p0, 4 * p0, 16 * p0 }) {
                if ((rc = build_segments(h, best_seg, pb))) return rc;
                if ((rc = trial(best_nb, &ms))) return rc;
                if (ms < best) { best = ms; best_pb = pb; }
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

int batched_prepare(dmd_scatt *h, int B)
{
    int rc;
    if ((rc = tiles4_prepare(h))) return rc;
    if ((rc = build_visits(h))) return rc;
    const int Bp = (B + 31) / 32 * 32;
    if (Bp > h->b_cap) {
        dfree(h->d_xb); dfree(h->d_Fb); dfree(h->d_rhoSB);
        h->d_xb = nullptr; h->d_Fb = nullptr; h->d_rhoSB = nullptr;
        h->b_cap = 0;
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
    const int nb2 = (Bp / 32) % 2 == 0 && nblk_req == 2 ? 2 : 1;
    dim3 g(nblk(h->nseg, 4), Bp / (32 * nb2));
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
    if (nb2 == 2)
        hipLaunchKernelGGL(k_batched<2>, g, dim3(256), 0, h->st, h->nseg, Bp, h->d_seg, h->d_vrec, h->d_vq,
                           (const uint16_t *)h->d_xb, h->d_Fb);
    else
        hipLaunchKernelGGL(k_batched<1>, g, dim3(256), 0, h->st, h->nseg, Bp, h->d_seg, h->d_vrec, h->d_vq,
                           (const uint16_t *)h->d_xb, h->d_Fb);
}

// ddm of B trajectories (rho: B x nk x 16 device, interaction picture at the
// clock time a) into out (B x nk x 16).  The X and row-sum buffers are sized
// for the padded B the last batched_prepare saw.
int launch_batched(dmd_scatt *h, int B, const double2 *rho, double a, double2 *out)
{
    const int Bp = (B + 31) / 32 * 32;
    if (Bp > h->b_cap) return fail(DMD_ERR_STATE, "batched buffers too small");
    const int64_t nx = (int64_t)Bp * h->nk * 16;
    hipLaunchKernelGGL(k_prepare_x, dim3(nblk(nx, 256)), dim3(256), 0, h->st, h->nk, B, Bp, rho, h->d_e, h->d_clk, a,
                       h->expt, h->d_fsplit, h->d_rhoSB, (uint16_t *)h->d_xb);
    if (h->nseg > 0) launch_batched_core(h, Bp, h->batch_nblk);
    const int64_t nout = (int64_t)Bp * h->nk;
    hipLaunchKernelGGL(k_assemble_b, dim3(nblk(nout, 256)), dim3(256), 0, h->st, h->nk, B, Bp, h->d_Fb, h->d_feq,
                       h->d_rhoSB, h->d_e, h->d_clk, a, h->expt, h->schrodinger, h->prefac,
                       h->has_eq ? h->d_eq : nullptr, out);
    DMD_HIP(hipGetLastError());
    return DMD_OK;
}

}  // namespace dmd
