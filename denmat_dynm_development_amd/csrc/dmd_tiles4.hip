// dmd_tiles4.hip -- nb = 4 tile kernel of the Lindblad scattering operator,
// in FP64 and in the reduced-precision forms (complex64 P, or BF16 re/im P
// with a per-pair FP32 scale).
//
// Same contraction as electronphonon::evolve (src_v4.5.7/evolve_eph.cpp:34-166)
// for nb = 4: per stored pair (k <= k') with P1, P2 (16 x 16 complex each)
//   forward  Pdm_k  += P1 vec(rho_k')          dm1P_k  += vec(1-rho_k')^T P2
//   reverse  Pdm_k' += conj(P2) vec(rho_k)     dm1P_k' += vec(1-rho_k)^T conj(P1)   (k < k')
// and per row C = (1-rho) mat(Pdm) - mat(dm1P) rho (zhemm upper), ddm =
// prefac/2 (C + C^+).
//
// Work mapping.  Sixteen lanes share a pair; lane l = 4a + b holds the 4 x 4
// block P[4a..4a+3][4b..4b+3] of P1 and of P2 (16 + 16 complex).  The pair
// record is stored "blocked": element e = 4i + j of lane l's block sits at
// ((e / U) * 16 + l) * U + e % U, U = 16 bytes / element size, so each load
// instruction of a 16-lane group reads 256 contiguous bytes.  With the block
// in registers every GEMV is lane-local up to a 4-lane reduction:
//   P vec(x)   rows 4a+i : partial over the lane's 4 columns, reduced over b (quad)
//   vec(y)^T P cols 4b+j : partial over the lane's 4 rows,    reduced over a
// (reduce-scatter: after two exchange steps lane (a, b) owns output 4a+b or
// 4b+a).  A wavefront holds 4 pairs; its four 16-lane groups walk the four
// quarters of a chunk, so the rows they add to usually differ.
//
// Rows are cut into blocks of R = DMD_TILE4_R; the pairs of tile (I, J) are
// contiguous.  A workgroup takes a group of tiles of one row block I; the
// row sums F1 = Pdm, F2 = dm1P of blocks I and J live in LDS (FP64), each
// tile's block J is assembled per row into the 16 reals of the Hermitian
// increment (C + C^+, linear in F1, F2 for a fixed rho_k) and flushed with one
// FP64 atomic line per row; block I is flushed at the end.
//
// Reduced-precision forms split rho = diag(f) + drho.  F1, F2 are linear in
// (rho, 1-rho) per pair, so F = F(diag f, 1 - diag f) + F(drho, -drho): the
// first term is computed once in FP64 from the stored (quantised) P, the
// second per evaluation with P and drho in FP32 products and FP64 row sums.
// The state, the assembly with rho_k and the epilogue stay FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dmd_internal.h"
#include "synth.h"

namespace dmd {

// ---------------------------------------------------------------------------
// element types of stored P
// ---------------------------------------------------------------------------
struct bf16c {  // BF16 re (low half) and im (high half) of one complex element
    uint32_t w;
};
__host__ __device__ inline uint32_t f32_to_bf16_bits(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return u >> 16;
}
__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }

template <typename PT> struct PTraits;
template <> struct PTraits<double2> { static constexpr int U = 1; };
template <> struct PTraits<float2> { static constexpr int U = 2; };
template <> struct PTraits<bf16c> { static constexpr int U = 4; };

// complex helpers in the compute type
template <typename CT> struct C2;
template <> struct C2<double> { using T = double2; };
template <> struct C2<float> { using T = float2; };
template <typename V> __device__ __forceinline__ V cz();
template <> __device__ __forceinline__ double2 cz<double2>() { return make_double2(0, 0); }
template <> __device__ __forceinline__ float2 cz<float2>() { return make_float2(0, 0); }
template <typename V> __device__ __forceinline__ void cfma_t(V &acc, V a, V b)  // acc += a b
{
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.y, b.x, acc.y);
    acc.y = fma(a.x, b.y, acc.y);
}
template <typename V> __device__ __forceinline__ void cfmac_t(V &acc, V a, V b)  // acc += a conj(b)
{
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(a.y, b.x, acc.y);
    acc.y = fma(-a.x, b.y, acc.y);
}
template <typename V> __device__ __forceinline__ V shx(V v, int m)
{
    V r;
    r.x = __shfl_xor(v.x, m);
    r.y = __shfl_xor(v.y, m);
    return r;
}
template <typename V> __device__ __forceinline__ V vadd(V a, V b)
{
    a.x += b.x;
    a.y += b.y;
    return a;
}
template <typename V> __device__ __forceinline__ V vsel(bool c, V a, V b) { return c ? a : b; }

// Reduce-scatter of v[0..3] over the two lane bits m_hi, m_lo of a 4-lane
// set: afterwards the lane with bits (h, o) holds sum over the set of v[2h+o].
template <typename V>
__device__ __forceinline__ V reduce4(const V v[4], int lane, int m_hi, int m_lo)
{
    const bool h = (lane & m_hi) != 0, o = (lane & m_lo) != 0;
    V k0 = vsel(h, v[2], v[0]), k1 = vsel(h, v[3], v[1]);
    const V s0 = vsel(h, v[0], v[2]), s1 = vsel(h, v[1], v[3]);
    k0 = vadd(k0, shx(s0, m_hi));
    k1 = vadd(k1, shx(s1, m_hi));
    const V keep = vsel(o, k1, k0), send = vsel(o, k0, k1);
    return vadd(keep, shx(send, m_lo));
}

// Loads the lane's 4 x 4 blocks of P1 and P2 of the pair record at rec,
// converted to the compute type.
template <typename PT, typename V> struct Loader;
template <> struct Loader<double2, double2> {
    static __device__ __forceinline__ void load(const double2 *rec, int l, double2 p1[16], double2 p2[16])
    {
#pragma unroll
        for (int e = 0; e < 16; e++) p1[e] = ld_stream(rec + e * 16 + l);
#pragma unroll
        for (int e = 0; e < 16; e++) p2[e] = ld_stream(rec + 256 + e * 16 + l);
    }
};
template <typename V> struct Loader<float2, V> {
    static __device__ __forceinline__ void load(const float2 *rec, int l, V p1[16], V p2[16])
    {
#pragma unroll
        for (int mat = 0; mat < 2; mat++)
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const uint4 q = ld_stream_u4(rec + mat * 256 + (u * 16 + l) * 2);
                V *p = mat ? p2 : p1;
                p[2 * u].x = __uint_as_float(q.x); p[2 * u].y = __uint_as_float(q.y);
                p[2 * u + 1].x = __uint_as_float(q.z); p[2 * u + 1].y = __uint_as_float(q.w);
            }
    }
};
template <typename V> struct Loader<bf16c, V> {
    static __device__ __forceinline__ void load(const bf16c *rec, int l, V p1[16], V p2[16])
    {
#pragma unroll
        for (int mat = 0; mat < 2; mat++)
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint4 q = ld_stream_u4(rec + mat * 256 + (u * 16 + l) * 4);
                V *p = mat ? p2 : p1;
                const uint32_t w[4] = { q.x, q.y, q.z, q.w };
#pragma unroll
                for (int t = 0; t < 4; t++) { p[4 * u + t].x = bf16_lo(w[t]); p[4 * u + t].y = bf16_hi(w[t]); }
            }
    }
};

// Hermitian element (i, j) of a 4 x 4 matrix from its upper triangle with a
// real diagonal (how zhemm reads A); one_minus: of 1 - A.
__device__ __forceinline__ double2 up4(const double2 *A, int i, int j, bool one_minus)
{
    double2 v;
    if (i < j) v = A[i * 4 + j];
    else if (i > j) v = cconj(A[j * 4 + i]);
    else v = make_double2(A[i * 5].x, 0.0);
    if (one_minus) v = make_double2((i == j ? 1.0 : 0.0) - v.x, -v.y);
    return v;
}

__device__ __forceinline__ double2 c_elem(const double2 *F, const double2 *rk, const double2 *r1k, int i, int j);
// X_ij of C = (1-rho) F1 - F2 rho for one row; F as [F1 re/im | F2 re/im];
// 1-rho from r1k when given (explicit dm1), else formed from rho.
__device__ __forceinline__ double2 c_elem(const double *F, const double2 *rk, const double2 *r1k, int i, int j)
{
    double2 x = make_double2(0, 0);
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const double2 f1 = make_double2(F[2 * (4 * m + j)], F[2 * (4 * m + j) + 1]);
        cfma(x, r1k ? up4(r1k, i, m, false) : up4(rk, i, m, true), f1);
    }
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const double2 f2 = make_double2(F[32 + 2 * (4 * i + m)], F[32 + 2 * (4 * i + m) + 1]);
        cfma(x, make_double2(-f2.x, -f2.y), up4(rk, m, j, false));
    }
    return x;
}

// the same with F as 32 complex [F1 | F2]
__device__ __forceinline__ double2 c_elem(const double2 *F, const double2 *rk, const double2 *r1k, int i, int j)
{
    return c_elem((const double *)F, rk, r1k, i, j);
}

// Packed Hermitian increment of element l = 4i + j: H_ii (i == j), Re H_ij
// (i < j), Im H_ji (i > j), from X = C of the row.
__device__ __forceinline__ double herm_pack(double2 xij, double2 xji, int i, int j)
{
    if (i == j) return 2.0 * xij.x;
    if (i < j) return xij.x + xji.x;
    return xji.y - xij.y;  // Im(H_ji) = Im(X_ji + conj X_ij)
}

constexpr int RS4 = 66;  // LDS doubles per row accumulator: F1 (32) | F2 (32) + 2 pad

// Flush of the rows [0, R) of one LDS block (rows base + r).  RAW: add F1|F2
// (64 doubles) to out[row * 64 ..]; else assemble and add the 16 packed
// Hermitian reals to out[row * 16 ..].  Clears the block.
template <bool RAW, bool DM1, int R, bool LIN = false>
__device__ __forceinline__ void flush_block4(double *acc, int64_t base, int64_t nk, const double2 *__restrict__ rhoS,
                                             const double2 *__restrict__ rho1S, double *__restrict__ out)
{
    if (RAW) {
        for (int t = threadIdx.x; t < R * 64; t += blockDim.x) {
            const int r = t >> 6, d = t & 63;
            const double v = acc[r * RS4 + d];
            acc[r * RS4 + d] = 0.0;
            if (base + r < nk && v != 0.0) unsafeAtomicAdd(out + (base + r) * 64 + d, v);
        }
        return;
    }
    const int l = threadIdx.x & 15, i = l >> 2, j = l & 3;
    for (int r0 = 0; r0 < R; r0 += 16) {
        const int r = r0 + (threadIdx.x >> 4);
        const int64_t row = base + r;
        double v = 0.0;
        if (LIN && row < nk) {  // H = X + X^+ of the row sums X (evolve_eph_linear.cpp:47-53)
            const double *F = acc + r * RS4;
            v = herm_pack(make_double2(F[2 * (4 * i + j)], F[2 * (4 * i + j) + 1]),
                          make_double2(F[2 * (4 * j + i)], F[2 * (4 * j + i) + 1]), i, j);
        } else if (row < nk) {
            // rho of the row read where it lies (an indexed private copy would sit in scratch)
            const double *F = acc + r * RS4;
            const double2 *rk = rhoS + row * 16, *r1 = DM1 ? rho1S + row * 16 : nullptr;
            v = herm_pack(c_elem(F, rk, r1, i, j), c_elem(F, rk, r1, j, i), i, j);
        }
        __syncthreads();  // every lane of the row has read F before it is cleared
#pragma unroll
        for (int q = 0; q < 4; q++) acc[r * RS4 + 4 * l + q] = 0.0;
        if (row < nk && v != 0.0) unsafeAtomicAdd(out + row * 16 + l, v);
    }
}

// The panel kernel.  A workgroup streams one chunk of the pairs of one row
// block I (k in [I R, I R + R)); X: nk x 16 rho_S (FP64 plain, !YNEG:
// y = 1 - x) or drho_S (reduced forms, YNEG: y = -x); DM1: y = Y, the
// caller's explicit 1 - rho (FP64 only).  Forward row sums F1|F2 of block I
// stay in LDS and are assembled once at the end; each pair's reverse
// contribution to row k' is assembled at once into the 16 reals of its
// Hermitian increment and added with one FP64 atomic line.  RAW writes the
// raw row sums F1|F2 instead (F_eq).  LIN: the record holds the linearized
// operator Lscij | Lscji and X is drho_S: forward X_k += Lscij drho_k',
// reverse Y_k' += Lscji drho_k, increments X + X^+ (evolve_eph_linear.cpp:26-53).
// FP32 compute: three waves per SIMD (at most 168 VGPRs); the FP64 form keeps
// its allocation (under that bound it spills).
template <typename PT, typename CT, bool RAW, bool YNEG, bool DM1, int R, bool LIN = false>
__global__ __launch_bounds__(256, sizeof(CT) == 4 ? 3 : 1) void k_panels4(int64_t nk, const PT *__restrict__ P,
                                                 const float *__restrict__ pscale, const int32_t *__restrict__ k1s,
                                                 const int32_t *__restrict__ k2s, const int64_t *__restrict__ cbeg,
                                                 const int32_t *__restrict__ clen, const int2 *__restrict__ cij,
                                                 const typename C2<CT>::T *__restrict__ X,
                                                 const typename C2<CT>::T *__restrict__ Y,
                                                 const double2 *__restrict__ rhoS, const double2 *__restrict__ rho1S,
                                                 double *__restrict__ out)
{
    using V = typename C2<CT>::T;
    constexpr int RY = DM1 ? R : 1;
    __shared__ double accI[R * RS4];
    __shared__ V xI[R * 16], yI[RY * 16];
    // per 16-lane group: the pair's reverse sums G1 | G2 and the rows rho_k', 1 - rho_k'
    __shared__ double2 gsc[16][32];
    __shared__ double2 rsc[16][16], r1sc[DM1 ? 16 : 1][16];
    __shared__ double2 csc[16][16];  // per 16-lane group: the pair's reverse C' elements
    __shared__ V xsc[16][16], ysc[DM1 ? 16 : 1][16];  // per 16-lane group: the partner row of X (and Y)
    const int c = blockIdx.x;
    const int I = cij[c].x;
    const int64_t I0 = (int64_t)I * R;
    for (int t = threadIdx.x; t < R * RS4; t += blockDim.x) accI[t] = 0.0;
    for (int t = threadIdx.x; t < R * 16; t += blockDim.x) {
        xI[t] = I0 * 16 + t < nk * 16 ? X[I0 * 16 + t] : cz<V>();
        if (DM1) yI[t] = I0 * 16 + t < nk * 16 ? Y[I0 * 16 + t] : cz<V>();
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, grp = lane >> 4, l = lane & 15;
    const int a = l >> 2, b = l & 3, gi = wv * 4 + grp;
    const int64_t sb = cbeg[c];
    const int Q = clen[c] >> 2;  // the chunk length is a multiple of 4
    for (int it = wv; it < Q; it += 4) {
        const int64_t s = sb + (int64_t)grp * Q + it;
        const int32_t kk = k1s[s];
        const bool ok = kk >= 0;
        const int64_t kp = ok ? k2s[s] : 0;
        const int lo1 = ok ? (int)(kk - I0) : 0;
        V p1[16], p2[16];
        Loader<PT, V>::load(P + s * 512, l, p1, p2);
        if (LIN) {
            // x slices: columns 4b.. of x (rows b of the matrices)
            V xkp[4], xk[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                xkp[q] = X[kp * 16 + 4 * b + q];
                xk[q] = xI[lo1 * 16 + 4 * b + q];
            }
            V f1[4], g1[4];
#pragma unroll
            for (int q = 0; q < 4; q++) { f1[q] = cz<V>(); g1[q] = cz<V>(); }
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    cfma_t(f1[i], p1[4 * i + j], xkp[j]);  // X_k[4a+i]  += Lscij[4a+i][4b+j] drho_k'[4b+j]
                    cfma_t(g1[i], p2[4 * i + j], xk[j]);   // Y_k'[4a+i] += Lscji[4a+i][4b+j] drho_k[4b+j]
                }
            const V F1 = reduce4(f1, lane, 2, 1);  // output 4a + b
            const V G1 = reduce4(g1, lane, 2, 1);
            if (ok) {
                double *r1 = accI + lo1 * RS4;
                atomicAdd(r1 + 2 * (4 * a + b), (double)F1.x);
                atomicAdd(r1 + 2 * (4 * a + b) + 1, (double)F1.y);
            }
            gsc[gi][4 * a + b] = make_double2((double)G1.x, (double)G1.y);
            __builtin_amdgcn_wave_barrier();
            const double v = herm_pack(gsc[gi][4 * a + b], gsc[gi][4 * b + a], a, b);
            __builtin_amdgcn_wave_barrier();  // read before the next pair writes the group's slots
            if (ok && kk < kp && v != 0.0) unsafeAtomicAdd(out + kp * 16 + l, v);
            continue;
        }
        // the partner row k' staged per group; the GEMVs read x and y from LDS
        // and run one after another, so few registers beyond P stay live
        const V xv = X[kp * 16 + l];
        __builtin_amdgcn_wave_barrier();  // the previous pair's reads of the group's rows come first
        xsc[gi][l] = xv;
        if (DM1) ysc[gi][l] = Y[kp * 16 + l];
        __builtin_amdgcn_wave_barrier();
        const V *xr = xsc[gi], *xl = xI + lo1 * 16, *yr = DM1 ? ysc[gi] : nullptr, *yl = DM1 ? yI + lo1 * 16 : nullptr;
        // y = 1 - x of the row (YNEG: -x; DM1: the given Y) at (4a + q)
        auto yof = [&](const V *x, const V *y, int q) {
            V r;
            if (DM1) {
                r = y[4 * a + q];
            } else {
                const V u = x[4 * a + q];
                r.x = (YNEG ? 0 : (a == q ? 1 : 0)) - u.x;
                r.y = -u.y;
            }
            return r;
        };
        V f[4];
        // Pdm_k[4a+i] += P1[4a+i][4b+j] rho_k'[4b+j]
#pragma unroll
        for (int q = 0; q < 4; q++) f[q] = cz<V>();
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const V x = xr[4 * b + j];
#pragma unroll
            for (int i = 0; i < 4; i++) cfma_t(f[i], p1[4 * i + j], x);
        }
        V F1 = reduce4(f, lane, 2, 1);  // output 4a + b
        // dm1P_k'[4b+j] += (1-rho_k)[4a+i] conj(P1[4a+i][4b+j])
#pragma unroll
        for (int q = 0; q < 4; q++) f[q] = cz<V>();
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const V y = yof(xl, yl, i);
#pragma unroll
            for (int j = 0; j < 4; j++) cfmac_t(f[j], y, p1[4 * i + j]);
        }
        V G2 = reduce4(f, lane, 8, 4);  // output 4b + a
        // dm1P_k[4b+j] += (1-rho_k')[4a+i] P2[4a+i][4b+j]
#pragma unroll
        for (int q = 0; q < 4; q++) f[q] = cz<V>();
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const V y = yof(xr, yr, i);
#pragma unroll
            for (int j = 0; j < 4; j++) cfma_t(f[j], y, p2[4 * i + j]);
        }
        V F2 = reduce4(f, lane, 8, 4);
        // Pdm_k'[4a+i] += rho_k[4b+j] conj(P2[4a+i][4b+j])
#pragma unroll
        for (int q = 0; q < 4; q++) f[q] = cz<V>();
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const V x = xl[4 * b + j];
#pragma unroll
            for (int i = 0; i < 4; i++) cfmac_t(f[i], x, p2[4 * i + j]);
        }
        V G1 = reduce4(f, lane, 2, 1);
        if (pscale) {  // a power of two: scaling the reduced sums is exact, 8 products instead of 32
            const CT sc = (CT)pscale[s];
            F1.x *= sc; F1.y *= sc; F2.x *= sc; F2.y *= sc;
            G1.x *= sc; G1.y *= sc; G2.x *= sc; G2.y *= sc;
        }
        if (ok) {
            double *r1 = accI + lo1 * RS4;
            atomicAdd(r1 + 2 * (4 * a + b), (double)F1.x);
            atomicAdd(r1 + 2 * (4 * a + b) + 1, (double)F1.y);
            atomicAdd(r1 + 32 + 2 * (4 * b + a), (double)F2.x);
            atomicAdd(r1 + 32 + 2 * (4 * b + a) + 1, (double)F2.y);
        }
        const bool rev = ok && kk < kp;
        if (RAW) {
            if (rev) {
                double *o = out + kp * 64;
                unsafeAtomicAdd(o + 2 * (4 * a + b), (double)G1.x);
                unsafeAtomicAdd(o + 2 * (4 * a + b) + 1, (double)G1.y);
                unsafeAtomicAdd(o + 32 + 2 * (4 * b + a), (double)G2.x);
                unsafeAtomicAdd(o + 32 + 2 * (4 * b + a) + 1, (double)G2.y);
            }
        } else {
            // the pair's reverse term, assembled: C' = (1-rho_k') G1 - G2 rho_k' (zhemm
            // upper), H' = C' + C'^+, lane (i, j) = (a, b) owns packed element l
            gsc[gi][4 * a + b] = make_double2((double)G1.x, (double)G1.y);
            gsc[gi][16 + 4 * b + a] = make_double2((double)G2.x, (double)G2.y);
            rsc[gi][l] = rhoS[kp * 16 + l];
            if (DM1) r1sc[gi][l] = rho1S[kp * 16 + l];
            __builtin_amdgcn_wave_barrier();
            const double2 *r1p = DM1 ? r1sc[gi] : nullptr;
            const double2 cab = c_elem(gsc[gi], rsc[gi], r1p, a, b);  // lane (a, b) forms C'_ab once
            csc[gi][l] = cab;
            __builtin_amdgcn_wave_barrier();
            const double v = herm_pack(cab, csc[gi][4 * b + a], a, b);  // C'_ba from lane (b, a)
            __builtin_amdgcn_wave_barrier();  // the group's LDS rows are read before the next pair writes them
            if (rev && v != 0.0) unsafeAtomicAdd(out + kp * 16 + l, v);
        }
    }
    __syncthreads();
    flush_block4<RAW, DM1, R, LIN>(accI, I0, nk, rhoS, rho1S, out);
}

// Epilogue: ddm = prefac/2 (H + H_eq), rotated back, minus ddmdt_eq.  H from
// the packed accumulator (cleared here); H_eq = C_eq + C_eq^+ with C_eq =
// (1-rho) F1eq - F2eq rho when feq is given.  One thread per (k, i, j).
__global__ __launch_bounds__(256) void k_assemble_h4(int64_t nk, double *__restrict__ hacc,
                                                     const double *__restrict__ feq,
                                                     const double2 *__restrict__ rhoS, const double *__restrict__ e,
                                                     const DClock *__restrict__ clk, double a, int expt,
                                                     int schrodinger, double prefac, const double2 *__restrict__ eq,
                                                     double2 *__restrict__ out)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t k = idx >> 4;
    const int l = (int)(idx & 15), i = l >> 2, j = l & 3;
    double2 d = make_double2(0, 0);
    if (k < nk) {
        const double *h = hacc + k * 16;
        if (i == j) d = make_double2(h[5 * i], 0.0);
        else if (i < j) d = make_double2(h[4 * i + j], h[4 * j + i]);
        else d = make_double2(h[4 * j + i], -h[4 * i + j]);
        if (feq) {
            double2 rk[16];
#pragma unroll
            for (int q = 0; q < 16; q++) rk[q] = rhoS[k * 16 + q];
            const double *F = feq + k * 64;
            const double2 xij = c_elem(F, rk, nullptr, i, j), xji = c_elem(F, rk, nullptr, j, i);
            d = cadd(d, cadd(xij, cconj(xji)));
        }
    }
    __syncthreads();  // all reads of the row before it is cleared
    if (k >= nk) return;
    hacc[k * 16 + l] = 0.0;
    d = make_double2(d.x * (prefac * 0.5), d.y * (prefac * 0.5));
    const double t = clock_t_at(clk, a);
    const double arg = (e[k * 4 + i] - e[k * 4 + j]) * t;
    if (expt && i != j) d = cmul(d, ccis(arg));
    if (eq) {
        double2 q = eq[k * 16 + l];
        if (i != j && !schrodinger) q = cmul(q, ccis(arg));
        d = csub(d, q);
    }
    out[k * 16 + l] = d;
}

// drho_S = rho_S - diag(f) in FP32 (the low-precision operand)
__global__ void k_split_x(int64_t nk, const double2 *__restrict__ rhoS, const double *__restrict__ f,
                          float2 *__restrict__ x)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nk * 16) return;
    const int64_t k = idx >> 4;
    const int l = (int)(idx & 15);
    double2 v = rhoS[idx];
    if (f && (l % 5) == 0) v.x -= f[k * 4 + l / 5];
    x[idx] = make_float2((float)v.x, (float)v.y);
}

// diag(f) as an nk x 16 complex state
__global__ void k_diag_state(int64_t nk, const double *__restrict__ f, double2 *__restrict__ x)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nk * 16) return;
    const int64_t k = idx >> 4;
    const int l = (int)(idx & 15);
    x[idx] = make_double2(f && (l % 5) == 0 ? f[k * 4 + l / 5] : 0.0, 0.0);
}

// ---------------------------------------------------------------------------
// P storage of the reduced forms: FP64 input (standard pair order) ->
// quantised blocked records; and back.  One 64-lane group per pair.
// ---------------------------------------------------------------------------
template <typename PT> __device__ __forceinline__ double2 pt_get(const PT *rec, int o, float sc);
template <> __device__ __forceinline__ double2 pt_get<float2>(const float2 *rec, int o, float)
{
    return make_double2(rec[o].x, rec[o].y);
}
template <> __device__ __forceinline__ double2 pt_get<bf16c>(const bf16c *rec, int o, float sc)
{
    return make_double2((double)bf16_lo(rec[o].w) * sc, (double)bf16_hi(rec[o].w) * sc);
}

template <typename PT>
__global__ __launch_bounds__(64) void k_place_q(int64_t n, int64_t first, const double2 *__restrict__ P1,
                                                const double2 *__restrict__ P2, const int64_t *__restrict__ inv,
                                                double scale, int add, PT *__restrict__ Pd,
                                                float *__restrict__ pscale)
{
    constexpr int U = PTraits<PT>::U;
    const int64_t f = blockIdx.x;
    if (f >= n) return;
    const int64_t pos = inv ? inv[first + f] : first + f;
    PT *rec = Pd + pos * 512;
    const float old_sc = pscale ? pscale[pos] : 1.0f;
    double2 v[8];
    double mx = 0.0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int el = threadIdx.x * 8 + q, mat = el >> 8, row = (el >> 4) & 15, col = el & 15;
        double2 x = mat ? P2[f * 256 + (el & 255)] : P1[f * 256 + (el & 255)];
        x = make_double2(x.x * scale, x.y * scale);
        if (add) x = cadd(x, pt_get<PT>(rec, blk_off(mat, row, col, U), old_sc));
        v[q] = x;
        mx = fmax(mx, fmax(fabs(x.x), fabs(x.y)));
    }
    float sc = 1.0f;
    if (pscale) {  // power-of-two scale: max |re|, |im| of the pair maps into [0.5, 1]
        for (int m = 32; m >= 1; m >>= 1) mx = fmax(mx, __shfl_xor(mx, m));
        int ex = 0;
        if (mx > 0.0) (void)frexp(mx, &ex);
        sc = ldexpf(1.0f, ex);
        __syncthreads();  // every lane has read the old record (add) before it is overwritten
        if (threadIdx.x == 0) pscale[pos] = sc;
    } else {
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int el = threadIdx.x * 8 + q, mat = el >> 8, row = (el >> 4) & 15, col = el & 15;
        const int o = blk_off(mat, row, col, U);
        if constexpr (U == 2) {
            rec[o] = make_float2((float)v[q].x, (float)v[q].y);
        } else {
            const uint32_t re = f32_to_bf16_bits((float)(v[q].x / sc)), im = f32_to_bf16_bits((float)(v[q].y / sc));
            rec[o].w = re | (im << 16);
        }
    }
}

template <typename PT>
__global__ __launch_bounds__(64) void k_fetch_q(int64_t n, int64_t first, const PT *__restrict__ Pd,
                                                const float *__restrict__ pscale, const int64_t *__restrict__ inv,
                                                double2 *__restrict__ P1, double2 *__restrict__ P2)
{
    constexpr int U = PTraits<PT>::U;
    const int64_t f = blockIdx.x;
    if (f >= n) return;
    const int64_t pos = inv ? inv[first + f] : first + f;
    const PT *rec = Pd + pos * 512;
    const float sc = pscale ? pscale[pos] : 1.0f;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int el = threadIdx.x * 8 + q, mat = el >> 8, row = (el >> 4) & 15, col = el & 15;
        const double2 v = pt_get<PT>(rec, blk_off(mat, row, col, U), sc);
        if (mat) P2[f * 256 + (el & 255)] = v;
        else P1[f * 256 + (el & 255)] = v;
    }
}

// On-device synthetic P in the reduced forms: one lane per slot writes the
// FP64 pair (standard order) into a scratch batch, then one 64-lane group per
// slot quantises it into the blocked record.
__global__ __launch_bounds__(64) void k_generate_scratch(int64_t s0, int64_t ns, dmd_synth_params p,
                                                         const int32_t *__restrict__ k1s,
                                                         const int32_t *__restrict__ k2s,
                                                         const double *__restrict__ e, double2 *__restrict__ scratch)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ns) return;
    const int64_t s = s0 + q;
    double2 *tmp = scratch + q * 512;
    const int64_t k = k1s[s], kp = k2s[s];
    if (k >= 0) {
        synth_pair_P_nb<4>(p, e + k * 4, e + kp * 4, k, kp, (double *)tmp, (double *)(tmp + 256));
    } else {
        for (int i = 0; i < 512; i++) tmp[i] = make_double2(0, 0);
    }
}

template <typename PT>
__global__ __launch_bounds__(64) void k_quantise_scratch(int64_t s0, int64_t ns, const double2 *__restrict__ scratch,
                                                         PT *__restrict__ Pd, float *__restrict__ pscale)
{
    constexpr int U = PTraits<PT>::U;
    const int64_t q0 = blockIdx.x;
    if (q0 >= ns) return;
    const int64_t s = s0 + q0;
    const double2 *tmp = scratch + q0 * 512;
    double2 v[8];
    double mx = 0.0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        v[q] = tmp[threadIdx.x * 8 + q];
        mx = fmax(mx, fmax(fabs(v[q].x), fabs(v[q].y)));
    }
    float sc = 1.0f;
    if (pscale) {
        for (int m = 32; m >= 1; m >>= 1) mx = fmax(mx, __shfl_xor(mx, m));
        int ex = 0;
        if (mx > 0.0) (void)frexp(mx, &ex);
        sc = ldexpf(1.0f, ex);
        if (threadIdx.x == 0) pscale[s] = sc;
    }
    PT *rec = Pd + s * 512;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int el = threadIdx.x * 8 + q, mat = el >> 8, row = (el >> 4) & 15, col = el & 15;
        const int o = blk_off(mat, row, col, U);
        if constexpr (U == 2) {
            rec[o] = make_float2((float)v[q].x, (float)v[q].y);
        } else {
            const uint32_t re = f32_to_bf16_bits((float)(v[q].x / sc)), im = f32_to_bf16_bits((float)(v[q].y / sc));
            rec[o].w = re | (im << 16);
        }
    }
}

// FP64 P: in-place permutation of every pair record between the standard
// order (P1 then P2, row-major) and the blocked order.
__global__ __launch_bounds__(256) void k_relayout4(int64_t nslots, int to_blocked, double2 *__restrict__ P)
{
    __shared__ double2 buf[512];
    const int64_t s = blockIdx.x;
    if (s >= nslots) return;
    double2 *rec = P + s * 512;
    for (int t = threadIdx.x; t < 512; t += blockDim.x) buf[t] = rec[t];
    __syncthreads();
    for (int el = threadIdx.x; el < 512; el += blockDim.x) {
        const int mat = el >> 8, row = (el >> 4) & 15, col = el & 15;
        const int o = blk_off(mat, row, col, 1);
        if (to_blocked) rec[o] = buf[el];
        else rec[el] = buf[o];
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int tiles4_place(dmd_scatt *h, const double2 *P1, const double2 *P2, int64_t n, int64_t first, double scale, int add)
{
    if (n <= 0) return DMD_OK;
    if (h->precision == DMD_PREC_C64)
        hipLaunchKernelGGL(k_place_q<float2>, dim3((unsigned)n), dim3(64), 0, h->st, n, first, P1, P2, h->d_inv, scale,
                           add, (float2 *)h->d_P, nullptr);
    else
        hipLaunchKernelGGL(k_place_q<bf16c>, dim3((unsigned)n), dim3(64), 0, h->st, n, first, P1, P2, h->d_inv, scale,
                           add, (bf16c *)h->d_P, h->d_pscale);
    DMD_HIP(hipGetLastError());
    h->feq_dirty = true;
    h->bvis_dirty = true;
    return DMD_OK;
}

int tiles4_fetch(dmd_scatt *h, int64_t first, int64_t n, double2 *P1, double2 *P2)
{
    if (n <= 0) return DMD_OK;
    if (h->precision == DMD_PREC_C64)
        hipLaunchKernelGGL(k_fetch_q<float2>, dim3((unsigned)n), dim3(64), 0, h->st, n, first, (const float2 *)h->d_P,
                           nullptr, h->d_inv, P1, P2);
    else
        hipLaunchKernelGGL(k_fetch_q<bf16c>, dim3((unsigned)n), dim3(64), 0, h->st, n, first, (const bf16c *)h->d_P,
                           h->d_pscale, h->d_inv, P1, P2);
    DMD_HIP(hipGetLastError());
    return DMD_OK;
}

int tiles4_generate(dmd_scatt *h, const dmd_synth_params &p)
{
    const int64_t batch = 1 << 16;
    double2 *scratch = nullptr;
    int rc;
    if ((rc = dalloc_t(&scratch, (size_t)batch * 512, "generator scratch"))) return rc;
    for (int64_t s0 = 0; s0 < h->nslots; s0 += batch) {
        const int64_t ns = std::min<int64_t>(batch, h->nslots - s0);
        hipLaunchKernelGGL(k_generate_scratch, dim3(nblk(ns, 64)), dim3(64), 0, h->st, s0, ns, p, h->d_k1, h->d_k2,
                           h->d_e, scratch);
        if (h->precision == DMD_PREC_C64)
            hipLaunchKernelGGL(k_quantise_scratch<float2>, dim3((unsigned)ns), dim3(64), 0, h->st, s0, ns, scratch,
                               (float2 *)h->d_P, nullptr);
        else
            hipLaunchKernelGGL(k_quantise_scratch<bf16c>, dim3((unsigned)ns), dim3(64), 0, h->st, s0, ns, scratch,
                               (bf16c *)h->d_P, h->d_pscale);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { dfree(scratch); return fail(DMD_ERR_HIP, hipGetErrorString(e)); }
    }
    hipError_t e = hipStreamSynchronize(h->st);
    dfree(scratch);
    if (e != hipSuccess) return fail(DMD_ERR_HIP, hipGetErrorString(e));
    h->feq_dirty = true;
    h->bvis_dirty = true;
    return DMD_OK;
}

int tiles4_relayout(dmd_scatt *h, int to_blocked)
{
    if (h->nslots > 0)
        hipLaunchKernelGGL(k_relayout4, dim3((unsigned)h->nslots), dim3(256), 0, h->st, h->nslots, to_blocked, h->d_P);
    DMD_HIP(hipGetLastError());
    DMD_HIP(hipStreamSynchronize(h->st));
    return DMD_OK;
}

template <typename PT, typename CT, bool RAW, bool YNEG, bool DM1 = false, bool LIN = false>
static void launch_k4(dmd_scatt *h, const void *X, const void *Y, double *out)
{
    if (h->nchunks <= 0) return;
    hipLaunchKernelGGL((k_panels4<PT, CT, RAW, YNEG, DM1, DMD_TILE4_R, LIN>), dim3((unsigned)h->nchunks), dim3(256), 0,
                       h->st, h->nk, (const PT *)h->d_P, h->d_pscale, h->d_k1, h->d_k2, h->d_cbeg, h->d_clen, h->d_cij,
                       (const typename C2<CT>::T *)X, (const typename C2<CT>::T *)Y, h->d_rhoS, h->d_rho1S, out);
}

int tiles4_prepare(dmd_scatt *h)
{
    if (h->precision == DMD_PREC_FP64 || !h->feq_dirty) return DMD_OK;
    int rc;
    const size_t ns = (size_t)h->nk * 16;
    if (!h->d_feq && (rc = dalloc_t(&h->d_feq, (size_t)h->nk * 64, "F_eq"))) return rc;
    if (!h->d_x32 && (rc = dalloc_t(&h->d_x32, ns, "drho FP32"))) return rc;
    double2 *x = nullptr;
    if ((rc = dalloc_t(&x, ns, "diag(f)"))) return rc;
    hipLaunchKernelGGL(k_diag_state, dim3(nblk((int64_t)ns, 256)), dim3(256), 0, h->st, h->nk, h->d_fsplit, x);
    hipError_t e = hipMemsetAsync(h->d_feq, 0, (size_t)h->nk * 64 * sizeof(double), h->st);
    if (e == hipSuccess) {
        // F(diag f, 1 - diag f) in FP64 arithmetic from the stored P
        if (h->precision == DMD_PREC_C64) launch_k4<float2, double, true, false>(h, x, nullptr, h->d_feq);
        else launch_k4<bf16c, double, true, false>(h, x, nullptr, h->d_feq);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    dfree(x);
    if (e != hipSuccess) return fail(DMD_ERR_HIP, hipGetErrorString(e));
    h->feq_dirty = false;
    return DMD_OK;
}

// contraction + epilogue for nb = 4 on the current rhoS (and rho1S when dm1;
// the reduced forms always form 1 - rho from rho) into out
int launch_tiles4(dmd_scatt *h, bool dm1, double a, double2 *out)
{
    const double2 *eq = h->has_eq ? h->d_eq : nullptr;
    const int64_t n16 = h->nk * 16;
    if (h->linear) {
        launch_k4<double2, double, false, false, false, true>(h, h->d_rhoS, nullptr, h->d_hacc16);
    } else if (h->precision == DMD_PREC_FP64) {
        if (dm1) launch_k4<double2, double, false, false, true>(h, h->d_rhoS, h->d_rho1S, h->d_hacc16);
        else launch_k4<double2, double, false, false>(h, h->d_rhoS, nullptr, h->d_hacc16);
    } else {
        hipLaunchKernelGGL(k_split_x, dim3(nblk(n16, 256)), dim3(256), 0, h->st, h->nk, h->d_rhoS, h->d_fsplit, h->d_x32);
        if (h->precision == DMD_PREC_C64) launch_k4<float2, float, false, true>(h, h->d_x32, nullptr, h->d_hacc16);
        else launch_k4<bf16c, float, false, true>(h, h->d_x32, nullptr, h->d_hacc16);
    }
    hipLaunchKernelGGL(k_assemble_h4, dim3(nblk(n16, 256)), dim3(256), 0, h->st, h->nk, h->d_hacc16,
                       h->precision == DMD_PREC_FP64 ? nullptr : h->d_feq, h->d_rhoS, h->d_e, h->d_clk, a, h->expt,
                       h->schrodinger, h->prefac, eq, out);
    DMD_HIP(hipGetLastError());
    return DMD_OK;
}

}  // namespace dmd
