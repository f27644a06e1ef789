// dmd_scatt.hip -- MI355X (gfx950) Lindblad scattering operator for DMD.
//
// Hot path: electronphonon::evolve (src_v4.5.7/evolve_eph.cpp:34-166).  For
// every stored pair (k <= k') with P1, P2 (nb^4 complex each) it accumulates
//   Pdm_k  += P1t . vec(rho_k')           dm1P_k  += vec(1-rho_k')^T . P2t
// and, for k < k', the reverse with P1_next = conj(P2t), P2_next = conj(P1t)
//   Pdm_k' += conj(P2t) . vec(rho_k)      dm1P_k' += vec(1-rho_k)^T . conj(P1t)
// then per k: C = (1-rho_k) mat(Pdm_k) - mat(dm1P_k) rho_k (zhemm, upper
// triangle), ddmdt_k = prefac/2 (C + C^dagger), minus ddmdt_eq (optional).
//
// MI355X design.
//  * P lives in HBM once, per stored pair [P1 | P2] = 2 nb^4 complex128
//    contiguous, pairs sorted by (k, k') (CSR rows = k).
//  * The interaction-picture phase of compute_Pt (evolve_eph.cpp:169-187)
//    factorises: P1t[(i1 i2),(i3 i4)] = P1 u_k[i1] conj(u_k[i2]) conj(u_k'[i3])
//    u_k'[i4] with u = e^{i e t}.  So rho is rotated to the Schroedinger
//    picture once per k (nk nb^2 sincos instead of 2 nb^4 per pair), P is
//    streamed phase free, and the result is rotated back in the epilogue.
//  * Row kernel: one 64-wide wavefront per destination row; lanes map to the
//    elements of [P1 | P2] so each load instruction moves 1 KiB of P fully
//    coalesced (nb = 2: two pairs per instruction); partner rho rows are
//    gathered from L2; per-lane FP64 partial sums are reduced once per row in
//    LDS and the Hermitian epilogue is fused.  Deterministic: fixed order.
//  * Time enters through a device-resident clock so a whole RK4 step is
//    captured once in a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "dmd_internal.h"
#include "synth.h"

namespace dmd {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

bool is_device_ptr(const void *p)
{
    if (!p) return false;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice;
}

// Hermitian completion from the upper triangle with a real diagonal, the
// way CBLAS zhemm(CblasUpper) reads A.
template <int NB>
__device__ __forceinline__ double2 herm_up(const double2 *A, int i, int j)
{
    if (i < j) return A[i * NB + j];
    if (i > j) return cconj(A[j * NB + i]);
    return make_double2(A[i * NB + i].x, 0.0);
}

// ---------------------------------------------------------------------------
// rho -> Schroedinger picture: rhoS[i,j] = rho[i,j] cis(-(e_i - e_j) t)
// (window rows/cols [b0, b0+NB) of an nbd x nbd state).  With a separate
// 1-rho array (the reference's dm1 argument) it is rotated alike.
// ---------------------------------------------------------------------------
// fsub (the linearized operator's f_eq, nk x NB): rhoS = rho_S - diag(fsub).
template <int NB>
__global__ void k_rotate_in(int64_t nk, int nbd, int b0, const double2 *__restrict__ rho,
                            const double2 *__restrict__ rho1, const double *__restrict__ e,
                            const DClock *__restrict__ clk, double a, int expt,
                            double2 *__restrict__ rhoS, double2 *__restrict__ rho1S, const double *__restrict__ fsub,
                            int semic)
{
    constexpr int NB2 = NB * NB;
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nk * NB2) return;
    int64_t k = idx / NB2;
    int ij = (int)(idx - k * NB2), i = ij / NB, j = ij - i * NB;
    int64_t src = k * (int64_t)nbd * nbd + (int64_t)(b0 + i) * nbd + (b0 + j);
    double2 v = rho[src];
    double2 ph = make_double2(1.0, 0.0);
    bool rot = expt && i != j;
    if (rot) ph = ccis(-(e[k * NB + i] - e[k * NB + j]) * clock_t_at(clk, a));
    v = rot ? cmul(v, ph) : v;
    if (fsub && i == j) v.x -= fsub[k * NB + i];
    if (semic && i != j) v = make_double2(0.0, 0.0);  // alg_semiclassical (DensityMatrix_Dynamics.h:249)
    rhoS[idx] = v;
    if (rho1) {
        double2 w = rho1[src];
        rho1S[idx] = rot ? cmul(w, ph) : w;
    }
}

// ---------------------------------------------------------------------------
// Row kernel (deterministic): one wavefront per destination row k.
// Lane roles: element e of a pair's [P1 | P2] (2 NB^4 complex).  For
// 2 NB^4 <= 64 a load instruction covers 64/(2 NB^4) pairs; otherwise a
// pair spans J = ceil(2 NB^4 / 64) registers per lane.
// Forward pair (k, k'):  P1 lane (i12,i34): af += P1 * rhoS_k'[i34]      -> Pdm[i12]
//                        P2 lane (i12,i34): af += (1-rhoS_k')[i12] * P2  -> dm1P[i34]
// Reverse pair (k'', k): P1 lane: ar += (1-rhoS_k'')[i12] * conj(P1)     -> dm1P[i34]
//                        P2 lane: ar += conj(P2) * rhoS_k''[i34]         -> Pdm[i12]
// ---------------------------------------------------------------------------
template <int NB>
struct Shape {
    static constexpr int NB2 = NB * NB, NB4 = NB2 * NB2, EP = 2 * NB4;
    static constexpr int PPI = EP <= 64 ? 64 / EP : 1;  // pairs per load instruction
    static constexpr int J = EP <= 64 ? 1 : (EP + 63) / 64;
    static constexpr int SLOTS = EP <= 64 ? PPI * EP : 64 * J;  // LDS slots per wave
};

template <int NB, bool DM1>
__device__ __forceinline__ double2 rho_elem(const double2 *__restrict__ rhoS, const double2 *__restrict__ rho1S,
                                            int64_t kp, int idx, bool one)
{
    constexpr int NB2 = NB * NB;
    if (DM1 && one) return rho1S[kp * NB2 + idx];
    double2 x = rhoS[kp * NB2 + idx];
    if (!one) return x;
    bool diag = (idx % (NB + 1)) == 0;
    return make_double2((diag ? 1.0 : 0.0) - x.x, -x.y);
}

template <int NB, bool DM1>
__global__ __launch_bounds__(256) void k_rows(int64_t nk, const double2 *__restrict__ P,
                                              const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                              const int64_t *__restrict__ fwdpos,
                                              const int64_t *__restrict__ revptr,
                                              const int64_t *__restrict__ revpos,
                                              const int32_t *__restrict__ revcol,
                                              const double2 *__restrict__ rhoS, const double2 *__restrict__ rho1S,
                                              const double *__restrict__ e, const DClock *__restrict__ clk,
                                              double a, int expt, int schrodinger, double prefac,
                                              const double2 *__restrict__ eq, double2 *__restrict__ out)
{
    using S = Shape<NB>;
    constexpr int NB2 = S::NB2, NB4 = S::NB4, EP = S::EP, PPI = S::PPI, J = S::J, SLOTS = S::SLOTS;
    __shared__ double2 red[4][2][SLOTS];
    __shared__ double2 vec[4][4][NB2];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t k = (int64_t)blockIdx.x * 4 + wv;
    const bool active = k < nk;

    const int slot = EP <= 64 ? lane / EP : 0;
    int eidx[J], fidx[J], ridx[J];
    bool fone[J], rone[J], valid[J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        int el = EP <= 64 ? lane % EP : lane + 64 * j;
        valid[j] = EP <= 64 ? true : el < EP;
        if (!valid[j]) el = 0;
        int which = el / NB4, r = el - which * NB4, i12 = r / NB2, i34 = r - i12 * NB2;
        eidx[j] = el;
        fidx[j] = which == 0 ? i34 : i12;
        fone[j] = which == 1;
        ridx[j] = which == 0 ? i12 : i34;
        rone[j] = which == 0;
    }
    double2 af[J], ar[J];
#pragma unroll
    for (int j = 0; j < J; j++) { af[j] = make_double2(0, 0); ar[j] = make_double2(0, 0); }

    if (active) {
        // forward pairs of row k: contiguous P
        const int64_t p0 = rowptr[k], p1 = rowptr[k + 1];
        constexpr int U = EP <= 64 ? 4 : 1;
        for (int64_t q = p0; q < p1; q += PPI * U) {
            double2 pv[U][J];
            double2 rv[U][J];
            bool ok[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                int64_t it = q + (int64_t)u * PPI + slot;
                ok[u] = it < p1;
                int64_t kp = ok[u] ? (int64_t)col[it] : 0;
                int64_t pair = ok[u] ? fwdpos[it] : 0;
#pragma unroll
                for (int j = 0; j < J; j++) {
                    bool v = ok[u] && valid[j];
                    pv[u][j] = v ? P[pair * EP + eidx[j]] : make_double2(0, 0);
                    rv[u][j] = v ? rho_elem<NB, DM1>(rhoS, rho1S, kp, fidx[j], fone[j]) : make_double2(0, 0);
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int j = 0; j < J; j++) af[j] = cadd(af[j], cmul(pv[u][j], rv[u][j]));
        }
        // reverse pairs (k'', k), k'' < k: P at their own rows
        const int64_t r0 = revptr[k], r1 = revptr[k + 1];
        for (int64_t q = r0; q < r1; q += PPI * U) {
            double2 pv[U][J];
            double2 rv[U][J];
#pragma unroll
            for (int u = 0; u < U; u++) {
                int64_t it = q + (int64_t)u * PPI + slot;
                bool ok = it < r1;
                int64_t pos = ok ? revpos[it] : 0;
                int64_t kpp = ok ? (int64_t)revcol[it] : 0;
#pragma unroll
                for (int j = 0; j < J; j++) {
                    bool v = ok && valid[j];
                    pv[u][j] = v ? P[pos * EP + eidx[j]] : make_double2(0, 0);
                    rv[u][j] = v ? rho_elem<NB, DM1>(rhoS, rho1S, kpp, ridx[j], rone[j]) : make_double2(0, 0);
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int j = 0; j < J; j++) ar[j] = cadd(ar[j], cmulc(rv[u][j], pv[u][j]));
        }
    }
    // per-row reduction of the lane partials
#pragma unroll
    for (int j = 0; j < J; j++) {
        int s = EP <= 64 ? lane : lane + 64 * j;
        if (s < SLOTS) { red[wv][0][s] = af[j]; red[wv][1][s] = ar[j]; }
    }
    __syncthreads();
    if (lane < NB2) {
        double2 pdm = make_double2(0, 0), d1p = make_double2(0, 0);
        for (int sl = 0; sl < PPI; sl++) {
            const int base = sl * EP;
            for (int m = 0; m < NB2; m++) {
                // Pdm[lane=i12]: forward P1 (i12, m) + reverse P2 (i12, m)
                pdm = cadd(pdm, cadd(red[wv][0][base + lane * NB2 + m], red[wv][1][base + NB4 + lane * NB2 + m]));
                // dm1P[lane=i34]: forward P2 (m, i34) + reverse P1 (m, i34)
                d1p = cadd(d1p, cadd(red[wv][0][base + NB4 + m * NB2 + lane], red[wv][1][base + m * NB2 + lane]));
            }
        }
        vec[wv][0][lane] = pdm;
        vec[wv][1][lane] = d1p;
        if (active) {
            vec[wv][2][lane] = rhoS[k * NB2 + lane];
            vec[wv][3][lane] = rho_elem<NB, DM1>(rhoS, rho1S, k, lane, true);
        }
    }
    __syncthreads();
    if (!active || lane >= NB2) return;
    // epilogue: C = (1-rho) Pdm - dm1P rho (zhemm upper), Hermitian part, rotate back
    const int i = lane / NB, j = lane - i * NB;
    const double2 *pdm = vec[wv][0], *d1p = vec[wv][1], *rk = vec[wv][2], *r1k = vec[wv][3];
    double2 cij = make_double2(0, 0), cji = make_double2(0, 0);
#pragma unroll
    for (int m = 0; m < NB; m++) {
        cij = cadd(cij, cmul(herm_up<NB>(r1k, i, m), pdm[m * NB + j]));
        cji = cadd(cji, cmul(herm_up<NB>(r1k, j, m), pdm[m * NB + i]));
    }
#pragma unroll
    for (int m = 0; m < NB; m++) {
        cij = csub(cij, cmul(d1p[i * NB + m], herm_up<NB>(rk, m, j)));
        cji = csub(cji, cmul(d1p[j * NB + m], herm_up<NB>(rk, m, i)));
    }
    double2 d = cadd(cij, cconj(cji));
    d = make_double2(d.x * (prefac * 0.5), d.y * (prefac * 0.5));
    const double t = clock_t_at(clk, a);
    const double phase_arg = (e[k * NB + i] - e[k * NB + j]) * t;
    if (expt && i != j) d = cmul(d, ccis(phase_arg));
    if (eq) {
        double2 q = eq[k * NB2 + lane];
        if (i != j && !schrodinger) q = cmul(q, ccis(phase_arg));
        d = csub(d, q);
    }
    out[k * NB2 + lane] = d;
}

// ---------------------------------------------------------------------------
// Single-read kernel: each stored pair's P is streamed once.  One wavefront
// per row k walks the forward pairs (k, k') (contiguous P):
//   forward contributions to row k stay in registers (as in the row kernel);
//   reverse contributions to row k' (k < k') are reduced across the pair's
//   lanes with shuffles to 2 NB^2 complex and added with FP64 atomics into a
//   per-row accumulator acc[k'] = [Pdm | dm1P] (NB = 2: one 128-byte line),
//   which stays resident in L2 / Infinity Cache.
// k_assemble then forms the Hermitian epilogue per row and re-zeros acc, so
// no memset is needed between evaluations (safe inside a captured graph).
// Summation order of the atomics is not fixed: results agree with the
// deterministic row kernel to rounding.
// ---------------------------------------------------------------------------
// Reverse contributions of one pair: summed over the pair's lanes, then added
// with FP64 atomics into row k' of the accumulator ([Pdm | dm1P], 4 NB^2
// doubles).  `row` is that row (null for lanes whose pair has no reverse term).
template <int NB>
struct RevAdd;
// NB = 1 (2 lanes per pair): lane e=0 holds P1 -> dm1P[0], e=1 holds P2 -> Pdm[0].
template <>
struct RevAdd<1> {
    __device__ static void run(double2 v, int lane, double *row)
    {
        const int e = lane & 1;
        if (!row) return;
        double *a = row + (e ? 0 : 2);
        unsafeAtomicAdd(a, v.x);
        unsafeAtomicAdd(a + 1, v.y);
    }
};
// NB = 2 (32 lanes per pair): P1 lanes (e<16) sum over i12 (xor 4, 8) ->
// dm1P[i34] held at e = i34; P2 lanes sum over i34 (xor 1, 2) -> Pdm[i12]
// held at e = 16 + 4 i12.  Lane L < 16 then takes component L&1 of target
// L>>1 so one atomic instruction covers the row's 128-byte line.
template <>
struct RevAdd<2> {
    __device__ static void run(double2 x, int lane, double *row)
    {
        const int e = lane & 31, base = lane & 32;
        const bool p1 = e < 16;
        const int m1 = p1 ? 4 : 1, m2 = p1 ? 8 : 2;
        x.x += __shfl(x.x, lane ^ m1); x.y += __shfl(x.y, lane ^ m1);
        x.x += __shfl(x.x, lane ^ m2); x.y += __shfl(x.y, lane ^ m2);
        const int t = (e & 15) >> 1;  // target complex 0..7: Pdm[0..3], dm1P[0..3]
        const int src = base + (t < 4 ? 16 + 4 * t : t - 4);
        double sx = __shfl(x.x, src), sy = __shfl(x.y, src);
        if (row && e < 16) unsafeAtomicAdd(row + e, (e & 1) ? sy : sx);
    }
};

template <int NB, bool DM1>
__global__ __launch_bounds__(256) void k_fwd_atomic(int64_t nk, const double2 *__restrict__ P,
                                                    const int64_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ col,
                                                    const int64_t *__restrict__ fwdpos,
                                                    const double2 *__restrict__ rhoS,
                                                    const double2 *__restrict__ rho1S, double *__restrict__ acc)
{
    using S = Shape<NB>;
    constexpr int NB2 = S::NB2, NB4 = S::NB4, EP = S::EP, PPI = S::PPI, SLOTS = S::SLOTS;
    static_assert(EP <= 64, "single-read kernel covers nb <= 2; nb = 4 uses k_fwd_atomic4");
    __shared__ double2 red[4][SLOTS];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t k = (int64_t)blockIdx.x * 4 + wv;
    const bool active = k < nk;
    const int slot = lane / EP, el = lane % EP;
    const int which = el / NB4, r = el - which * NB4, i12 = r / NB2, i34 = r - i12 * NB2;
    const int fidx = which == 0 ? i34 : i12, ridx = which == 0 ? i12 : i34;
    const bool fone = which == 1, rone = which == 0;
    double2 af = make_double2(0, 0);
    if (active) {
        // reverse factor: rho of this row (the pair's k), fixed for the row
        const double2 rrow = rho_elem<NB, DM1>(rhoS, rho1S, k, ridx, rone);
        const int64_t p0 = rowptr[k], p1 = rowptr[k + 1];
        constexpr int U = 4;
        for (int64_t q = p0; q < p1; q += PPI * U) {
            double2 pv[U], rv[U];
            int64_t kp[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                int64_t it = q + (int64_t)u * PPI + slot;
                bool ok = it < p1;
                kp[u] = ok ? (int64_t)col[it] : -1;
                int64_t pair = ok ? fwdpos[it] : 0;
                pv[u] = ok ? P[pair * EP + el] : make_double2(0, 0);
                rv[u] = ok ? rho_elem<NB, DM1>(rhoS, rho1S, kp[u], fidx, fone) : make_double2(0, 0);
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                af = cadd(af, cmul(pv[u], rv[u]));
                // reverse: P1 lane -> (1-rho_k)[i12] conj(P1), P2 lane -> conj(P2) rho_k[i34]
                double2 vr = kp[u] > k ? cmulc(rrow, pv[u]) : make_double2(0, 0);
                RevAdd<NB>::run(vr, lane, kp[u] > k ? acc + kp[u] * (4 * NB2) : nullptr);
            }
        }
    }
    red[wv][lane] = af;  // SLOTS == 64 for EP <= 64
    __syncthreads();
    if (!active || lane >= 2 * NB2) return;
    // forward sums: lane < NB2 -> Pdm[lane] (P1 rows), else dm1P[lane-NB2] (P2 columns)
    double2 s = make_double2(0, 0);
    for (int sl = 0; sl < PPI; sl++)
        for (int m = 0; m < NB2; m++)
            s = cadd(s, lane < NB2 ? red[wv][sl * EP + lane * NB2 + m] : red[wv][sl * EP + NB4 + m * NB2 + (lane - NB2)]);
    double *a = acc + k * (4 * NB2) + 2 * lane;
    unsafeAtomicAdd(a, s.x);
    unsafeAtomicAdd(a + 1, s.y);
}

// ---------------------------------------------------------------------------
// Tile kernel (nb = 2), lane per pair.  Rows are cut into blocks of
// R = DMD_TILE_R; the stored pairs of tile (I, J) (k in block I, k' in block
// J) are contiguous in HBM, padded to granules of 8 pairs, and each granule
// is stored structure-of-arrays: element j of its 8 pairs is one 128-byte
// line.  A wavefront takes 64 consecutive pairs (8 granules), one per lane,
// and loads them with 32 fully coalesced 1 KiB instructions (32 KiB in flight
// per wave).  Each lane then forms its pair's four GEMVs in registers
//   forward  Pdm_k[i]  = sum_m P1[i,m] rho_k'[m]   dm1P_k[j]  = sum_m (1-rho_k')[m] P2[m,j]
//   reverse  Pdm_k'[i] = sum_m conj(P2[i,m]) rho_k[m]  dm1P_k'[j] = sum_m (1-rho_k)[m] conj(P1[m,j])
// (no cross-lane traffic) and adds them into LDS row accumulators of block I
// and block J (ds_add_f64).  The chunk's two R-row blocks are flushed once
// with global FP64 atomics (~2R lines per chunk instead of one per pair).
// Padding pairs carry P = 0 and add exact zeros.
// ---------------------------------------------------------------------------
// Hermitian increment of one pair's contribution to one row (nb = 2): with
// pdm, d1p the pair's Pdm and dm1P terms for row k and r, r1 = rho_k, 1-rho_k,
// C = (1-rho) mat(pdm) - mat(d1p) rho (zhemm upper) and H = C + C^+ as
// [H00, H11, Re H01, Im H01].  C is linear in (Pdm, dm1P) for a fixed rho_k,
// so the sum of the pairs' H is the row's H (the reference sums Pdm, dm1P
// first and assembles once, evolve_eph.cpp:140-151; the initializer assembles
// per pair, lindbladInit.../ElectronPhonon.h:855-862).
__device__ __forceinline__ void herm_inc(const double2 *pdm, const double2 *d1p, const double2 *r,
                                         const double2 *r1, double hv[4])
{
    constexpr int NB = 2;
    double2 r1loc[4];
    if (!r1) {
#pragma unroll
        for (int q = 0; q < 4; q++) r1loc[q] = one_minus(r, q);
        r1 = r1loc;
    }
    double2 c[4];
#pragma unroll
    for (int i = 0; i < NB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) {
            double2 x = make_double2(0, 0);
#pragma unroll
            for (int q = 0; q < NB; q++) cfma(x, herm_up<NB>(r1, i, q), pdm[q * NB + j]);
#pragma unroll
            for (int q = 0; q < NB; q++) cfma(x, make_double2(-d1p[i * NB + q].x, -d1p[i * NB + q].y), herm_up<NB>(r, q, j));
            c[i * NB + j] = x;
        }
    hv[0] = 2.0 * c[0].x;     // H00 = C00 + conj(C00)
    hv[1] = 2.0 * c[3].x;     // H11
    hv[2] = c[1].x + c[2].x;  // H01 = C01 + conj(C10)
    hv[3] = c[1].y - c[2].y;
}

// Epilogue from the Hermitian accumulator (nb = 2): ddm = prefac/2 H,
// rotated back to the interaction picture, minus ddmdt_eq; zeroes hacc[k].
__global__ __launch_bounds__(256) void k_assemble_h(int64_t nk, double *__restrict__ hacc,
                                                    const double *__restrict__ e, const DClock *__restrict__ clk,
                                                    double a, int expt, int schrodinger, double prefac,
                                                    const double2 *__restrict__ eq, double2 *__restrict__ out)
{
    constexpr int NB = 2, NB2 = 4;
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    double *h = hacc + k * 4;
    const double h00 = h[0], h11 = h[1], hr = h[2], hi = h[3];
    h[0] = 0.0; h[1] = 0.0; h[2] = 0.0; h[3] = 0.0;
    const double s = prefac * 0.5;
    double2 d[NB2] = { make_double2(h00 * s, 0.0), make_double2(hr * s, hi * s), make_double2(hr * s, -hi * s),
                       make_double2(h11 * s, 0.0) };
    const double t = clock_t_at(clk, a);
    const double arg = (e[k * NB] - e[k * NB + 1]) * t;  // (e_0 - e_1) t for element (0,1)
    if (expt) {
        d[1] = cmul(d[1], ccis(arg));
        d[2] = cmul(d[2], ccis(-arg));
    }
    if (eq) {
#pragma unroll
        for (int ij = 0; ij < NB2; ij++) {
            double2 q = eq[k * NB2 + ij];
            if (ij == 1 && !schrodinger) q = cmul(q, ccis(arg));
            if (ij == 2 && !schrodinger) q = cmul(q, ccis(-arg));
            d[ij] = csub(d[ij], q);
        }
    }
#pragma unroll
    for (int ij = 0; ij < NB2; ij++) out[k * NB2 + ij] = d[ij];
}

// alg_linearize_dPee (evolve_eph.cpp:144-151): out[k][ij] += H[k][ij]
// cis((e_i - e_j) t), H = prefac/2 (D + D^+) the Hermitian part of the row term
// D_ij = (1 - f_i) dPdm_ij - dm1dP_ij f_j formed when dP was built; the phase
// is applied in either picture, as the reference does.
__global__ __launch_bounds__(256) void k_dP_term(int64_t n, int nb, const double2 *__restrict__ H,
                                                 const double *__restrict__ e, const DClock *__restrict__ clk,
                                                 double a, double2 *__restrict__ out)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int nb2 = nb * nb;
    const int64_t k = q / nb2;
    const int ij = (int)(q - k * nb2), i = ij / nb, j = ij % nb;
    const double t = clock_t_at(clk, a);
    const double2 v = i == j ? H[q] : cmul(H[q], ccis((e[k * nb + i] - e[k * nb + j]) * t));
    out[q] = cadd(out[q], v);
}

// LIN: the slot holds the linearized operator Lscij | Lscji and rhoS holds
// drho_S; forward X = Lscij drho_k', reverse Y = Lscji drho_k, increments
// X + X^+ and Y + Y^+ (evolve_eph_linear.cpp:26-37).
// SP: P held compressed (dmd_scatt_compress_P): per slot a mask of its
// nonzero elements (bit el < 16: P1 element el, bit 16 + el: P2 element el),
// per granule of 8 slots the offset of its values, stored element-major (for
// each element, the values of the granule's slots that have it, in slot
// order), so the lanes of a granule read one element's values together.  The
// lane rebuilds its pair's dense P1, P2 in registers (zeros where absent) and
// runs the dense GEMVs.
template <bool DM1, bool LIN, bool SP = false>
__global__ __launch_bounds__(256, SP ? 2 : 1) void k_tiles_soa(int64_t nk, const double2 *__restrict__ P,
                                                   const uint32_t *__restrict__ pk,
                                                   const int64_t *__restrict__ cbeg,
                                                   const int32_t *__restrict__ clen, const int2 *__restrict__ cij,
                                                   const int32_t *__restrict__ gbeg,
                                                   const double2 *__restrict__ rhoS,
                                                   const double2 *__restrict__ rho1S, double *__restrict__ hacc,
                                                   const uint32_t *__restrict__ pmask = nullptr,
                                                   const uint32_t *__restrict__ gofs = nullptr)
{
    constexpr int NB2 = 4, NB4 = 16, EP = 32, R = DMD_TILE_R;  // nb = 2
    constexpr int W = 4, RPT = R / 256;  // rows per thread in the flush
    // LDS: Hermitian row increments of blocks I and J, structure of arrays
    // (double d of row r at d*R + r, conflict-free for the flush), and the
    // rho rows of block J, gathered per pair from LDS instead of L2.
    __shared__ double accA[4 * R], accB[4 * R];
    __shared__ double2 rhoJ[R * NB2];
    const int c0 = gbeg[blockIdx.x], c1 = gbeg[blockIdx.x + 1];
    const int I = cij[c0].x;
    const int64_t I0 = (int64_t)I * R;
    for (int i = threadIdx.x; i < 4 * R; i += blockDim.x) { accA[i] = 0.0; accB[i] = 0.0; }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int c = c0; c < c1; c++) {
        const int64_t b = cbeg[c];
        const int n = clen[c];
        const int m = (n + 63) / 64;  // blocks of 64 slots; lane l owns chunk pairs l*m .. l*m+m-1
        const int J = cij[c].y;
        const int64_t J0 = (int64_t)J * R;
        double *aJ = (I == J) ? accA : accB;
        if (!DM1) {
            for (int i = threadIdx.x; i < R * NB2; i += blockDim.x)
                rhoJ[i] = J0 * NB2 + i < nk * NB2 ? rhoS[J0 * NB2 + i] : make_double2(0, 0);
        }
        __syncthreads();
        const int mq = (m + W - 1) / W;
        const int i0 = wv * mq, i1 = min(m, i0 + mq);
        uint32_t code_next = (i0 < i1 && lane * m + i0 < n) ? ld_stream(pk + b + (int64_t)i0 * 64 + lane) : 0u;
        // SP: the slot's mask and its granule's value offset, one iteration ahead
        uint32_t mask_next = 0u, gofs_next = 0u;
        if (SP && i0 < i1) {
            const int64_t s0 = b + (int64_t)i0 * 64 + lane;
            mask_next = ld_stream(pmask + s0);
            gofs_next = ld_stream(gofs + (s0 >> 3));
        }
        // rho of the lane's current row k: a lane walks a run of one k, so it is
        // reloaded only where k changes
        int lo1_cur = -1;
        double2 rk[NB2], r1k[NB2];
#pragma unroll
        for (int q = 0; q < NB2; q++) { rk[q] = make_double2(0, 0); r1k[q] = make_double2(0, 0); }
        for (int it = i0; it < i1; it++) {
            const bool ok = lane * m + it < n;
            // idle lanes read the chunk's first slot (always valid) and add
            // nothing; with SP every lane keeps its own slot (a padding slot,
            // mask 0, for idle lanes), so the 8 lanes of a granule are its 8 slots
            const int64_t s = b + ((ok || SP) ? (int64_t)it * 64 + lane : 0);
            const uint32_t code = code_next;
            code_next = (it + 1 < i1 && lane * m + it + 1 < n) ? ld_stream(pk + b + (int64_t)(it + 1) * 64 + lane) : 0u;
            uint32_t mask = 0u, vofs = 0u;
            if (SP) {
                mask = mask_next;
                vofs = gofs_next;
                if (it + 1 < i1) {
                    const int64_t s1 = b + (int64_t)(it + 1) * 64 + lane;
                    mask_next = ld_stream(pmask + s1);
                    gofs_next = ld_stream(gofs + (s1 >> 3));
                }
            }
            const int lo1 = code % R, lo2 = code / R;
            const int64_t k = I0 + lo1, kp = J0 + lo2;
            const double2 *g = P + (s >> 3) * (8 * EP) + (s & 7);
            double2 rkp[NB2];
#pragma unroll
            for (int q = 0; q < NB2; q++) rkp[q] = DM1 ? rhoS[kp * NB2 + q] : rhoJ[lo2 * NB2 + q];
            if (ok && lo1 != lo1_cur) {
#pragma unroll
                for (int q = 0; q < NB2; q++) rk[q] = rhoS[k * NB2 + q];
                if (DM1) {
#pragma unroll
                    for (int q = 0; q < NB2; q++) r1k[q] = rho1S[k * NB2 + q];
                }
                lo1_cur = lo1;
            }
            double2 r1kp[NB2];
            if (DM1) {
#pragma unroll
                for (int q = 0; q < NB2; q++) r1kp[q] = rho1S[kp * NB2 + q];
            }
            // 1-rho: explicit rows with DM1, else formed from rho where used
            auto om_k = [&](int q) { return DM1 ? r1k[q] : one_minus(rk, q); };
            auto om_kp = [&](int q) { return DM1 ? r1kp[q] : one_minus(rkp, q); };
            const bool rev = ok && k < kp;
            // forward Pdm_k = P1 rho_k', dm1P_k = (1-rho_k')^T P2 (f1, f2);
            // reverse Pdm_k' = conj(P2) rho_k, dm1P_k' = (1-rho_k)^T conj(P1) (g1, g2)
            double2 f1[NB2], f2[NB2], g1[NB2], g2[NB2];
            double2 p1[NB4], p2[NB4];
            if (SP) {
                // element-major values of the granule: for element el = 0 .. 31
                // (P1 then P2) the values of the granule's slots that have it,
                // in slot order.  One ballot per element gives each lane its
                // place; an element a slot does not have is zero, so the dense
                // GEMVs below run unchanged.
                const double2 *gv = P + vofs;
                const int gsh = lane & 56;
                const uint32_t below = (1u << (lane & 7)) - 1u;
                int base = 0;
#pragma unroll
                for (int el = 0; el < 2 * NB4; el++) {
                    const bool has = (mask >> el) & 1u;
                    const uint32_t gb = (uint32_t)(__ballot(has) >> gsh) & 0xffu;
                    double2 v = make_double2(0, 0);
                    if (has) v = ld_stream(gv + base + __popc(gb & below));
                    base += __popc(gb);
                    if (el < NB4) p1[el] = v;
                    else p2[el - NB4] = v;
                }
            } else {
#pragma unroll
                for (int el = 0; el < NB4; el++) p1[el] = ld_stream(g + el * 8);
#pragma unroll
                for (int el = 0; el < NB4; el++) p2[el] = ld_stream(g + (NB4 + el) * 8);
            }
            {
                if (LIN) {
                    double2 x[NB2], y[NB2];
#pragma unroll
                    for (int i = 0; i < NB2; i++) {
                        x[i] = make_double2(0, 0);
                        y[i] = make_double2(0, 0);
#pragma unroll
                        for (int q = 0; q < NB2; q++) {
                            cfma(x[i], p1[i * NB2 + q], rkp[q]);  // X_k[i]  += Lscij[i,q] drho_k'[q]
                            cfma(y[i], p2[i * NB2 + q], rk[q]);   // Y_k'[i] += Lscji[i,q] drho_k[q]
                        }
                    }
                    double hf[4], hr[4];
                    herm_of(x, hf);
                    herm_of(y, hr);
                    if (ok) {
#pragma unroll
                        for (int d = 0; d < 4; d++) atomicAdd(&accA[d * R + lo1], hf[d]);
                    }
                    if (rev) {
#pragma unroll
                        for (int d = 0; d < 4; d++) atomicAdd(&aJ[d * R + lo2], hr[d]);
                    }
                    continue;
                }
                // forward (row k)
                {
#pragma unroll
                    for (int i = 0; i < NB2; i++) {
                        f1[i] = make_double2(0, 0);
                        f2[i] = make_double2(0, 0);
#pragma unroll
                        for (int q = 0; q < NB2; q++) {
                            cfma(f1[i], p1[i * NB2 + q], rkp[q]);      // Pdm_k[i]  += P1[i,q] rho_k'[q]
                            cfma(f2[i], om_kp(q), p2[q * NB2 + i]);    // dm1P_k[i] += (1-rho_k')[q] P2[q,i]
                        }
                    }
                    double hf[4];
                    herm_inc(f1, f2, rk, DM1 ? r1k : nullptr, hf);
                    if (ok) {
#pragma unroll
                        for (int d = 0; d < 4; d++) atomicAdd(&accA[d * R + lo1], hf[d]);
                    }
                }
                // reverse (row k')
                {
#pragma unroll
                    for (int i = 0; i < NB2; i++) {
                        g1[i] = make_double2(0, 0);
                        g2[i] = make_double2(0, 0);
#pragma unroll
                        for (int q = 0; q < NB2; q++) {
                            cfmac(g1[i], rk[q], p2[i * NB2 + q]);      // Pdm_k'[i]  += rho_k[q] conj(P2[i,q])
                            cfmac(g2[i], om_k(q), p1[q * NB2 + i]);    // dm1P_k'[i] += (1-rho_k)[q] conj(P1[q,i])
                        }
                    }
                    double hr[4];
                    herm_inc(g1, g2, rkp, DM1 ? r1kp : nullptr, hr);
                    if (rev) {
#pragma unroll
                        for (int d = 0; d < 4; d++) atomicAdd(&aJ[d * R + lo2], hr[d]);
                    }
                }
            }
        }
        __syncthreads();  // block J done
        if (I != J) {  // flush and clear block J
#pragma unroll
            for (int t = 0; t < RPT; t++) {
                const int r = t * 256 + threadIdx.x;
                double hv[4];
#pragma unroll
                for (int d = 0; d < 4; d++) { hv[d] = accB[d * R + r]; accB[d * R + r] = 0.0; }
                flush_herm_wave(hv, lane, J0 + (r & ~63), nk, hacc);
            }
        }
        __syncthreads();  // rhoJ and accB free for the next chunk
    }
#pragma unroll
    for (int t = 0; t < RPT; t++) {
        const int r = t * 256 + threadIdx.x;
        double hv[4];
#pragma unroll
        for (int d = 0; d < 4; d++) hv[d] = accA[d * R + r];
        flush_herm_wave(hv, lane, I0 + (r & ~63), nk, hacc);
    }
}

// In-place transpose of every 8-pair granule between pair-contiguous
// ([pair][element]) and structure-of-arrays ([element][pair]) storage.
__global__ void k_relayout(int64_t ngran, int ep, int to_soa, double2 *__restrict__ P)
{
    extern __shared__ double2 buf[];
    const int64_t g = blockIdx.x;
    if (g >= ngran) return;
    double2 *base = P + g * 8 * ep;
    const int n = 8 * ep;
    for (int i = threadIdx.x; i < n; i += blockDim.x) buf[i] = base[i];
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        // i indexes the destination layout
        int pr, el;
        if (to_soa) { el = i / 8; pr = i - el * 8; }
        else { pr = i / ep; el = i - pr * ep; }
        base[i] = to_soa ? buf[pr * ep + el] : buf[el * 8 + pr];
    }
}

// Epilogue from the accumulator: C = (1-rho) Pdm - dm1P rho (zhemm upper),
// ddm = prefac/2 (C + C^+), rotated back, minus ddmdt_eq; zeroes acc[k].
template <int NB, bool DM1>
__global__ __launch_bounds__(256) void k_assemble(int64_t nk, double *__restrict__ acc,
                                                  const double2 *__restrict__ rhoS,
                                                  const double2 *__restrict__ rho1S, const double *__restrict__ e,
                                                  const DClock *__restrict__ clk, double a, int expt, int schrodinger,
                                                  double prefac, const double2 *__restrict__ eq,
                                                  double2 *__restrict__ out)
{
    constexpr int NB2 = NB * NB;
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    double2 pdm[NB2], d1p[NB2], rk[NB2], r1k[NB2];
    double *ak = acc + k * (4 * NB2);
#pragma unroll
    for (int m = 0; m < NB2; m++) {
        pdm[m] = make_double2(ak[2 * m], ak[2 * m + 1]);
        d1p[m] = make_double2(ak[2 * (NB2 + m)], ak[2 * (NB2 + m) + 1]);
        rk[m] = rhoS[k * NB2 + m];
        r1k[m] = rho_elem<NB, DM1>(rhoS, rho1S, k, m, true);
    }
#pragma unroll
    for (int m = 0; m < 4 * NB2; m++) ak[m] = 0.0;
    const double t = clock_t_at(clk, a);
#pragma unroll
    for (int i = 0; i < NB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) {
            double2 cij = make_double2(0, 0), cji = make_double2(0, 0);
#pragma unroll
            for (int m = 0; m < NB; m++) {
                cij = cadd(cij, cmul(herm_up<NB>(r1k, i, m), pdm[m * NB + j]));
                cji = cadd(cji, cmul(herm_up<NB>(r1k, j, m), pdm[m * NB + i]));
            }
#pragma unroll
            for (int m = 0; m < NB; m++) {
                cij = csub(cij, cmul(d1p[i * NB + m], herm_up<NB>(rk, m, j)));
                cji = csub(cji, cmul(d1p[j * NB + m], herm_up<NB>(rk, m, i)));
            }
            double2 d = cadd(cij, cconj(cji));
            d = make_double2(d.x * (prefac * 0.5), d.y * (prefac * 0.5));
            const double arg = (e[k * NB + i] - e[k * NB + j]) * t;
            if (expt && i != j) d = cmul(d, ccis(arg));
            if (eq) {
                double2 q = eq[k * NB2 + i * NB + j];
                if (i != j && !schrodinger) q = cmul(q, ccis(arg));
                d = csub(d, q);
            }
            out[k * NB2 + i * NB + j] = d;
        }
}

// ---------------------------------------------------------------------------
// P placement: file order -> sorted layout, scaled; mode 0 set, 1 add.
// ---------------------------------------------------------------------------
__global__ void k_place_P(int64_t n, int64_t first, int nb4, const double2 *__restrict__ P1,
                          const double2 *__restrict__ P2, const int64_t *__restrict__ inv, double scale, int add,
                          int soa, double2 *__restrict__ Pd)
{
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t tot = n * 2 * nb4;
    if (idx >= tot) return;
    int64_t f = idx / (2 * nb4);
    int el = (int)(idx - f * 2 * nb4);
    int64_t pos = inv ? inv[first + f] : first + f;
    double2 v = el < nb4 ? P1[f * nb4 + el] : P2[f * nb4 + el - nb4];
    v = make_double2(v.x * scale, v.y * scale);
    double2 *d = Pd + paddr(pos, el, 2 * nb4, soa);
    *d = add ? cadd(*d, v) : v;
}
__global__ void k_fetch_P(int64_t n, int64_t first, int nb4, const double2 *__restrict__ Pd,
                          const int64_t *__restrict__ inv, int soa, double2 *__restrict__ P1,
                          double2 *__restrict__ P2)
{
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t tot = n * 2 * nb4;
    if (idx >= tot) return;
    int64_t f = idx / (2 * nb4);
    int el = (int)(idx - f * 2 * nb4);
    int64_t pos = inv ? inv[first + f] : first + f;
    double2 v = Pd[paddr(pos, el, 2 * nb4, soa)];
    if (el < nb4) P1[f * nb4 + el] = v;
    else P2[f * nb4 + el - nb4] = v;
}

template <int NB>
__global__ void k_generate_P(int64_t nslots, dmd_synth_params p, const int32_t *__restrict__ k1s,
                             const int32_t *__restrict__ k2s, const double *__restrict__ e, int soa,
                             double2 *__restrict__ Pd)
{
    constexpr int NB4 = NB * NB * NB * NB;
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    int64_t k = k1s[s], kp = k2s[s];
    double P1[2 * NB4], P2[2 * NB4];
    if (k >= 0) {
        synth_pair_P_nb<NB>(p, e + k * NB, e + kp * NB, k, kp, P1, P2);
    } else {  // padding slot
        for (int i = 0; i < 2 * NB4; i++) { P1[i] = 0.0; P2[i] = 0.0; }
    }
    for (int i = 0; i < NB4; i++) {
        Pd[paddr(s, i, 2 * NB4, soa)] = make_double2(P1[2 * i], P1[2 * i + 1]);
        Pd[paddr(s, NB4 + i, 2 * NB4, soa)] = make_double2(P2[2 * i], P2[2 * i + 1]);
    }
}

}  // namespace dmd

using namespace dmd;

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
namespace dmd {
void dfree(void *p)
{
    if (p) (void)hipFree(p);
}
}  // namespace dmd
dmd_scatt::~dmd_scatt()
{
    dfree(d_e); dfree(d_edm); dfree(d_P); dfree(d_inv); dfree(d_rowptr); dfree(d_revptr); dfree(d_revpos);
    dfree(d_col); dfree(d_revcol); dfree(d_k1); dfree(d_k2); dfree(d_fwdpos); dfree(d_pk); dfree(d_cbeg);
    dfree(d_clen); dfree(d_cij); dfree(d_gbeg); dfree(d_rhoS); dfree(d_rho1S); dfree(d_out); dfree(d_eq); dfree(d_dPH);
    dfree(d_in); dfree(d_in1); dfree(d_outw); dfree(d_clk);
    dfree(d_H); dfree(d_s); dfree(d_f); dfree(d_obs); dfree(d_part); dfree(d_acc_rows); dfree(d_acc_herm);
    dfree(d_y0); dfree(d_ya); dfree(d_yb); dfree(d_acc); dfree(d_w);
    dfree(d_pscale); dfree(d_fsplit); dfree(d_feq); dfree(d_x32); dfree(d_hacc16);
    dfree(d_vrec); dfree(d_bxmax); dfree(d_bqueue); dfree(d_vq); dfree(d_vsd); dfree(d_seg); dfree(d_xb); dfree(d_Fb); dfree(d_rhoSB); dfree(d_bin); dfree(d_bout);
    dfree(d_yB0); dfree(d_yBa); dfree(d_yBb); dfree(d_accB); dfree(d_wB); dfree(d_adj); dfree(d_dotpart);
    dfree(d_dpos); dfree(d_dk); dfree(d_flin); dfree(d_ad); dfree(d_adpart); dfree(d_phfeq);
    dfree(d_spmask); dfree(d_spgofs); dfree(d_spv);
    for (void *p : extra) dfree(p);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev2) (void)hipEventDestroy(ev2);
    if (st) (void)hipStreamDestroy(st);
}

namespace dmd {
// Copies n elements of host-or-device src into device dst (stream ordered).
int to_device(dmd_scatt *h, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return DMD_OK;
    DMD_HIP(hipMemcpyAsync(dst, src, bytes, is_device_ptr(src) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                           h->st));
    return DMD_OK;
}
int from_device(dmd_scatt *h, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return DMD_OK;
    DMD_HIP(hipMemcpyAsync(dst, src, bytes, is_device_ptr(dst) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                           h->st));
    return DMD_OK;
}
int set_clock(dmd_scatt *h, double t0, double dt, long long step)
{
    h->h_clk.t0 = t0; h->h_clk.dt = dt; h->h_clk.step = step; h->h_clk.pad = 0;
    DMD_HIP(hipMemcpyAsync(h->d_clk, &h->h_clk, sizeof(DClock), hipMemcpyHostToDevice, h->st));
    return DMD_OK;
}


}  // namespace dmd

extern "C" const char *dmd_scatt_last_error(void) { return g_err.c_str(); }
extern "C" int dmd_scatt_version(void) { return 1; }

extern "C" int dmd_scatt_create(int64_t nk, int nb, int64_t n_pairs, const uint64_t *k1st, const uint64_t *k2nd,
                                const double *e, double nk_full, int precision, dmd_scatt **out)
{
    if (!out) return fail(DMD_ERR_ARG, "dmd_scatt_create: out is null");
    *out = nullptr;
    if (nk < 1 || nk >= ((int64_t)1 << 31)) return fail(DMD_ERR_ARG, "dmd_scatt_create: nk out of range");
    if (nb < 1 || nb > 4) return fail(DMD_ERR_UNSUPPORTED, "dmd_scatt_create: nb must be 1..4 (got " + std::to_string(nb) + ")");
    if (n_pairs < 0 || (n_pairs > 0 && (!k1st || !k2nd))) return fail(DMD_ERR_ARG, "dmd_scatt_create: bad k-pair lists");
    if (!e) return fail(DMD_ERR_ARG, "dmd_scatt_create: e is null");
    if (!(nk_full > 0)) return fail(DMD_ERR_ARG, "dmd_scatt_create: nk_full must be positive");
    if (precision != DMD_PREC_FP64 && precision != DMD_PREC_C64 && precision != DMD_PREC_BF16)
        return fail(DMD_ERR_ARG, "dmd_scatt_create: unknown precision");
    if (precision != DMD_PREC_FP64 && nb != 4)
        return fail(DMD_ERR_UNSUPPORTED, "dmd_scatt_create: the reduced-precision forms are built for nb = 4");
    for (int64_t i = 0; i < n_pairs; i++)
        if (k1st[i] > k2nd[i] || k2nd[i] >= (uint64_t)nk)
            return fail(DMD_ERR_ARG, "dmd_scatt_create: pair " + std::to_string(i) + " is not k1st <= k2nd < nk");

    std::unique_ptr<dmd_scatt> h(new dmd_scatt());
    DMD_HIP(hipGetDevice(&h->dev));
    DMD_HIP(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
    DMD_HIP(hipEventCreate(&h->ev0));
    DMD_HIP(hipEventCreate(&h->ev1));
    DMD_HIP(hipEventCreate(&h->ev2));
    h->nk = nk; h->nb = nb; h->nb2 = nb * nb; h->nb4 = nb * nb * nb * nb; h->np = n_pairs;
    h->prefac = 2.0 * M_PI / nk_full;  // ElectronPhonon.h:61
    h->precision = precision;
    h->nbd = nb; h->b0 = 0;

    // Layouts (all O(n_pairs) counting sorts):
    //  kk order   stored pairs sorted by (k, k'): CSR rows of the forward lists
    //  physical   order of P in HBM; for nb <= 2 sorted by (k/R, k'/R, k, k')
    //             (R = DMD_TILE_R row tiles) so each tile's pairs are
    //             contiguous for the tile kernel; otherwise the kk order.
    std::vector<int64_t> cnt(nk + 1, 0);
    bool sorted = true;
    for (int64_t i = 0; i < n_pairs; i++) {
        cnt[k1st[i] + 1]++;
        if (i && (k1st[i] < k1st[i - 1] || (k1st[i] == k1st[i - 1] && k2nd[i] < k2nd[i - 1]))) sorted = false;
    }
    for (int64_t k = 0; k < nk; k++) cnt[k + 1] += cnt[k];
    std::vector<int64_t> kk;  // kk position -> file index (empty: identity)
    if (!sorted) {
        kk.resize(n_pairs);
        std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
        for (int64_t i = 0; i < n_pairs; i++) kk[cur[k1st[i]]++] = i;
        for (int64_t k = 0; k < nk; k++)
            std::stable_sort(kk.begin() + cnt[k], kk.begin() + cnt[k + 1],
                             [&](int64_t a, int64_t b) { return k2nd[a] < k2nd[b]; });
    }
    auto KK = [&](int64_t i) { return sorted ? i : kk[i]; };
    const int R = nb <= 2 ? DMD_TILE_R : (nb == 4 ? DMD_TILE4_R : 0);
    const bool lane_perm = nb <= 2;  // nb = 4: a chunk is walked by quarters, no slot permutation
    h->tile_R = R;
    // Slots: physical positions of P.  With tiles, each tile's pairs (in
    // (k, k') order) are cut into chunks of at most cmax pairs; a chunk of n
    // pairs occupies m = ceil(n/64) blocks of 64 slots, pair L of the chunk
    // going to slot (L % m) * 64 + L / m.  So lane l of a wavefront walking
    // the chunk block by block meets pairs l*m, l*m+1, ... (one k for long
    // runs; the 64 lanes sit on different rows), while each block of 64 slots
    // is still read as 8 coalesced SoA granules.  Unused slots hold P = 0 and
    // are never read.
    std::vector<int64_t> phys;  // slot -> file index, -1 for padding
    std::vector<int64_t> cbeg;
    std::vector<int32_t> clen;
    std::vector<int2> cij;
    int64_t nslots = n_pairs;
    if (R) {
        const int64_t nbk = (nk + R - 1) / R;
        const int64_t cmax = lane_perm ? std::min<int64_t>(16384, std::max<int64_t>(512, n_pairs / 4096)) / 64 * 64
                                       : DMD_TILE4_CMAX;
        const int64_t pad = lane_perm ? 64 : 4;
        // nb <= 2: tiles (K/R, K'/R); nb = 4: row panels K/R (all k' of the rows)
        const int64_t nbc = lane_perm ? nbk : 1, ntile = nbk * nbc;
        auto tile_of = [&](int64_t f) {
            return (int64_t)(k1st[f] / R) * nbc + (lane_perm ? (int64_t)(k2nd[f] / R) : 0);
        };
        std::vector<int64_t> tc(ntile + 1, 0);
        for (int64_t i = 0; i < n_pairs; i++) tc[tile_of(KK(i)) + 1]++;
        for (int64_t b = 0; b < ntile; b++) tc[b + 1] += tc[b];
        std::vector<int64_t> byt(n_pairs);  // file indices grouped by tile, (k, k') order inside
        {
            std::vector<int64_t> cur(tc.begin(), tc.end() - 1);
            for (int64_t i = 0; i < n_pairs; i++) {
                int64_t f = KK(i);
                byt[cur[tile_of(f)]++] = f;
            }
        }
        nslots = 0;
        for (int64_t t = 0; t < ntile; t++)
            for (int64_t a = tc[t]; a < tc[t + 1]; a += cmax) {
                int64_t n = std::min<int64_t>(cmax, tc[t + 1] - a);
                cbeg.push_back(nslots);
                clen.push_back((int32_t)n);
                cij.push_back(make_int2((int)(t / nbc), (int)(t % nbc)));
                nslots += (n + pad - 1) / pad * pad;
            }
        phys.assign(nslots, -1);
        int64_t c = 0;
        for (int64_t t = 0; t < ntile; t++)
            for (int64_t a = tc[t]; a < tc[t + 1]; a += cmax, c++) {
                const int64_t n = clen[c], m = (n + 63) / 64;
                for (int64_t L = 0; L < n; L++)
                    phys[cbeg[c] + (lane_perm ? (L % m) * 64 + L / m : L)] = byt[a + L];
                if (!lane_perm) clen[c] = (int32_t)((n + pad - 1) / pad * pad);  // padded slots carry no pair
            }
    }
    h->nslots = nslots;
    auto PH = [&](int64_t pos) { return R ? phys[pos] : KK(pos); };
    const bool identity = !R && sorted;
    std::vector<int64_t> inv;
    if (!identity) {
        inv.resize(n_pairs);
        for (int64_t pos = 0; pos < nslots; pos++)
            if (PH(pos) >= 0) inv[PH(pos)] = pos;
    }
    auto INV = [&](int64_t f) { return identity ? f : inv[f]; };
    h->identity_order = identity;
    // forward lists (CSR by k) and reverse lists (by k', then k)
    std::vector<int64_t> fwdpos(n_pairs);
    std::vector<int32_t> col(n_pairs), k1p(nslots), k2p(nslots);
    std::vector<int64_t> rcnt(nk + 1, 0);
    for (int64_t i = 0; i < n_pairs; i++) {
        int64_t f = KK(i);
        fwdpos[i] = INV(f);
        col[i] = (int32_t)k2nd[f];
        if (k1st[f] < k2nd[f]) rcnt[k2nd[f] + 1]++;
    }
    for (int64_t pos = 0; pos < nslots; pos++) {
        int64_t f = PH(pos);
        k1p[pos] = f >= 0 ? (int32_t)k1st[f] : -1;
        k2p[pos] = f >= 0 ? (int32_t)k2nd[f] : -1;
    }
    for (int64_t k = 0; k < nk; k++) rcnt[k + 1] += rcnt[k];
    h->nrev = rcnt[nk];
    std::vector<int64_t> revpos(h->nrev);
    std::vector<int32_t> revcol(h->nrev);
    {
        std::vector<int64_t> cur(rcnt.begin(), rcnt.end() - 1);
        for (int64_t i = 0; i < n_pairs; i++) {
            int64_t f = KK(i);
            if (k1st[f] < k2nd[f]) {
                int64_t c = cur[k2nd[f]]++;
                revpos[c] = INV(f);
                revcol[c] = (int32_t)k1st[f];
            }
        }
    }
    std::vector<int64_t> dpos;  // pairs with k = k' (the adjoint treats them apart)
    std::vector<int32_t> dk;
    for (int64_t i = 0; i < n_pairs; i++) {
        int64_t f = KK(i);
        if (k1st[f] == k2nd[f]) { dpos.push_back(fwdpos[i]); dk.push_back((int32_t)k1st[f]); }
    }
    std::vector<int32_t> gbeg;  // workgroup -> first chunk (groups of chunks of one row block)
    const int group = lane_perm ? DMD_TILE_GROUP : DMD_TILE4_GROUP;
    for (int64_t c = 0; c < (int64_t)cbeg.size(); c++)
        if (gbeg.empty() || cij[c].x != cij[gbeg.back()].x || c - gbeg.back() >= group)
            gbeg.push_back((int32_t)c);
    gbeg.push_back((int32_t)cbeg.size());
    h->ngroups = (int64_t)gbeg.size() - 1;
    std::vector<uint32_t> pk;
    if (R) {
        // padding slots: (0, 0) for nb <= 2 (never read), a marker for nb = 4
        pk.assign(nslots, lane_perm ? 0u : 0xFFFFFFFFu);
        for (int64_t pos = 0; pos < nslots; pos++)
            if (k1p[pos] >= 0) pk[pos] = (uint32_t)(k1p[pos] % R) + (uint32_t)(k2p[pos] % R) * (uint32_t)R;
    }
    h->nchunks = (int64_t)cbeg.size();
    int rc;
    const size_t ns = (size_t)nk * h->nb2;
    h->p_esize = precision == DMD_PREC_FP64 ? 16 : (precision == DMD_PREC_C64 ? 8 : 4);
    const size_t p_words = ((size_t)nslots * 2 * h->nb4 * h->p_esize + 15) / 16;
    if ((rc = dalloc_t(&h->d_e, (size_t)nk * nb, "e")) || (rc = dalloc_t(&h->d_edm, (size_t)nk * nb, "e_dm")) ||
        (rc = dalloc_t(&h->d_P, p_words, "P")) ||
        (rc = dalloc_t(&h->d_rowptr, nk + 1, "rowptr")) || (rc = dalloc_t(&h->d_revptr, nk + 1, "revptr")) ||
        (rc = dalloc_t(&h->d_revpos, h->nrev, "revpos")) || (rc = dalloc_t(&h->d_revcol, h->nrev, "revcol")) ||
        (rc = dalloc_t(&h->d_col, n_pairs, "col")) || (rc = dalloc_t(&h->d_fwdpos, n_pairs, "fwdpos")) ||
        (rc = dalloc_t(&h->d_k1, nslots, "k1")) || (rc = dalloc_t(&h->d_k2, nslots, "k2")) ||
        (rc = dalloc_t(&h->d_rhoS, ns, "rhoS")) || (rc = dalloc_t(&h->d_rho1S, ns, "rho1S")) ||
        (rc = dalloc_t(&h->d_out, ns, "out")) || (rc = dalloc_t(&h->d_in, ns, "in")) ||
        (rc = dalloc_t(&h->d_in1, ns, "in1")) || (rc = dalloc_t(&h->d_clk, 1, "clock")) ||
        (rc = dalloc_t(&h->d_acc_rows, (size_t)nk * 4 * h->nb2, "row accumulator")) ||
        (rc = dalloc_t(&h->d_acc_herm, (size_t)nk * h->nb2, "Hermitian accumulator")))
        return rc;
    if (R && ((rc = dalloc_t(&h->d_pk, nslots, "tile offsets")) || (rc = dalloc_t(&h->d_cbeg, h->nchunks, "chunks")) ||
              (rc = dalloc_t(&h->d_clen, h->nchunks, "chunks")) || (rc = dalloc_t(&h->d_cij, h->nchunks, "chunks")) ||
              (rc = dalloc_t(&h->d_gbeg, h->ngroups + 1, "chunk groups"))))
        return rc;
    h->ndiag = (int64_t)dpos.size();
    if (h->ndiag) {
        if ((rc = dalloc_t(&h->d_dpos, h->ndiag, "diagonal pairs")) || (rc = dalloc_t(&h->d_dk, h->ndiag, "diagonal rows")))
            return rc;
        DMD_HIP(hipMemcpy(h->d_dpos, dpos.data(), h->ndiag * 8, hipMemcpyHostToDevice));
        DMD_HIP(hipMemcpy(h->d_dk, dk.data(), h->ndiag * 4, hipMemcpyHostToDevice));
    }
    if (precision == DMD_PREC_BF16 && (rc = dalloc_t(&h->d_pscale, nslots, "P scales"))) return rc;
    if (nb == 4 && (rc = dalloc_t(&h->d_hacc16, (size_t)nk * 16, "Hermitian accumulator"))) return rc;
    if (!identity) {
        if ((rc = dalloc_t(&h->d_inv, n_pairs, "inv"))) return rc;
        DMD_HIP(hipMemcpy(h->d_inv, inv.data(), n_pairs * 8, hipMemcpyHostToDevice));
    }
    DMD_HIP(hipMemcpy(h->d_e, e, (size_t)nk * nb * 8, is_device_ptr(e) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
    DMD_HIP(hipMemcpy(h->d_edm, h->d_e, (size_t)nk * nb * 8, hipMemcpyDeviceToDevice));
    DMD_HIP(hipMemcpy(h->d_rowptr, cnt.data(), (nk + 1) * 8, hipMemcpyHostToDevice));
    DMD_HIP(hipMemcpy(h->d_revptr, rcnt.data(), (nk + 1) * 8, hipMemcpyHostToDevice));
    if (h->nrev) {
        DMD_HIP(hipMemcpy(h->d_revpos, revpos.data(), h->nrev * 8, hipMemcpyHostToDevice));
        DMD_HIP(hipMemcpy(h->d_revcol, revcol.data(), h->nrev * 4, hipMemcpyHostToDevice));
    }
    if (n_pairs) {
        DMD_HIP(hipMemcpy(h->d_col, col.data(), n_pairs * 4, hipMemcpyHostToDevice));
        DMD_HIP(hipMemcpy(h->d_fwdpos, fwdpos.data(), n_pairs * 8, hipMemcpyHostToDevice));
        DMD_HIP(hipMemcpy(h->d_k1, k1p.data(), nslots * 4, hipMemcpyHostToDevice));
        DMD_HIP(hipMemcpy(h->d_k2, k2p.data(), nslots * 4, hipMemcpyHostToDevice));
    }
    if (R && n_pairs) {
        DMD_HIP(hipMemcpy(h->d_pk, pk.data(), nslots * 4, hipMemcpyHostToDevice));
        DMD_HIP(hipMemcpy(h->d_cbeg, cbeg.data(), h->nchunks * 8, hipMemcpyHostToDevice));
        DMD_HIP(hipMemcpy(h->d_clen, clen.data(), h->nchunks * 4, hipMemcpyHostToDevice));
        DMD_HIP(hipMemcpy(h->d_cij, cij.data(), h->nchunks * 8, hipMemcpyHostToDevice));
        DMD_HIP(hipMemcpy(h->d_gbeg, gbeg.data(), (h->ngroups + 1) * 4, hipMemcpyHostToDevice));
    }
    if (nslots) DMD_HIP(hipMemset(h->d_P, 0, p_words * sizeof(double2)));
    if (h->d_pscale) DMD_HIP(hipMemset(h->d_pscale, 0, (size_t)nslots * sizeof(float)));
    if (h->d_hacc16) DMD_HIP(hipMemset(h->d_hacc16, 0, (size_t)nk * 16 * sizeof(double)));
    // the default kernel (tiles) streams SoA granules (nb = 2) or blocked records (nb = 4)
    h->soa = (nb == 2 && R) ? 1 : ((nb == 4 && R) ? 2 : 0);
    for (double2 *b : {h->d_rhoS, h->d_rho1S, h->d_out, h->d_in, h->d_in1}) DMD_HIP(hipMemset(b, 0, ns * sizeof(double2)));
    DMD_HIP(hipMemset(h->d_acc_rows, 0, (size_t)nk * 4 * h->nb2 * sizeof(double)));
    DMD_HIP(hipMemset(h->d_acc_herm, 0, (size_t)nk * h->nb2 * sizeof(double)));
    DMD_HIP(hipDeviceSynchronize());
    *out = h.release();
    return DMD_OK;
}

extern "C" int dmd_scatt_destroy(dmd_scatt *h)
{
    if (!h) return DMD_OK;
    if (h->st) (void)hipStreamSynchronize(h->st);
    delete h;
    return DMD_OK;
}

// Places stored pairs [first, first+n) (file order) from device buffers a, b
// (n x nb^4 complex128 each), scaled; add = 0 sets, 1 adds.  Stream-ordered.
int dmd::place_P_dev(dmd_scatt *h, const double2 *a, const double2 *b, int64_t n, int64_t first, double scale, int add)
{
    {
        const int sg = sp_guard(h, "P update");
        if (sg) return sg;
    }
    if (n <= 0) return DMD_OK;
    if (h->precision != DMD_PREC_FP64) return tiles4_place(h, a, b, n, first, scale, add);
    const int64_t tot = n * 2 * (int64_t)h->nb4;
    hipLaunchKernelGGL(k_place_P, dim3(nblk(tot, 256)), dim3(256), 0, h->st, n, first, h->nb4, a, b, h->d_inv, scale,
                       add, h->soa, h->d_P);
    DMD_HIP(hipGetLastError());
    return DMD_OK;
}

// P of stored pairs [first, first+count) from P1, P2 (count x nb^4 complex,
// host or device memory).
static int put_P(dmd_scatt *h, int64_t first, int64_t count, const double *P1, const double *P2, double scale,
                 int add)
{
    if (!h) return fail(DMD_ERR_ARG, "null handle");
    if (const int sg = sp_guard(h, "P upload")) return sg;
    if (first < 0 || count < 0 || first + count > h->np) return fail(DMD_ERR_ARG, "pair range outside the stored pairs");
    if (h->linear && (add || first != 0 || count != h->np))
        return fail(DMD_ERR_STATE, "the handle holds the linearized operator: upload the whole P first");
    if (count == 0) {
        if (h->np == 0) h->has_P = true;
        return DMD_OK;
    }
    if (!P1 || !P2) return fail(DMD_ERR_ARG, "P1/P2 null");
    const size_t per = (size_t)h->nb4;  // complex per P array per pair
    const bool dev1 = is_device_ptr(P1), dev2 = is_device_ptr(P2);
    // staging chunk: 64 Mi complex per array at most
    const int64_t chunk = std::max<int64_t>(1, (int64_t)((64ll << 20) / per));
    double2 *s1 = nullptr, *s2 = nullptr;
    int rc = DMD_OK;
    if (!dev1 || !dev2) {
        int64_t c = std::min<int64_t>(chunk, count);
        if ((rc = dalloc_t(&s1, (size_t)c * per, "P staging"))) return rc;
        if ((rc = dalloc_t(&s2, (size_t)c * per, "P staging"))) { dfree(s1); return rc; }
    }
    for (int64_t f0 = 0; f0 < count && rc == DMD_OK; f0 += chunk) {
        int64_t n = std::min<int64_t>(chunk, count - f0);
        const double2 *a = (const double2 *)P1 + f0 * per, *b = (const double2 *)P2 + f0 * per;
        hipError_t e = hipSuccess;
        if (!dev1) {
            e = hipMemcpyAsync(s1, a, n * per * 16, hipMemcpyHostToDevice, h->st);
            a = s1;
        }
        if (e == hipSuccess && !dev2) {
            e = hipMemcpyAsync(s2, b, n * per * 16, hipMemcpyHostToDevice, h->st);
            b = s2;
        }
        if (e != hipSuccess) { rc = fail(DMD_ERR_HIP, hipGetErrorString(e)); break; }
        rc = place_P_dev(h, a, b, n, first + f0, scale, add);
        // pageable sources: the staging buffers are reused by the next chunk
        if (rc == DMD_OK && (!dev1 || !dev2) && (e = hipStreamSynchronize(h->st)) != hipSuccess)
            rc = fail(DMD_ERR_HIP, hipGetErrorString(e));
    }
    hipError_t e = hipStreamSynchronize(h->st);
    dfree(s1); dfree(s2);
    if (rc == DMD_OK && e != hipSuccess) rc = fail(DMD_ERR_HIP, hipGetErrorString(e));
    if (rc) {
        h->has_P = false;  // a failure after the first chunk leaves P partly written
        return rc;
    }
    h->has_P = true;
    h->linear = false;
    return DMD_OK;
}

extern "C" int dmd_scatt_upload_P(dmd_scatt *h, const double *P1, const double *P2, double scale)
{
    return put_P(h, 0, h ? h->np : 0, P1, P2, scale, 0);
}
extern "C" int dmd_scatt_add_P(dmd_scatt *h, const double *P1, const double *P2, double scale)
{
    return put_P(h, 0, h ? h->np : 0, P1, P2, scale, 1);
}
extern "C" int dmd_scatt_upload_P_range(dmd_scatt *h, int64_t first, int64_t count, const double *P1,
                                        const double *P2, double scale, int add)
{
    return put_P(h, first, count, P1, P2, scale, add ? 1 : 0);
}

extern "C" int dmd_scatt_download_P(dmd_scatt *h, int64_t first, int64_t count, double *P1, double *P2)
{
    {
        const int sg = sp_guard(h, "dmd_scatt_download_P");
        if (sg) return sg;
    }
    if (!h || !P1 || !P2 || first < 0 || count < 0 || first + count > h->np)
        return fail(DMD_ERR_ARG, "dmd_scatt_download_P: bad arguments");
    if (count == 0) return DMD_OK;
    const size_t per = (size_t)h->nb4;
    int rc;
    double2 *s1 = nullptr, *s2 = nullptr;
    if ((rc = dalloc_t(&s1, count * per, "download"))) return rc;
    if ((rc = dalloc_t(&s2, count * per, "download"))) { dfree(s1); return rc; }
    int64_t tot = count * 2 * (int64_t)per;
    if (h->precision != DMD_PREC_FP64) {
        if ((rc = tiles4_fetch(h, first, count, s1, s2))) { dfree(s1); dfree(s2); return rc; }
    } else {
        hipLaunchKernelGGL(k_fetch_P, dim3(nblk(tot, 256)), dim3(256), 0, h->st, count, first, h->nb4, h->d_P, h->d_inv,
                           h->soa, s1, s2);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(P1, s1, count * per * 16, hipMemcpyDefault, h->st);
    if (e == hipSuccess) e = hipMemcpyAsync(P2, s2, count * per * 16, hipMemcpyDefault, h->st);
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    dfree(s1); dfree(s2);
    if (e != hipSuccess) return fail(DMD_ERR_HIP, hipGetErrorString(e));
    return DMD_OK;
}

// dst's P (set, or added to with add = 1) = scale x the P src stores, pair by
// pair in file order, converted on the device between precisions and layouts
// (a reduced store's values as complex128 into an FP64 handle, or an FP64 P
// quantised into a reduced one).  Both handles hold the same stored pairs.
extern "C" int dmd_scatt_copy_P(dmd_scatt *dst, dmd_scatt *src, double scale, int add)
{
    {
        const int sg = sp_guard(src, "dmd_scatt_copy_P");
        if (sg) return sg;
    }
    if (!dst || !src || dst == src) return fail(DMD_ERR_ARG, "dmd_scatt_copy_P: two distinct handles are needed");
    if (const int sg = sp_guard(dst, "dmd_scatt_copy_P")) return sg;
    if (dst->nk != src->nk || dst->nb != src->nb || dst->np != src->np || dst->dev != src->dev)
        return fail(DMD_ERR_ARG, "dmd_scatt_copy_P: the handles differ in nk, nb, stored pairs or device");
    if (!src->has_P) return fail(DMD_ERR_STATE, "dmd_scatt_copy_P: the source holds no P");
    if (src->linear) return fail(DMD_ERR_STATE, "dmd_scatt_copy_P: the source holds the linearized operator");
    if (dst->linear && add) return fail(DMD_ERR_STATE, "dmd_scatt_copy_P: the destination holds the linearized operator");
    if (add && !dst->has_P) return fail(DMD_ERR_STATE, "dmd_scatt_copy_P: nothing to add to");
    if (src->np == 0) {
        dst->has_P = true;
        dst->linear = false;
        return DMD_OK;
    }
    const size_t per = (size_t)src->nb4;
    const int64_t chunk = std::min<int64_t>(src->np, std::max<int64_t>(1, (int64_t)((64ll << 20) / per)));
    double2 *s1 = nullptr, *s2 = nullptr;
    int rc;
    if ((rc = dalloc_t(&s1, (size_t)chunk * per, "P copy staging"))) return rc;
    if ((rc = dalloc_t(&s2, (size_t)chunk * per, "P copy staging"))) { dfree(s1); return rc; }
    for (int64_t f0 = 0; f0 < src->np && rc == DMD_OK; f0 += chunk) {
        const int64_t n = std::min<int64_t>(chunk, src->np - f0);
        if (src->precision != DMD_PREC_FP64) {
            rc = tiles4_fetch(src, f0, n, s1, s2);
        } else {
            const int64_t tot = n * 2 * (int64_t)per;
            hipLaunchKernelGGL(k_fetch_P, dim3(nblk(tot, 256)), dim3(256), 0, src->st, n, f0, src->nb4, src->d_P,
                               src->d_inv, src->soa, s1, s2);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) rc = fail(DMD_ERR_HIP, hipGetErrorString(e));
        }
        hipError_t e = hipStreamSynchronize(src->st);
        if (rc == DMD_OK && e != hipSuccess) rc = fail(DMD_ERR_HIP, hipGetErrorString(e));
        if (rc == DMD_OK) rc = place_P_dev(dst, s1, s2, n, f0, scale, add);
        e = hipStreamSynchronize(dst->st);  // the staging is reused by the next chunk
        if (rc == DMD_OK && e != hipSuccess) rc = fail(DMD_ERR_HIP, hipGetErrorString(e));
    }
    dfree(s1);
    dfree(s2);
    if (rc) {
        dst->has_P = false;
        return rc;
    }
    dst->has_P = true;
    dst->linear = false;
    return DMD_OK;
}

// Zero the elements with |P| <= thr (|.| as comm/scalar.h:102-103, no
// contraction); a and b hold n complex each.
__global__ void k_threshold(int64_t n, double2 *__restrict__ a, double2 *__restrict__ b, double thr)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 2 * n) return;
    double2 *p = idx < n ? a + idx : b + (idx - n);
    const double2 v = *p;
    if (!(sqrt(__dadd_rn(__dmul_rn(v.x, v.x), __dmul_rn(v.y, v.y))) > thr)) *p = make_double2(0.0, 0.0);
}

// set_sparseP's sparsification (ElectronPhonon.cpp:115-129) on the stored P,
// in chunks of 64 MiB per array: fetch into pair order, zero, place back.
extern "C" int dmd_scatt_threshold_P(dmd_scatt *h, double thr)
{
    {
        const int sg = sp_guard(h, "dmd_scatt_threshold_P");
        if (sg) return sg;
    }
    if (!h || !(thr >= 0)) return fail(DMD_ERR_ARG, "dmd_scatt_threshold_P: bad arguments");
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    if (h->linear) return fail(DMD_ERR_STATE, "the handle holds the linearized operator, not P");
    if (h->np == 0) return DMD_OK;
    const int64_t per = h->nb4;
    const int64_t chunk = std::min<int64_t>(h->np, std::max<int64_t>(1, (64ll << 20) / (per * 16)));
    double2 *s1 = nullptr, *s2 = nullptr;
    int rc;
    if ((rc = dalloc_t(&s1, (size_t)(chunk * per), "threshold staging"))) return rc;
    if ((rc = dalloc_t(&s2, (size_t)(chunk * per), "threshold staging"))) { dfree(s1); return rc; }
    for (int64_t first = 0; first < h->np && rc == DMD_OK; first += chunk) {
        const int64_t n = std::min(chunk, h->np - first), tot = n * 2 * per;
        if (h->precision != DMD_PREC_FP64) {
            if ((rc = tiles4_fetch(h, first, n, s1, s2))) break;
        } else {
            hipLaunchKernelGGL(k_fetch_P, dim3(nblk(tot, 256)), dim3(256), 0, h->st, n, first, h->nb4, h->d_P,
                               h->d_inv, h->soa, s1, s2);
        }
        hipLaunchKernelGGL(k_threshold, dim3(nblk(tot, 256)), dim3(256), 0, h->st, n * per, s1, s2, thr);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { rc = fail(DMD_ERR_HIP, hipGetErrorString(e)); break; }
        rc = place_P_dev(h, s1, s2, n, first, 1.0, 0);
    }
    hipError_t e = hipStreamSynchronize(h->st);
    dfree(s1); dfree(s2);
    if (rc == DMD_OK && e != hipSuccess) rc = fail(DMD_ERR_HIP, hipGetErrorString(e));
    if (rc) h->has_P = false;  // partly thresholded: not to be used
    return rc;
}

// P of the pairs with keep = 0 zeroed; a and b hold n pairs of per complex each
__global__ void k_zero_pairs(int64_t n, int per, const uint8_t *__restrict__ keep, double2 *__restrict__ a,
                             double2 *__restrict__ b)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, m = n * per;
    if (idx >= 2 * m) return;
    const int64_t e = idx < m ? idx : idx - m;
    if (!keep[e / per]) (idx < m ? a : b)[e] = make_double2(0.0, 0.0);
}

// the pair filter of evolve (evolve_eph.cpp:52-58) applied to the stored P
extern "C" int dmd_scatt_mask_pairs(dmd_scatt *h, const uint8_t *keep)
{
    {
        const int sg = sp_guard(h, "dmd_scatt_mask_pairs");
        if (sg) return sg;
    }
    if (!h || (!keep && h->np > 0)) return fail(DMD_ERR_ARG, "dmd_scatt_mask_pairs: bad arguments");
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    if (h->linear) return fail(DMD_ERR_STATE, "the handle holds the linearized operator, not P");
    if (h->np == 0) return DMD_OK;
    const int64_t per = h->nb4;
    const int64_t chunk = std::min<int64_t>(h->np, std::max<int64_t>(1, (64ll << 20) / (per * 16)));
    double2 *s1 = nullptr, *s2 = nullptr;
    uint8_t *dk = nullptr;
    int rc;
    if ((rc = dalloc_t(&s1, (size_t)(chunk * per), "mask staging"))) return rc;
    if ((rc = dalloc_t(&s2, (size_t)(chunk * per), "mask staging"))) { dfree(s1); return rc; }
    if ((rc = dalloc_t(&dk, (size_t)chunk, "mask staging"))) { dfree(s1); dfree(s2); return rc; }
    for (int64_t first = 0; first < h->np && rc == DMD_OK; first += chunk) {
        const int64_t n = std::min(chunk, h->np - first), tot = n * 2 * per;
        bool any = false;
        for (int64_t q = 0; q < n && !any; q++) any = keep[first + q] == 0;
        if (!any) continue;  // nothing to drop in this chunk
        hipError_t e = hipMemcpyAsync(dk, keep + first, (size_t)n, hipMemcpyHostToDevice, h->st);
        if (e != hipSuccess) { rc = fail(DMD_ERR_HIP, hipGetErrorString(e)); break; }
        if (h->precision != DMD_PREC_FP64) {
            if ((rc = tiles4_fetch(h, first, n, s1, s2))) break;
        } else {
            hipLaunchKernelGGL(k_fetch_P, dim3(nblk(tot, 256)), dim3(256), 0, h->st, n, first, h->nb4, h->d_P,
                               h->d_inv, h->soa, s1, s2);
        }
        hipLaunchKernelGGL(k_zero_pairs, dim3(nblk(tot, 256)), dim3(256), 0, h->st, n, (int)per, dk, s1, s2);
        e = hipGetLastError();
        if (e != hipSuccess) { rc = fail(DMD_ERR_HIP, hipGetErrorString(e)); break; }
        if ((rc = place_P_dev(h, s1, s2, n, first, 1.0, 0))) break;
        // keep + first is host memory the next chunk's copy may not overwrite before this one ran
        if ((e = hipStreamSynchronize(h->st)) != hipSuccess) { rc = fail(DMD_ERR_HIP, hipGetErrorString(e)); break; }
    }
    hipError_t e = hipStreamSynchronize(h->st);
    dfree(s1); dfree(s2); dfree(dk);
    if (rc == DMD_OK && e != hipSuccess) rc = fail(DMD_ERR_HIP, hipGetErrorString(e));
    if (rc) h->has_P = false;
    return rc;
}

extern "C" int dmd_scatt_generate_P(dmd_scatt *h, const dmd_synth_params *p)
{
    {
        const int sg = sp_guard(h, "dmd_scatt_generate_P");
        if (sg) return sg;
    }
    if (!h || !p) return fail(DMD_ERR_ARG, "dmd_scatt_generate_P: null argument");
    if (p->nb != h->nb || p->nk != h->nk) return fail(DMD_ERR_ARG, "dmd_scatt_generate_P: params do not match the handle");
    if (h->np == 0) { h->has_P = true; return DMD_OK; }
    if (h->precision != DMD_PREC_FP64) {
        int rc = tiles4_generate(h, *p);
        if (rc) return rc;
        h->has_P = true;
        return DMD_OK;
    }
    const unsigned g = nblk(h->nslots, 64);
    switch (h->nb) {
    case 1: hipLaunchKernelGGL(k_generate_P<1>, dim3(g), dim3(64), 0, h->st, h->nslots, *p, h->d_k1, h->d_k2, h->d_e, h->soa, h->d_P); break;
    case 2: hipLaunchKernelGGL(k_generate_P<2>, dim3(g), dim3(64), 0, h->st, h->nslots, *p, h->d_k1, h->d_k2, h->d_e, h->soa, h->d_P); break;
    case 3: hipLaunchKernelGGL(k_generate_P<3>, dim3(g), dim3(64), 0, h->st, h->nslots, *p, h->d_k1, h->d_k2, h->d_e, h->soa, h->d_P); break;
    case 4: hipLaunchKernelGGL(k_generate_P<4>, dim3(g), dim3(64), 0, h->st, h->nslots, *p, h->d_k1, h->d_k2, h->d_e, h->soa, h->d_P); break;
    }
    DMD_HIP(hipGetLastError());
    DMD_HIP(hipStreamSynchronize(h->st));
    h->has_P = true;
    h->linear = false;
    return DMD_OK;
}

extern "C" int dmd_scatt_set_flags(dmd_scatt *h, int expt, int schrodinger, int kernel)
{
    if (!h) return fail(DMD_ERR_ARG, "null handle");
    if (h->d_spv && kernel != DMD_KERNEL_AUTO && kernel != DMD_KERNEL_TILES)
        return fail(DMD_ERR_STATE, "dmd_scatt_set_flags: the compressed P runs on the tile kernel only");
    if (kernel < DMD_KERNEL_AUTO || kernel > DMD_KERNEL_TILES) return fail(DMD_ERR_ARG, "unknown kernel");
    if (schrodinger && expt) return fail(DMD_ERR_ARG, "in the schrodinger picture expt must be 0 (parameters.cpp:452-453)");
    if (h->precision != DMD_PREC_FP64 && kernel != DMD_KERNEL_AUTO && kernel != DMD_KERNEL_TILES)
        return fail(DMD_ERR_UNSUPPORTED, "the reduced-precision forms run on the tile kernel only");
    h->expt = expt ? 1 : 0;
    h->schrodinger = schrodinger ? 1 : 0;
    h->kernel = kernel;
    // the tile kernels stream SoA granules (nb = 2) or blocked records (nb = 4),
    // the row kernels pair-contiguous P
    const bool tiles = contract_kind(h) == DMD_KERNEL_TILES;
    if (h->nb == 4) {
        const int want = tiles ? 2 : 0;
        if (want != h->soa) {
            int rc = tiles4_relayout(h, want == 2);
            if (rc) return rc;
            h->soa = want;
        }
        return DMD_OK;
    }
    const int want_soa = tiles ? 1 : 0;
    if (want_soa != h->soa) {
        const int64_t ngran = h->nslots / 8;
        const int ep = 2 * h->nb4;
        if (ngran > 0)
            hipLaunchKernelGGL(k_relayout, dim3((unsigned)ngran), dim3(256), 8 * ep * sizeof(double2), h->st, ngran, ep,
                               want_soa, h->d_P);
        DMD_HIP(hipGetLastError());
        DMD_HIP(hipStreamSynchronize(h->st));
        h->soa = want_soa;
    }
    return DMD_OK;
}

extern "C" int dmd_scatt_set_split(dmd_scatt *h, const double *f)
{
    if (!h) return fail(DMD_ERR_ARG, "null handle");
    if (h->precision == DMD_PREC_FP64) return DMD_OK;  // the FP64 form does not split rho
    int rc;
    if (!f) {
        dfree(h->d_fsplit);
        h->d_fsplit = nullptr;
    } else {
        if (!h->d_fsplit && (rc = dalloc_t(&h->d_fsplit, (size_t)h->nk * h->nb, "split occupations"))) return rc;
        if ((rc = to_device(h, h->d_fsplit, f, (size_t)h->nk * h->nb * 8))) return rc;
        DMD_HIP(hipStreamSynchronize(h->st));
    }
    h->feq_dirty = true;
    return DMD_OK;
}

extern "C" int dmd_scatt_set_ddmdt_eq(dmd_scatt *h, const double *eq)
{
    if (!h) return fail(DMD_ERR_ARG, "null handle");
    if (!eq) { h->has_eq = false; return DMD_OK; }
    int rc;
    if (!h->d_eq && (rc = dalloc_t(&h->d_eq, (size_t)h->nk * h->nb2, "ddmdt_eq"))) return rc;
    if ((rc = to_device(h, h->d_eq, eq, (size_t)h->nk * h->nb2 * 16))) return rc;
    DMD_HIP(hipStreamSynchronize(h->st));
    h->has_eq = true;
    return DMD_OK;
}

namespace dmd {
// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
// rotate the window of state `in` (nbd x nbd per k, offset b0) to rhoS
int launch_rotate(dmd_scatt *h, const double2 *in, const double2 *in1, int nbd, int b0, double a, int semic)
{
    int64_t n = h->nk * h->nb2;
    const double *fs = h->linear ? h->d_flin : nullptr;  // the linearized operator acts on drho
    dim3 g(nblk(n, 256)), b(256);
    switch (h->nb) {
    case 1: hipLaunchKernelGGL(k_rotate_in<1>, g, b, 0, h->st, h->nk, nbd, b0, in, in1, h->d_e, h->d_clk, a, h->expt, h->d_rhoS, h->d_rho1S, fs, semic); break;
    case 2: hipLaunchKernelGGL(k_rotate_in<2>, g, b, 0, h->st, h->nk, nbd, b0, in, in1, h->d_e, h->d_clk, a, h->expt, h->d_rhoS, h->d_rho1S, fs, semic); break;
    case 3: hipLaunchKernelGGL(k_rotate_in<3>, g, b, 0, h->st, h->nk, nbd, b0, in, in1, h->d_e, h->d_clk, a, h->expt, h->d_rhoS, h->d_rho1S, fs, semic); break;
    case 4: hipLaunchKernelGGL(k_rotate_in<4>, g, b, 0, h->st, h->nk, nbd, b0, in, in1, h->d_e, h->d_clk, a, h->expt, h->d_rhoS, h->d_rho1S, fs, semic); break;
    }
    DMD_HIP(hipGetLastError());
    return DMD_OK;
}

template <int NB>
static void launch_rows_nb(dmd_scatt *h, bool dm1, double a, double2 *out)
{
    dim3 g(nblk(h->nk, 4)), b(256);
    const double2 *eq = h->has_eq ? h->d_eq : nullptr;
    if (dm1)
        hipLaunchKernelGGL((k_rows<NB, true>), g, b, 0, h->st, h->nk, h->d_P, h->d_rowptr, h->d_col, h->d_fwdpos, h->d_revptr,
                           h->d_revpos, h->d_revcol, h->d_rhoS, h->d_rho1S, h->d_e, h->d_clk, a, h->expt,
                           h->schrodinger, h->prefac, eq, out);
    else
        hipLaunchKernelGGL((k_rows<NB, false>), g, b, 0, h->st, h->nk, h->d_P, h->d_rowptr, h->d_col, h->d_fwdpos, h->d_revptr,
                           h->d_revpos, h->d_revcol, h->d_rhoS, h->d_rho1S, h->d_e, h->d_clk, a, h->expt,
                           h->schrodinger, h->prefac, eq, out);
}

template <int NB>
static void launch_single_nb(dmd_scatt *h, bool dm1, double a, double2 *out)
{
    const double2 *eq = h->has_eq ? h->d_eq : nullptr;
    dim3 g(nblk(h->nk, 4)), b(256), ga(nblk(h->nk, 256));
    if (dm1) {
        hipLaunchKernelGGL((k_fwd_atomic<NB, true>), g, b, 0, h->st, h->nk, h->d_P, h->d_rowptr, h->d_col, h->d_fwdpos, h->d_rhoS,
                           h->d_rho1S, h->d_acc_rows);
        hipLaunchKernelGGL((k_assemble<NB, true>), ga, b, 0, h->st, h->nk, h->d_acc_rows, h->d_rhoS, h->d_rho1S, h->d_e,
                           h->d_clk, a, h->expt, h->schrodinger, h->prefac, eq, out);
    } else {
        hipLaunchKernelGGL((k_fwd_atomic<NB, false>), g, b, 0, h->st, h->nk, h->d_P, h->d_rowptr, h->d_col, h->d_fwdpos, h->d_rhoS,
                           h->d_rho1S, h->d_acc_rows);
        hipLaunchKernelGGL((k_assemble<NB, false>), ga, b, 0, h->st, h->nk, h->d_acc_rows, h->d_rhoS, h->d_rho1S,
                           h->d_e, h->d_clk, a, h->expt, h->schrodinger, h->prefac, eq, out);
    }
}

int launch_raw_rows(dmd_scatt *h, const double2 *X, const double2 *Y)
{
    if (h->nb > 2 || h->soa != 0) return fail(DMD_ERR_STATE, "raw row sums need nb <= 2 and the single-read layout");
    dim3 g(nblk(h->nk, 4)), b(256);
    if (h->nb == 1) {
        if (Y) hipLaunchKernelGGL((k_fwd_atomic<1, true>), g, b, 0, h->st, h->nk, h->d_P, h->d_rowptr, h->d_col, h->d_fwdpos, X, Y, h->d_acc_rows);
        else hipLaunchKernelGGL((k_fwd_atomic<1, false>), g, b, 0, h->st, h->nk, h->d_P, h->d_rowptr, h->d_col, h->d_fwdpos, X, X, h->d_acc_rows);
    } else {
        if (Y) hipLaunchKernelGGL((k_fwd_atomic<2, true>), g, b, 0, h->st, h->nk, h->d_P, h->d_rowptr, h->d_col, h->d_fwdpos, X, Y, h->d_acc_rows);
        else hipLaunchKernelGGL((k_fwd_atomic<2, false>), g, b, 0, h->st, h->nk, h->d_P, h->d_rowptr, h->d_col, h->d_fwdpos, X, X, h->d_acc_rows);
    }
    DMD_HIP(hipGetLastError());
    return DMD_OK;
}

static void launch_tiles(dmd_scatt *h, bool dm1, double a, double2 *out)
{
    const double2 *eq = h->has_eq ? h->d_eq : nullptr;
    if (h->ngroups > 0) {
        if (h->linear)
            hipLaunchKernelGGL((k_tiles_soa<false, true>), dim3((unsigned)h->ngroups), dim3(256), 0, h->st, h->nk, h->d_P,
                               h->d_pk, h->d_cbeg, h->d_clen, h->d_cij, h->d_gbeg, h->d_rhoS, h->d_rho1S, h->d_acc_herm);
        else if (h->d_spv && dm1)
            hipLaunchKernelGGL((k_tiles_soa<true, false, true>), dim3((unsigned)h->ngroups), dim3(256), 0, h->st, h->nk,
                               h->d_spv, h->d_pk, h->d_cbeg, h->d_clen, h->d_cij, h->d_gbeg, h->d_rhoS, h->d_rho1S,
                               h->d_acc_herm, h->d_spmask, h->d_spgofs);
        else if (h->d_spv)
            hipLaunchKernelGGL((k_tiles_soa<false, false, true>), dim3((unsigned)h->ngroups), dim3(256), 0, h->st, h->nk,
                               h->d_spv, h->d_pk, h->d_cbeg, h->d_clen, h->d_cij, h->d_gbeg, h->d_rhoS, h->d_rho1S,
                               h->d_acc_herm, h->d_spmask, h->d_spgofs);
        else if (dm1)
            hipLaunchKernelGGL((k_tiles_soa<true, false>), dim3((unsigned)h->ngroups), dim3(256), 0, h->st, h->nk, h->d_P,
                               h->d_pk, h->d_cbeg, h->d_clen, h->d_cij, h->d_gbeg, h->d_rhoS, h->d_rho1S, h->d_acc_herm);
        else
            hipLaunchKernelGGL((k_tiles_soa<false, false>), dim3((unsigned)h->ngroups), dim3(256), 0, h->st, h->nk, h->d_P,
                               h->d_pk, h->d_cbeg, h->d_clen, h->d_cij, h->d_gbeg, h->d_rhoS, h->d_rho1S, h->d_acc_herm);
    }
    hipLaunchKernelGGL(k_assemble_h, dim3(nblk(h->nk, 256)), dim3(256), 0, h->st, h->nk, h->d_acc_herm, h->d_e,
                       h->d_clk, a, h->expt, h->schrodinger, h->prefac, eq, out);
}

// Which contraction kernel runs.  AUTO: the tile kernel for nb = 2, the
// per-pair atomic kernel for nb = 1, the row kernel for nb >= 3.
int contract_kind(const dmd_scatt *h)
{
    if (h->nb == 4 && h->tile_R && (h->kernel == DMD_KERNEL_AUTO || h->kernel == DMD_KERNEL_TILES))
        return DMD_KERNEL_TILES;
    if (h->kernel == DMD_KERNEL_ROWS || h->nb > 2) return DMD_KERNEL_ROWS;
    if (h->kernel == DMD_KERNEL_SINGLE_READ) return DMD_KERNEL_SINGLE_READ;
    return h->nb == 2 && h->tile_R ? DMD_KERNEL_TILES : DMD_KERNEL_SINGLE_READ;
}

// the contraction + epilogue on the current rhoS (and rho1S when dm1)
int launch_assemble_herm(dmd_scatt *h, double a, double prefac, double2 *out)
{
    if (h->nb != 2) return fail(DMD_ERR_STATE, "the packed Hermitian rows are an nb = 2 form");
    hipLaunchKernelGGL(k_assemble_h, dim3(nblk(h->nk, 256)), dim3(256), 0, h->st, h->nk, h->d_acc_herm, h->d_e,
                       h->d_clk, a, h->expt, 0, prefac, nullptr, out);
    DMD_HIP(hipGetLastError());
    return DMD_OK;
}

int launch_dP_term(dmd_scatt *h, double a, double2 *out)
{
    const int64_t n = h->nk * h->nb2;
    hipLaunchKernelGGL(k_dP_term, dim3(nblk(n, 256)), dim3(256), 0, h->st, n, h->nb, h->d_dPH, h->d_e, h->d_clk, a,
                       out);
    DMD_HIP(hipGetLastError());
    return DMD_OK;
}

static int launch_contract_core(dmd_scatt *h, bool dm1, double a, double2 *out);

int launch_contract(dmd_scatt *h, bool dm1, double a, double2 *out)
{
    int rc = launch_contract_core(h, dm1, a, out);
    if (rc == DMD_OK && h->has_dP && !h->linear) rc = launch_dP_term(h, a, out);
    return rc;
}

static int launch_contract_core(dmd_scatt *h, bool dm1, double a, double2 *out)
{
    const int kind = contract_kind(h);
    if (h->linear && kind != DMD_KERNEL_TILES) return launch_linear(h, a, out);
    if (kind == DMD_KERNEL_TILES && h->nb == 4) {
        return launch_tiles4(h, dm1, a, out);
    } else if (kind == DMD_KERNEL_TILES) {
        launch_tiles(h, dm1, a, out);
    } else if (kind == DMD_KERNEL_SINGLE_READ) {
        if (h->nb == 1) launch_single_nb<1>(h, dm1, a, out);
        else launch_single_nb<2>(h, dm1, a, out);
    } else {
        switch (h->nb) {
        case 1: launch_rows_nb<1>(h, dm1, a, out); break;
        case 2: launch_rows_nb<2>(h, dm1, a, out); break;
        case 3: launch_rows_nb<3>(h, dm1, a, out); break;
        case 4: launch_rows_nb<4>(h, dm1, a, out); break;
        }
    }
    DMD_HIP(hipGetLastError());
    return DMD_OK;
}

}  // namespace dmd

static int eval_core(dmd_scatt *h, double t, const double *rho, const double *rho1, double *ddmdt, bool timed)
{
    if (!h) return fail(DMD_ERR_ARG, "null handle");
    if (!rho || !ddmdt) return fail(DMD_ERR_ARG, "rho/ddmdt null");
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    const size_t bytes = (size_t)h->nk * h->nb2 * 16;
    int rc;
    const double2 *in = (const double2 *)rho, *in1 = (const double2 *)rho1;
    // the state always lands in d_in (dmd_scatt_time_eval replays it)
    if ((rc = to_device(h, h->d_in, rho, bytes))) return rc;
    in = h->d_in;
    if (rho1 && !is_device_ptr(rho1)) {
        if ((rc = to_device(h, h->d_in1, rho1, bytes))) return rc;
        in1 = h->d_in1;
    }
    if ((rc = tiles4_prepare(h))) return rc;
    if ((rc = set_clock(h, t, 0.0, 0))) return rc;
    if ((rc = launch_rotate(h, in, in1, h->nb, 0, 0.0))) return rc;
    double2 *out = is_device_ptr(ddmdt) ? (double2 *)ddmdt : h->d_out;
    if (timed) DMD_HIP(hipEventRecord(h->ev0, h->st));
    if ((rc = launch_contract(h, rho1 != nullptr, 0.0, out))) return rc;
    if (timed) DMD_HIP(hipEventRecord(h->ev1, h->st));
    if (out == h->d_out && (rc = from_device(h, ddmdt, h->d_out, bytes))) return rc;
    DMD_HIP(hipStreamSynchronize(h->st));
    if (timed) {
        float ms = 0;
        DMD_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        h->last_ms = ms;
    }
    return DMD_OK;
}

extern "C" int dmd_scatt_eval(dmd_scatt *h, double t, const double *rho, const double *rho1, double *ddmdt)
{
    return eval_core(h, t, rho, rho1, ddmdt, true);
}

extern "C" int dmd_scatt_compute_ddm(dmd_scatt *h, const double *rho, const double *rho1, double *ddm)
{
    if (!h || !ddm) return fail(DMD_ERR_ARG, "dmd_scatt_compute_ddm: bad arguments");
    int save_expt = h->expt;
    bool save_eq = h->has_eq, save_dP = h->has_dP;
    h->expt = 0;
    h->has_eq = false;
    h->has_dP = false;
    const size_t n = (size_t)h->nk * h->nb2 * 2;
    std::vector<double> tmp(n);
    int rc = eval_core(h, 0.0, rho, rho1, tmp.data(), false);
    h->expt = save_expt;
    h->has_eq = save_eq;
    h->has_dP = save_dP;
    if (rc) return rc;
    if (is_device_ptr(ddm)) {
        std::vector<double> cur(n);
        DMD_HIP(hipMemcpy(cur.data(), ddm, n * 8, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) cur[i] += tmp[i];
        DMD_HIP(hipMemcpy(ddm, cur.data(), n * 8, hipMemcpyHostToDevice));
    } else {
        for (size_t i = 0; i < n; i++) ddm[i] += tmp[i];
    }
    return DMD_OK;
}

extern "C" int dmd_scatt_compute_ddmdt_eq(dmd_scatt *h, const double *f)
{
    // compute_ddmdt_eq (evolve_eph.cpp:3-19): evolve at t=0 with diag(f), diag(1-f)
    if (!h || !f) return fail(DMD_ERR_ARG, "dmd_scatt_compute_ddmdt_eq: bad arguments");
    const size_t n = (size_t)h->nk * h->nb2;
    std::vector<double> fh((size_t)h->nk * h->nb);
    if (is_device_ptr(f)) DMD_HIP(hipMemcpy(fh.data(), f, fh.size() * 8, hipMemcpyDeviceToHost));
    else std::memcpy(fh.data(), f, fh.size() * 8);
    std::vector<double> dm(2 * n, 0.0), dm1(2 * n, 0.0), out(2 * n);
    for (int64_t k = 0; k < h->nk; k++)
        for (int i = 0; i < h->nb; i++) {
            size_t o = 2 * (k * h->nb2 + i * h->nb + i);
            dm[o] = fh[k * h->nb + i];
            dm1[o] = 1 - fh[k * h->nb + i];
        }
    // evolve with compute_eq: no ddmdt_eq and no alg_linearize_dPee term (:107,144)
    bool save_eq = h->has_eq, save_dP = h->has_dP;
    h->has_eq = false;
    h->has_dP = false;
    int rc = eval_core(h, 0.0, dm.data(), dm1.data(), out.data(), false);
    h->has_eq = save_eq;
    h->has_dP = save_dP;
    if (rc) return rc;
    return dmd_scatt_set_ddmdt_eq(h, out.data());
}

extern "C" int dmd_scatt_info(dmd_scatt *h, int64_t *n_pairs, int64_t *p_bytes, double *last_kernel_ms)
{
    if (!h) return fail(DMD_ERR_ARG, "null handle");
    if (n_pairs) *n_pairs = h->np;
    if (p_bytes) *p_bytes = h->np * 2 * h->nb4 * (int64_t)h->p_esize + (h->d_pscale ? h->np * 4 : 0);
    if (p_bytes && h->d_spv)  // compressed: values, per-slot masks and offsets
        *p_bytes = h->sp_values * 16 + h->nslots * 4 + (h->nslots / 8) * 4;
    if (last_kernel_ms) *last_kernel_ms = h->last_ms;
    return DMD_OK;
}

extern "C" int dmd_scatt_time_eval(dmd_scatt *h, double t, int n, double *kernel_ms, double *eval_ms)
{
    if (!h || n < 1) return fail(DMD_ERR_ARG, "dmd_scatt_time_eval: bad arguments");
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    int rc;
    if ((rc = tiles4_prepare(h))) return rc;
    if ((rc = set_clock(h, t, 0.0, 0))) return rc;
    // d_in holds the state of the last dmd_scatt_eval (zeros before any)
    if ((rc = launch_rotate(h, h->d_in, nullptr, h->nb, 0, 0.0))) return rc;
    if ((rc = launch_contract(h, false, 0.0, h->d_out))) return rc;  // warm
    DMD_HIP(hipStreamSynchronize(h->st));
    float kms = 0, ems = 0;
    DMD_HIP(hipEventRecord(h->ev2, h->st));
    for (int i = 0; i < n; i++) {
        if ((rc = launch_rotate(h, h->d_in, nullptr, h->nb, 0, 0.0))) return rc;
        if (i == 0) DMD_HIP(hipEventRecord(h->ev0, h->st));
        if ((rc = launch_contract(h, false, 0.0, h->d_out))) return rc;
    }
    DMD_HIP(hipEventRecord(h->ev1, h->st));
    DMD_HIP(hipEventSynchronize(h->ev1));
    DMD_HIP(hipEventElapsedTime(&ems, h->ev2, h->ev1));
    // kernel-only: time n back-to-back contractions
    DMD_HIP(hipEventRecord(h->ev0, h->st));
    for (int i = 0; i < n; i++)
        if ((rc = launch_contract(h, false, 0.0, h->d_out))) return rc;
    DMD_HIP(hipEventRecord(h->ev1, h->st));
    DMD_HIP(hipEventSynchronize(h->ev1));
    DMD_HIP(hipEventElapsedTime(&kms, h->ev0, h->ev1));
    if (kernel_ms) *kernel_ms = kms / n;
    if (eval_ms) *eval_ms = ems / n;
    h->last_ms = kms / n;
    return DMD_OK;
}
