// dmd_adjoint.hip -- vector-Jacobian product of the scattering right-hand
// side and the checkpointed discrete adjoint of the RK4 integrator.
//
// The right-hand side is f(rho) = prefac/2 (C + C^+) with C_k = (1-rho_k) F1_k
// - F2_k rho_k, F1_k = mat(Pdm_k), F2_k = mat(dm1P_k) (evolve_eph.cpp:99-151),
// quadratic in rho.  Under the real pairing <a, b> = Re sum conj(a) b and for
// Hermitian rho, the cotangent v of f gives, with u = v + v^+ (Schroedinger
// picture):
//   g_k = prefac/2 [ -(u_k F1_k^+ + F2_k^+ u_k)                   (rho_k in C_k)
//                    + Pdm_k[X = u rho] + dm1P_k[Y = (1-rho) u] ]  (rho_k' in F1, F2)
// where Pdm[X], dm1P[Y] are the forward contraction's own row sums (both
// directions per stored pair) with rho_k' replaced by X_k' = u_k' rho_k' and
// 1 - rho_k' by Y_k' = (1 - rho_k') u_k': the reverse direction of a pair
// carries the gradient of its forward direction and vice versa.  Pairs with
// k = k' have one direction only and get a correction.  So the VJP is two
// passes of the single-read contraction (raw row sums) and a few element-wise
// kernels; the result is projected on Hermitian matrices (the dynamics stays
// Hermitian).
// Phases: rho_S = rho e^{-i w t}, f = f_S e^{i w t} element-wise, so the
// cotangent enters as v e^{-i w t} and the gradient leaves as g_S e^{i w t}.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dmd_internal.h"

namespace dmd {

template <int NB>
__device__ __forceinline__ double2 phase_of(const double *e, int64_t k, int i, int j, double t, int expt, double sgn)
{
    if (!expt || i == j) return make_double2(1.0, 0.0);
    return ccis(sgn * (e[k * NB + i] - e[k * NB + j]) * t);
}

// Per row: reads F1|F2 from acc (and clears it), forms u = v_S + v_S^+,
// X' = u rho_S, Y' = (1 - rho_S) u and the local gradient -(u F1^+ + F2^+ u).
template <int NB>
__global__ void k_vjp_prep(int64_t nk, double *__restrict__ acc, const double2 *__restrict__ rhoS,
                           const double2 *__restrict__ v, const double *__restrict__ e, const DClock *__restrict__ clk,
                           int expt, double2 *__restrict__ X, double2 *__restrict__ Y, double2 *__restrict__ gloc)
{
    constexpr int NB2 = NB * NB;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    const double t = clock_t_at(clk, 0.0);
    double2 vs[NB2], u[NB2], r[NB2], F1[NB2], F2[NB2];
    double *a = acc + k * (4 * NB2);
#pragma unroll
    for (int m = 0; m < NB2; m++) {
        vs[m] = cmul(v[k * NB2 + m], phase_of<NB>(e, k, m / NB, m % NB, t, expt, -1.0));
        r[m] = rhoS[k * NB2 + m];
        F1[m] = make_double2(a[2 * m], a[2 * m + 1]);
        F2[m] = make_double2(a[2 * NB2 + 2 * m], a[2 * NB2 + 2 * m + 1]);
    }
#pragma unroll
    for (int m = 0; m < 4 * NB2; m++) a[m] = 0.0;
#pragma unroll
    for (int i = 0; i < NB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) u[i * NB + j] = cadd(vs[i * NB + j], cconj(vs[j * NB + i]));
#pragma unroll
    for (int i = 0; i < NB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) {
            double2 x = make_double2(0, 0), y = make_double2(0, 0), g = make_double2(0, 0);
#pragma unroll
            for (int m = 0; m < NB; m++) {
                cfma(x, u[i * NB + m], r[m * NB + j]);                                     // (u rho)_ij
                const double2 om = make_double2((i == m ? 1.0 : 0.0) - r[i * NB + m].x, -r[i * NB + m].y);
                cfma(y, om, u[m * NB + j]);                                                // ((1-rho) u)_ij
                cfma(g, u[i * NB + m], cconj(F1[j * NB + m]));                             // (u F1^+)_ij
                cfma(g, cconj(F2[m * NB + i]), u[m * NB + j]);                             // (F2^+ u)_ij
            }
            X[k * NB2 + i * NB + j] = x;
            Y[k * NB2 + i * NB + j] = y;
            gloc[k * NB2 + i * NB + j] = make_double2(-g.x, -g.y);
        }
}

// Pairs with k = k' feed their own row in the forward direction only, so the
// raw sums of the second pass hold P1 X + P2^T Y for them where the gradient
// needs conj(P2) X + P1^H Y: add the difference.  One thread per such pair (a
// row has at most one), after the second pass.
template <int NB>
__global__ void k_vjp_diag(int64_t ndiag, const double2 *__restrict__ P, const int64_t *__restrict__ dpos,
                           const int32_t *__restrict__ dk, const double2 *__restrict__ X,
                           const double2 *__restrict__ Y, double *__restrict__ acc)
{
    constexpr int NB2 = NB * NB, NB4 = NB2 * NB2;
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= ndiag) return;
    const double2 *p1 = P + dpos[d] * (2 * NB4), *p2 = p1 + NB4;
    const int64_t k = dk[d];
    double *a = acc + k * (4 * NB2);
#pragma unroll
    for (int i = 0; i < NB2; i++) {
        double2 c1 = make_double2(0, 0), c2 = make_double2(0, 0);
#pragma unroll
        for (int q = 0; q < NB2; q++) {
            const double2 x = X[k * NB2 + q], y = Y[k * NB2 + q];
            cfmac(c1, x, p2[i * NB2 + q]);                                   // conj(P2) X
            cfma(c1, make_double2(-p1[i * NB2 + q].x, -p1[i * NB2 + q].y), x);  // - P1 X
            cfmac(c2, y, p1[q * NB2 + i]);                                   // P1^H Y
            cfma(c2, make_double2(-y.x, -y.y), p2[q * NB2 + i]);             // - P2^T Y
        }
        a[2 * i] += c1.x; a[2 * i + 1] += c1.y;
        a[2 * NB2 + 2 * i] += c2.x; a[2 * NB2 + 2 * i + 1] += c2.y;
    }
}

// g = Herm( prefac/2 (Pdm[X'] + dm1P[Y'] + gloc) e^{i w t} ); clears acc.
template <int NB>
__global__ void k_vjp_final(int64_t nk, double *__restrict__ acc, const double2 *__restrict__ gloc, double prefac,
                            const double *__restrict__ e, const DClock *__restrict__ clk, int expt,
                            double2 *__restrict__ g)
{
    constexpr int NB2 = NB * NB;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    const double t = clock_t_at(clk, 0.0);
    double2 s[NB2];
    double *a = acc + k * (4 * NB2);
#pragma unroll
    for (int m = 0; m < NB2; m++) {
        double2 x = make_double2(a[2 * m] + a[2 * NB2 + 2 * m], a[2 * m + 1] + a[2 * NB2 + 2 * m + 1]);
        x = cadd(x, gloc[k * NB2 + m]);
        x = make_double2(x.x * (0.5 * prefac), x.y * (0.5 * prefac));
        s[m] = cmul(x, phase_of<NB>(e, k, m / NB, m % NB, t, expt, 1.0));
    }
#pragma unroll
    for (int m = 0; m < 4 * NB2; m++) a[m] = 0.0;
#pragma unroll
    for (int i = 0; i < NB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) {
            const double2 p = s[i * NB + j], q = s[j * NB + i];
            g[k * NB2 + i * NB + j] = make_double2(0.5 * (p.x + q.x), 0.5 * (p.y - q.y));
        }
}

// out[idx] += Re <a, b> over n complex numbers (fixed-order block tree, then one block)
__global__ __launch_bounds__(256) void k_dot_partial(int64_t n, const double2 *__restrict__ a,
                                                     const double2 *__restrict__ b, double *__restrict__ part)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        s += a[i].x * b[i].x + a[i].y * b[i].y;
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ __launch_bounds__(256) void k_dot_final(int nparts, const double *__restrict__ part, double *__restrict__ out)
{
    __shared__ double red[256];
    double s = 0;
    for (int i = threadIdx.x; i < nparts; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] += red[0];
}

// y = a + c b (complex arrays of n)
__global__ void k_axpy(int64_t n, const double2 *__restrict__ a, double c, const double2 *__restrict__ b,
                       double2 *__restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = make_double2(a[i].x + c * b[i].x, a[i].y + c * b[i].y);
}
// y += c x
__global__ void k_acc(int64_t n, double c, const double2 *__restrict__ x, double2 *__restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = make_double2(y[i].x + c * x[i].x, y[i].y + c * x[i].y);
}
// y = Herm(x) per nb x nb block; one thread owns both (i, j) and (j, i), so
// x == y is allowed
__global__ void k_herm(int64_t nk, int nb, const double2 *x, double2 *y)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int n2 = nb * nb;
    if (idx >= nk * n2) return;
    const int64_t k = idx / n2;
    const int ij = (int)(idx - k * n2), i = ij / nb, j = ij - i * nb;
    if (i > j) return;
    const double2 p = x[idx], q = x[k * n2 + j * nb + i];
    const double2 hh = make_double2(0.5 * (p.x + q.x), 0.5 * (p.y - q.y));
    y[idx] = hh;
    y[k * n2 + j * nb + i] = cconj(hh);
}
// observable cotangent: y[k][b][i] += w conj(s_t[i][b]) for i in [i0, i1) (f only shifts S)
__global__ void k_obs_adj(int64_t nk, int n, int i0, int i1, const double2 *__restrict__ s,
                          const double *__restrict__ edm, const DClock *__restrict__ clk, int interaction, double w,
                          double2 *__restrict__ y)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int n2 = n * n;
    if (idx >= nk * n2) return;
    const int64_t k = idx / n2;
    const int bi = (int)(idx - k * n2), b = bi / n, i = bi - b * n;
    if (i < i0 || i >= i1) return;
    const double t = clock_t_at(clk, 0.0);
    double2 sk = s[k * n2 + i * n + b];
    if (i != b && interaction) sk = cmul(sk, ccis((edm[k * n + i] - edm[k * n + b]) * t));
    y[idx] = make_double2(y[idx].x + w * sk.x, y[idx].y - w * sk.y);
}

// ---------------------------------------------------------------------------
// nb = 2 tile form of the VJP: one pass over P.
//
// In the pair's SoA tile layout (the forward tile kernel's), each lane forms
// for its stored pair (k, k') both row increments of the gradient at once:
// row k gets
//   -(u_k F1^+ + F2^+ u_k) + P1 X'_k' + P2^T Y'_k'
// with F1 = P1 rho_k', F2 = (1 - rho_k')^T P2 the pair's forward Pdm / dm1P
// terms, X' = u rho, Y' = (1 - rho) u; row k' (k < k') gets the same with the
// reverse direction's matrices conj(P2), conj(P1), and a pair with k = k'
// gets its single direction's partner term in the reverse form
// conj(P2) X'_k + P1^H Y'_k.  The row sums only enter through their
// Hermitian part, which commutes with the element-wise phase, so the lanes
// add packed Hermitian increments (4 doubles per row) into LDS exactly as the
// forward kernel does; launch_assemble_herm then scales by prefac/4 and
// rotates (g = Herm(prefac/2 s) e^{i w t}).  Partner rows' rho_S and u_S
// come from LDS, the lane's own row's from HBM once per run of pairs.
// ---------------------------------------------------------------------------
// rho_S = rho e^{-i w t} and u_S = v_S + v_S^+ (v_S = v e^{-i w t}) at stage a
template <int NB>
__global__ void k_vjp_stage_in(int64_t nk, const double2 *__restrict__ rho, const double2 *__restrict__ v,
                               const double *__restrict__ e, const DClock *__restrict__ clk, double a, int expt,
                               double2 *__restrict__ rS, double2 *__restrict__ uS)
{
    constexpr int NB2 = NB * NB;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    const double t = clock_t_at(clk, a);
    double2 vs[NB2];
#pragma unroll
    for (int m = 0; m < NB2; m++) {
        const double2 ph = phase_of<NB>(e, k, m / NB, m % NB, t, expt, -1.0);
        rS[k * NB2 + m] = cmul(rho[k * NB2 + m], ph);
        vs[m] = cmul(v[k * NB2 + m], ph);
    }
#pragma unroll
    for (int i = 0; i < NB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) uS[k * NB2 + i * NB + j] = cadd(vs[i * NB + j], cconj(vs[j * NB + i]));
}

// 2 x 2 products of the VJP: X' = u r, Y' = (1 - r) u
__device__ __forceinline__ void xy_of(const double2 *r, const double2 *u, double2 *x, double2 *y)
{
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            double2 p = make_double2(0, 0), q = make_double2(0, 0);
#pragma unroll
            for (int m = 0; m < 2; m++) {
                cfma(p, u[i * 2 + m], r[m * 2 + j]);
                cfma(q, one_minus(r, i * 2 + m), u[m * 2 + j]);
            }
            x[i * 2 + j] = p;
            y[i * 2 + j] = q;
        }
}
// packed Hermitian part of s - (u F1^+ + F2^+ u) (2 x 2)
__device__ __forceinline__ void vjp_row_inc(const double2 *s, const double2 *u, const double2 *F1, const double2 *F2,
                                            double hv[4])
{
    double2 x[4];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            double2 g = s[i * 2 + j];
#pragma unroll
            for (int m = 0; m < 2; m++) {
                const double2 a = cmulc(u[i * 2 + m], F1[j * 2 + m]);        // (u F1^+)_ij
                const double2 b = cmul(cconj(F2[m * 2 + i]), u[m * 2 + j]);  // (F2^+ u)_ij
                g = make_double2(g.x - a.x - b.x, g.y - a.y - b.y);
            }
            x[i * 2 + j] = g;
        }
    herm_of(x, hv);
}

__global__ __launch_bounds__(256) void k_tiles_vjp(int64_t nk, const double2 *__restrict__ P,
                                                   const uint32_t *__restrict__ pk, const int64_t *__restrict__ cbeg,
                                                   const int32_t *__restrict__ clen, const int2 *__restrict__ cij,
                                                   const int32_t *__restrict__ gbeg, const double2 *__restrict__ rS,
                                                   const double2 *__restrict__ uS, double *__restrict__ hacc)
{
    constexpr int NB2 = 4, NB4 = 16, EP = 32, R = DMD_TILE_R;
    constexpr int W = 4, RPT = R / 256;
    __shared__ double accA[4 * R], accB[4 * R];
    __shared__ double2 rJ[R * NB2], uJ[R * NB2];
    const int c0 = gbeg[blockIdx.x], c1 = gbeg[blockIdx.x + 1];
    const int I = cij[c0].x;
    const int64_t I0 = (int64_t)I * R;
    for (int i = threadIdx.x; i < 4 * R; i += blockDim.x) { accA[i] = 0.0; accB[i] = 0.0; }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int c = c0; c < c1; c++) {
        const int64_t b = cbeg[c];
        const int n = clen[c];
        const int m = (n + 63) / 64;
        const int J = cij[c].y;
        const int64_t J0 = (int64_t)J * R;
        double *aJ = (I == J) ? accA : accB;
        for (int i = threadIdx.x; i < R * NB2; i += blockDim.x) {
            const bool in = J0 * NB2 + i < nk * NB2;
            rJ[i] = in ? rS[J0 * NB2 + i] : make_double2(0, 0);
            uJ[i] = in ? uS[J0 * NB2 + i] : make_double2(0, 0);
        }
        __syncthreads();
        const int mq = (m + W - 1) / W;
        const int i0 = wv * mq, i1 = min(m, i0 + mq);
        uint32_t code_next = (i0 < i1 && lane * m + i0 < n) ? ld_stream(pk + b + (int64_t)i0 * 64 + lane) : 0u;
        int lo1_cur = -1;
        double2 rk[NB2], uk[NB2], xk[NB2], yk[NB2];
#pragma unroll
        for (int q = 0; q < NB2; q++) {
            rk[q] = make_double2(0, 0); uk[q] = make_double2(0, 0);
            xk[q] = make_double2(0, 0); yk[q] = make_double2(0, 0);
        }
        for (int it = i0; it < i1; it++) {
            const bool ok = lane * m + it < n;
            const int64_t s = b + (ok ? (int64_t)it * 64 + lane : 0);
            const uint32_t code = code_next;
            code_next = (it + 1 < i1 && lane * m + it + 1 < n) ? ld_stream(pk + b + (int64_t)(it + 1) * 64 + lane) : 0u;
            const int lo1 = code % R, lo2 = code / R;
            const int64_t k = I0 + lo1, kp = J0 + lo2;
            const double2 *g = P + (s >> 3) * (8 * EP) + (s & 7);
            double2 rp[NB2], up[NB2], xp[NB2], yp[NB2];
#pragma unroll
            for (int q = 0; q < NB2; q++) { rp[q] = rJ[lo2 * NB2 + q]; up[q] = uJ[lo2 * NB2 + q]; }
            xy_of(rp, up, xp, yp);
            if (ok && lo1 != lo1_cur) {
#pragma unroll
                for (int q = 0; q < NB2; q++) { rk[q] = rS[k * NB2 + q]; uk[q] = uS[k * NB2 + q]; }
                xy_of(rk, uk, xk, yk);
                lo1_cur = lo1;
            }
            const bool diag = k == kp, rev = ok && k < kp;
            double2 p1[NB4], p2[NB4];
#pragma unroll
            for (int el = 0; el < NB4; el++) p1[el] = ld_stream(g + el * 8);
#pragma unroll
            for (int el = 0; el < NB4; el++) p2[el] = ld_stream(g + (NB4 + el) * 8);
            // the reverse-form partner term conj(P2) X'_k + P1^H Y'_k: row k' of a
            // k < k' pair, and row k of a k = k' pair
            double2 sr[NB2];
#pragma unroll
            for (int i = 0; i < NB2; i++) {
                sr[i] = make_double2(0, 0);
#pragma unroll
                for (int q = 0; q < NB2; q++) {
                    cfmac(sr[i], xk[q], p2[i * NB2 + q]);
                    cfmac(sr[i], yk[q], p1[q * NB2 + i]);
                }
            }
            {  // forward direction, row k
                double2 f1[NB2], f2[NB2], sf[NB2];
#pragma unroll
                for (int i = 0; i < NB2; i++) {
                    f1[i] = make_double2(0, 0);
                    f2[i] = make_double2(0, 0);
                    sf[i] = make_double2(0, 0);
#pragma unroll
                    for (int q = 0; q < NB2; q++) {
                        cfma(f1[i], p1[i * NB2 + q], rp[q]);         // P1 rho_k'
                        cfma(f2[i], one_minus(rp, q), p2[q * NB2 + i]);  // (1 - rho_k')^T P2
                        cfma(sf[i], p1[i * NB2 + q], xp[q]);         // P1 X'_k'
                        cfma(sf[i], yp[q], p2[q * NB2 + i]);         // P2^T Y'_k'
                    }
                }
                double hf[4];
                vjp_row_inc(diag ? sr : sf, uk, f1, f2, hf);
                if (ok) {
#pragma unroll
                    for (int d = 0; d < 4; d++) atomicAdd(&accA[d * R + lo1], hf[d]);
                }
            }
            {  // reverse direction, row k'
                double2 g1[NB2], g2[NB2];
#pragma unroll
                for (int i = 0; i < NB2; i++) {
                    g1[i] = make_double2(0, 0);
                    g2[i] = make_double2(0, 0);
#pragma unroll
                    for (int q = 0; q < NB2; q++) {
                        cfmac(g1[i], rk[q], p2[i * NB2 + q]);             // conj(P2) rho_k
                        cfmac(g2[i], one_minus(rk, q), p1[q * NB2 + i]);  // (1 - rho_k)^T conj(P1)
                    }
                }
                double hr[4];
                vjp_row_inc(sr, up, g1, g2, hr);
                if (rev) {
#pragma unroll
                    for (int d = 0; d < 4; d++) atomicAdd(&aJ[d * R + lo2], hr[d]);
                }
            }
        }
        __syncthreads();
        if (I != J) {
#pragma unroll
            for (int t = 0; t < RPT; t++) {
                const int r = t * 256 + threadIdx.x;
                double hv[4];
#pragma unroll
                for (int d = 0; d < 4; d++) { hv[d] = accB[d * R + r]; accB[d * R + r] = 0.0; }
                flush_herm_wave(hv, lane, J0 + (r & ~63), nk, hacc);
            }
        }
        __syncthreads();
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

// The parameter cotangents through P (nb = 2): the scattering term is linear
// in P, and Re <v, f_P(rho)> = prefac/2 Re sum_pairs sum_mn (P1[m,n] G1[m,n] +
// P2[m,n] G2[m,n]) with, in the Schroedinger picture (vec index m = (i1 i2),
// n = (i3 i4)),
//   G1 = conj(Y'_k)[m] rho_k'[n] - [k < k'] conj(1 - rho_k)[m] X'_k'[n]
//   G2 = -(1 - rho_k')[m] conj(X'_k)[n] + [k < k'] Y'_k'[m] conj(rho_k)[n].
// Gc = conj(G) of every stage is summed per stored slot, in the slot layout
// of P, so that at the end dtheta_c = prefac_c/2 Re <Gc, P_c> for every
// component c in one read of each.  One thread per slot; nst stages of
// rho_S, u_S rows (stride nk x 4 between stages).
__global__ __launch_bounds__(256) void k_adj_G(int64_t nslots, const int32_t *__restrict__ k1,
                                               const int32_t *__restrict__ k2, int nst, int64_t stride,
                                               const double2 *__restrict__ rS, const double2 *__restrict__ uS,
                                               double2 *__restrict__ Gc)
{
    constexpr int NB2 = 4, NB4 = 16, EP = 32;
    // Workgroups run on XCD w % 8 (the hardware's round-robin), so XCD x
    // (workgroups x, x + 8, ...) walks the x-th eighth of the slots in order
    // and its L2 holds the stage rows of the few tiles it is on.  The grid
    // is a multiple of 8 workgroups, which makes the remap a bijection.
    const int64_t per = gridDim.x / 8;
    const int64_t bid = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    const int64_t s = bid * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const int32_t ka = k1[s];
    if (ka < 0) return;  // padding slot
    const int64_t a = ka, bk = k2[s];
    const bool rev = a < bk;
    double2 *g = Gc + (s >> 3) * (8 * EP) + (s & 7);
    double2 g1[NB4], g2[NB4];
#pragma unroll
    for (int el = 0; el < NB4; el++) { g1[el] = g[el * 8]; g2[el] = g[(NB4 + el) * 8]; }
    for (int st = 0; st < nst; st++) {
        double2 rk[NB2], uk[NB2], rp[NB2], up[NB2], xk[NB2], yk[NB2], xp[NB2], yp[NB2];
#pragma unroll
        for (int q = 0; q < NB2; q++) {
            rk[q] = rS[st * stride + a * NB2 + q];
            uk[q] = uS[st * stride + a * NB2 + q];
            rp[q] = rS[st * stride + bk * NB2 + q];
            up[q] = uS[st * stride + bk * NB2 + q];
        }
        xy_of(rk, uk, xk, yk);
        xy_of(rp, up, xp, yp);
#pragma unroll
        for (int mm = 0; mm < NB2; mm++)
#pragma unroll
            for (int nn = 0; nn < NB2; nn++) {
                const int el = mm * NB2 + nn;
                cfmac(g1[el], yk[mm], rp[nn]);                                    // Y'_k conj(rho_k')
                cfma(g2[el], make_double2(-one_minus(rp, mm).x, one_minus(rp, mm).y), xk[nn]);  // -conj(1-rho_k') X'_k
                if (rev) {
                    const double2 om = one_minus(rk, mm);
                    cfmac(g1[el], make_double2(-om.x, -om.y), xp[nn]);           // -(1-rho_k) conj(X'_k')
                    cfma(g2[el], cconj(yp[mm]), rk[nn]);                         // conj(Y'_k') rho_k
                }
            }
    }
#pragma unroll
    for (int el = 0; el < NB4; el++) { g[el * 8] = g1[el]; g[(NB4 + el) * 8] = g2[el]; }
}

// y[0] += c x[0]; x[0] = 0
__global__ void k_take_scaled(double *__restrict__ x, double c, double *__restrict__ y)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        y[0] += c * x[0];
        x[0] = 0.0;
    }
}
// y = 0 (fresh allocations may hold any bit pattern, NaN included)
__global__ void k_zero2(int64_t n, double2 *__restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = make_double2(0.0, 0.0);
}
// y = c x
__global__ void k_scale(int64_t n, double c, const double2 *__restrict__ x, double2 *__restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = make_double2(c * x[i].x, c * x[i].y);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int vjp_ready(dmd_scatt *h)
{
    if (h->nb > 2) return fail(DMD_ERR_UNSUPPORTED, "the adjoint is built for nb <= 2");
    if (h->linear) return fail(DMD_ERR_UNSUPPORTED, "the adjoint is built for the P form, not the linearized operator");
    if (h->nbd != h->nb || h->has_H || h->has_eq || h->has_dP || h->rhs_hook || h->schrodinger)
        return fail(DMD_ERR_UNSUPPORTED, "the adjoint covers the scattering term alone (no window, coherent term, "
                                         "ddmdt_eq, linearize_dPee term or right-hand-side hook)");
    if (h->precision != DMD_PREC_FP64) return fail(DMD_ERR_UNSUPPORTED, "the adjoint is FP64");
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    // nb = 2 with tiles: the one-pass tile VJP on the forward kernel's layout;
    // otherwise the raw row sums of the single-read contraction
    const int want = (h->nb == 2 && h->tile_R) ? DMD_KERNEL_TILES : DMD_KERNEL_SINGLE_READ;
    if (contract_kind(h) != want) {
        int rc = dmd_scatt_set_flags(h, h->expt, h->schrodinger, want);
        if (rc) return rc;
    }
    const size_t n = (size_t)h->nk * h->nb2;
    int rc;
    if (!h->d_adj && (rc = dalloc_t(&h->d_adj, 3 * n, "vjp scratch"))) return rc;
    return DMD_OK;
}
static bool vjp_tiles(const dmd_scatt *h) { return h->nb == 2 && h->soa == 1; }

template <int NB>
static int launch_vjp_nb(dmd_scatt *h, const double2 *y, const double2 *v, double2 *g)
{
    const int64_t nk = h->nk;
    const size_t n = (size_t)nk * h->nb2;
    double2 *X = h->d_adj, *Y = h->d_adj + n, *gl = h->d_adj + 2 * n;
    int rc;
    if ((rc = launch_rotate(h, y, nullptr, h->nb, 0, 0.0))) return rc;
    dim3 b256(256), gk(nblk(nk, 256));
    if ((rc = launch_raw_rows(h, h->d_rhoS, nullptr))) return rc;  // F1 | F2 at rho
    hipLaunchKernelGGL(k_vjp_prep<NB>, gk, b256, 0, h->st, nk, h->d_acc_rows, h->d_rhoS, v, h->d_e, h->d_clk, h->expt,
                       X, Y, gl);
    if ((rc = launch_raw_rows(h, X, Y))) return rc;  // Pdm[X'] | dm1P[Y']
    if (h->ndiag)
        hipLaunchKernelGGL(k_vjp_diag<NB>, dim3(nblk(h->ndiag, 256)), b256, 0, h->st, h->ndiag, h->d_P, h->d_dpos,
                           h->d_dk, X, Y, h->d_acc_rows);
    hipLaunchKernelGGL(k_vjp_final<NB>, gk, b256, 0, h->st, nk, h->d_acc_rows, gl, h->prefac, h->d_e, h->d_clk,
                       h->expt, g);
    DMD_HIP(hipGetLastError());
    return DMD_OK;
}

// One-pass tile VJP at stage a of the clock: rho_S, u_S into rS, uS (kept
// for the parameter cotangents), then the pass and the epilogue into g.
static int launch_vjp_tiles(dmd_scatt *h, const double2 *y, const double2 *v, double a, double2 *rS, double2 *uS,
                            double2 *g)
{
    hipLaunchKernelGGL(k_vjp_stage_in<2>, dim3(nblk(h->nk, 256)), dim3(256), 0, h->st, h->nk, y, v, h->d_e, h->d_clk, a,
                       h->expt, rS, uS);
    if (h->ngroups > 0)
        hipLaunchKernelGGL(k_tiles_vjp, dim3((unsigned)h->ngroups), dim3(256), 0, h->st, h->nk, h->d_P, h->d_pk,
                           h->d_cbeg, h->d_clen, h->d_cij, h->d_gbeg, rS, uS, h->d_acc_herm);
    DMD_HIP(hipGetLastError());
    return launch_assemble_herm(h, a, 0.5 * h->prefac, g);
}

// g = (df/drho)^T v at the state y (device arrays nk x nb^2) at stage a of the clock
static int launch_vjp(dmd_scatt *h, const double2 *y, const double2 *v, double a, double2 *g)
{
    if (vjp_tiles(h)) {
        const size_t n = (size_t)h->nk * h->nb2;
        return launch_vjp_tiles(h, y, v, a, h->d_adj, h->d_adj + n, g);
    }
    // the raw-row form reads the clock at a = 0: move t0 to the stage time
    if (a != 0.0) {
        h->h_clk.t0 = h->h_clk.t0 + (double)h->h_clk.step * h->h_clk.dt + a * h->h_clk.dt;
        h->h_clk.dt = 0.0;
        h->h_clk.step = 0;
        DMD_HIP(hipMemcpyAsync(h->d_clk, &h->h_clk, sizeof(DClock), hipMemcpyHostToDevice, h->st));
    }
    return h->nb == 1 ? launch_vjp_nb<1>(h, y, v, g) : launch_vjp_nb<2>(h, y, v, g);
}

// dot += Re <a, b> over n complex
static int launch_dot(dmd_scatt *h, int64_t n, const double2 *a, const double2 *b, double *dot)
{
    const int nparts = (int)std::min<int64_t>(1024, std::max<int64_t>(1, nblk(n, 256)));
    hipLaunchKernelGGL(k_dot_partial, dim3(nparts), dim3(256), 0, h->st, n, a, b, h->d_dotpart);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, h->st, nparts, h->d_dotpart, dot);
    DMD_HIP(hipGetLastError());
    return DMD_OK;
}

}  // namespace dmd

using namespace dmd;

extern "C" int dmd_scatt_vjp(dmd_scatt *h, double t, const double *rho, const double *v, double *g)
{
    {
        const int sg = sp_guard(h, "dmd_scatt_vjp");
        if (sg) return sg;
    }
    if (!h || !rho || !v || !g) return fail(DMD_ERR_ARG, "dmd_scatt_vjp: bad arguments");
    int rc;
    if ((rc = vjp_ready(h))) return rc;
    const size_t n = (size_t)h->nk * h->nb2, bytes = n * 16;
    double2 *tmp = nullptr;
    if ((rc = dalloc_t(&tmp, 3 * n, "vjp io"))) return rc;
    struct Guard { double2 *p; ~Guard() { dfree(p); } } guard{ tmp };
    if ((rc = to_device(h, tmp, rho, bytes)) || (rc = to_device(h, tmp + n, v, bytes))) return rc;
    if ((rc = set_clock(h, t, 0.0, 0))) return rc;
    if ((rc = launch_vjp(h, tmp, tmp + n, 0.0, tmp + 2 * n))) return rc;
    if ((rc = from_device(h, g, tmp + 2 * n, bytes))) return rc;
    DMD_HIP(hipStreamSynchronize(h->st));
    return DMD_OK;
}

extern "C" int dmd_scatt_combine_P(dmd_scatt *h, int ncomp, dmd_scatt *const *comps, const double *weights)
{
    {
        const int sg = sp_guard(h, "dmd_scatt_combine_P");
        if (sg) return sg;
    }
    if (!h || ncomp < 1 || !comps || !weights) return fail(DMD_ERR_ARG, "dmd_scatt_combine_P: bad arguments");
    for (int c = 0; c < ncomp; c++) {
        const dmd_scatt *o = comps[c];
        if (!o || o->nk != h->nk || o->nb != h->nb || o->np != h->np || o->nslots != h->nslots ||
            o->precision != DMD_PREC_FP64 || !o->has_P || o->linear)
            return fail(DMD_ERR_ARG, "dmd_scatt_combine_P: components must hold FP64 P of the same pairs");
    }
    if (h->precision != DMD_PREC_FP64) return fail(DMD_ERR_UNSUPPORTED, "dmd_scatt_combine_P: FP64 handles only");
    // every component in the handle's layout (same pairs give the same slot order)
    for (int c = 0; c < ncomp; c++)
        if (comps[c]->soa != h->soa) {
            int rc = dmd_scatt_set_flags(comps[c], comps[c]->expt, comps[c]->schrodinger, h->kernel);
            if (rc) return rc;
            if (comps[c]->soa != h->soa) return fail(DMD_ERR_STATE, "dmd_scatt_combine_P: layouts differ");
        }
    const int64_t n = h->nslots * 2 * h->nb4;
    for (int c = 0; c < ncomp; c++) {
        if (c == 0)
            hipLaunchKernelGGL(k_axpy, dim3(nblk(n, 256)), dim3(256), 0, h->st, n, comps[0]->d_P, weights[0] - 1.0,
                               comps[0]->d_P, h->d_P);
        else
            hipLaunchKernelGGL(k_acc, dim3(nblk(n, 256)), dim3(256), 0, h->st, n, weights[c], comps[c]->d_P, h->d_P);
    }
    DMD_HIP(hipGetLastError());
    DMD_HIP(hipStreamSynchronize(h->st));
    h->has_P = true;
    h->linear = false;
    return DMD_OK;
}

// Discrete adjoint of RK4 over segments of C = checkpoint_every steps: the
// forward sweep keeps the state at every segment start and, for the last
// segment, every stage state Y_{j,s}; the reverse sweep rebuilds the stage
// states of each earlier segment from its checkpoint (4 evaluations per
// step), then walks the steps backwards: per stage one VJP and, for the
// parameter cotangents, either (nb = 2 tiles) the stages' rho_S / u_S rows
// folded every 4 steps into the per-slot sum Gc that is dotted with every
// component P_c at the end, or (other layouts) one evaluation of every
// component per stage.  With C >= n_steps nothing is recomputed.
extern "C" int dmd_scatt_adjoint_rk4(dmd_scatt *h, double t0, double dt, int64_t n_steps, int report_every,
                                     int checkpoint_every, const double *rho0, const double *dL_dobs, int ncomp,
                                     dmd_scatt *const *comps, double *dL_dtheta, double *dL_drho0)
{
    {
        const int sg = sp_guard(h, "dmd_scatt_adjoint_rk4");
        if (sg) return sg;
    }
    if (!h || !rho0 || n_steps < 1 || checkpoint_every < 1 || (ncomp > 0 && (!comps || !dL_dtheta)))
        return fail(DMD_ERR_ARG, "dmd_scatt_adjoint_rk4: bad arguments");
    if (h->semiclassical || h->ph_tau_inv != 0.0)
        return fail(DMD_ERR_UNSUPPORTED, "the adjoint has no semiclassical mode or phenomenological relaxation");
    int rc;
    if ((rc = vjp_ready(h))) return rc;
    const bool obs = report_every > 0 && dL_dobs;
    if (obs && !h->has_obs) return fail(DMD_ERR_STATE, "dmd_scatt_adjoint_rk4: set_observable first");
    for (int c = 0; c < ncomp; c++) {
        if (!comps[c] || comps[c]->nk != h->nk || comps[c]->nb != h->nb || comps[c]->nbd != comps[c]->nb)
            return fail(DMD_ERR_ARG, "dmd_scatt_adjoint_rk4: components must match the handle (no window)");
        if (const int sg = sp_guard(comps[c], "dmd_scatt_adjoint_rk4 (component)")) return sg;
    }
    const bool tiles = vjp_tiles(h);
    if (tiles)
        for (int c = 0; c < ncomp; c++) {
            dmd_scatt *o = comps[c];
            if (o->np != h->np || o->nslots != h->nslots || o->precision != DMD_PREC_FP64 || !o->has_P || o->linear ||
                o->expt != h->expt)
                return fail(DMD_ERR_ARG, "dmd_scatt_adjoint_rk4: components must hold FP64 P of the handle's pairs "
                                         "with the same phase flag");
            if (o->soa != h->soa && (rc = dmd_scatt_set_flags(o, o->expt, o->schrodinger, DMD_KERNEL_TILES))) return rc;
            if (o->soa != h->soa) return fail(DMD_ERR_STATE, "dmd_scatt_adjoint_rk4: component layouts differ");
        }
    const int64_t nk = h->nk;
    const size_t n = (size_t)nk * h->nb2, bytes = n * 16;
    const int64_t C = std::min<int64_t>(checkpoint_every, n_steps), ncp = (n_steps + C - 1) / C;
    const int64_t nobs = obs ? n_steps / report_every + 1 : 0;
    std::vector<void *> mem;
    struct Guard { std::vector<void *> *m; ~Guard() { for (void *p : *m) dfree(p); } } guard{ &mem };
    auto alloc = [&](double2 **p, size_t cnt, const char *what) {
        int r = dalloc_t(p, cnt, what);
        if (!r) mem.push_back(*p);
        return r;
    };
    // cp: ncp segment starts and the final state; stg: the segment's stage
    // states, 4 per step; rS/uS: rho_S, u_S rows of the stages not yet folded
    double2 *cp, *stg, *Ks, *ybar, *kbar, *ynew, *tmp, *rS = nullptr, *uS = nullptr, *Gc = nullptr;
    double *d_dth = nullptr;
    if ((rc = alloc(&cp, (size_t)(ncp + 1) * n, "checkpoints")) || (rc = alloc(&stg, (size_t)C * 4 * n, "stage states")) ||
        (rc = alloc(&Ks, n, "stage slope")) || (rc = alloc(&ybar, n, "adjoint state")) ||
        (rc = alloc(&kbar, 4 * n, "stage cotangents")) || (rc = alloc(&ynew, n, "adjoint next")) ||
        (rc = alloc(&tmp, 2 * n, "scratch")))
        return rc;
    const int64_t gwords = tiles && ncomp > 0 ? h->nslots * 2 * h->nb4 : 0;
    if (gwords) {
        if ((rc = alloc(&Gc, (size_t)gwords, "parameter cotangent sums"))) return rc;
        hipLaunchKernelGGL(k_zero2, dim3(nblk(gwords, 256)), dim3(256), 0, h->st, gwords, Gc);
    }
    // the parameter fold runs once per kGroup steps over their 4 kGroup stages
    constexpr int kGroup = 4;
    if (tiles && ((rc = alloc(&rS, 4 * kGroup * n, "stage rho_S")) || (rc = alloc(&uS, 4 * kGroup * n, "stage u_S"))))
        return rc;
    int pending = 0;  // stages held in rS / uS
    auto fold_G = [&]() -> int {
        if (gwords && pending > 0 && h->nslots > 0)
            hipLaunchKernelGGL(k_adj_G, dim3((nblk(h->nslots, 256) + 7) / 8 * 8), dim3(256), 0, h->st, h->nslots,
                               h->d_k1, h->d_k2, pending, (int64_t)n, rS, uS, Gc);
        pending = 0;
        DMD_HIP(hipGetLastError());
        return DMD_OK;
    };
    if (ncomp > 0) {
        if ((rc = dalloc_t(&d_dth, ncomp, "dtheta"))) return rc;
        mem.push_back(d_dth);
        DMD_HIP(hipMemsetAsync(d_dth, 0, ncomp * sizeof(double), h->st));
    }
    if (!h->d_dotpart && (rc = dalloc_t(&h->d_dotpart, 1024, "dot partials"))) return rc;
    std::vector<double> hobs(nobs, 0.0);
    if (nobs) {
        if (is_device_ptr(dL_dobs)) DMD_HIP(hipMemcpy(hobs.data(), dL_dobs, nobs * 8, hipMemcpyDeviceToHost));
        else std::memcpy(hobs.data(), dL_dobs, nobs * 8);
    }
    const dim3 gn(nblk((int64_t)n, 256)), b256(256);
    const double a_s[4] = { 0.0, 0.5, 0.5, 1.0 };
    // RHS at state y, stage time a of step j: out = f(y)
    auto rhs = [&](dmd_scatt *op, int64_t j, double a, const double2 *y, double2 *out) -> int {
        int r;
        if ((r = set_clock(op, t0, dt, j))) return r;
        if ((r = launch_rotate(op, y, nullptr, op->nb, 0, a))) return r;
        return launch_contract(op, false, a, out);
    };
    // one forward step j from y: the stage states Y1..Y3 into Y[1..3] (Y[0]
    // is y itself), the Hermitized next state into ynext
    auto step = [&](int64_t j, const double2 *y, double2 *const Y[3], double2 *ynext) -> int {
        int r;
        const double cy[3] = { 0.5 * dt, 0.5 * dt, dt };
        const double wk[4] = { dt / 6.0, dt / 3.0, dt / 3.0, dt / 6.0 };
        const double2 *cur = y;
        for (int s = 0; s < 4; s++) {
            if ((r = rhs(h, j, a_s[s], cur, Ks))) return r;
            if (s < 3) {
                hipLaunchKernelGGL(k_axpy, gn, b256, 0, h->st, (int64_t)n, y, cy[s], Ks, Y[s]);
                cur = Y[s];
            }
            if (s == 0) hipLaunchKernelGGL(k_axpy, gn, b256, 0, h->st, (int64_t)n, y, wk[0], Ks, ynext);
            else hipLaunchKernelGGL(k_acc, gn, b256, 0, h->st, (int64_t)n, wk[s], Ks, ynext);
        }
        hipLaunchKernelGGL(k_herm, gn, b256, 0, h->st, nk, h->nb, ynext, ynext);
        DMD_HIP(hipGetLastError());
        return DMD_OK;
    };
    auto stage = [&](int64_t jj, int s) { return stg + ((size_t)jj * 4 + s) * n; };
    // the steps of segment c from its checkpoint, keeping every stage state
    auto run_segment = [&](int64_t c) -> int {
        const int64_t j0 = c * C, j1 = std::min<int64_t>(n_steps, (c + 1) * C);
        DMD_HIP(hipMemcpyAsync(stage(0, 0), cp + c * n, bytes, hipMemcpyDeviceToDevice, h->st));
        for (int64_t j = j0; j < j1; j++) {
            double2 *const Y[3] = { stage(j - j0, 1), stage(j - j0, 2), stage(j - j0, 3) };
            double2 *next = j + 1 < j1 ? stage(j + 1 - j0, 0) : cp + (c + 1) * n;
            int r;
            if ((r = step(j, stage(j - j0, 0), Y, next))) return r;
        }
        return DMD_OK;
    };
    // forward sweep: checkpoints at steps 0, C, 2C, ...; the last segment keeps its stages
    if ((rc = to_device(h, cp, rho0, bytes))) return rc;
    hipLaunchKernelGGL(k_herm, gn, b256, 0, h->st, nk, h->nb, cp, cp);
    for (int64_t c = 0; c < ncp; c++) {
        if (c + 1 == ncp) {
            if ((rc = run_segment(c))) return rc;
            break;
        }
        // states ping-pong between tmp and tmp + n; stage states are scratch
        DMD_HIP(hipMemcpyAsync(tmp, cp + c * n, bytes, hipMemcpyDeviceToDevice, h->st));
        for (int64_t j = c * C; j < (c + 1) * C; j++) {
            double2 *const Y[3] = { stage(0, 1), stage(0, 2), stage(0, 3) };
            const int64_t q = j - c * C;
            const double2 *y = tmp + (q % 2 ? n : 0);
            double2 *next = j + 1 < (c + 1) * C ? tmp + (q % 2 ? 0 : n) : cp + (c + 1) * n;
            if ((rc = step(j, y, Y, next))) return rc;
        }
    }
    // reverse sweep
    hipLaunchKernelGGL(k_zero2, gn, b256, 0, h->st, (int64_t)n, ybar);
    auto add_obs = [&](int64_t jstep) -> int {  // cotangent of S at step jstep
        if (!obs || jstep % report_every) return DMD_OK;
        const double w = hobs[jstep / report_every] * h->obs_prefac;
        if (w == 0.0) return DMD_OK;
        int r;
        if ((r = set_clock(h, t0, dt, jstep))) return r;
        hipLaunchKernelGGL(k_obs_adj, gn, b256, 0, h->st, nk, h->nbd, h->obs_i0, h->obs_i1, h->d_s, h->d_edm,
                           h->d_clk, h->schrodinger ? 0 : 1, w, ybar);
        DMD_HIP(hipGetLastError());
        return DMD_OK;
    };
    if ((rc = add_obs(n_steps))) return rc;
    const double ws[4] = { dt / 6.0, dt / 3.0, dt / 3.0, dt / 6.0 };
    const double back[4] = { 0.0, 0.5 * dt, 0.5 * dt, dt };  // Y_s = y + back_s K_{s-1}
    for (int64_t c = ncp - 1; c >= 0; c--) {
        const int64_t j0 = c * C, j1 = std::min<int64_t>(n_steps, (c + 1) * C);
        if (c + 1 < ncp && (rc = run_segment(c))) return rc;
        for (int64_t j = j1 - 1; j >= j0; j--) {
            // zbar = Herm(ybar); Kbar_s = w_s zbar
            double2 *Kb[4] = { kbar, kbar + n, kbar + 2 * n, kbar + 3 * n };
            hipLaunchKernelGGL(k_herm, gn, b256, 0, h->st, nk, h->nb, ybar, ynew);
            for (int s = 0; s < 4; s++) hipLaunchKernelGGL(k_scale, gn, b256, 0, h->st, (int64_t)n, ws[s], ynew, Kb[s]);
            if ((rc = set_clock(h, t0, dt, j))) return rc;
            for (int s = 3; s >= 0; s--) {
                const double2 *Y = stage(j - j0, s);
                if (!tiles) {
                    // parameter cotangents Re <Kbar_s, f_c(Y_s)>: each component
                    // evaluates on its own stream, the dot runs on the handle's
                    for (int cc = 0; cc < ncomp; cc++) {
                        DMD_HIP(hipStreamSynchronize(h->st));
                        if ((rc = rhs(comps[cc], j, a_s[s], Y, tmp))) return rc;
                        DMD_HIP(hipStreamSynchronize(comps[cc]->st));
                        if ((rc = launch_dot(h, (int64_t)n, Kb[s], tmp, d_dth + cc))) return rc;
                    }
                    if ((rc = set_clock(h, t0, dt, j))) return rc;
                    if ((rc = launch_vjp(h, Y, Kb[s], a_s[s], tmp))) return rc;
                } else {
                    const size_t slot = (size_t)(pending + s) * n;
                    if ((rc = launch_vjp_tiles(h, Y, Kb[s], a_s[s], rS + slot, uS + slot, tmp))) return rc;
                }
                // Ybar_s = J_s^T Kbar_s
                hipLaunchKernelGGL(k_acc, gn, b256, 0, h->st, (int64_t)n, 1.0, tmp, ynew);
                if (s > 0) hipLaunchKernelGGL(k_acc, gn, b256, 0, h->st, (int64_t)n, back[s], tmp, Kb[s - 1]);
            }
            if (tiles) {
                pending += 4;
                if (pending == 4 * kGroup && (rc = fold_G())) return rc;
            }
            DMD_HIP(hipGetLastError());
            DMD_HIP(hipMemcpyAsync(ybar, ynew, bytes, hipMemcpyDeviceToDevice, h->st));
            if ((rc = add_obs(j))) return rc;
        }
    }
    if (tiles && (rc = fold_G())) return rc;
    // dtheta_c = prefac_c / 2 Re <Gc, P_c>
    if (gwords) DMD_HIP(hipMemsetAsync(tmp, 0, sizeof(double), h->st));
    for (int cc = 0; gwords && cc < ncomp; cc++) {
        if ((rc = launch_dot(h, gwords, Gc, comps[cc]->d_P, (double *)tmp))) return rc;
        hipLaunchKernelGGL(k_take_scaled, dim3(1), dim3(64), 0, h->st, (double *)tmp, 0.5 * comps[cc]->prefac,
                           d_dth + cc);
    }
    DMD_HIP(hipGetLastError());
    DMD_HIP(hipStreamSynchronize(h->st));
    if (dL_drho0) {
        hipLaunchKernelGGL(k_herm, gn, b256, 0, h->st, nk, h->nb, ybar, ybar);
        if ((rc = from_device(h, dL_drho0, ybar, bytes))) return rc;
    }
    if (ncomp > 0) DMD_HIP(hipMemcpyAsync(dL_dtheta, d_dth, ncomp * sizeof(double), hipMemcpyDeviceToHost, h->st));
    DMD_HIP(hipStreamSynchronize(h->st));
    return DMD_OK;
}
