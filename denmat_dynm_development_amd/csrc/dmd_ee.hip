// dmd_ee.hip -- the model electron-electron scattering P generated on the
// device: elecelec_model::calc_P (src_v4.5.7/scattering_model/
// ElecElec_Model.h:185-262) at the occupations f (eeMode = Pee_fixed_at_eq) or
// at the current density matrix (eeMode = Pee_update, rebuilt by reset_scatt,
// ElectronPhonon.cpp:15-42), folded into P as add_scatt_contrib("ee") does
// (:44-113); under alg_linearize_dPee the change dP of the P at f_eq towards
// the state (calc_A1rho_dP, ElecElec_Model.h:213-217,271-281) and its row term
// in evolve (evolve_eph.cpp:107-148); and the e-e terms of the linearized operator, calc_P3P4 /
// calc_P5P6 (ElecElec_Model_linear.cpp:3-111) turned into Lscij / Lscji as
// add_Lsc_contrib("ee") does (ElectronPhonon_linear.cpp:38-104,183-253).
//
// For a stored pair (k1, k2) and every k3 with k4 = k1 + k3 - k2 a kept k
// point (the first k of its mesh cell), calc_A (ElecElec_Model.h:144-166):
//   mee(p q x y) = v(k1 - k2) <k1 p|k2 q> <k3 x|k4 y>
//                  (exchange: (mee - v(k1 - k4) <k1 p|k4 y> <k3 x|k2 q>) / 2),
//   A2 = mee exp(pexp de^2) inside |de| < ethr, de = e1p - e2q + e3x - e4y,
//   A1 = A2 (lindblad) or mee;
//   P1[(i1 i2),(i3 i4)] += sum_xy [dm1_k3 A1(i1 i3) dm_k4]_xy conj(A2(i2 i4))_xy
//   P2[(i1 i2),(i3 i4)] += sum_xy [dm_k3 A1(i3 i1) dm1_k4]_xy conj(A2(i4 i2))_xy
// (dm = diag f, dm1 = diag(1 - f) at equilibrium), times prefac_gauss /
// nk_full, doubled with exchange.  The density matrices are Hermitian and
// read from their upper triangle, as the reference's zhemm does.
//
// MI355X mapping: nb <= 2, one wavefront per stored pair (k_ee_pair2): the
// lanes test 64 k3 at a time against the energy window and queue the
// survivors in LDS, so every term runs with all lanes busy, and one term forms
// the overlaps and A2 blocks once for P1 and P2; nb > 2, one thread per P
// element (k_ee_P), which skips, before any overlap work, every k3 whose band
// energies cannot bring the A2 entries it multiplies inside the window.  The chunk's P goes into
// the handle's layout through place_P_dev (every precision), or, for the
// linearized operator, the chunk's Lscij / Lscji increments are added into
// the converted store.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "dmd_internal.h"
#include "dmd_model.h"

namespace dmd {

// Device view of the e-e inputs.
struct EeArgs {
    const double2 *U;        // nk x nw x NB: U_k[w, b]
    const int *ik3;          // nk x 3 mesh cells
    const int *grid;         // mesh cell -> first k (INT_MAX: none)
    const double *e, *f;     // nk x NB
    const double2 *dm, *dm1; // nk x NB^2, Schroedinger picture (null: diag f)
    const double *kvec;      // nk x 3
    const double *vq_cell;   // v(q) per mesh cell (static RPA) or null (closed form)
    const double *vq_pair;   // v(k1 - k2) of each pair of the chunk, or null
    int64_t nk;
    int nw, m0, m1, m2;
    double pfac;             // prefac_gauss / nk_full (x 2 with exchange)
    int dP;                  // with dm: the change dP (calc_A1rho_dP at f = f_eq) instead of P(dm)
};

// mesh cell of k_a + sc (k_c - k_b)
__device__ __forceinline__ int ee_cell(const EeArgs &a, int64_t ka, int64_t kb, int64_t kc, int sc)
{
    const int x = ((a.ik3[3 * ka] + sc * (a.ik3[3 * kc] - a.ik3[3 * kb])) % a.m0 + a.m0) % a.m0;
    const int y = ((a.ik3[3 * ka + 1] + sc * (a.ik3[3 * kc + 1] - a.ik3[3 * kb + 1])) % a.m1 + a.m1) % a.m1;
    const int z = ((a.ik3[3 * ka + 2] + sc * (a.ik3[3 * kc + 2] - a.ik3[3 * kb + 2])) % a.m2 + a.m2) % a.m2;
    return (x * a.m1 + y) * a.m2 + z;
}
// k4 = ka + kc - kb on the mesh, or -1
__device__ __forceinline__ int64_t ee_k4(const EeArgs &a, int64_t ka, int64_t kb, int64_t kc)
{
    const int g = a.grid[ee_cell(a, ka, kb, kc, 1)];
    return g == INT_MAX ? -1 : (int64_t)g;
}
// v(k_a - k_b): the per-cell table, else the closed-form screening
__device__ __forceinline__ double ee_v(const EeArgs &a, const dmd_eimp_model &m, int64_t ka, int64_t kb)
{
    if (a.vq_cell) {
        const int x = ((a.ik3[3 * ka] - a.ik3[3 * kb]) % a.m0 + a.m0) % a.m0;
        const int y = ((a.ik3[3 * ka + 1] - a.ik3[3 * kb + 1]) % a.m1 + a.m1) % a.m1;
        const int z = ((a.ik3[3 * ka + 2] - a.ik3[3 * kb + 2]) % a.m2 + a.m2) % a.m2;
        return a.vq_cell[((int64_t)x * a.m1 + y) * a.m2 + z];
    }
    return eimp_vq(m, a.kvec + 3 * ka, a.kvec + 3 * kb);
}
// <k_a p|k_b q> = sum_w conj(U_a[w, p]) U_b[w, q]
template <int NB>
__device__ __forceinline__ double2 ee_ov(const EeArgs &a, int64_t ka, int64_t kb, int p, int q)
{
    const double2 *Ua = a.U + ka * a.nw * NB, *Ub = a.U + kb * a.nw * NB;
    double2 o = make_double2(0.0, 0.0);
    for (int w = 0; w < a.nw; w++) o = cadd(o, cmulc(Ub[w * NB + q], Ua[w * NB + p]));
    return o;
}
__device__ __forceinline__ double ee_dl(const EimpConst &c, double de)
{
    return fabs(de) < c.m.ethr ? exp(c.prefac_exp * (de * de)) : 0.0;
}
// Hermitian completion of a density matrix from its upper triangle (real diagonal)
template <int NB>
__device__ __forceinline__ double2 ee_hu(const double2 *M, int i, int j)
{
    if (i < j) return M[i * NB + j];
    if (i > j) return cconj(M[j * NB + i]);
    return make_double2(M[i * NB + i].x, 0.0);
}
// The row-subtracted Hermitian read of calc_A1rho_dP (ElecElec_Model.h:271-281):
// there ddm[i][j] = dm[i][j] - g[i] for every j (row i's occupation taken from
// each element of row i), and the zhemm that follows reads the upper triangle,
// so element (i, j) is dm[i][j] - g[i] above the diagonal and its conjugate
// mirror dm[i][j] - g[j] below (real diagonal).
template <int NB>
__device__ __forceinline__ double2 ee_hus(const double2 *M, const double *g, int i, int j)
{
    if (i < j) return make_double2(M[i * NB + j].x - g[i], M[i * NB + j].y);
    if (i > j) return make_double2(M[j * NB + i].x - g[j], -M[j * NB + i].y);
    return make_double2(M[i * NB + i].x - g[i], 0.0);
}
// calc_A1rho_dP: out = fl o (A1 ddmr) + (ddml A1) o fr, fl scaling rows and fr
// columns, with ddml = L - fl and ddmr = R - fr read as above.
template <int NB>
__device__ __forceinline__ void ee_a1rho_dP(const double2 *A1, const double2 *L, const double *fl, const double2 *R,
                                            const double *fr, double2 *out)
{
#pragma unroll
    for (int x = 0; x < NB; x++)
#pragma unroll
        for (int y = 0; y < NB; y++) {
            double2 s = make_double2(0.0, 0.0), u = make_double2(0.0, 0.0);
#pragma unroll
            for (int k = 0; k < NB; k++) {
                s = cadd(s, cmul(A1[x * NB + k], ee_hus<NB>(R, fr, k, y)));
                u = cadd(u, cmul(ee_hus<NB>(L, fl, x, k), A1[k * NB + y]));
            }
            out[x * NB + y] = make_double2(fl[x] * s.x + u.x * fr[y], fl[x] * s.y + u.y * fr[y]);
        }
}
// can e_p - e_q + (e3 - e4 over the bands) fall inside (-ethr, ethr)?
template <int NB>
__device__ __forceinline__ bool ee_window(double d, const double *e3, const double *e4, double ethr)
{
    double lo = 1e300, hi = -1e300;
#pragma unroll
    for (int x = 0; x < NB; x++)
#pragma unroll
        for (int y = 0; y < NB; y++) {
            const double t = e3[x] - e4[y];
            lo = fmin(lo, t);
            hi = fmax(hi, t);
        }
    return d + lo < ethr && d + hi > -ethr;
}

// P1, P2 of n stored pairs (file order), one thread per element.
template <int NB>
__global__ __launch_bounds__(256) void k_ee_P(int64_t n, const uint64_t *__restrict__ k1,
                                              const uint64_t *__restrict__ k2, EeArgs a, EimpConst c,
                                              double2 *__restrict__ P1, double2 *__restrict__ P2)
{
    constexpr int NB2 = NB * NB, NB4 = NB2 * NB2;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * NB4) return;
    const int64_t fp = idx / NB4;
    const int el = (int)(idx - fp * NB4);
    const int i1 = el / (NB * NB2), i2 = (el / NB2) % NB, i3 = (el / NB) % NB, i4 = el % NB;
    const int64_t A = (int64_t)k1[fp], B = (int64_t)k2[fp];
    const bool ex = c.m.exchange != 0, lb = c.m.lindblad != 0;
    double2 acc1 = make_double2(0.0, 0.0), acc2 = make_double2(0.0, 0.0);
    if (A != B || ex) {
        const double v12 = a.vq_pair ? a.vq_pair[fp] : ee_v(a, c.m, A, B);
        auto vo = [&](int p, int q) {
            const double2 o = ee_ov<NB>(a, A, B, p, q);
            return make_double2(o.x * v12, o.y * v12);
        };
        const double2 o13 = vo(i1, i3), o24 = vo(i2, i4), o31 = vo(i3, i1), o42 = vo(i4, i2);
        const double *eA = a.e + A * NB, *eB = a.e + B * NB;
        const double d13 = eA[i1] - eB[i3], d24 = eA[i2] - eB[i4], d31 = eA[i3] - eB[i1], d42 = eA[i4] - eB[i2];
        for (int64_t c3 = 0; c3 < a.nk; c3++) {
            if (ex && A == B && B == c3) continue;
            const int64_t c4 = ee_k4(a, A, B, c3);
            if (c4 < 0) continue;
            const double *e3 = a.e + c3 * NB, *e4 = a.e + c4 * NB;
            // P1 needs an A2(i2 i4) entry inside the window, P2 an A2(i4 i2) one
            const bool w1 = ee_window<NB>(d24, e3, e4, c.m.ethr), w2 = ee_window<NB>(d42, e3, e4, c.m.ethr);
            if (!w1 && !w2) continue;
            double2 o34[NB2], o14[NB2], o32[NB2];
#pragma unroll
            for (int x = 0; x < NB; x++)
#pragma unroll
                for (int y = 0; y < NB; y++) o34[x * NB + y] = ee_ov<NB>(a, c3, c4, x, y);
            if (ex) {
                const double v14 = ee_v(a, c.m, A, c4);
#pragma unroll
                for (int p = 0; p < NB; p++)
#pragma unroll
                    for (int q = 0; q < NB; q++) {
                        const double2 o = ee_ov<NB>(a, A, c4, p, q);
                        o14[p * NB + q] = make_double2(o.x * v14, o.y * v14);
                        o32[p * NB + q] = ee_ov<NB>(a, c3, B, p, q);
                    }
            }
            // the four A blocks this element multiplies: (p q) = (i1 i3), (i2 i4), (i3 i1), (i4 i2)
            auto mee = [&](double2 opq, int p, int q, int x, int y) {
                double2 m = cmul(opq, o34[x * NB + y]);
                if (ex) {
                    const double2 t = cmul(o14[p * NB + y], o32[x * NB + q]);
                    m = make_double2(0.5 * m.x - 0.5 * t.x, 0.5 * m.y - 0.5 * t.y);
                }
                return m;
            };
            double2 b13[NB2], b24[NB2], b31[NB2], b42[NB2];
#pragma unroll
            for (int x = 0; x < NB; x++)
#pragma unroll
                for (int y = 0; y < NB; y++) {
                    const int xy = x * NB + y;
                    const double d34 = e3[x] - e4[y];
                    double2 m = mee(o13, i1, i3, x, y);
                    double dl = ee_dl(c, d13 + d34);
                    b13[xy] = lb ? make_double2(m.x * dl, m.y * dl) : m;
                    m = mee(o24, i2, i4, x, y);
                    dl = ee_dl(c, d24 + d34);
                    b24[xy] = make_double2(m.x * dl, m.y * dl);
                    m = mee(o31, i3, i1, x, y);
                    dl = ee_dl(c, d31 + d34);
                    b31[xy] = lb ? make_double2(m.x * dl, m.y * dl) : m;
                    m = mee(o42, i4, i2, x, y);
                    dl = ee_dl(c, d42 + d34);
                    b42[xy] = make_double2(m.x * dl, m.y * dl);
                }
            double2 t1[NB2], t2[NB2];
            if (a.dm && a.dP) {  // calc_A1rho_dP (:216-217): (dm1_k3, 1 - f3 | dm_k4, f4), (dm_k3, f3 | dm1_k4, 1 - f4)
                const double *f3 = a.f + c3 * NB, *f4 = a.f + c4 * NB;
                double g3[NB], g4[NB];
#pragma unroll
                for (int x = 0; x < NB; x++) {
                    g3[x] = 1 - f3[x];
                    g4[x] = 1 - f4[x];
                }
                ee_a1rho_dP<NB>(b13, a.dm1 + c3 * NB2, g3, a.dm + c4 * NB2, f4, t1);
                ee_a1rho_dP<NB>(b31, a.dm + c3 * NB2, f3, a.dm1 + c4 * NB2, g4, t2);
            } else if (a.dm) {  // calc_A1rho (ElecElec_Model.h:266-270): dm1_k3 A1 dm_k4, dm_k3 A1 dm1_k4
                const double2 *r3 = a.dm + c3 * NB2, *r4 = a.dm + c4 * NB2, *s3 = a.dm1 + c3 * NB2,
                              *s4 = a.dm1 + c4 * NB2;
                double2 u1[NB2], u2[NB2];
#pragma unroll
                for (int x = 0; x < NB; x++)
#pragma unroll
                    for (int y = 0; y < NB; y++) {
                        double2 s = make_double2(0.0, 0.0), t = make_double2(0.0, 0.0);
#pragma unroll
                        for (int k = 0; k < NB; k++) {
                            s = cadd(s, cmul(b13[x * NB + k], ee_hu<NB>(r4, k, y)));
                            t = cadd(t, cmul(b31[x * NB + k], ee_hu<NB>(s4, k, y)));
                        }
                        u1[x * NB + y] = s;
                        u2[x * NB + y] = t;
                    }
#pragma unroll
                for (int x = 0; x < NB; x++)
#pragma unroll
                    for (int y = 0; y < NB; y++) {
                        double2 s = make_double2(0.0, 0.0), t = make_double2(0.0, 0.0);
#pragma unroll
                        for (int k = 0; k < NB; k++) {
                            s = cadd(s, cmul(ee_hu<NB>(s3, x, k), u1[k * NB + y]));
                            t = cadd(t, cmul(ee_hu<NB>(r3, x, k), u2[k * NB + y]));
                        }
                        t1[x * NB + y] = s;
                        t2[x * NB + y] = t;
                    }
            } else {  // calc_A1f / calc_A1fbar (:282-293)
                const double *f3 = a.f + c3 * NB, *f4 = a.f + c4 * NB;
#pragma unroll
                for (int x = 0; x < NB; x++)
#pragma unroll
                    for (int y = 0; y < NB; y++) {
                        const double2 p = b13[x * NB + y], q = b31[x * NB + y];
                        t1[x * NB + y] = make_double2(p.x * (1 - f3[x]) * f4[y], p.y * (1 - f3[x]) * f4[y]);
                        t2[x * NB + y] = make_double2(q.x * f3[x] * (1 - f4[y]), q.y * f3[x] * (1 - f4[y]));
                    }
            }
#pragma unroll
            for (int xy = 0; xy < NB2; xy++) {
                acc1 = cadd(acc1, cmulc(t1[xy], b24[xy]));
                acc2 = cadd(acc2, cmulc(t2[xy], b42[xy]));
            }
        }
    }
    P1[idx] = make_double2(acc1.x * a.pfac, acc1.y * a.pfac);
    P2[idx] = make_double2(acc2.x * a.pfac, acc2.y * a.pfac);
}

// One (pair, k3) term for nb <= 2 (the body of k_ee_pair2): the overlaps
// <k3 x|k4 y> (and the exchange overlaps), the A2 blocks of every (p, q) with
// their Gaussians, then per form and (p, q) the A1 block times the density
// matrices (or f), added into every element.  W: 0 P1, 1 P2, 2 both (the A2
// blocks serve both forms).
template <int NB, int W>
__device__ __forceinline__ void ee_pair_term(const EeArgs &a, const EimpConst &c, int64_t A, int64_t B,
                                             const double2 *o12, const double *d12, int64_t c3, int64_t c4,
                                             double2 *acc1, double2 *acc2)
{
    constexpr int NB2 = NB * NB, NB3 = NB2 * NB;
    const bool ex = c.m.exchange != 0, lb = c.m.lindblad != 0;
    const double *e3 = a.e + c3 * NB, *e4 = a.e + c4 * NB;
    double2 o34[NB2], o14[NB2], o32[NB2];
#pragma unroll
    for (int x = 0; x < NB; x++)
#pragma unroll
        for (int y = 0; y < NB; y++) o34[x * NB + y] = ee_ov<NB>(a, c3, c4, x, y);
    if (ex) {
        const double v14 = ee_v(a, c.m, A, c4);
#pragma unroll
        for (int p = 0; p < NB; p++)
#pragma unroll
            for (int q = 0; q < NB; q++) {
                const double2 o = ee_ov<NB>(a, A, c4, p, q);
                o14[p * NB + q] = make_double2(o.x * v14, o.y * v14);
                o32[p * NB + q] = ee_ov<NB>(a, c3, B, p, q);
            }
    }
    auto mee = [&](int p, int q, int x, int y) {
        double2 m = cmul(o12[p * NB + q], o34[x * NB + y]);
        if (ex) {
            const double2 t = cmul(o14[p * NB + y], o32[x * NB + q]);
            m = make_double2(0.5 * m.x - 0.5 * t.x, 0.5 * m.y - 0.5 * t.y);
        }
        return m;
    };
    double2 a2[NB2][NB2];
#pragma unroll
    for (int pq = 0; pq < NB2; pq++)
#pragma unroll
        for (int xy = 0; xy < NB2; xy++) {
            const int p = pq / NB, q = pq % NB, x = xy / NB, y = xy % NB;
            const double2 m = mee(p, q, x, y);
            const double dl = ee_dl(c, d12[pq] + e3[x] - e4[y]);
            a2[pq][xy] = make_double2(m.x * dl, m.y * dl);
        }
    const double *f3 = a.f + c3 * NB, *f4 = a.f + c4 * NB;
#pragma unroll
    for (int which = (W == 1 ? 1 : 0); which < (W == 0 ? 1 : 2); which++) {
        // P1: dm1_k3 A1 dm_k4, P2: dm_k3 A1 dm1_k4 (upper-triangle Hermitian reads); null: diag f
        const double2 *L = nullptr, *Rm = nullptr;
        if (a.dm) {
            L = which ? a.dm + c3 * NB2 : a.dm1 + c3 * NB2;
            Rm = which ? a.dm1 + c4 * NB2 : a.dm + c4 * NB2;
        }
        double2 *acc = (which && W == 2) ? acc2 : acc1;
#pragma unroll
        for (int fv = 0; fv < NB * NB; fv++) {
            // P1: (i1, i3) = (fix, v); P2: (i3, i1) = (v, fix)
            const int fix = fv / NB, v = fv % NB;
            const int p = which ? v : fix, q = which ? fix : v;
            double2 a1[NB2], t[NB2];
#pragma unroll
            for (int xy = 0; xy < NB2; xy++) a1[xy] = lb ? a2[p * NB + q][xy] : mee(p, q, xy / NB, xy % NB);
            if (L && a.dP) {  // calc_A1rho_dP: P1 (dm1_k3, 1 - f3 | dm_k4, f4), P2 (dm_k3, f3 | dm1_k4, 1 - f4)
                double fl[NB], fr[NB];
#pragma unroll
                for (int x = 0; x < NB; x++) {
                    fl[x] = which ? f3[x] : 1 - f3[x];
                    fr[x] = which ? 1 - f4[x] : f4[x];
                }
                ee_a1rho_dP<NB>(a1, L, fl, Rm, fr, t);
            } else if (L) {
                double2 u[NB2];
#pragma unroll
                for (int x = 0; x < NB; x++)
#pragma unroll
                    for (int y = 0; y < NB; y++) {
                        double2 sm = make_double2(0.0, 0.0);
#pragma unroll
                        for (int k = 0; k < NB; k++) sm = cadd(sm, cmul(a1[x * NB + k], ee_hu<NB>(Rm, k, y)));
                        u[x * NB + y] = sm;
                    }
#pragma unroll
                for (int x = 0; x < NB; x++)
#pragma unroll
                    for (int y = 0; y < NB; y++) {
                        double2 sm = make_double2(0.0, 0.0);
#pragma unroll
                        for (int k = 0; k < NB; k++) sm = cadd(sm, cmul(ee_hu<NB>(L, x, k), u[k * NB + y]));
                        t[x * NB + y] = sm;
                    }
            } else {
#pragma unroll
                for (int x = 0; x < NB; x++)
#pragma unroll
                    for (int y = 0; y < NB; y++) {
                        const double w = which ? f3[x] * (1 - f4[y]) : (1 - f3[x]) * f4[y];
                        t[x * NB + y] = make_double2(a1[x * NB + y].x * w, a1[x * NB + y].y * w);
                    }
            }
            // P1 element (fix, r, v, s) with A2 of (i2, i4) = (r, s); P2 element
            // (fix, s, v, r) with A2 of (i4, i2) = (r, s)
#pragma unroll
            for (int r = 0; r < NB; r++)
#pragma unroll
                for (int s = 0; s < NB; s++) {
                    const int li = fix * NB3 + (which ? (s * NB + v) * NB + r : (r * NB + v) * NB + s);
                    double2 sum = acc[li];
#pragma unroll
                    for (int xy = 0; xy < NB2; xy++) sum = cadd(sum, cmulc(t[xy], a2[r * NB + s][xy]));
                    acc[li] = sum;
                }
        }
    }
}

// P1 and P2 for nb <= 2, one wavefront per stored pair.  The lanes test 64
// k3 at a time (k4 on the mesh, some A2 entry inside the energy window) and
// append the survivors to a per-wavefront LDS queue (ballot and popcount);
// every 64 queued k3 run the term with all lanes busy, so the lanes do not
// idle through the terms of k3 outside the window, and the overlaps and A2
// blocks serve both forms.  A wavefront reduction ends the pair.
template <int NB, int W>
__global__ __launch_bounds__(256) void k_ee_pair2(int64_t n, const uint64_t *__restrict__ k1,
                                                  const uint64_t *__restrict__ k2, EeArgs a, EimpConst c,
                                                  double2 *__restrict__ P1, double2 *__restrict__ P2)
{
    constexpr int NB2 = NB * NB, NB4 = NB2 * NB2;
    __shared__ int2 queue[4][128];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t fp = (int64_t)blockIdx.x * 4 + wv;
    if (fp >= n) return;  // whole wavefronts; only wavefront-level synchronisation below
    const int64_t A = (int64_t)k1[fp], B = (int64_t)k2[fp];
    const bool ex = c.m.exchange != 0;
    constexpr int NA2 = W == 2 ? NB4 : 1;  // acc1: P1 (W = 0, 2) or P2 (W = 1); acc2: P2 when W = 2
    double2 acc1[NB4], acc2[NA2];
#pragma unroll
    for (int e = 0; e < NB4; e++) acc1[e] = make_double2(0.0, 0.0);
#pragma unroll
    for (int e = 0; e < NA2; e++) acc2[e] = make_double2(0.0, 0.0);
    if (A != B || ex) {
        const double v12 = a.vq_pair ? a.vq_pair[fp] : ee_v(a, c.m, A, B);
        const double *eA = a.e + A * NB, *eB = a.e + B * NB;
        double2 o12[NB2];
        double d12[NB2], dmin = 1e300, dmax = -1e300;
#pragma unroll
        for (int p = 0; p < NB; p++)
#pragma unroll
            for (int q = 0; q < NB; q++) {
                const double2 o = ee_ov<NB>(a, A, B, p, q);
                o12[p * NB + q] = make_double2(o.x * v12, o.y * v12);
                d12[p * NB + q] = eA[p] - eB[q];
                dmin = fmin(dmin, d12[p * NB + q]);
                dmax = fmax(dmax, d12[p * NB + q]);
            }
        int qlen = 0;  // wavefront-uniform
        const uint64_t below = (1ull << lane) - 1ull;
        for (int64_t base = 0; base < a.nk; base += 64) {
            const int64_t c3 = base + lane;
            int64_t c4 = -1;
            bool ok = false;
            if (c3 < a.nk && !(ex && A == B && B == c3)) {
                c4 = ee_k4(a, A, B, c3);
                if (c4 >= 0) {
                    const double *e3 = a.e + c3 * NB, *e4 = a.e + c4 * NB;
                    double tlo = 1e300, thi = -1e300;
#pragma unroll
                    for (int x = 0; x < NB; x++)
#pragma unroll
                        for (int y = 0; y < NB; y++) {
                            tlo = fmin(tlo, e3[x] - e4[y]);
                            thi = fmax(thi, e3[x] - e4[y]);
                        }
                    // some A2 entry inside the window
                    ok = dmin + tlo < c.m.ethr && dmax + thi > -c.m.ethr;
                }
            }
            const uint64_t m = __ballot(ok);
            if (ok) queue[wv][qlen + __popcll(m & below)] = make_int2((int)c3, (int)c4);
            qlen += __popcll(m);
            __builtin_amdgcn_wave_barrier();
            if (qlen >= 64) {
                const int2 q = queue[wv][lane];
                ee_pair_term<NB, W>(a, c, A, B, o12, d12, q.x, q.y, acc1, acc2);
                const int rest = qlen - 64;
                int2 r = make_int2(0, 0);
                if (lane < rest) r = queue[wv][64 + lane];
                __builtin_amdgcn_wave_barrier();
                if (lane < rest) queue[wv][lane] = r;
                __builtin_amdgcn_wave_barrier();
                qlen = rest;
            }
        }
        if (lane < qlen) {
            const int2 q = queue[wv][lane];
            ee_pair_term<NB, W>(a, c, A, B, o12, d12, q.x, q.y, acc1, acc2);
        }
    }
#pragma unroll
    for (int e = 0; e < NB4; e++)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            acc1[e].x += __shfl_down(acc1[e].x, off, 64);
            acc1[e].y += __shfl_down(acc1[e].y, off, 64);
        }
#pragma unroll
    for (int e = 0; e < NA2; e++)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            acc2[e].x += __shfl_down(acc2[e].x, off, 64);
            acc2[e].y += __shfl_down(acc2[e].y, off, 64);
        }
    if (lane == 0) {
        double2 *o1 = (W == 1 ? P2 : P1) + fp * NB4;
#pragma unroll
        for (int e = 0; e < NB4; e++) o1[e] = make_double2(acc1[e].x * a.pfac, acc1[e].y * a.pfac);
        if (W == 2) {
#pragma unroll
            for (int e = 0; e < NA2; e++) P2[fp * NB4 + e] = make_double2(acc2[e].x * a.pfac, acc2[e].y * a.pfac);
        }
    }
}

// P3..P6 of n stored pairs (k1, ka) at f, one thread per element (i1 i2 ia ib):
//   P3/P4: k3 over k, k4 = k1 + k3 - ka, A from calc_A(k1, k4, k3, ka),
//     P3 += sum_{i4 i3} A1[(i1 i4),(i3 ia)] (1-f3) f4 conj(A2[(i2 i4),(i3 ib)]),
//     P4 the same with f3 (1-f4); skipped with exchange, as the reference does;
//   P5/P6: k3 over k, k4 = k1 + ka - k3, A from calc_A(k1, k3, ka, k4),
//     P5 += sum_{i3 i4} A1[(i1 i3),(ib i4)] f3 f4 conj(A2[(i2 i3),(ia i4)]),
//     P6 the same with (1-f3)(1-f4).
template <int NB>
__global__ __launch_bounds__(256) void k_ee_P3456(int64_t n, const uint64_t *__restrict__ k1,
                                                  const uint64_t *__restrict__ k2, EeArgs a, EimpConst c,
                                                  double2 *__restrict__ P3, double2 *__restrict__ P4,
                                                  double2 *__restrict__ P5, double2 *__restrict__ P6)
{
    constexpr int NB2 = NB * NB, NB4 = NB2 * NB2;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * NB4) return;
    const int64_t fp = idx / NB4;
    const int el = (int)(idx - fp * NB4);
    const int i1 = el / (NB * NB2), i2 = (el / NB2) % NB, ia = (el / NB) % NB, ib = el % NB;
    const int64_t A = (int64_t)k1[fp], B = (int64_t)k2[fp];
    const bool ex = c.m.exchange != 0, lb = c.m.lindblad != 0;
    const double *eA = a.e + A * NB, *eB = a.e + B * NB;
    double2 acc3 = make_double2(0.0, 0.0), acc4 = acc3, acc5 = acc3, acc6 = acc3;
    if (!ex) {
        for (int64_t c3 = 0; c3 < a.nk; c3++) {
            if (B == c3) continue;
            const int64_t c4 = ee_k4(a, A, B, c3);
            if (c4 < 0) continue;
            const double *e3 = a.e + c3 * NB, *e4 = a.e + c4 * NB, *f3 = a.f + c3 * NB, *f4 = a.f + c4 * NB;
            // A2(p = i2, q = i4 at k4, x = i3 at k3, y = ib at ka): de = eA[i2] - e4[i4] + e3[i3] - eB[ib]
            {
                double lo = 1e300, hi = -1e300;
#pragma unroll
                for (int q = 0; q < NB; q++)
#pragma unroll
                    for (int x = 0; x < NB; x++) {
                        const double t = e3[x] - e4[q];
                        lo = fmin(lo, t);
                        hi = fmax(hi, t);
                    }
                const double d = eA[i2] - eB[ib];
                if (!(d + lo < c.m.ethr && d + hi > -c.m.ethr)) continue;
            }
            const double v = ee_v(a, c.m, A, c4);
            double2 oa1[NB], oa2[NB], ob1[NB], ob2[NB];  // v <k1 i1|k4 q>, v <k1 i2|k4 q>, <k3 x|ka ia>, <k3 x|ka ib>
#pragma unroll
            for (int q = 0; q < NB; q++) {
                double2 o = ee_ov<NB>(a, A, c4, i1, q);
                oa1[q] = make_double2(o.x * v, o.y * v);
                o = ee_ov<NB>(a, A, c4, i2, q);
                oa2[q] = make_double2(o.x * v, o.y * v);
                ob1[q] = ee_ov<NB>(a, c3, B, q, ia);
                ob2[q] = ee_ov<NB>(a, c3, B, q, ib);
            }
#pragma unroll
            for (int j4 = 0; j4 < NB; j4++)
#pragma unroll
                for (int j3 = 0; j3 < NB; j3++) {
                    double2 m1 = cmul(oa1[j4], ob1[j3]);
                    const double dl1 = ee_dl(c, eA[i1] - e4[j4] + e3[j3] - eB[ia]);
                    if (lb) m1 = make_double2(m1.x * dl1, m1.y * dl1);
                    double2 m2 = cmul(oa2[j4], ob2[j3]);
                    const double dl2 = ee_dl(c, eA[i2] - e4[j4] + e3[j3] - eB[ib]);
                    m2 = make_double2(m2.x * dl2, m2.y * dl2);
                    const double w3 = (1 - f3[j3]) * f4[j4], w4 = f3[j3] * (1 - f4[j4]);
                    const double2 r3 = make_double2(m1.x * w3, m1.y * w3), r4 = make_double2(m1.x * w4, m1.y * w4);
                    acc3 = cadd(acc3, cmulc(r3, m2));
                    acc4 = cadd(acc4, cmulc(r4, m2));
                }
        }
    }
    for (int64_t c3 = 0; c3 < a.nk; c3++) {
        if (A == c3 && !ex) continue;
        if (ex && A == c3 && B == c3) continue;
        const int64_t c4 = ee_k4(a, A, c3, B);  // k1 + ka - k3
        if (c4 < 0) continue;
        const double *e3 = a.e + c3 * NB, *e4 = a.e + c4 * NB, *f3 = a.f + c3 * NB, *f4 = a.f + c4 * NB;
        // A2(p = i2, q = i3 at k3, x = ia at ka, y = i4 at k4): de = eA[i2] - e3[i3] + eB[ia] - e4[i4]
        {
            double lo = 1e300, hi = -1e300;
#pragma unroll
            for (int q = 0; q < NB; q++)
#pragma unroll
                for (int y = 0; y < NB; y++) {
                    const double t = -e3[q] - e4[y];
                    lo = fmin(lo, t);
                    hi = fmax(hi, t);
                }
            const double d = eA[i2] + eB[ia];
            if (!(d + lo < c.m.ethr && d + hi > -c.m.ethr)) continue;
        }
        const double v = ee_v(a, c.m, A, c3);
        const double vx = ex ? ee_v(a, c.m, A, c4) : 0.0;
        // direct: v <k1 p|k3 q> <ka x|k4 y>; exchange: v(k1 - k4) <k1 p|k4 y> <ka x|k3 q>
        double2 p1[NB], p2[NB], qb[NB], qa[NB], x1[NB], x2[NB], yb[NB], ya[NB];
#pragma unroll
        for (int j = 0; j < NB; j++) {
            double2 o = ee_ov<NB>(a, A, c3, i1, j);
            p1[j] = make_double2(o.x * v, o.y * v);
            o = ee_ov<NB>(a, A, c3, i2, j);
            p2[j] = make_double2(o.x * v, o.y * v);
            qb[j] = ee_ov<NB>(a, B, c4, ib, j);
            qa[j] = ee_ov<NB>(a, B, c4, ia, j);
            if (ex) {
                o = ee_ov<NB>(a, A, c4, i1, j);
                x1[j] = make_double2(o.x * vx, o.y * vx);
                o = ee_ov<NB>(a, A, c4, i2, j);
                x2[j] = make_double2(o.x * vx, o.y * vx);
                yb[j] = ee_ov<NB>(a, B, c3, ib, j);
                ya[j] = ee_ov<NB>(a, B, c3, ia, j);
            }
        }
#pragma unroll
        for (int j3 = 0; j3 < NB; j3++)
#pragma unroll
            for (int j4 = 0; j4 < NB; j4++) {
                // A1[(i1 j3),(ib j4)] and A2[(i2 j3),(ia j4)]
                double2 m1 = cmul(p1[j3], qb[j4]);
                double2 m2 = cmul(p2[j3], qa[j4]);
                if (ex) {
                    const double2 t1 = cmul(x1[j4], yb[j3]), t2 = cmul(x2[j4], ya[j3]);
                    m1 = make_double2(0.5 * m1.x - 0.5 * t1.x, 0.5 * m1.y - 0.5 * t1.y);
                    m2 = make_double2(0.5 * m2.x - 0.5 * t2.x, 0.5 * m2.y - 0.5 * t2.y);
                }
                const double dl1 = ee_dl(c, eA[i1] - e3[j3] + eB[ib] - e4[j4]);
                if (lb) m1 = make_double2(m1.x * dl1, m1.y * dl1);
                const double dl2 = ee_dl(c, eA[i2] - e3[j3] + eB[ia] - e4[j4]);
                m2 = make_double2(m2.x * dl2, m2.y * dl2);
                const double w5 = f3[j3] * f4[j4], w6 = (1 - f3[j3]) * (1 - f4[j4]);
                const double2 r5 = make_double2(m1.x * w5, m1.y * w5), r6 = make_double2(m1.x * w6, m1.y * w6);
                acc5 = cadd(acc5, cmulc(r5, m2));
                acc6 = cadd(acc6, cmulc(r6, m2));
            }
    }
    const double pf = a.pfac;
    P3[idx] = make_double2(acc3.x * pf, acc3.y * pf);
    P4[idx] = make_double2(acc4.x * pf, acc4.y * pf);
    P5[idx] = make_double2(acc5.x * pf, acc5.y * pf);
    P6[idx] = make_double2(acc6.x * pf, acc6.y * pf);
}

// Lscij / Lscji increments of n stored pairs (file order) at f_eq
// (ElectronPhonon_linear.cpp:183-253): extra x the P1/P2 form of Pee (the
// doubling of fac2 = ee_antisymmetry beyond what set_Lsc already holds),
// -coef34 x the P3/P4 form (coef34 = 1 is the reference's sign), and minus
// the P5/P6 form.  Any of Pee, P3, P5 may be null.
template <int NB>
__global__ __launch_bounds__(256) void k_lsc_ee(int64_t n, const uint64_t *__restrict__ k1,
                                                const uint64_t *__restrict__ k2, const double *__restrict__ f,
                                                double extra, double coef34, const double2 *__restrict__ Pe1,
                                                const double2 *__restrict__ Pe2, const double2 *__restrict__ P3,
                                                const double2 *__restrict__ P4, const double2 *__restrict__ P5,
                                                const double2 *__restrict__ P6, double2 *__restrict__ dLij,
                                                double2 *__restrict__ dLji)
{
    constexpr int NB2 = NB * NB, NB4 = NB2 * NB2;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * NB4) return;
    const int64_t fp = idx / NB4;
    const int el = (int)(idx - fp * NB4);
    const int i1 = el / (NB * NB2), i2 = (el / NB2) % NB, ia = (el / NB) % NB, ib = el % NB;
    const int i12 = i1 * NB + i2, iab = ia * NB + ib;
    const int64_t base = fp * NB4;
    const int64_t i12ab = base + i12 * NB2 + iab, iab12 = base + iab * NB2 + i12,
                  iba21 = base + (ib * NB + ia) * NB2 + i2 * NB + i1;
    const double *fi = f + (int64_t)k1[fp] * NB, *fj = f + (int64_t)k2[fp] * NB;
    double2 x = make_double2(0.0, 0.0), y = make_double2(0.0, 0.0);
    auto axpy = [](double2 &acc, double s, double2 v) { acc = make_double2(acc.x + s * v.x, acc.y + s * v.y); };
    if (Pe1 && extra != 0.0) {
        axpy(x, extra * (1 - fi[i1]), Pe1[i12ab]);
        axpy(x, extra * fi[i2], Pe2[iab12]);
        axpy(y, extra * (1 - fj[i1]), cconj(Pe2[i12ab]));
        axpy(y, extra * fj[i2], cconj(Pe1[iab12]));
    }
    if (P3 && coef34 != 0.0) {
        axpy(x, -coef34 * (1 - fi[i1]), P3[i12ab]);
        axpy(x, -coef34 * fi[i2], P4[i12ab]);
        axpy(y, -coef34 * (1 - fj[i1]), cconj(P4[iab12]));
        axpy(y, -coef34 * fj[i2], cconj(P3[iab12]));
    }
    if (P5) {
        axpy(x, -(1 - fi[i1]), P5[i12ab]);
        axpy(x, -fi[i2], P6[i12ab]);
        axpy(y, -(1 - fj[i1]), P5[iba21]);
        axpy(y, -fj[i2], P6[iba21]);
    }
    dLij[idx] = x;
    dLji[idx] = y;
}

// alg_linearize_dPee's row sums of n stored pairs' dP (evolve_eph.cpp:107-131):
//   dPdm_k[i12]  += sum_i3 dP1[(i12),(i3 i3)] f_k'[i3]
//   dm1dP_k[i12] += sum_i3 (1 - f_k'[i3]) dP2[(i3 i3),(i12)]
// and, for k < k', with P1_next = conj(dP2) and P2_next = conj(dP1),
//   dPdm_k'[i12]  += sum_i3 conj(dP2[(i12),(i3 i3)]) f_k[i3]
//   dm1dP_k'[i12] += sum_i3 (1 - f_k[i3]) conj(dP1[(i3 i3),(i12)]).
// One thread per (pair, i12); acc holds dPdm then dm1dP (nk x nb^2 complex each).
template <int NB>
__global__ __launch_bounds__(256) void k_dP_rows(int64_t n, const uint64_t *__restrict__ k1,
                                                 const uint64_t *__restrict__ k2, const double2 *__restrict__ dP1,
                                                 const double2 *__restrict__ dP2, const double *__restrict__ f,
                                                 int64_t nk, double *__restrict__ acc)
{
    constexpr int NB2 = NB * NB, NB4 = NB2 * NB2;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * NB2) return;
    const int64_t fp = idx / NB2;
    const int i12 = (int)(idx - fp * NB2);
    const int64_t k = (int64_t)k1[fp], kp = (int64_t)k2[fp];
    const double2 *D1 = dP1 + fp * NB4, *D2 = dP2 + fp * NB4;
    double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
#pragma unroll
    for (int i3 = 0; i3 < NB; i3++) {
        const int d = i3 * NB + i3;
        const double fq = f[kp * NB + i3];
        const double2 x = D1[i12 * NB2 + d], y = D2[d * NB2 + i12];
        a = make_double2(a.x + x.x * fq, a.y + x.y * fq);
        b = make_double2(b.x + y.x * (1 - fq), b.y + y.y * (1 - fq));
    }
    double *pd = acc + 2 * (k * NB2 + i12), *pm = acc + 2 * ((nk + k) * NB2 + i12);
    unsafeAtomicAdd(pd, a.x);
    unsafeAtomicAdd(pd + 1, a.y);
    unsafeAtomicAdd(pm, b.x);
    unsafeAtomicAdd(pm + 1, b.y);
    if (k < kp) {
        a = make_double2(0.0, 0.0);
        b = make_double2(0.0, 0.0);
#pragma unroll
        for (int i3 = 0; i3 < NB; i3++) {
            const int d = i3 * NB + i3;
            const double fq = f[k * NB + i3];
            const double2 x = D2[i12 * NB2 + d], y = D1[d * NB2 + i12];
            a = make_double2(a.x + x.x * fq, a.y - x.y * fq);
            b = make_double2(b.x + y.x * (1 - fq), b.y - y.y * (1 - fq));
        }
        pd = acc + 2 * (kp * NB2 + i12);
        pm = acc + 2 * ((nk + kp) * NB2 + i12);
        unsafeAtomicAdd(pd, a.x);
        unsafeAtomicAdd(pd + 1, a.y);
        unsafeAtomicAdd(pm, b.x);
        unsafeAtomicAdd(pm + 1, b.y);
    }
}

// The row term of evolve_eph.cpp:144-151 without its phase: D_ij = (1 - f_i)
// dPdm_ij - dm1dP_ij f_j, and H = prefac/2 (D + D^+), the part the assembly's
// Hermitian symmetrization keeps (cis((e_i - e_j) t) is common to D_ij and
// conj(D_ji cis((e_j - e_i) t)), so the phase multiplies H afterwards).
template <int NB>
__global__ __launch_bounds__(256) void k_dP_herm(int64_t nk, const double *__restrict__ acc,
                                                 const double *__restrict__ f, double prefac, double2 *__restrict__ H)
{
    constexpr int NB2 = NB * NB;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    const double2 *pd = (const double2 *)acc + k * NB2, *pm = (const double2 *)acc + (nk + k) * NB2;
    const double *fk = f + k * NB;
    double2 D[NB2];
#pragma unroll
    for (int i = 0; i < NB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) {
            const double2 x = pd[i * NB + j], y = pm[i * NB + j];
            D[i * NB + j] = make_double2((1 - fk[i]) * x.x - y.x * fk[j], (1 - fk[i]) * x.y - y.y * fk[j]);
        }
    const double s = 0.5 * prefac;
#pragma unroll
    for (int i = 0; i < NB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++)
            H[k * NB2 + i * NB + j] =
                make_double2(s * (D[i * NB + j].x + D[j * NB + i].x), s * (D[i * NB + j].y - D[j * NB + i].y));
}

namespace {

// Device copies of what every e-e call needs, released together.
struct EeSetup {
    double2 *dU = nullptr, *ddm = nullptr, *ddm1 = nullptr;
    double *dk = nullptr, *df = nullptr, *dvc = nullptr;
    int *d_ik3 = nullptr, *d_grid = nullptr;
    ~EeSetup()
    {
        dfree(dU); dfree(ddm); dfree(ddm1); dfree(dk); dfree(df); dfree(dvc); dfree(d_ik3); dfree(d_grid);
    }
};

int ee_check(dmd_scatt *h, const uint64_t *k1st, const uint64_t *k2nd, const double *U, const double *kvec,
             const double *f, const dmd_eimp_model *m, const char *who)
{
    if (!h || !m || (h->np > 0 && (!k1st || !k2nd || !U || !kvec || !f)))
        return fail(DMD_ERR_ARG, std::string(who) + ": null argument");
    if (const int sg = sp_guard(h, who)) return sg;
    if (m->nb_wannier <= 0 || !(m->degauss > 0) || m->screening < DMD_SCR_UNSCREENED || m->screening > DMD_SCR_RPA ||
        m->detail_balance)
        return fail(DMD_ERR_ARG, std::string(who) + ": bad model (nb_wannier, degauss, screening, detail_balance)");
    const int *km = m->kmesh;
    if (km[0] <= 0 || km[1] <= 0 || km[2] <= 0 || (int64_t)km[0] * km[1] * km[2] > (1ll << 31))
        return fail(DMD_ERR_ARG, std::string(who) + ": the e-e sum needs the k mesh");
    if (h->nb < 1 || h->nb > 4) return fail(DMD_ERR_UNSUPPORTED, "the model e-e P is built for nb = 1..4");
    // nb = 3 is switched off: a device run of the nb = 3 instance failed and its
    // cause is not found (the restatement matches the oracle on the CPU)
    if (h->nb == 3) return fail(DMD_ERR_UNSUPPORTED, "the model e-e P is not available for nb = 3 in this version");
    for (int64_t q = 0; q < h->np; q++)
        if (k1st[q] >= (uint64_t)h->nk || k2nd[q] >= (uint64_t)h->nk)
            return fail(DMD_ERR_ARG, std::string(who) + ": k index outside [0, nk)");
    return DMD_OK;
}

// Uploads U, kvec, f (and the state), builds the mesh map, fills the view.
int ee_setup(dmd_scatt *h, const double *U, const double *kvec, const double *f, const double *dm, const double *dm1,
             const double *vq_cell, const dmd_eimp_model &m, EeSetup &s, EeArgs &a)
{
    const int *km = m.kmesh;
    const int64_t nk = h->nk, nw = m.nb_wannier, cells = (int64_t)km[0] * km[1] * km[2];
    std::vector<int> ik3((size_t)nk * 3);
    for (int64_t k = 0; k < nk; k++) mesh_cell(kvec + 3 * k, km, &ik3[3 * k]);
    int rc;
    if ((rc = dalloc_t(&s.dU, (size_t)(nk * nw * h->nb), "U")) || (rc = dalloc_t(&s.dk, (size_t)nk * 3, "kvec")) ||
        (rc = dalloc_t(&s.df, (size_t)nk * h->nb, "f")) || (rc = dalloc_t(&s.d_ik3, (size_t)nk * 3, "k cells")) ||
        (rc = dalloc_t(&s.d_grid, (size_t)cells, "k mesh map")))
        return rc;
    if (dm && ((rc = dalloc_t(&s.ddm, (size_t)nk * h->nb2, "dm")) || (rc = dalloc_t(&s.ddm1, (size_t)nk * h->nb2, "dm1"))))
        return rc;
    if (vq_cell && (rc = dalloc_t(&s.dvc, (size_t)cells, "v(q) per cell"))) return rc;
    hipError_t e = hipMemcpyAsync(s.dU, U, (size_t)(nk * nw * h->nb) * 16, hipMemcpyDefault, h->st);
    if (e == hipSuccess) e = hipMemcpyAsync(s.dk, kvec, (size_t)nk * 3 * 8, hipMemcpyDefault, h->st);
    if (e == hipSuccess) e = hipMemcpyAsync(s.df, f, (size_t)nk * h->nb * 8, hipMemcpyDefault, h->st);
    if (e == hipSuccess) e = hipMemcpyAsync(s.d_ik3, ik3.data(), ik3.size() * 4, hipMemcpyHostToDevice, h->st);
    if (e == hipSuccess && dm) e = hipMemcpyAsync(s.ddm, dm, (size_t)nk * h->nb2 * 16, hipMemcpyDefault, h->st);
    if (e == hipSuccess && dm) e = hipMemcpyAsync(s.ddm1, dm1, (size_t)nk * h->nb2 * 16, hipMemcpyDefault, h->st);
    if (e == hipSuccess && vq_cell) e = hipMemcpyAsync(s.dvc, vq_cell, (size_t)cells * 8, hipMemcpyDefault, h->st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fill_int, dim3(nblk(cells, 256)), dim3(256), 0, h->st, cells, s.d_grid, INT_MAX);
        hipLaunchKernelGGL(k_kgrid, dim3(nblk(nk, 256)), dim3(256), 0, h->st, nk, s.d_ik3, km[1], km[2], s.d_grid);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);  // the host vectors go out of scope
    if (e != hipSuccess) return fail(DMD_ERR_HIP, hipGetErrorString(e));
    a = EeArgs{};
    a.U = s.dU;
    a.ik3 = s.d_ik3;
    a.grid = s.d_grid;
    a.e = h->d_e;
    a.f = s.df;
    a.dm = s.ddm;
    a.dm1 = s.ddm1;
    a.kvec = s.dk;
    a.vq_cell = s.dvc;
    a.vq_pair = nullptr;
    a.nk = nk;
    a.nw = (int)nw;
    a.m0 = km[0];
    a.m1 = km[1];
    a.m2 = km[2];
    a.pfac = (m.exchange ? 2.0 : 1.0) / (m.degauss * std::sqrt(2. * M_PI)) / m.nk_full;
    return DMD_OK;
}

EimpConst ee_const(const dmd_eimp_model &m)
{
    EimpConst c;
    c.m = m;
    c.prefac_delta = 1.0;
    c.prefac_exp = m.lindblad ? -0.25 / std::pow(m.degauss, 2) : -0.5 / std::pow(m.degauss, 2);
    return c;
}

// nb <= 2: a wavefront per pair and form; nb > 2: a thread per element
template <int NB>
void launch_ee_P(dmd_scatt *h, int64_t n, const uint64_t *d1, const uint64_t *d2, const EeArgs &a, const EimpConst &c,
                 double2 *s1, double2 *s2)
{
    if constexpr (NB <= 2) {
        const dim3 g(nblk(n, 4)), b(256);
        // both forms in one pass (the A2 blocks serve both; twice the sums held);
        // DMD_EE_FORMS=1: one pass per form
        const char *env = std::getenv("DMD_EE_FORMS");
        if (!(env && env[0] == '1')) {
            hipLaunchKernelGGL((k_ee_pair2<NB, 2>), g, b, 0, h->st, n, d1, d2, a, c, s1, s2);
        } else {
            hipLaunchKernelGGL((k_ee_pair2<NB, 0>), g, b, 0, h->st, n, d1, d2, a, c, s1, s2);
            hipLaunchKernelGGL((k_ee_pair2<NB, 1>), g, b, 0, h->st, n, d1, d2, a, c, s1, s2);
        }
    } else
        hipLaunchKernelGGL(k_ee_P<NB>, dim3(nblk(n * NB * NB * NB * NB, 256)), dim3(256), 0, h->st, n, d1, d2, a, c, s1,
                           s2);
}

// The fold of scale x P_ee into P of every stored pair (add = 0: set), chunked.
// vq_pair_host: v(k1 - k2) of every stored pair (or null: a.vq_cell / closed form).
// rows (device, 2 nk nb^2 complex, zeroed): the chunks' dP go into the
// alg_linearize_dPee row sums (k_dP_rows) instead, and P is left as it is.
int ee_fold(dmd_scatt *h, const uint64_t *k1st, const uint64_t *k2nd, EeArgs a, const EimpConst &c,
            const std::vector<double> *vq_pair_host, double scale, int add, double *rows = nullptr)
{
    const int64_t per = h->nb4;
    const int64_t chunk = std::min<int64_t>(h->np, std::max<int64_t>(1, (64ll << 20) / per));
    double2 *s1 = nullptr, *s2 = nullptr;
    uint64_t *d1 = nullptr, *d2 = nullptr;
    double *dvq = nullptr;
    int rc;
    auto release = [&]() { dfree(s1); dfree(s2); dfree(d1); dfree(d2); dfree(dvq); };
    if ((rc = dalloc_t(&s1, (size_t)(chunk * per), "e-e P staging")) ||
        (rc = dalloc_t(&s2, (size_t)(chunk * per), "e-e P staging")) || (rc = dalloc_t(&d1, (size_t)chunk, "k1")) ||
        (rc = dalloc_t(&d2, (size_t)chunk, "k2")) ||
        (vq_pair_host && (rc = dalloc_t(&dvq, (size_t)chunk, "v(q)")))) {
        release();
        return rc;
    }
    hipError_t e = hipSuccess;
    for (int64_t first = 0; first < h->np && e == hipSuccess && rc == DMD_OK; first += chunk) {
        const int64_t n = std::min(chunk, h->np - first);
        e = hipMemcpyAsync(d1, k1st + first, (size_t)n * 8, hipMemcpyDefault, h->st);
        if (e == hipSuccess) e = hipMemcpyAsync(d2, k2nd + first, (size_t)n * 8, hipMemcpyDefault, h->st);
        if (e == hipSuccess && vq_pair_host)
            e = hipMemcpyAsync(dvq, vq_pair_host->data() + first, (size_t)n * 8, hipMemcpyHostToDevice, h->st);
        if (e != hipSuccess) break;
        a.vq_pair = vq_pair_host ? dvq : nullptr;
        switch (h->nb) {
        case 1: launch_ee_P<1>(h, n, d1, d2, a, c, s1, s2); break;
        case 2: launch_ee_P<2>(h, n, d1, d2, a, c, s1, s2); break;
        case 3: launch_ee_P<3>(h, n, d1, d2, a, c, s1, s2); break;
        default: launch_ee_P<4>(h, n, d1, d2, a, c, s1, s2); break;
        }
        if ((e = hipGetLastError()) != hipSuccess) break;
        if (rows) {
            const dim3 g(nblk(n * h->nb2, 256)), b(256);
            switch (h->nb) {
            case 1: hipLaunchKernelGGL(k_dP_rows<1>, g, b, 0, h->st, n, d1, d2, s1, s2, a.f, h->nk, rows); break;
            case 2: hipLaunchKernelGGL(k_dP_rows<2>, g, b, 0, h->st, n, d1, d2, s1, s2, a.f, h->nk, rows); break;
            case 3: hipLaunchKernelGGL(k_dP_rows<3>, g, b, 0, h->st, n, d1, d2, s1, s2, a.f, h->nk, rows); break;
            default: hipLaunchKernelGGL(k_dP_rows<4>, g, b, 0, h->st, n, d1, d2, s1, s2, a.f, h->nk, rows); break;
            }
            e = hipGetLastError();
        } else {
            rc = place_P_dev(h, s1, s2, n, first, scale, add ? 1 : 0);
        }
        // the next chunk's pair copies come from pageable host memory
        if (rc == DMD_OK && e == hipSuccess) e = hipStreamSynchronize(h->st);
    }
    hipError_t e2 = hipStreamSynchronize(h->st);
    release();
    if (rc == DMD_OK && e == hipSuccess) e = e2;
    if (rc == DMD_OK && e != hipSuccess) rc = fail(DMD_ERR_HIP, hipGetErrorString(e));
    if (rows) return rc;
    if (rc) {
        h->has_P = false;
        return rc;
    }
    h->has_P = true;
    return DMD_OK;
}

}  // namespace
}  // namespace dmd

using namespace dmd;

// The e-e fold-in of set_eph, add_scatt_contrib("ee") (ElectronPhonon.cpp:10-11,
// 44-113) for eeMode = Pee_fixed_at_eq: P += scale * P_ee(f); m carries the
// screening (as for the e-i term; prefac_A unused), degauss = degauss_ee, ethr,
// lindblad, nb_wannier, kmesh, nk_full and exchange.  The static RPA v(q) of
// the direct term is summed for the stored pairs' q; with exchange it comes
// from the per-cell table (every k-point difference).
extern "C" int dmd_scatt_add_ee_model(dmd_scatt *h, const uint64_t *k1st, const uint64_t *k2nd, const double *U,
                                      const double *kvec, const double *f, const dmd_eimp_model *m, double scale,
                                      int add)
{
    int rc;
    if ((rc = ee_check(h, k1st, k2nd, U, kvec, f, m, "dmd_scatt_add_ee_model"))) return rc;
    if (h->linear) return fail(DMD_ERR_STATE, "the handle holds the linearized operator: add the e-e P before set_Lsc");
    if (add && !h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated (add = 0 sets it)");
    if (h->np == 0) { h->has_P = true; return DMD_OK; }
    EeSetup s;
    EeArgs a;
    if ((rc = ee_setup(h, U, kvec, f, nullptr, nullptr, nullptr, *m, s, a))) return rc;
    std::vector<double> vq, vq_cell;
    if (m->screening == DMD_SCR_RPA) {
        if (m->exchange) {
            if ((rc = rpa_vq_cells(h, s.dU, kvec, f, *m, vq_cell))) return rc;
            if ((rc = dalloc_t(&s.dvc, vq_cell.size(), "v(q) per cell"))) return rc;
            DMD_HIP(hipMemcpy(s.dvc, vq_cell.data(), vq_cell.size() * 8, hipMemcpyHostToDevice));
            a.vq_cell = s.dvc;
        } else if ((rc = rpa_vq(h, k1st, k2nd, s.dU, kvec, f, *m, vq))) {
            return rc;
        }
    }
    return ee_fold(h, k1st, k2nd, a, ee_const(*m), vq.empty() ? nullptr : &vq, scale, add);
}

extern "C" int dmd_scatt_ee_vq_cells(dmd_scatt *h, const double *U, const double *kvec, const double *f,
                                     const dmd_eimp_model *m, double *vq_cell)
{
    if (!h || !U || !kvec || !f || !m || !vq_cell) return fail(DMD_ERR_ARG, "dmd_scatt_ee_vq_cells: null argument");
    if (m->nb_wannier <= 0 || !(m->temperature > 0) || !(m->nk_full > 0))
        return fail(DMD_ERR_ARG, "dmd_scatt_ee_vq_cells: bad model (nb_wannier, temperature, nk_full)");
    const int64_t nw = m->nb_wannier, nk = h->nk;
    double2 *dU = nullptr;
    int rc;
    if ((rc = dalloc_t(&dU, (size_t)(nk * nw * h->nb), "U"))) return rc;
    hipError_t e = hipMemcpy(dU, U, (size_t)(nk * nw * h->nb) * 16, hipMemcpyDefault);
    if (e != hipSuccess) {
        dfree(dU);
        return fail(DMD_ERR_HIP, hipGetErrorString(e));
    }
    std::vector<double> cells;
    rc = rpa_vq_cells(h, dU, kvec, f, *m, cells);
    dfree(dU);
    if (rc) return rc;
    std::memcpy(vq_cell, cells.data(), cells.size() * 8);
    return DMD_OK;
}

extern "C" int dmd_scatt_add_ee_model_state(dmd_scatt *h, const uint64_t *k1st, const uint64_t *k2nd, const double *U,
                                            const double *kvec, const double *f, const double *dm, const double *dm1,
                                            const double *vq_cell, const dmd_eimp_model *m, double scale, int add)
{
    int rc;
    if ((rc = ee_check(h, k1st, k2nd, U, kvec, f, m, "dmd_scatt_add_ee_model_state"))) return rc;
    if ((dm == nullptr) != (dm1 == nullptr)) return fail(DMD_ERR_ARG, "dmd_scatt_add_ee_model_state: dm and dm1 go together");
    if (m->screening == DMD_SCR_RPA && !vq_cell)
        return fail(DMD_ERR_ARG, "dmd_scatt_add_ee_model_state: the static RPA needs the per-cell v(q) table "
                                 "(dmd_scatt_ee_vq_cells)");
    if (h->linear) return fail(DMD_ERR_STATE, "the handle holds the linearized operator: add the e-e P before set_Lsc");
    if (add && !h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated (add = 0 sets it)");
    if (h->np == 0) { h->has_P = true; return DMD_OK; }
    EeSetup s;
    EeArgs a;
    if ((rc = ee_setup(h, U, kvec, f, dm, dm1, m->screening == DMD_SCR_RPA ? vq_cell : nullptr, *m, s, a))) return rc;
    return ee_fold(h, k1st, k2nd, a, ee_const(*m), nullptr, scale, add);
}

// alg_linearize_dPee: the inputs of the dP build (add_scatt_contrib("ee") with
// dP1ee, dP2ee, ElectronPhonon.cpp:84; elecelec_model::calc_P with dP,
// ElecElec_Model.h:185-262,271-281), checked and uploaded.
static int ee_dP_setup(dmd_scatt *h, const uint64_t *k1st, const uint64_t *k2nd, const double *U, const double *kvec,
                       const double *f_eq, const double *dm, const double *dm1, const double *vq_cell,
                       const dmd_eimp_model *m, const char *who, EeSetup &s, EeArgs &a)
{
    int rc;
    if ((rc = ee_check(h, k1st, k2nd, U, kvec, f_eq, m, who))) return rc;
    if (!dm || !dm1) return fail(DMD_ERR_ARG, std::string(who) + ": dP needs the state (dm and dm1)");
    if (m->screening == DMD_SCR_RPA && !vq_cell)
        return fail(DMD_ERR_ARG, std::string(who) + ": the static RPA needs the per-cell v(q) table (dmd_scatt_ee_vq_cells)");
    if (h->linear) return fail(DMD_ERR_STATE, std::string(who) + ": alg_linearize_dPee excludes the linearized operator");
    if ((rc = ee_setup(h, U, kvec, f_eq, dm, dm1, m->screening == DMD_SCR_RPA ? vq_cell : nullptr, *m, s, a))) return rc;
    a.dP = 1;
    return DMD_OK;
}

extern "C" int dmd_scatt_add_ee_dP(dmd_scatt *h, const uint64_t *k1st, const uint64_t *k2nd, const double *U,
                                   const double *kvec, const double *f_eq, const double *dm, const double *dm1,
                                   const double *vq_cell, const dmd_eimp_model *m, double scale, int add)
{
    if (h && add && !h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated (add = 0 sets it)");
    if (h && h->np == 0 && h->nb >= 1) {
        int rc = ee_check(h, k1st, k2nd, U, kvec, f_eq, m, "dmd_scatt_add_ee_dP");
        if (rc == DMD_OK) h->has_P = true;
        return rc;
    }
    EeSetup s;
    EeArgs a;
    int rc;
    if ((rc = ee_dP_setup(h, k1st, k2nd, U, kvec, f_eq, dm, dm1, vq_cell, m, "dmd_scatt_add_ee_dP", s, a))) return rc;
    return ee_fold(h, k1st, k2nd, a, ee_const(*m), nullptr, scale, add);
}

extern "C" int dmd_scatt_set_ee_dP(dmd_scatt *h, const uint64_t *k1st, const uint64_t *k2nd, const double *U,
                                   const double *kvec, const double *f_eq, const double *dm, const double *dm1,
                                   const double *vq_cell, const dmd_eimp_model *m)
{
    if (!h) return fail(DMD_ERR_ARG, "null handle");
    if (!dm && !dm1) {
        h->has_dP = false;
        return DMD_OK;
    }
    EeSetup s;
    EeArgs a;
    int rc;
    if ((rc = ee_dP_setup(h, k1st, k2nd, U, kvec, f_eq, dm, dm1, vq_cell, m, "dmd_scatt_set_ee_dP", s, a))) return rc;
    const size_t nrow = (size_t)h->nk * h->nb2;
    if (!h->d_dPH && (rc = dalloc_t(&h->d_dPH, nrow, "linearize_dPee rows"))) return rc;
    double *acc = nullptr;
    if ((rc = dalloc_t(&acc, 4 * nrow, "linearize_dPee row sums"))) return rc;
    hipError_t e = hipMemsetAsync(acc, 0, 4 * nrow * 8, h->st);
    if (e != hipSuccess) {
        dfree(acc);
        return fail(DMD_ERR_HIP, hipGetErrorString(e));
    }
    if (h->np > 0) rc = ee_fold(h, k1st, k2nd, a, ee_const(*m), nullptr, 1.0, 1, acc);
    if (rc == DMD_OK) {
        const dim3 g(nblk(h->nk, 256)), b(256);
        switch (h->nb) {
        case 1: hipLaunchKernelGGL(k_dP_herm<1>, g, b, 0, h->st, h->nk, acc, a.f, h->prefac, h->d_dPH); break;
        case 2: hipLaunchKernelGGL(k_dP_herm<2>, g, b, 0, h->st, h->nk, acc, a.f, h->prefac, h->d_dPH); break;
        case 3: hipLaunchKernelGGL(k_dP_herm<3>, g, b, 0, h->st, h->nk, acc, a.f, h->prefac, h->d_dPH); break;
        default: hipLaunchKernelGGL(k_dP_herm<4>, g, b, 0, h->st, h->nk, acc, a.f, h->prefac, h->d_dPH); break;
        }
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(h->st);
        if (e != hipSuccess) rc = fail(DMD_ERR_HIP, hipGetErrorString(e));
    }
    dfree(acc);
    h->has_dP = rc == DMD_OK;
    return rc;
}

extern "C" int dmd_scatt_get_ee_dP(dmd_scatt *h, double *H)
{
    if (!h || !H) return fail(DMD_ERR_ARG, "dmd_scatt_get_ee_dP: null argument");
    if (!h->has_dP) return fail(DMD_ERR_STATE, "dmd_scatt_get_ee_dP: no linearize_dPee term is set");
    DMD_HIP(hipMemcpy(H, h->d_dPH, (size_t)h->nk * h->nb2 * 16, hipMemcpyDefault));
    return DMD_OK;
}

extern "C" int dmd_scatt_add_Lsc_ee(dmd_scatt *h, const uint64_t *k1st, const uint64_t *k2nd, const double *U,
                                    const double *kvec, const double *f_eq, const double *vq_cell,
                                    const dmd_eimp_model *m, double scale, double extra, double coef34, int dP_terms)
{
    int rc;
    if ((rc = ee_check(h, k1st, k2nd, U, kvec, f_eq, m, "dmd_scatt_add_Lsc_ee"))) return rc;
    if (!h->linear) return fail(DMD_ERR_STATE, "dmd_scatt_add_Lsc_ee: set_Lsc first (the handle holds P, not Lsc)");
    if (m->screening == DMD_SCR_RPA && !vq_cell)
        return fail(DMD_ERR_ARG, "dmd_scatt_add_Lsc_ee: the static RPA needs the per-cell v(q) table (dmd_scatt_ee_vq_cells)");
    if (h->np == 0 || (extra == 0.0 && !dP_terms)) return DMD_OK;
    EeSetup s;
    EeArgs a;
    if ((rc = ee_setup(h, U, kvec, f_eq, nullptr, nullptr, m->screening == DMD_SCR_RPA ? vq_cell : nullptr, *m, s, a)))
        return rc;
    const EimpConst c = ee_const(*m);
    const int64_t per = h->nb4;
    const int64_t chunk = std::min<int64_t>(h->np, std::max<int64_t>(1, (32ll << 20) / per));
    double2 *buf = nullptr;
    uint64_t *d1 = nullptr, *d2 = nullptr;
    const int nbuf = 8;  // Pee1, Pee2, P3..P6, dLij, dLji
    if ((rc = dalloc_t(&buf, (size_t)(nbuf * chunk * per), "e-e Lsc staging")) || (rc = dalloc_t(&d1, (size_t)chunk, "k1")) ||
        (rc = dalloc_t(&d2, (size_t)chunk, "k2"))) {
        dfree(buf); dfree(d1); dfree(d2);
        return rc;
    }
    double2 *Pe1 = buf, *Pe2 = buf + chunk * per, *P3 = buf + 2 * chunk * per, *P4 = buf + 3 * chunk * per,
            *P5 = buf + 4 * chunk * per, *P6 = buf + 5 * chunk * per, *L1 = buf + 6 * chunk * per,
            *L2 = buf + 7 * chunk * per;
    hipError_t e = hipSuccess;
    for (int64_t first = 0; first < h->np && e == hipSuccess && rc == DMD_OK; first += chunk) {
        const int64_t n = std::min(chunk, h->np - first);
        e = hipMemcpyAsync(d1, k1st + first, (size_t)n * 8, hipMemcpyDefault, h->st);
        if (e == hipSuccess) e = hipMemcpyAsync(d2, k2nd + first, (size_t)n * 8, hipMemcpyDefault, h->st);
        if (e != hipSuccess) break;
        const dim3 g(nblk(n * per, 256)), b(256);
#define DMD_LSC_EE(NB_)                                                                                              \
    do {                                                                                                             \
        if (extra != 0.0) launch_ee_P<NB_>(h, n, d1, d2, a, c, Pe1, Pe2);                                             \
        if (dP_terms) hipLaunchKernelGGL(k_ee_P3456<NB_>, g, b, 0, h->st, n, d1, d2, a, c, P3, P4, P5, P6);         \
        hipLaunchKernelGGL(k_lsc_ee<NB_>, g, b, 0, h->st, n, d1, d2, a.f, extra, coef34, extra != 0.0 ? Pe1 : nullptr, \
                           Pe2, (dP_terms && !m->exchange) ? P3 : nullptr, P4, dP_terms ? P5 : nullptr, P6, L1, L2);    \
    } while (0)
        switch (h->nb) {
        case 1: DMD_LSC_EE(1); break;
        case 2: DMD_LSC_EE(2); break;
        case 3: DMD_LSC_EE(3); break;
        default: DMD_LSC_EE(4); break;
        }
#undef DMD_LSC_EE
        if ((e = hipGetLastError()) != hipSuccess) break;
        rc = place_P_dev(h, L1, L2, n, first, scale, 1);
        if (rc == DMD_OK) e = hipStreamSynchronize(h->st);
    }
    hipError_t e2 = hipStreamSynchronize(h->st);
    dfree(buf); dfree(d1); dfree(d2);
    if (rc == DMD_OK && e == hipSuccess) e = e2;
    if (rc == DMD_OK && e != hipSuccess) rc = fail(DMD_ERR_HIP, hipGetErrorString(e));
    return rc;
}
