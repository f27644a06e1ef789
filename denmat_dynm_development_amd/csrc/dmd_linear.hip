// dmd_linear.hip -- the linearized scattering operator (alg_linearize):
// electronphonon::set_Lsc (src_v4.5.7/ElectronPhonon_linear.cpp:3-36,147-207)
// builds, per stored pair, the matrices Lscij and Lscji from P1, P2 and the
// occupations f_eq, and evolve_linear (src_v4.5.7/evolve_eph_linear.cpp:3-53)
// applies them: ddmdt_k = prefac/2 sum (X + X^+), X = Lsc (rho_k' - f_eq_k').
//
// The operator replaces P in the handle's own store, in place (the slot of a
// pair keeps Lscij where P1 was and Lscji where P2 was), so the memory is the
// reference's, which never holds P and L together.  Lscii, the part acting on
// the row's own drho, is gathered per row as two nb x nb matrices,
//   SA_b2 = sum_pairs sum_3 A[(b,2),(3,3)] f'_3,  SB_1a = sum_pairs sum_3 (1 - f'_3) B[(3,3),(1,a)]
// (A, B = P1, P2 forward; conj(P2), conj(P1) reverse; f' the partner's),
// Lscii_{12,ab} = -delta_1a SA_b2 - delta_2b SB_1a, and folded into the
// diagonal pairs (k, k) as the reference does: a row without its diagonal
// pair drops it.  Evaluation is a deterministic row kernel; the interaction-
// picture phase of compute_Pt (P1 type) factorises like P1's, so drho enters
// in the Schroedinger picture and the row result is rotated back.
#include <hip/hip_runtime.h>

#include <vector>

#include "dmd_internal.h"

namespace dmd {

// Lscii per row: lane t < NB2 sums SA entry t = (b, 2), lane NB2 + t sums SB
// entry t = (1, a).  One wavefront per row.
template <int NB>
__global__ __launch_bounds__(256) void k_lin_sums(int64_t nk, const double2 *__restrict__ P,
                                                  const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                                  const int64_t *__restrict__ fwdpos,
                                                  const int64_t *__restrict__ revptr,
                                                  const int64_t *__restrict__ revpos,
                                                  const int32_t *__restrict__ revcol, const double *__restrict__ f,
                                                  int soa, double2 *__restrict__ SA, double2 *__restrict__ SB)
{
    constexpr int NB2 = NB * NB, NB4 = NB2 * NB2, EP = 2 * NB4;
    const int lane = threadIdx.x & 63;
    const int64_t k = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (k >= nk || lane >= 2 * NB2) return;
    const bool isA = lane < NB2;
    const int t = isA ? lane : lane - NB2, x = t / NB, y = t % NB;
    double2 acc = make_double2(0, 0);
    // forward pairs (k, k'): A = P1, B = P2, f' = f_k'
    for (int64_t it = rowptr[k]; it < rowptr[k + 1]; it++) {
        const int64_t s = fwdpos[it], kp = col[it];
        for (int i3 = 0; i3 < NB; i3++) {
            const double fp = f[kp * NB + i3];
            if (isA) {  // A[(b=x, 2=y), (3,3)] f'_3
                const double2 v = P[paddr(s, (x * NB + y) * NB2 + i3 * NB + i3, EP, soa)];
                acc = make_double2(acc.x + v.x * fp, acc.y + v.y * fp);
            } else {    // (1 - f'_3) B[(3,3), (1=x, a=y)]
                const double2 v = P[paddr(s, NB4 + (i3 * NB + i3) * NB2 + x * NB + y, EP, soa)];
                acc = make_double2(acc.x + v.x * (1 - fp), acc.y + v.y * (1 - fp));
            }
        }
    }
    // reverse pairs (k'', k), k'' < k: A = conj(P2), B = conj(P1), f' = f_k''
    for (int64_t it = revptr[k]; it < revptr[k + 1]; it++) {
        const int64_t s = revpos[it], kpp = revcol[it];
        for (int i3 = 0; i3 < NB; i3++) {
            const double fp = f[kpp * NB + i3];
            if (isA) {
                const double2 v = P[paddr(s, NB4 + (x * NB + y) * NB2 + i3 * NB + i3, EP, soa)];
                acc = make_double2(acc.x + v.x * fp, acc.y - v.y * fp);
            } else {
                const double2 v = P[paddr(s, (i3 * NB + i3) * NB2 + x * NB + y, EP, soa)];
                acc = make_double2(acc.x + v.x * (1 - fp), acc.y - v.y * (1 - fp));
            }
        }
    }
    (isA ? SA : SB)[k * NB2 + t] = acc;
}

// P -> (Lscij, Lscji) of one slot in place (add_Lscij_from_P1_P2,
// ElectronPhonon_linear.cpp:183-207), plus Lscii for a diagonal pair.  One
// 64-thread workgroup per slot; the record is staged in LDS first.
template <int NB>
__global__ __launch_bounds__(64) void k_lin_build(int64_t nslots, double2 *__restrict__ P,
                                                  const int32_t *__restrict__ k1s, const int32_t *__restrict__ k2s,
                                                  const double *__restrict__ f, const double2 *__restrict__ SA,
                                                  const double2 *__restrict__ SB, int soa)
{
    constexpr int NB2 = NB * NB, NB4 = NB2 * NB2, EP = 2 * NB4;
    __shared__ double2 rec[EP];
    const int64_t s = blockIdx.x;
    if (s >= nslots) return;
    const int64_t ik = k1s[s], jk = k2s[s];
    if (ik < 0) return;  // padding slot: P and L are zero
    for (int el = threadIdx.x; el < EP; el += 64) rec[el] = P[paddr(s, el, EP, soa)];
    __syncthreads();
    for (int el = threadIdx.x; el < EP; el += 64) {
        const bool ji = el >= NB4;
        const int r = ji ? el - NB4 : el, i12 = r / NB2, iab = r % NB2;
        const int i1 = i12 / NB, i2 = i12 % NB, ia = iab / NB, ib = iab % NB;
        const int64_t kk = ji ? jk : ik;
        const double f1 = f[kk * NB + i1], f2 = f[kk * NB + i2];
        double2 v;
        if (!ji) {  // (1 - f^ik_1) P1_{12,ab} + P2_{ab,12} f^ik_2
            const double2 p1 = rec[r], p2 = rec[NB4 + iab * NB2 + i12];
            v = make_double2((1 - f1) * p1.x + p2.x * f2, (1 - f1) * p1.y + p2.y * f2);
        } else {    // (1 - f^jk_1) conj(P2_{12,ab}) + conj(P1_{ab,12}) f^jk_2
            const double2 p2 = rec[NB4 + r], p1 = rec[iab * NB2 + i12];
            v = make_double2((1 - f1) * p2.x + p1.x * f2, -((1 - f1) * p2.y) - p1.y * f2);
        }
        if (ik == jk) {  // + Lscii_{12,ab} = -delta_1a SA_b2 - delta_2b SB_1a
            double2 l = make_double2(0, 0);
            if (i1 == ia) l = csub(l, SA[ik * NB2 + ib * NB + i2]);
            if (i2 == ib) l = csub(l, SB[ik * NB2 + i1 * NB + ia]);
            v = cadd(v, l);
        }
        P[paddr(s, el, EP, soa)] = v;
    }
}

// evolve_linear per destination row (deterministic): X_k = sum over forward
// pairs Lscij drho_k' + over reverse pairs (k'', k) Lscji drho_k'', then
// ddmdt_k = prefac/2 (X + X^+) rotated back.  One wavefront per row; a lane
// owns element e of the pair's L (J per lane for NB^4 > 64, else 64/NB^4
// pairs side by side) and the element products are reduced through LDS.
template <int NB>
__global__ __launch_bounds__(256) void k_lin_rows(int64_t nk, const double2 *__restrict__ P,
                                                  const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                                  const int64_t *__restrict__ fwdpos,
                                                  const int64_t *__restrict__ revptr,
                                                  const int64_t *__restrict__ revpos,
                                                  const int32_t *__restrict__ revcol,
                                                  const double2 *__restrict__ rhoS, const double *__restrict__ f,
                                                  const double *__restrict__ e, const DClock *__restrict__ clk,
                                                  double a, int expt, double prefac, int soa,
                                                  double2 *__restrict__ out)
{
    constexpr int NB2 = NB * NB, NB4 = NB2 * NB2, EP = 2 * NB4;
    constexpr int PPI = NB4 <= 64 ? 64 / NB4 : 1, J = NB4 <= 64 ? 1 : (NB4 + 63) / 64;
    __shared__ double2 red[4][64 * J];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t k = (int64_t)blockIdx.x * 4 + wv;
    const bool active = k < nk;
    const int slot = NB4 <= 64 ? lane / NB4 : 0;
    int el[J];
    bool valid[J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        el[j] = NB4 <= 64 ? lane % NB4 : lane + 64 * j;
        valid[j] = (NB4 <= 64 ? slot < PPI : true) && el[j] < NB4;
        if (!valid[j]) el[j] = 0;
    }
    double2 acc[J];
#pragma unroll
    for (int j = 0; j < J; j++) acc[j] = make_double2(0, 0);
    (void)f;
    auto drho = [&](int64_t q, int ab) { return rhoS[q * NB2 + ab]; };  // drho_S (rotation subtracted f_eq)
    if (active) {
        for (int pass = 0; pass < 2; pass++) {
            const int64_t p0 = pass ? revptr[k] : rowptr[k], p1 = pass ? revptr[k + 1] : rowptr[k + 1];
            for (int64_t q = p0; q < p1; q += PPI) {
                const int64_t it = q + slot;
                const bool ok = it < p1;
                const int64_t s = ok ? (pass ? revpos[it] : fwdpos[it]) : 0;
                const int64_t kq = ok ? (int64_t)(pass ? revcol[it] : col[it]) : 0;
#pragma unroll
                for (int j = 0; j < J; j++) {
                    if (!(ok && valid[j])) continue;
                    const double2 l = P[paddr(s, (pass ? NB4 : 0) + el[j], EP, soa)];
                    acc[j] = cadd(acc[j], cmul(l, drho(kq, el[j] % NB2)));
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < J; j++) red[wv][lane + 64 * j] = valid[j] ? acc[j] : make_double2(0, 0);
    __syncthreads();
    if (!active || lane >= NB2) return;
    // X[i12] = sum over the lanes (and side-by-side pairs) holding row i12
    // every lane < NB2 forms the whole X (NB2 sums of NB2 terms) to build X + X^+
    double2 x12[NB2];
#pragma unroll
    for (int r = 0; r < NB2; r++) {
        double2 sum = make_double2(0, 0);
        for (int p = 0; p < PPI; p++)
            for (int c = 0; c < NB2; c++) {
                const int idx = p * NB4 + r * NB2 + c;  // NB4 <= 64: lane p*NB4 + el; else lane + 64 j = el
                sum = cadd(sum, red[wv][idx]);
            }
        x12[r] = sum;
    }
    const int i = lane / NB, jj = lane % NB;
    double2 d = cadd(x12[i * NB + jj], cconj(x12[jj * NB + i]));
    d = make_double2(d.x * (prefac * 0.5), d.y * (prefac * 0.5));
    if (expt && i != jj) d = cmul(d, ccis((e[k * NB + i] - e[k * NB + jj]) * clock_t_at(clk, a)));
    out[k * NB2 + lane] = d;
}

template <int NB>
static void lin_build_nb(dmd_scatt *h, double2 *SA, double2 *SB)
{
    hipLaunchKernelGGL(k_lin_sums<NB>, dim3(nblk(h->nk, 4)), dim3(256), 0, h->st, h->nk, h->d_P, h->d_rowptr, h->d_col,
                       h->d_fwdpos, h->d_revptr, h->d_revpos, h->d_revcol, h->d_flin, h->soa, SA, SB);
    if (h->nslots > 0)
        hipLaunchKernelGGL(k_lin_build<NB>, dim3((unsigned)h->nslots), dim3(64), 0, h->st, h->nslots, h->d_P, h->d_k1,
                           h->d_k2, h->d_flin, SA, SB, h->soa);
}

template <int NB>
static void lin_rows_nb(dmd_scatt *h, double a, double2 *out)
{
    hipLaunchKernelGGL(k_lin_rows<NB>, dim3(nblk(h->nk, 4)), dim3(256), 0, h->st, h->nk, h->d_P, h->d_rowptr, h->d_col,
                       h->d_fwdpos, h->d_revptr, h->d_revpos, h->d_revcol, h->d_rhoS, h->d_flin, h->d_e, h->d_clk, a,
                       h->expt, h->prefac, h->soa, out);
}

int launch_linear(dmd_scatt *h, double a, double2 *out)
{
    switch (h->nb) {
    case 1: lin_rows_nb<1>(h, a, out); break;
    case 2: lin_rows_nb<2>(h, a, out); break;
    case 3: lin_rows_nb<3>(h, a, out); break;
    case 4: lin_rows_nb<4>(h, a, out); break;
    default: return fail(DMD_ERR_UNSUPPORTED, "linearized operator: nb > 4");
    }
    DMD_HIP(hipGetLastError());
    return DMD_OK;
}

}  // namespace dmd

using namespace dmd;

// electronphonon::set_Lsc (ElectronPhonon_linear.cpp:3-36) for the e-ph P in
// the handle (with any e-i P already folded in, as add_Lsc_contrib("eimp")
// adds its own P's Lsc, :38-104): P becomes (Lscij, Lscji) at f_eq (nk x nb,
// the e-ph window), and eval / evolve_driver / integrate_rk4 evaluate
// evolve_linear (evolve_eph_linear.cpp:3-45) from then on.  FP64 only.
extern "C" int dmd_scatt_set_Lsc(dmd_scatt *h, const double *f_eq)
{
    {
        const int sg = sp_guard(h, "dmd_scatt_set_Lsc");
        if (sg) return sg;
    }
    if (!h || !f_eq) return fail(DMD_ERR_ARG, "dmd_scatt_set_Lsc: bad arguments");
    if (h->precision != DMD_PREC_FP64) return fail(DMD_ERR_UNSUPPORTED, "the linearized operator is FP64 only");
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    if (h->linear) return fail(DMD_ERR_STATE, "the handle already holds a linearized operator; upload P again");
    int rc;
    const size_t nf = (size_t)h->nk * h->nb;
    if (!h->d_flin && (rc = dalloc_t(&h->d_flin, nf, "f_eq of the linearized operator"))) return rc;
    DMD_HIP(hipMemcpyAsync(h->d_flin, f_eq, nf * 8, is_device_ptr(f_eq) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                           h->st));
    double2 *SA = nullptr, *SB = nullptr;
    if ((rc = dalloc_t(&SA, (size_t)h->nk * h->nb2, "Lscii sums"))) return rc;
    if ((rc = dalloc_t(&SB, (size_t)h->nk * h->nb2, "Lscii sums"))) { dfree(SA); return rc; }
    switch (h->nb) {
    case 1: lin_build_nb<1>(h, SA, SB); break;
    case 2: lin_build_nb<2>(h, SA, SB); break;
    case 3: lin_build_nb<3>(h, SA, SB); break;
    case 4: lin_build_nb<4>(h, SA, SB); break;
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    dfree(SA);
    dfree(SB);
    if (e != hipSuccess) return fail(DMD_ERR_HIP, std::string("set_Lsc: ") + hipGetErrorString(e));
    h->linear = true;
    h->has_eq = false;  // evolve_linear subtracts no ddmdt_eq
    h->has_dP = false;  // nor adds a linearize_dPee term (the two modes exclude each other)
    return DMD_OK;
}
