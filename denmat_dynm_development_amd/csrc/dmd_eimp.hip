// dmd_eimp.hip -- model ionized-impurity scattering (impMode = model_ionized)
// generated on the device: elecimp_model::calc_P (src_v4.5.7/scattering_model/
// ElecImp_Model.h:185-213) for every stored pair, with the matrix element of
// calc_eimp / calc_A (:67-137) and the static screened Coulomb potential of
// coulomb_model::vq (Coulomb_Model.h:632-662), then folded into P as
// add_scatt_contrib("eimp") does (ElectronPhonon.cpp:44-113).
//
// Per stored pair (k, k'), bands b1 of k and b2 of k':
//   <k|k'>_{b1 b2} = sum_w conj(U_k[w, b1]) U_k'[w, b2]          (calc_ovlp)
//   M_{b1 b2}      = prefac_A v(q) <k|k'>_{b1 b2},  q = wrap(k - k')
//   A2 = M sqrt-or-full Gaussian(e_k,b1 - e_k',b2) inside |de| < ethr,
//   A1 = A2 (lindblad) or M (conventional)
//   P1[(i1 i2),(i3 i4)] = A1[i1,i3] conj(A2[i2,i4])
//   P2[(i1 i2),(i3 i4)] = A1[i3,i1] conj(A2[i4,i2])
// with the emission/absorption pair of A at de -+ omegaL and the detailed-
// balance factors under detailBalance.  The nb^2 A entries of a pair are
// formed once (one thread each, into LDS) and shared by its nb^4 elements,
// one thread per element, so the writes coalesce.  The chunk goes into the
// handle's P layout through place_P_dev (all precisions).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <numeric>
#include <vector>

#include "dmd_internal.h"
#include "dmd_model.h"

namespace dmd {

// The A entries of one (b1, b2): without detailed balance a[0] = A1, a[1] = A2;
// with it a[0..7] = A1pp, A1pm, A2pp, A2pm, A1mp, A1mm, A2mp, A2mm.
__device__ void eimp_A(const EimpConst &c, double pv, const double2 *__restrict__ Uk, const double2 *__restrict__ Ukp,
                       int nb, int b1, int b2, double de, double2 *a)
{
    double2 o = make_double2(0.0, 0.0);
    for (int w = 0; w < c.m.nb_wannier; w++) o = cadd(o, cmulc(Ukp[w * nb + b2], Uk[w * nb + b1]));
    const double2 M = make_double2(pv * o.x, pv * o.y);
    const double ethr = c.m.ethr;
    if (!c.m.detail_balance) {
        double2 A2 = make_double2(0.0, 0.0);
        if (fabs(de) < ethr) {
            const double dl = c.prefac_delta * exp(c.prefac_exp * (de * de));
            A2 = make_double2(M.x * dl, M.y * dl);
        }
        a[0] = c.m.lindblad ? A2 : M;
        a[1] = A2;
        return;
    }
    const double wL = c.m.omegaL;
    double2 G2p = make_double2(0.0, 0.0), G2m = make_double2(0.0, 0.0);
    if (fabs(de + wL) < ethr) {
        const double dl = c.prefac_delta * exp(c.prefac_exp * ((de + wL) * (de + wL)));
        G2p = make_double2(M.x * dl, M.y * dl);
    }
    if (fabs(de - wL) < ethr) {
        const double dl = c.prefac_delta * exp(c.prefac_exp * ((de - wL) * (de - wL)));
        G2m = make_double2(M.x * dl, M.y * dl);
    }
    const double2 G1p = c.m.lindblad ? G2p : M, G1m = c.m.lindblad ? G2m : M;
    const double dEbyT = de / c.m.temperature;
    const double fp = (-dEbyT < 46) ? exp(-dEbyT / 2) : 1.0, fm = (dEbyT < 46) ? exp(dEbyT / 2) : 1.0;
    a[0] = make_double2(G1p.x * fp, G1p.y * fp);
    a[1] = G1p;
    a[2] = make_double2(G2p.x * fp, G2p.y * fp);
    a[3] = G2p;
    a[4] = make_double2(G1m.x * fm, G1m.y * fm);
    a[5] = G1m;
    a[6] = make_double2(G2m.x * fm, G2m.y * fm);
    a[7] = G2m;
}

// The static RPA / lindhard screening sum of one q of the mesh
// (calc_qscr2_static_RPA, Coulomb_Model.h:721-744): over every k whose k - q
// is a kept k point (kgrid: mesh cell -> smallest k index, INT_MAX if none),
// sum_b1b2 dfde |<k b1|k-q b2>|^2 (RPA) or sum_b dfde (lindhard), dfde =
// (f_k-q,b2 - f_k,b1) / (e_k,b1 - e_k-q,b2), f(1-f)/T when |de| < 1e-8.
// Grid: x over k, y over q (strided).  All terms are >= 0 (f falls with e),
// so the atomic sum is well conditioned.
__global__ __launch_bounds__(256) void k_rpa_qscr2(int64_t nq, const int *__restrict__ q3, int64_t nk,
                                                   const int *__restrict__ ik3, const int *__restrict__ kgrid, int m0,
                                                   int m1, int m2, const double2 *__restrict__ U, int nw, int nb,
                                                   const double *__restrict__ e, const double *__restrict__ f,
                                                   double T, int lindhard, double *__restrict__ qsum)
{
    __shared__ double red[4];
    const int64_t ik = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t iq = blockIdx.y; iq < nq; iq += gridDim.y) {
        double acc = 0.0;
        if (ik < nk) {
            const int x = ((ik3[3 * ik] - q3[3 * iq]) % m0 + m0) % m0, y = ((ik3[3 * ik + 1] - q3[3 * iq + 1]) % m1 + m1) % m1,
                      z = ((ik3[3 * ik + 2] - q3[3 * iq + 2]) % m2 + m2) % m2;
            const int jk = kgrid[((int64_t)x * m1 + y) * m2 + z];
            if (jk != INT_MAX) {
                const double2 *Uk = U + ik * nw * nb, *Uj = U + (int64_t)jk * nw * nb;
                for (int b1 = 0; b1 < nb; b1++)
                    for (int b2 = 0; b2 < nb; b2++) {
                        if (lindhard && b1 != b2) continue;
                        const double de = e[ik * nb + b1] - e[(int64_t)jk * nb + b2];
                        double dfde;
                        if (fabs(de) < 1e-8) {
                            const double fa = 0.5 * (f[ik * nb + b1] + f[(int64_t)jk * nb + b2]);
                            dfde = (1 - fa) * fa / T;
                        } else {
                            dfde = (f[(int64_t)jk * nb + b2] - f[ik * nb + b1]) / de;
                        }
                        if (lindhard) {
                            acc += dfde;
                        } else {
                            double2 o = make_double2(0.0, 0.0);
                            for (int w = 0; w < nw; w++) o = cadd(o, cmulc(Uj[w * nb + b2], Uk[w * nb + b1]));
                            acc += dfde * (o.x * o.x + o.y * o.y);
                        }
                    }
            }
        }
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            const double t = red[0] + red[1] + red[2] + red[3];
            if (t != 0.0) atomicAdd(qsum + iq, t);
        }
        __syncthreads();
    }
}
// P1, P2 of n stored pairs (k1[f], k2[f]) in the ldbd_P element order; vq,
// when given, holds v(q) of every pair of the chunk (the k-summed screenings).
// A workgroup takes PP = 256 / NB^4 pairs (3 for NB = 3): first one thread per
// A entry (b1, b2) of each pair forms that entry (overlap, v(q), delta) into
// LDS, then one thread per P element multiplies two entries.
template <int NB>
__global__ __launch_bounds__(256) void k_eimp_model(int64_t n, const uint64_t *__restrict__ k1,
                                                    const uint64_t *__restrict__ k2, const double2 *__restrict__ U,
                                                    const double *__restrict__ kvec, const double *__restrict__ e,
                                                    const double *__restrict__ vq, EimpConst c,
                                                    double2 *__restrict__ P1, double2 *__restrict__ P2)
{
    constexpr int NB2 = NB * NB, NB4 = NB2 * NB2, PP = 256 / NB4 > 0 ? 256 / NB4 : 1;
    constexpr int NA = 8;  // A slots per entry (2 used without detailed balance)
    __shared__ double2 A[PP][NB2][NA];
    const int t = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * PP;
    if (t < PP * NB2) {
        const int lp = t / NB2, q = t % NB2;
        const int64_t f = p0 + lp;
        if (f < n) {
            const int64_t k = (int64_t)k1[f], kp = (int64_t)k2[f];
            const double pv = c.m.prefac_A * (vq ? vq[f] : eimp_vq(c.m, kvec + 3 * k, kvec + 3 * kp));
            const int nw = c.m.nb_wannier, b1 = q / NB, b2 = q % NB;
            eimp_A(c, pv, U + k * nw * NB, U + kp * nw * NB, NB, b1, b2, e[k * NB + b1] - e[kp * NB + b2], A[lp][q]);
        }
    }
    __syncthreads();
    if (t >= PP * NB4) return;
    const int lp = t / NB4, el = t % NB4;
    const int64_t f = p0 + lp;
    if (f >= n) return;
    const int i1 = el / (NB * NB2), i2 = (el / NB2) % NB, i3 = (el / NB) % NB, i4 = el % NB;
    const double2 *a13 = A[lp][i1 * NB + i3], *a24 = A[lp][i2 * NB + i4], *a31 = A[lp][i3 * NB + i1],
                  *a42 = A[lp][i4 * NB + i2];
    double2 p1, p2;
    if (!c.m.detail_balance) {
        p1 = cmulc(a13[0], a24[1]);
        p2 = cmulc(a31[0], a42[1]);
    } else {
        p1 = cadd(cmulc(a13[0], a24[2]), cmulc(a13[5], a24[7]));  // A1pp conj(A2pp) + A1mm conj(A2mm)
        p2 = cadd(cmulc(a31[4], a42[6]), cmulc(a31[1], a42[3]));  // A1mp conj(A2mp) + A1pm conj(A2pm)
    }
    P1[f * NB4 + el] = p1;
    P2[f * NB4 + el] = p2;
}

// v(q) of every stored pair for the k-summed screenings: coulomb_model::vq
// (Coulomb_Model.h:632-683) with vq_RPA = prefac_vq / (|q|^2 + q_scr^2(q))
// (:820-828); q_scr^2 is summed on the device for each distinct q of the
// stored pairs only.  Host vectors in, host vector out.
int rpa_vq(dmd_scatt *h, const uint64_t *k1st, const uint64_t *k2nd, const double2 *dU, const double *kvec,
                  const double *f, const dmd_eimp_model &m, std::vector<double> &vq)
{
    const int *km = m.kmesh;
    if (km[0] <= 0 || km[1] <= 0 || km[2] <= 0) return fail(DMD_ERR_ARG, "RPA/lindhard screening needs the k mesh");
    if (!f) return fail(DMD_ERR_ARG, "RPA/lindhard screening needs the occupations f");
    const int64_t cells = (int64_t)km[0] * km[1] * km[2];
    if (cells > (1ll << 31)) return fail(DMD_ERR_UNSUPPORTED, "k mesh too large for the mesh map");
    const int64_t nk = h->nk, np = h->np;
    std::vector<int> ik3((size_t)nk * 3);
    for (int64_t k = 0; k < nk; k++) mesh_cell(kvec + 3 * k, km, &ik3[3 * k]);
    // distinct q cells of the stored pairs, q = k - k' (qIndexMap::q2iq)
    std::vector<int64_t> code((size_t)np);
    for (int64_t p = 0; p < np; p++) {
        double q[3];
        int v[3];
        for (int d = 0; d < 3; d++) q[d] = kvec[3 * k1st[p] + d] - kvec[3 * k2nd[p] + d];
        mesh_cell(q, km, v);
        code[p] = ((int64_t)v[0] * km[1] + v[1]) * km[2] + v[2];
    }
    // distinct cells numbered in order of first appearance: a dense cell table
    // (one int per mesh cell) for meshes up to 2^28 cells, else a sort
    std::vector<int32_t> uq((size_t)np);
    std::vector<int> q3;
    std::vector<int64_t> rep;  // first pair of each distinct q
    auto add_q = [&](int64_t p) {
        const int64_t c = code[p];
        q3.push_back((int)(c / ((int64_t)km[1] * km[2])));
        q3.push_back((int)((c / km[2]) % km[1]));
        q3.push_back((int)(c % km[2]));
        rep.push_back(p);
    };
    if (cells <= (1ll << 28)) {
        std::vector<int32_t> cell_q((size_t)cells, -1);
        for (int64_t p = 0; p < np; p++) {
            int32_t &u = cell_q[(size_t)code[p]];
            if (u < 0) {
                u = (int32_t)rep.size();
                add_q(p);
            }
            uq[p] = u;
        }
    } else {
        std::vector<int64_t> order((size_t)np);
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(),
                  [&](int64_t a, int64_t b) { return code[a] < code[b] || (code[a] == code[b] && a < b); });
        for (int64_t i = 0; i < np; i++) {
            const int64_t p = order[i];
            if (i == 0 || code[p] != code[order[i - 1]]) add_q(p);
            uq[p] = (int32_t)(rep.size() - 1);
        }
    }
    const int64_t nq = (int64_t)rep.size();
    int *d_ik3 = nullptr, *d_q3 = nullptr, *d_grid = nullptr;
    double *d_f = nullptr, *d_sum = nullptr;
    int rc;
    auto release = [&]() { dfree(d_ik3); dfree(d_q3); dfree(d_grid); dfree(d_f); dfree(d_sum); };
    if ((rc = dalloc_t(&d_ik3, (size_t)nk * 3, "k cells")) || (rc = dalloc_t(&d_q3, (size_t)nq * 3, "q cells")) ||
        (rc = dalloc_t(&d_grid, (size_t)cells, "k mesh map")) || (rc = dalloc_t(&d_f, (size_t)nk * h->nb, "f")) ||
        (rc = dalloc_t(&d_sum, (size_t)nq, "q_scr^2"))) {
        release();
        return rc;
    }
    std::vector<double> sum((size_t)nq, 0.0);
    hipError_t e = hipMemcpyAsync(d_ik3, ik3.data(), ik3.size() * 4, hipMemcpyHostToDevice, h->st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_q3, q3.data(), q3.size() * 4, hipMemcpyHostToDevice, h->st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_f, f, (size_t)nk * h->nb * 8, hipMemcpyDefault, h->st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_sum, sum.data(), (size_t)nq * 8, hipMemcpyHostToDevice, h->st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fill_int, dim3(nblk(cells, 256)), dim3(256), 0, h->st, cells, d_grid, INT_MAX);
        hipLaunchKernelGGL(k_kgrid, dim3(nblk(nk, 256)), dim3(256), 0, h->st, nk, d_ik3, km[1], km[2], d_grid);
        const dim3 g(nblk(nk, 256), (unsigned)std::min<int64_t>(nq, 65535));
        hipLaunchKernelGGL(k_rpa_qscr2, g, dim3(256), 0, h->st, nq, d_q3, nk, d_ik3, d_grid, km[0], km[1], km[2], dU,
                           m.nb_wannier, h->nb, h->d_e, d_f, m.temperature, m.screening == DMD_SCR_LINDHARD ? 1 : 0,
                           d_sum);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(sum.data(), d_sum, (size_t)nq * 8, hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    release();
    if (e != hipSuccess) return fail(DMD_ERR_HIP, hipGetErrorString(e));
    // vq per distinct q with the |q|^2 of its first pair, then per pair (Gamma: 0)
    auto q2_of = [&](int64_t p) {
        double v[3];
        for (int d = 0; d < 3; d++) {
            v[d] = kvec[3 * k1st[p] + d] - kvec[3 * k2nd[p] + d];
            v[d] -= std::floor(0.5 + v[d]);
            if (std::fabs(v[d] - 0.5) < 1e-6) v[d] = -0.5;
        }
        const double *G = m.GGT;
        return v[0] * v[0] * G[0] + v[1] * v[1] * G[4] + v[2] * v[2] * G[8] +
               2 * (v[0] * v[1] * G[1] + v[0] * v[2] * G[2] + v[1] * v[2] * G[5]);
    };
    std::vector<double> vq_u((size_t)nq);
    for (int64_t u = 0; u < nq; u++) vq_u[u] = m.prefac_vq / (q2_of(rep[u]) + m.prefac_vq / m.nk_full * sum[u]);
    vq.resize((size_t)np);
    for (int64_t p = 0; p < np; p++) vq[p] = q2_of(p) < 1e-20 ? 0.0 : vq_u[uq[p]];
    return DMD_OK;
}

// First (ik, jk) in (ik, jk) order whose k difference falls in each mesh cell
// (qIndexMap::build over all (ik, jk), comm/kmap.h:43-55): code ik * nk + jk.
// Grid: x over jk, y over ik (strided).
__global__ __launch_bounds__(256) void k_qcells_first(int64_t nk, const int *__restrict__ ik3, int m0, int m1, int m2,
                                                      unsigned long long *__restrict__ first)
{
    const int64_t jk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (jk >= nk) return;
    for (int64_t ik = blockIdx.y; ik < nk; ik += gridDim.y) {
        const int x = ((ik3[3 * ik] - ik3[3 * jk]) % m0 + m0) % m0, y = ((ik3[3 * ik + 1] - ik3[3 * jk + 1]) % m1 + m1) % m1,
                  z = ((ik3[3 * ik + 2] - ik3[3 * jk + 2]) % m2 + m2) % m2;
        atomicMin(first + ((int64_t)x * m1 + y) * m2 + z, (unsigned long long)(ik * nk + jk));
    }
}
__global__ void k_fill_u64(int64_t n, unsigned long long *__restrict__ a, unsigned long long v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

int rpa_vq_cells(dmd_scatt *h, const double2 *dU, const double *kvec, const double *f, const dmd_eimp_model &m,
                 std::vector<double> &vq_cell)
{
    const int *km = m.kmesh;
    if (km[0] <= 0 || km[1] <= 0 || km[2] <= 0) return fail(DMD_ERR_ARG, "RPA screening needs the k mesh");
    if (!f) return fail(DMD_ERR_ARG, "RPA screening needs the occupations f");
    const int64_t cells = (int64_t)km[0] * km[1] * km[2], nk = h->nk;
    if (cells > (1ll << 26)) return fail(DMD_ERR_UNSUPPORTED, "k mesh too large for the per-cell RPA table");
    std::vector<int> ik3((size_t)nk * 3);
    for (int64_t k = 0; k < nk; k++) mesh_cell(kvec + 3 * k, km, &ik3[3 * k]);
    int *d_ik3 = nullptr, *d_q3 = nullptr, *d_grid = nullptr;
    double *d_f = nullptr, *d_sum = nullptr;
    unsigned long long *d_first = nullptr;
    int rc;
    auto release = [&]() { dfree(d_ik3); dfree(d_q3); dfree(d_grid); dfree(d_f); dfree(d_sum); dfree(d_first); };
    if ((rc = dalloc_t(&d_ik3, (size_t)nk * 3, "k cells")) || (rc = dalloc_t(&d_grid, (size_t)cells, "k mesh map")) ||
        (rc = dalloc_t(&d_f, (size_t)nk * h->nb, "f")) || (rc = dalloc_t(&d_first, (size_t)cells, "q cells"))) {
        release();
        return rc;
    }
    std::vector<unsigned long long> first((size_t)cells);
    hipError_t e = hipMemcpyAsync(d_ik3, ik3.data(), ik3.size() * 4, hipMemcpyHostToDevice, h->st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_f, f, (size_t)nk * h->nb * 8, hipMemcpyDefault, h->st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fill_u64, dim3(nblk(cells, 256)), dim3(256), 0, h->st, cells, d_first, ~0ull);
        const dim3 g(nblk(nk, 256), (unsigned)std::min<int64_t>(nk, 65535));
        hipLaunchKernelGGL(k_qcells_first, g, dim3(256), 0, h->st, nk, d_ik3, km[0], km[1], km[2], d_first);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(first.data(), d_first, (size_t)cells * 8, hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    if (e != hipSuccess) {
        release();
        return fail(DMD_ERR_HIP, hipGetErrorString(e));
    }
    std::vector<int> q3;
    std::vector<int64_t> used;
    for (int64_t c = 0; c < cells; c++)
        if (first[c] != ~0ull) {
            used.push_back(c);
            q3.push_back((int)(c / ((int64_t)km[1] * km[2])));
            q3.push_back((int)((c / km[2]) % km[1]));
            q3.push_back((int)(c % km[2]));
        }
    const int64_t nq = (int64_t)used.size();
    std::vector<double> sum((size_t)nq, 0.0);
    if ((rc = dalloc_t(&d_q3, (size_t)nq * 3, "q cells")) || (rc = dalloc_t(&d_sum, (size_t)nq, "q_scr^2"))) {
        release();
        return rc;
    }
    e = hipMemcpyAsync(d_q3, q3.data(), q3.size() * 4, hipMemcpyHostToDevice, h->st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_sum, sum.data(), (size_t)nq * 8, hipMemcpyHostToDevice, h->st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fill_int, dim3(nblk(cells, 256)), dim3(256), 0, h->st, cells, d_grid, INT_MAX);
        hipLaunchKernelGGL(k_kgrid, dim3(nblk(nk, 256)), dim3(256), 0, h->st, nk, d_ik3, km[1], km[2], d_grid);
        const dim3 g(nblk(nk, 256), (unsigned)std::min<int64_t>(nq, 65535));
        hipLaunchKernelGGL(k_rpa_qscr2, g, dim3(256), 0, h->st, nq, d_q3, nk, d_ik3, d_grid, km[0], km[1], km[2], dU,
                           m.nb_wannier, h->nb, h->d_e, d_f, m.temperature, 0, d_sum);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(sum.data(), d_sum, (size_t)nq * 8, hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    release();
    if (e != hipSuccess) return fail(DMD_ERR_HIP, hipGetErrorString(e));
    vq_cell.assign((size_t)cells, 0.0);
    for (int64_t u = 0; u < nq; u++) {
        const int64_t code = (int64_t)first[used[u]], ik = code / nk, jk = code % nk;
        const double q2 = model_q2(m.GGT, kvec + 3 * ik, kvec + 3 * jk);
        vq_cell[used[u]] = q2 < 1e-20 ? 0.0 : m.prefac_vq / (q2 + m.prefac_vq / m.nk_full * sum[u]);
    }
    return DMD_OK;
}

}  // namespace dmd

using namespace dmd;

extern "C" int dmd_scatt_add_eimp_model(dmd_scatt *h, const uint64_t *k1st, const uint64_t *k2nd, const double *U,
                                        const double *kvec, const double *f, const dmd_eimp_model *m, double scale,
                                        int add)
{
    if (!h || !m || (h->np > 0 && (!k1st || !k2nd || !U || !kvec)))
        return fail(DMD_ERR_ARG, "dmd_scatt_add_eimp_model: null argument");
    if (const int sg = sp_guard(h, "dmd_scatt_add_eimp_model")) return sg;
    if (m->nb_wannier <= 0 || !(m->degauss > 0) || !(m->temperature > 0) || m->screening < DMD_SCR_UNSCREENED ||
        m->screening > DMD_SCR_LINDHARD)
        return fail(DMD_ERR_ARG, "dmd_scatt_add_eimp_model: bad model (nb_wannier, degauss, temperature, screening)");
    if (h->linear) return fail(DMD_ERR_STATE, "the handle holds the linearized operator: add the e-i P before set_Lsc");
    if (add && !h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated (add = 0 sets it)");
    if (h->nb < 1 || h->nb > 4) return fail(DMD_ERR_UNSUPPORTED, "the model e-i P is built for nb = 1..4");
    for (int64_t q = 0; q < h->np; q++)
        if (k1st[q] >= (uint64_t)h->nk || k2nd[q] >= (uint64_t)h->nk)
            return fail(DMD_ERR_ARG, "dmd_scatt_add_eimp_model: k index outside [0, nk)");
    if (h->np == 0) { h->has_P = true; return DMD_OK; }
    EimpConst c;
    c.m = *m;
    const double pg = 1. / (m->degauss * std::sqrt(2. * M_PI));  // ElecImp_Model.h:36-39
    c.prefac_delta = m->lindblad ? std::sqrt(pg) : pg;
    if (m->detail_balance) c.prefac_delta *= 0.5;
    c.prefac_exp = m->lindblad ? -0.25 / std::pow(m->degauss, 2) : -0.5 / std::pow(m->degauss, 2);

    const int64_t per = h->nb4, nw = m->nb_wannier;
    const int64_t chunk = std::min<int64_t>(h->np, std::max<int64_t>(1, (64ll << 20) / per));
    double2 *dU = nullptr, *s1 = nullptr, *s2 = nullptr;
    double *dk = nullptr, *dvq = nullptr;
    uint64_t *d1 = nullptr, *d2 = nullptr;
    int rc;
    auto release = [&]() { dfree(dU); dfree(dk); dfree(s1); dfree(s2); dfree(d1); dfree(d2); dfree(dvq); };
    const bool ksum = m->screening == DMD_SCR_RPA || m->screening == DMD_SCR_LINDHARD;
    std::vector<double> vq;
    if ((rc = dalloc_t(&dU, (size_t)(h->nk * nw * h->nb), "U")) || (rc = dalloc_t(&dk, (size_t)h->nk * 3, "kvec")) ||
        (rc = dalloc_t(&s1, (size_t)(chunk * per), "e-i P staging")) ||
        (rc = dalloc_t(&s2, (size_t)(chunk * per), "e-i P staging")) || (rc = dalloc_t(&d1, (size_t)chunk, "k1")) ||
        (rc = dalloc_t(&d2, (size_t)chunk, "k2"))) {
        release();
        return rc;
    }
    hipError_t e = hipMemcpyAsync(dU, U, (size_t)(h->nk * nw * h->nb) * 16, hipMemcpyDefault, h->st);
    if (e == hipSuccess) e = hipMemcpyAsync(dk, kvec, (size_t)h->nk * 3 * 8, hipMemcpyDefault, h->st);
    if (e == hipSuccess && ksum) {
        if ((rc = rpa_vq(h, k1st, k2nd, dU, kvec, f, *m, vq)) || (rc = dalloc_t(&dvq, (size_t)chunk, "v(q)"))) {
            (void)hipStreamSynchronize(h->st);
            release();
            return rc;
        }
    }
    for (int64_t first = 0; first < h->np && e == hipSuccess && rc == DMD_OK; first += chunk) {
        const int64_t n = std::min(chunk, h->np - first);
        e = hipMemcpyAsync(d1, k1st + first, (size_t)n * 8, hipMemcpyDefault, h->st);
        if (e == hipSuccess) e = hipMemcpyAsync(d2, k2nd + first, (size_t)n * 8, hipMemcpyDefault, h->st);
        if (e == hipSuccess && ksum) e = hipMemcpyAsync(dvq, vq.data() + first, (size_t)n * 8, hipMemcpyHostToDevice, h->st);
        if (e != hipSuccess) break;
        const int pp = std::max<int>(1, 256 / (int)per);
        const unsigned g = nblk(n, pp);
        switch (h->nb) {
        case 1: hipLaunchKernelGGL(k_eimp_model<1>, dim3(g), dim3(256), 0, h->st, n, d1, d2, dU, dk, h->d_e, ksum ? dvq : nullptr, c, s1, s2); break;
        case 2: hipLaunchKernelGGL(k_eimp_model<2>, dim3(g), dim3(256), 0, h->st, n, d1, d2, dU, dk, h->d_e, ksum ? dvq : nullptr, c, s1, s2); break;
        case 3: hipLaunchKernelGGL(k_eimp_model<3>, dim3(g), dim3(256), 0, h->st, n, d1, d2, dU, dk, h->d_e, ksum ? dvq : nullptr, c, s1, s2); break;
        default: hipLaunchKernelGGL(k_eimp_model<4>, dim3(g), dim3(256), 0, h->st, n, d1, d2, dU, dk, h->d_e, ksum ? dvq : nullptr, c, s1, s2); break;
        }
        if ((e = hipGetLastError()) != hipSuccess) break;
        rc = place_P_dev(h, s1, s2, n, first, scale, add ? 1 : 0);
        // the next chunk's pair copies come from pageable host memory
        if (rc == DMD_OK) e = hipStreamSynchronize(h->st);
    }
    hipError_t e2 = hipStreamSynchronize(h->st);
    release();
    if (rc == DMD_OK && e == hipSuccess) e = e2;
    if (rc == DMD_OK && e != hipSuccess) rc = fail(DMD_ERR_HIP, hipGetErrorString(e));
    if (rc) {
        h->has_P = false;  // P partly updated
        return rc;
    }
    h->has_P = true;
    return DMD_OK;
}

