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
// host side
// ---------------------------------------------------------------------------
static int vjp_ready(dmd_scatt *h)
{
    if (h->nb > 2) return fail(DMD_ERR_UNSUPPORTED, "the adjoint is built for nb <= 2");
    if (h->linear) return fail(DMD_ERR_UNSUPPORTED, "the adjoint is built for the P form, not the linearized operator");
    if (h->nbd != h->nb || h->has_H || h->has_eq || h->schrodinger)
        return fail(DMD_ERR_UNSUPPORTED, "the adjoint covers the scattering term alone (no window, coherent term or ddmdt_eq)");
    if (h->precision != DMD_PREC_FP64) return fail(DMD_ERR_UNSUPPORTED, "the adjoint is FP64");
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    // the VJP reads raw row sums of the single-read contraction
    if (contract_kind(h) != DMD_KERNEL_SINGLE_READ) {
        int rc = dmd_scatt_set_flags(h, h->expt, h->schrodinger, DMD_KERNEL_SINGLE_READ);
        if (rc) return rc;
    }
    const size_t n = (size_t)h->nk * h->nb2;
    int rc;
    if (!h->d_adj && (rc = dalloc_t(&h->d_adj, 3 * n, "vjp scratch"))) return rc;
    return DMD_OK;
}

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

// g = (df/drho)^T v at the state y (device arrays nk x nb^2), clock already set
static int launch_vjp(dmd_scatt *h, const double2 *y, const double2 *v, double2 *g)
{
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
    if ((rc = launch_vjp(h, tmp, tmp + n, tmp + 2 * n))) return rc;
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
    for (int c = 0; c < ncomp; c++)
        if (!comps[c] || comps[c]->nk != h->nk || comps[c]->nb != h->nb || comps[c]->nbd != comps[c]->nb)
            return fail(DMD_ERR_ARG, "dmd_scatt_adjoint_rk4: components must match the handle (no window)");
    const int64_t nk = h->nk;
    const size_t n = (size_t)nk * h->nb2, bytes = n * 16;
    const int64_t C = checkpoint_every, ncp = (n_steps + C - 1) / C;
    const int64_t nobs = obs ? n_steps / report_every + 1 : 0;
    // device memory: checkpoints, the segment's states, stages, cotangents
    std::vector<void *> mem;
    struct Guard { std::vector<void *> *m; ~Guard() { for (void *p : *m) dfree(p); } } guard{ &mem };
    auto alloc = [&](double2 **p, size_t cnt, const char *what) {
        int r = dalloc_t(p, cnt, what);
        if (!r) mem.push_back(*p);
        return r;
    };
    double2 *cp, *seg, *Ys, *Ks, *ybar, *kbar, *ynew, *tmp;
    double *d_dth = nullptr, *d_dobs = nullptr;
    if ((rc = alloc(&cp, (size_t)ncp * n, "checkpoints")) ||
        (rc = alloc(&seg, (size_t)std::max<int64_t>(C, 2) * n, "segment states")) ||
        (rc = alloc(&Ys, 4 * n, "stage states")) || (rc = alloc(&Ks, 3 * n, "stage slopes")) ||
        (rc = alloc(&ybar, n, "adjoint state")) || (rc = alloc(&kbar, 4 * n, "stage cotangents")) ||
        (rc = alloc(&ynew, n, "adjoint next")) || (rc = alloc(&tmp, 2 * n, "scratch")))
        return rc;
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
    (void)d_dobs;
    const dim3 gn(nblk((int64_t)n, 256)), b256(256);
    const double a_s[4] = { 0.0, 0.5, 0.5, 1.0 };
    // RHS at state y, stage time a of step j: out = f(y)
    auto rhs = [&](dmd_scatt *op, int64_t j, double a, const double2 *y, double2 *out) -> int {
        int r;
        if ((r = set_clock(op, t0, dt, j))) return r;
        if ((r = launch_rotate(op, y, nullptr, op->nb, 0, a))) return r;
        return launch_contract(op, false, a, out);
    };
    // one forward step j: y -> ynext (Hermitized)
    auto step = [&](int64_t j, const double2 *y, double2 *ynext) -> int {
        int r;
        double2 *K = Ks, *Y = tmp;
        if ((r = rhs(h, j, 0.0, y, K))) return r;                                 // K1
        hipLaunchKernelGGL(k_axpy, gn, b256, 0, h->st, (int64_t)n, y, 0.5 * dt, K, Y);
        hipLaunchKernelGGL(k_axpy, gn, b256, 0, h->st, (int64_t)n, y, dt / 6.0, K, ynext);
        if ((r = rhs(h, j, 0.5, Y, K))) return r;                                 // K2
        hipLaunchKernelGGL(k_axpy, gn, b256, 0, h->st, (int64_t)n, y, 0.5 * dt, K, Y);
        hipLaunchKernelGGL(k_acc, gn, b256, 0, h->st, (int64_t)n, dt / 3.0, K, ynext);
        if ((r = rhs(h, j, 0.5, Y, K))) return r;                                 // K3
        hipLaunchKernelGGL(k_axpy, gn, b256, 0, h->st, (int64_t)n, y, dt, K, Y);
        hipLaunchKernelGGL(k_acc, gn, b256, 0, h->st, (int64_t)n, dt / 3.0, K, ynext);
        if ((r = rhs(h, j, 1.0, Y, K))) return r;                                 // K4
        hipLaunchKernelGGL(k_acc, gn, b256, 0, h->st, (int64_t)n, dt / 6.0, K, ynext);
        hipLaunchKernelGGL(k_herm, gn, b256, 0, h->st, nk, h->nb, ynext, ynext);
        DMD_HIP(hipGetLastError());
        return DMD_OK;
    };
    // forward pass: checkpoints at steps 0, C, 2C, ...
    if ((rc = to_device(h, cp, rho0, bytes))) return rc;
    hipLaunchKernelGGL(k_herm, gn, b256, 0, h->st, nk, h->nb, cp, cp);
    for (int64_t c = 0; c + 1 < ncp; c++) {
        DMD_HIP(hipMemcpyAsync(seg, cp + c * n, bytes, hipMemcpyDeviceToDevice, h->st));
        for (int64_t j = c * C; j < (c + 1) * C; j++) {
            if ((rc = step(j, seg, seg + n))) return rc;
            DMD_HIP(hipMemcpyAsync(seg, seg + n, bytes, hipMemcpyDeviceToDevice, h->st));
        }
        DMD_HIP(hipMemcpyAsync(cp + (c + 1) * n, seg, bytes, hipMemcpyDeviceToDevice, h->st));
    }
    // reverse pass
    DMD_HIP(hipMemsetAsync(ybar, 0, bytes, h->st));
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
    for (int64_t c = ncp - 1; c >= 0; c--) {
        const int64_t j0 = c * C, j1 = std::min<int64_t>(n_steps, (c + 1) * C);
        // states y_j of the segment
        DMD_HIP(hipMemcpyAsync(seg, cp + c * n, bytes, hipMemcpyDeviceToDevice, h->st));
        for (int64_t j = j0; j + 1 < j1; j++)
            if ((rc = step(j, seg + (j - j0) * n, seg + (j + 1 - j0) * n))) return rc;
        for (int64_t j = j1 - 1; j >= j0; j--) {
            const double2 *y = seg + (j - j0) * n;
            // stage states Y1..Y4
            double2 *Y[4] = { Ys, Ys + n, Ys + 2 * n, Ys + 3 * n };
            DMD_HIP(hipMemcpyAsync(Y[0], y, bytes, hipMemcpyDeviceToDevice, h->st));
            const double cy[3] = { 0.5 * dt, 0.5 * dt, dt };
            for (int s = 0; s < 3; s++) {
                if ((rc = rhs(h, j, a_s[s], Y[s], Ks))) return rc;
                hipLaunchKernelGGL(k_axpy, gn, b256, 0, h->st, (int64_t)n, y, cy[s], Ks, Y[s + 1]);
            }
            // zbar = Herm(ybar); Kbar_s = dt/6 w_s zbar
            double2 *Kb[4] = { kbar, kbar + n, kbar + 2 * n, kbar + 3 * n };
            hipLaunchKernelGGL(k_herm, gn, b256, 0, h->st, nk, h->nb, ybar, ynew);
            const double ws[4] = { dt / 6.0, dt / 3.0, dt / 3.0, dt / 6.0 };
            for (int s = 0; s < 4; s++) {
                DMD_HIP(hipMemsetAsync(Kb[s], 0, bytes, h->st));
                hipLaunchKernelGGL(k_acc, gn, b256, 0, h->st, (int64_t)n, ws[s], ynew, Kb[s]);
            }
            const double back[4] = { 0.0, 0.5 * dt, 0.5 * dt, dt };  // Y_s = y + back_s K_{s-1}
            for (int s = 3; s >= 0; s--) {
                // parameter cotangents: Re <Kbar_s, f_c(Y_s)>; the component
                // evaluates on its own stream, the dot runs on the handle's
                for (int cc = 0; cc < ncomp; cc++) {
                    DMD_HIP(hipStreamSynchronize(h->st));
                    if ((rc = rhs(comps[cc], j, a_s[s], Y[s], tmp))) return rc;
                    DMD_HIP(hipStreamSynchronize(comps[cc]->st));
                    if ((rc = launch_dot(h, (int64_t)n, Kb[s], tmp, d_dth + cc))) return rc;
                }
                // Ybar_s = J_s^T Kbar_s
                if ((rc = set_clock(h, t0, dt, j))) return rc;
                h->h_clk.t0 = t0 + (double)j * dt + a_s[s] * dt;  // VJP phases at the stage time
                h->h_clk.dt = 0.0;
                h->h_clk.step = 0;
                DMD_HIP(hipMemcpyAsync(h->d_clk, &h->h_clk, sizeof(DClock), hipMemcpyHostToDevice, h->st));
                if ((rc = launch_vjp(h, Y[s], Kb[s], tmp))) return rc;
                hipLaunchKernelGGL(k_acc, gn, b256, 0, h->st, (int64_t)n, 1.0, tmp, ynew);
                if (s > 0) hipLaunchKernelGGL(k_acc, gn, b256, 0, h->st, (int64_t)n, back[s], tmp, Kb[s - 1]);
            }
            DMD_HIP(hipMemcpyAsync(ybar, ynew, bytes, hipMemcpyDeviceToDevice, h->st));
            if ((rc = add_obs(j))) return rc;
        }
    }
    DMD_HIP(hipStreamSynchronize(h->st));
    if (dL_drho0) {
        hipLaunchKernelGGL(k_herm, gn, b256, 0, h->st, nk, h->nb, ybar, ybar);
        if ((rc = from_device(h, dL_drho0, ybar, bytes))) return rc;
    }
    if (ncomp > 0) DMD_HIP(hipMemcpyAsync(dL_dtheta, d_dth, ncomp * sizeof(double), hipMemcpyDeviceToHost, h->st));
    DMD_HIP(hipStreamSynchronize(h->st));
    return DMD_OK;
}
