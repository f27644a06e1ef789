// dmd_driver.hip -- driver-side device code around the scattering operator:
// the band window of evolve_driver, the coherent term, the spin-type
// observable and the on-device fixed-step RK4 integrator (one step captured
// in a HIP graph and replayed).
//
// Reference: dm_dynamics::compute (src_v4.5.7/DensityMatrix_Dynamics.h:243-278),
// singdenmat_k::evolve_coh (src_v4.5.7/DenMat.cpp:29-55), update_dm_euler's
// Hermitization (DenMat.cpp:236-248), ob_1dmk::measure_brange
// (src_v4.5.7/observable.h:301-366), evolve_driver (evolve_eph.cpp:21-32).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dmd_internal.h"

namespace dmd {

// out (nk x nbd^2) = window ddm w (nk x nb^2) embedded at [b0, b0+nb), zero elsewhere
__global__ void k_expand(int64_t nk, int nb, int nbd, int b0, const double2 *__restrict__ w, double2 *__restrict__ out)
{
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nbd2 = nbd * nbd;
    if (idx >= nk * nbd2) return;
    int64_t k = idx / nbd2;
    int ij = (int)(idx - k * nbd2), i = ij / nbd, j = ij - i * nbd;
    int wi = i - b0, wj = j - b0;
    bool in = wi >= 0 && wi < nb && wj >= 0 && wj < nb;
    out[idx] = in ? w[k * nb * nb + wi * nb + wj] : make_double2(0.0, 0.0);
}

// Element (i,j) of -i[H(t), y] for one k (zhemm semantics on H's upper
// triangle; interaction-picture phases off the diagonal; in the Schroedinger
// picture H + diag(e) without phases, or -i[e, y] when no H is set).
__device__ __forceinline__ double2 herm_up_rt(const double2 *A, int n, int i, int j)
{
    if (i < j) return A[i * n + j];
    if (i > j) return cconj(A[j * n + i]);
    return make_double2(A[i * n + i].x, 0.0);
}
__device__ __forceinline__ double2 coh_H(const double2 *Hk, const double *ek, int n, int i, int j, double t,
                                         int schrod)
{
    double2 v = herm_up_rt(Hk, n, i, j);
    if (schrod) {
        if (i == j) v.x += ek[i];
        return v;
    }
    if (i != j) v = cmul(v, ccis((ek[i] - ek[j]) * t));
    return v;
}
__device__ double2 coherent_elem(const double2 *__restrict__ Hk, const double *__restrict__ ek,
                                 const double2 *__restrict__ yk, int n, int i, int j, double t, int schrod)
{
    double2 c = make_double2(0, 0);
    if (Hk) {
        for (int m = 0; m < n; m++) c = cadd(c, cmul(yk[i * n + m], coh_H(Hk, ek, n, m, j, t, schrod)));  // y H
        double2 hy = make_double2(0, 0);
        for (int m = 0; m < n; m++) hy = cadd(hy, cmul(coh_H(Hk, ek, n, i, m, t, schrod), yk[m * n + j]));  // H y
        c = csub(hy, c);
    } else {  // Schroedinger picture without H: -i (e_i - e_j) y_ij
        double d = ek[i] - ek[j];
        c = make_double2(d * yk[i * n + j].x, d * yk[i * n + j].y);
    }
    return make_double2(c.y, -c.x);  // times -i
}

struct StageArgs {
    int64_t nk;
    int64_t nkb;        // states: nk, or B nk for B trajectories stored one after another
    int nbd, nb, b0;
    const double2 *H;   // nk x nbd^2 or null
    int coh;            // coherent term active
    int schrod;         // Schroedinger picture
    int expt;           // phases on the scattering window
    const double *edm;  // nk x nbd
    const double *e;    // nk x nb (window)
    const DClock *clk;
    double dt;
    const double *fsub; // linearized operator: rhoS = rho_S - diag(f_eq) (nk x nb), else null
};

// right-hand side element (i, j) of state kk (row k = kk mod nk)
__device__ __forceinline__ double2 rhs_elem(const StageArgs &A, int64_t k, int64_t kk, int i, int j,
                                            const double2 *yk, const double2 *__restrict__ w, double t)
{
    double2 kv = make_double2(0, 0);
    const int n = A.nbd;
    if (A.coh) kv = coherent_elem(A.H ? A.H + k * n * n : nullptr, A.edm + k * n, yk, n, i, j, t, A.schrod);
    int wi = i - A.b0, wj = j - A.b0;
    if (wi >= 0 && wi < A.nb && wj >= 0 && wj < A.nb) kv = cadd(kv, w[kk * A.nb * A.nb + wi * A.nb + wj]);
    return kv;
}

__device__ __forceinline__ void rotate_out(const StageArgs &A, int64_t k, int i, int j, double2 v, double t,
                                           double2 *__restrict__ rhoS)
{
    int wi = i - A.b0, wj = j - A.b0;
    if (!rhoS || wi < 0 || wi >= A.nb || wj < 0 || wj >= A.nb) return;
    if (A.expt && wi != wj) v = cmul(v, ccis(-(A.e[k * A.nb + wi] - A.e[k * A.nb + wj]) * t));
    if (A.fsub && wi == wj) v.x -= A.fsub[k * A.nb + wi];
    rhoS[k * A.nb * A.nb + wi * A.nb + wj] = v;
}

// RK4 stages 0..2: kv = f(t_s, yin); acc (+)= w_s kv; yout = y0 + a_{s+1} dt kv;
// window of yout rotated to rhoS at t_{s+1}.  One thread per (k, i, j).
__global__ void k_rk_stage(StageArgs A, int s, const double2 *__restrict__ yin, const double2 *__restrict__ w,
                           const double2 *__restrict__ y0, double2 *__restrict__ acc, double2 *__restrict__ yout,
                           double2 *__restrict__ rhoS)
{
    const double a_s[4] = { 0.0, 0.5, 0.5, 1.0 }, w_s[4] = { 1.0, 2.0, 2.0, 1.0 };
    const int n = A.nbd, n2 = n * n;
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= A.nkb * n2) return;
    int64_t kk = idx / n2, k = kk % A.nk;
    int ij = (int)(idx - kk * n2), i = ij / n, j = ij - i * n;
    double2 kv = rhs_elem(A, k, kk, i, j, yin + kk * n2, w, clock_t_at(A.clk, a_s[s]));
    acc[idx] = s == 0 ? kv : cadd(acc[idx], make_double2(w_s[s] * kv.x, w_s[s] * kv.y));
    const double c = a_s[s + 1] * A.dt;
    double2 y = make_double2(y0[idx].x + c * kv.x, y0[idx].y + c * kv.y);
    yout[idx] = y;
    rotate_out(A, k, i, j, y, clock_t_at(A.clk, a_s[s + 1]), rhoS);
}

// RK4 stage 3: y = y0 + dt/6 (acc + kv), Hermitized, into ynew.
__global__ void k_rk_final(StageArgs A, const double2 *__restrict__ yin, const double2 *__restrict__ w,
                           const double2 *__restrict__ y0, const double2 *__restrict__ acc,
                           double2 *__restrict__ ynew)
{
    const int n = A.nbd, n2 = n * n;
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= A.nkb * n2) return;
    int64_t kk = idx / n2, k = kk % A.nk;
    int ij = (int)(idx - kk * n2), i = ij / n, j = ij - i * n;
    const double t = clock_t_at(A.clk, 1.0);
    const double c = A.dt / 6.0;
    const double2 *yk = yin + kk * n2;
    double2 kij = rhs_elem(A, k, kk, i, j, yk, w, t);
    double2 aij = cadd(acc[idx], kij);
    double2 yij = make_double2(y0[idx].x + c * aij.x, y0[idx].y + c * aij.y);
    if (i == j) { ynew[idx] = make_double2(yij.x, 0.0); return; }
    int64_t jdx = kk * n2 + j * n + i;
    double2 kji = rhs_elem(A, k, kk, j, i, yk, w, t);
    double2 aji = cadd(acc[jdx], kji);
    double2 yji = make_double2(y0[jdx].x + c * aji.x, y0[jdx].y + c * aji.y);
    double2 u = i < j ? yij : yji, l = i < j ? yji : yij;
    double2 hh = make_double2(0.5 * (u.x + l.x), 0.5 * (u.y - l.y));
    ynew[idx] = i < j ? hh : cconj(hh);
}

// y0 = ynew; window rotated to rhoS at the next step's first stage time.
__global__ void k_rk_finish(StageArgs A, const double2 *__restrict__ ynew, double2 *__restrict__ y0,
                            double2 *__restrict__ rhoS)
{
    const int n = A.nbd, n2 = n * n;
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= A.nkb * n2) return;
    int64_t kk = idx / n2, k = kk % A.nk;
    int ij = (int)(idx - kk * n2), i = ij / n, j = ij - i * n;
    double2 v = ynew[idx];
    y0[idx] = v;
    const double tn = (A.clk->t0 + (double)(A.clk->step + 1) * A.clk->dt) + 0.0 * A.clk->dt;
    rotate_out(A, k, i, j, v, tn, rhoS);
}

__global__ void k_tick(DClock *clk)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) clk->step += 1;
}

// Observable partial sums: one thread per k, fixed-order block tree.
__global__ __launch_bounds__(256) void k_obs_partial(int64_t nk, int n, int i0, int i1, const double2 *__restrict__ s,
                                                     const double *__restrict__ edm, const double *__restrict__ f,
                                                     const double2 *__restrict__ y, const DClock *__restrict__ clk,
                                                     int interaction, double *__restrict__ part)
{
    __shared__ double red[256];
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double tot = 0;
    if (k < nk) {
        const double t = clock_t_at(clk, 0.0);
        const int n2 = n * n;
        for (int i = i0; i < i1; i++) {
            double ob = 0;
            for (int b = 0; b < n; b++) {
                double2 sk = s[k * n2 + i * n + b];
                if (i != b && interaction) sk = cmul(sk, ccis((edm[k * n + i] - edm[k * n + b]) * t));
                double2 d = y[k * n2 + b * n + i];
                if (i == b) d.x -= f[k * n + i];
                ob += cmul(sk, d).x;
            }
            tot += ob;
        }
    }
    red[threadIdx.x] = tot;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ __launch_bounds__(256) void k_obs_final(int nparts, const double *__restrict__ part, double prefac,
                                                   double *__restrict__ out, int idx)
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
    if (threadIdx.x == 0) out[idx] = red[0] * prefac;
}

static StageArgs stage_args(dmd_scatt *h, double dt)
{
    StageArgs A;
    A.nk = h->nk; A.nkb = h->nk; A.nbd = h->nbd; A.nb = h->nb; A.b0 = h->b0;
    A.H = h->has_H ? h->d_H : nullptr;
    A.coh = (h->has_H || h->schrodinger) ? 1 : 0;
    A.schrod = h->schrodinger;
    A.expt = h->expt;
    A.edm = h->d_edm; A.e = h->d_e; A.clk = h->d_clk; A.dt = dt;
    A.fsub = h->linear ? h->d_flin : nullptr;
    return A;
}

static int ensure_rk(dmd_scatt *h)
{
    const int64_t n = h->nk * (int64_t)h->nbd * h->nbd;
    if (h->rk_cap >= n) return DMD_OK;
    for (double2 **p : {&h->d_y0, &h->d_ya, &h->d_yb, &h->d_acc}) { dfree(*p); *p = nullptr; }
    dfree(h->d_w); h->d_w = nullptr;
    dfree(h->d_part); h->d_part = nullptr;
    int rc;
    for (double2 **p : {&h->d_y0, &h->d_ya, &h->d_yb, &h->d_acc})
        if ((rc = dalloc_t(p, n, "rk state"))) return rc;
    if ((rc = dalloc_t(&h->d_w, (size_t)h->nk * h->nb2, "rk window"))) return rc;
    if ((rc = dalloc_t(&h->d_part, nblk(h->nk, 256), "obs partials"))) return rc;
    h->rk_cap = n;
    return DMD_OK;
}

static int launch_obs(dmd_scatt *h, const double2 *y, double *out, int idx)
{
    unsigned nb = nblk(h->nk, 256);
    hipLaunchKernelGGL(k_obs_partial, dim3(nb), dim3(256), 0, h->st, h->nk, h->nbd, h->obs_i0, h->obs_i1, h->d_s,
                       h->d_edm, h->d_f, y, h->d_clk, h->schrodinger ? 0 : 1, h->d_part);
    hipLaunchKernelGGL(k_obs_final, dim3(1), dim3(256), 0, h->st, (int)nb, h->d_part, h->obs_prefac, out, idx);
    DMD_HIP(hipGetLastError());
    return DMD_OK;
}

// one RK4 step on the stream (captured into a graph by the caller)
static int launch_step(dmd_scatt *h, const StageArgs &A)
{
    const int64_t n = h->nk * (int64_t)h->nbd * h->nbd;
    dim3 g(nblk(n, 256)), b(256);
    const double a_s[4] = { 0.0, 0.5, 0.5, 1.0 };
    double2 *yin[4] = { h->d_y0, h->d_ya, h->d_yb, h->d_ya };
    double2 *yout[3] = { h->d_ya, h->d_yb, h->d_ya };
    int rc;
    for (int s = 0; s < 4; s++) {
        if ((rc = launch_contract(h, false, a_s[s], h->d_w))) return rc;
        if (s < 3)
            hipLaunchKernelGGL(k_rk_stage, g, b, 0, h->st, A, s, yin[s], h->d_w, h->d_y0, h->d_acc, yout[s], h->d_rhoS);
        else
            hipLaunchKernelGGL(k_rk_final, g, b, 0, h->st, A, yin[3], h->d_w, h->d_y0, h->d_acc, h->d_yb);
    }
    hipLaunchKernelGGL(k_rk_finish, g, b, 0, h->st, A, h->d_yb, h->d_y0, h->d_rhoS);
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, h->st, h->d_clk);
    DMD_HIP(hipGetLastError());
    return DMD_OK;
}

}  // namespace dmd

using namespace dmd;

extern "C" int dmd_scatt_set_window(dmd_scatt *h, int nb_dm, int b0, const double *e_dm)
{
    if (!h || !e_dm) return fail(DMD_ERR_ARG, "dmd_scatt_set_window: bad arguments");
    if (nb_dm < h->nb || nb_dm > 8 || b0 < 0 || b0 + h->nb > nb_dm)
        return fail(DMD_ERR_ARG, "dmd_scatt_set_window: need 0 <= b0, b0 + nb <= nb_dm <= 8");
    std::vector<double> edm((size_t)h->nk * nb_dm), e((size_t)h->nk * h->nb);
    if (is_device_ptr(e_dm)) DMD_HIP(hipMemcpy(edm.data(), e_dm, edm.size() * 8, hipMemcpyDeviceToHost));
    else std::memcpy(edm.data(), e_dm, edm.size() * 8);
    for (int64_t k = 0; k < h->nk; k++)
        for (int i = 0; i < h->nb; i++) e[k * h->nb + i] = edm[k * nb_dm + b0 + i];
    int rc;
    if (nb_dm != h->nbd) {
        dfree(h->d_edm); h->d_edm = nullptr;
        for (double2 **p : {&h->d_H, &h->d_s}) { dfree(*p); *p = nullptr; }
        dfree(h->d_f); h->d_f = nullptr;
        h->has_H = h->has_obs = false;
        if ((rc = dalloc_t(&h->d_edm, edm.size(), "e_dm"))) return rc;
        // staging for expanded states
        const size_t nx = (size_t)h->nk * nb_dm * nb_dm;
        for (double2 **p : {&h->d_in, &h->d_in1, &h->d_out}) { dfree(*p); *p = nullptr; }
        for (double2 **p : {&h->d_in, &h->d_in1, &h->d_out}) {
            if ((rc = dalloc_t(p, nx, "state staging"))) return rc;
            DMD_HIP(hipMemset(*p, 0, nx * sizeof(double2)));
        }
        h->rk_cap = 0;
    }
    DMD_HIP(hipMemcpy(h->d_edm, edm.data(), edm.size() * 8, hipMemcpyHostToDevice));
    DMD_HIP(hipMemcpy(h->d_e, e.data(), e.size() * 8, hipMemcpyHostToDevice));
    h->nbd = nb_dm;
    h->b0 = b0;
    return DMD_OK;
}

extern "C" int dmd_scatt_evolve_driver(dmd_scatt *h, double t, const double *dm, const double *dm1, double *ddmdt)
{
    if (!h || !dm || !ddmdt) return fail(DMD_ERR_ARG, "dmd_scatt_evolve_driver: bad arguments");
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    const size_t bytes = (size_t)h->nk * h->nbd * h->nbd * 16;
    int rc;
    const double2 *in = (const double2 *)dm, *in1 = (const double2 *)dm1;
    if (!is_device_ptr(dm)) { if ((rc = to_device(h, h->d_in, dm, bytes))) return rc; in = h->d_in; }
    if (dm1 && !is_device_ptr(dm1)) { if ((rc = to_device(h, h->d_in1, dm1, bytes))) return rc; in1 = h->d_in1; }
    if (!h->d_w && (rc = dalloc_t(&h->d_w, (size_t)h->nk * h->nb2, "window ddm"))) return rc;
    if ((rc = tiles4_prepare(h))) return rc;
    if ((rc = set_clock(h, t, 0.0, 0))) return rc;
    if ((rc = launch_rotate(h, in, in1, h->nbd, h->b0, 0.0))) return rc;
    if ((rc = launch_contract(h, dm1 != nullptr, 0.0, h->d_w))) return rc;
    double2 *out = is_device_ptr(ddmdt) ? (double2 *)ddmdt : h->d_out;
    int64_t n = h->nk * (int64_t)h->nbd * h->nbd;
    hipLaunchKernelGGL(k_expand, dim3(nblk(n, 256)), dim3(256), 0, h->st, h->nk, h->nb, h->nbd, h->b0, h->d_w, out);
    DMD_HIP(hipGetLastError());
    if (out == h->d_out && (rc = from_device(h, ddmdt, h->d_out, bytes))) return rc;
    DMD_HIP(hipStreamSynchronize(h->st));
    return DMD_OK;
}

extern "C" int dmd_scatt_set_coherent(dmd_scatt *h, const double *H)
{
    if (!h) return fail(DMD_ERR_ARG, "null handle");
    if (!H) { h->has_H = false; return DMD_OK; }
    const size_t n = (size_t)h->nk * h->nbd * h->nbd;
    int rc;
    if (!h->d_H && (rc = dalloc_t(&h->d_H, n, "H"))) return rc;
    if ((rc = to_device(h, h->d_H, H, n * 16))) return rc;
    DMD_HIP(hipStreamSynchronize(h->st));
    h->has_H = true;
    return DMD_OK;
}

extern "C" int dmd_scatt_set_observable(dmd_scatt *h, const double *s, const double *f, int i0, int i1, double prefac)
{
    if (!h) return fail(DMD_ERR_ARG, "null handle");
    if (!s) { h->has_obs = false; return DMD_OK; }
    if (!f || i0 < 0 || i1 > h->nbd || i0 > i1) return fail(DMD_ERR_ARG, "dmd_scatt_set_observable: bad arguments");
    const size_t n = (size_t)h->nk * h->nbd * h->nbd;
    int rc;
    if (!h->d_s && (rc = dalloc_t(&h->d_s, n, "s"))) return rc;
    if (!h->d_f && (rc = dalloc_t(&h->d_f, (size_t)h->nk * h->nbd, "f"))) return rc;
    if (!h->d_obs && (rc = dalloc_t(&h->d_obs, 1, "obs"))) return rc;
    if ((rc = to_device(h, h->d_s, s, n * 16))) return rc;
    if ((rc = to_device(h, h->d_f, f, (size_t)h->nk * h->nbd * 8))) return rc;
    DMD_HIP(hipStreamSynchronize(h->st));
    h->obs_i0 = i0; h->obs_i1 = i1; h->obs_prefac = prefac;
    h->has_obs = true;
    return DMD_OK;
}

extern "C" int dmd_scatt_measure(dmd_scatt *h, double t, const double *rho, double *value)
{
    if (!h || !rho || !value) return fail(DMD_ERR_ARG, "dmd_scatt_measure: bad arguments");
    if (!h->has_obs) return fail(DMD_ERR_STATE, "no observable set");
    int rc;
    if ((rc = ensure_rk(h))) return rc;
    const size_t bytes = (size_t)h->nk * h->nbd * h->nbd * 16;
    const double2 *y = (const double2 *)rho;
    if (!is_device_ptr(rho)) { if ((rc = to_device(h, h->d_ya, rho, bytes))) return rc; y = h->d_ya; }
    if ((rc = set_clock(h, t, 0.0, 0))) return rc;
    if ((rc = launch_obs(h, y, h->d_obs, 0))) return rc;
    DMD_HIP(hipMemcpyAsync(value, h->d_obs, 8, hipMemcpyDeviceToHost, h->st));
    DMD_HIP(hipStreamSynchronize(h->st));
    return DMD_OK;
}

extern "C" int dmd_scatt_integrate_rk4(dmd_scatt *h, double t0, double dt, int64_t n_steps, int report_every,
                                       double *rho, double *obs_out)
{
    if (!h || !rho || n_steps < 0) return fail(DMD_ERR_ARG, "dmd_scatt_integrate_rk4: bad arguments");
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    const bool want_obs = report_every > 0 && h->has_obs;
    if (want_obs && !obs_out) return fail(DMD_ERR_ARG, "obs_out is null");
    int rc;
    if ((rc = ensure_rk(h))) return rc;
    if ((rc = tiles4_prepare(h))) return rc;
    const int64_t nobs = want_obs ? n_steps / report_every + 1 : 0;
    double *d_obs = nullptr;
    if (nobs && (rc = dalloc_t(&d_obs, nobs, "obs series"))) return rc;
    struct Guard { double *p; ~Guard() { dfree(p); } } guard{ d_obs };
    const size_t n = (size_t)h->nk * h->nbd * h->nbd;
    if ((rc = to_device(h, h->d_y0, rho, n * 16))) return rc;
    if ((rc = set_clock(h, t0, dt, 0))) return rc;
    StageArgs A = stage_args(h, dt);
    if ((rc = launch_rotate(h, h->d_y0, nullptr, h->nbd, h->b0, 0.0))) return rc;
    int iobs = 0;
    if (want_obs && (rc = launch_obs(h, h->d_y0, d_obs, iobs++))) return rc;
    const char *env = std::getenv("DMD_GRAPH");
    const bool use_graph = !(env && env[0] == '0') && n_steps > 1;
    hipGraphExec_t exec = nullptr;
    if (use_graph) {
        hipGraph_t graph = nullptr;
        DMD_HIP(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
        rc = launch_step(h, A);
        hipError_t e = hipStreamEndCapture(h->st, &graph);
        if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
        if (e != hipSuccess) return fail(DMD_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(DMD_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    for (int64_t it = 0; it < n_steps; it++) {
        if (exec) {
            hipError_t e = hipGraphLaunch(exec, h->st);
            if (e != hipSuccess) { (void)hipGraphExecDestroy(exec); return fail(DMD_ERR_HIP, hipGetErrorString(e)); }
        } else if ((rc = launch_step(h, A))) {
            return rc;
        }
        if (want_obs && (it + 1) % report_every == 0 && (rc = launch_obs(h, h->d_y0, d_obs, iobs++))) {
            if (exec) (void)hipGraphExecDestroy(exec);
            return rc;
        }
    }
    hipError_t e = hipStreamSynchronize(h->st);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (e != hipSuccess) return fail(DMD_ERR_HIP, hipGetErrorString(e));
    if ((rc = from_device(h, rho, h->d_y0, n * 16))) return rc;
    if (nobs) DMD_HIP(hipMemcpyAsync(obs_out, d_obs, nobs * 8, hipMemcpyDeviceToHost, h->st));
    DMD_HIP(hipStreamSynchronize(h->st));
    return DMD_OK;
}

namespace dmd {
// B x nk x nb^2 staging of batched inputs and outputs
static int ensure_bio(dmd_scatt *h, int B)
{
    const int64_t n = (int64_t)B * h->nk * h->nb2;
    if (h->bio_cap >= n) return DMD_OK;
    dfree(h->d_bin); dfree(h->d_bout);
    h->d_bin = h->d_bout = nullptr;
    h->bio_cap = 0;
    int rc;
    if ((rc = dalloc_t(&h->d_bin, n, "batched input")) || (rc = dalloc_t(&h->d_bout, n, "batched output"))) return rc;
    h->bio_cap = n;
    return DMD_OK;
}
static int ensure_rkb(dmd_scatt *h, int B)
{
    const int64_t n = (int64_t)B * h->nk * h->nb2;
    if (h->rkb_cap >= n) return DMD_OK;
    for (double2 **p : {&h->d_yB0, &h->d_yBa, &h->d_yBb, &h->d_accB, &h->d_wB}) { dfree(*p); *p = nullptr; }
    h->rkb_cap = 0;
    int rc;
    for (double2 **p : {&h->d_yB0, &h->d_yBa, &h->d_yBb, &h->d_accB, &h->d_wB})
        if ((rc = dalloc_t(p, n, "batched rk state"))) return rc;
    if (!h->d_part && (rc = dalloc_t(&h->d_part, nblk(h->nk, 256), "obs partials"))) return rc;
    h->rkb_cap = n;
    return DMD_OK;
}
static bool batched_native(const dmd_scatt *h)
{
    return (h->precision == DMD_PREC_BF16 || h->precision == DMD_PREC_C64) && h->nb == 4;
}

// one RK4 step of B trajectories (stage kernels over B nk states)
static int launch_step_b(dmd_scatt *h, int B, const StageArgs &A)
{
    const int64_t n = A.nkb * (int64_t)h->nb2;
    dim3 g(nblk(n, 256)), b(256);
    const double a_s[4] = { 0.0, 0.5, 0.5, 1.0 };
    double2 *yin[4] = { h->d_yB0, h->d_yBa, h->d_yBb, h->d_yBa };
    double2 *yout[3] = { h->d_yBa, h->d_yBb, h->d_yBa };
    int rc;
    for (int s = 0; s < 4; s++) {
        if ((rc = launch_batched(h, B, yin[s], a_s[s], h->d_wB))) return rc;
        if (s < 3)
            hipLaunchKernelGGL(k_rk_stage, g, b, 0, h->st, A, s, yin[s], h->d_wB, h->d_yB0, h->d_accB, yout[s],
                               (double2 *)nullptr);
        else
            hipLaunchKernelGGL(k_rk_final, g, b, 0, h->st, A, yin[3], h->d_wB, h->d_yB0, h->d_accB, h->d_yBb);
    }
    hipLaunchKernelGGL(k_rk_finish, g, b, 0, h->st, A, h->d_yBb, h->d_yB0, (double2 *)nullptr);
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, h->st, h->d_clk);
    DMD_HIP(hipGetLastError());
    return DMD_OK;
}
}  // namespace dmd

extern "C" int dmd_scatt_eval_batched(dmd_scatt *h, double t, int B, const double *rho_B, double *ddmdt_B)
{
    if (!h || B < 1 || !rho_B || !ddmdt_B) return fail(DMD_ERR_ARG, "dmd_scatt_eval_batched: bad arguments");
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    if (h->nbd != h->nb) return fail(DMD_ERR_UNSUPPORTED, "dmd_scatt_eval_batched: states are nk x nb^2 (no band window)");
    const size_t per = (size_t)h->nk * h->nb2 * 2;  // doubles per trajectory
    int rc;
    if (!batched_native(h)) {  // one evaluation per trajectory
        for (int b = 0; b < B; b++)
            if ((rc = dmd_scatt_eval(h, t, rho_B + b * per, nullptr, ddmdt_B + b * per))) return rc;
        return DMD_OK;
    }
    if ((rc = batched_prepare(h, B)) || (rc = ensure_bio(h, B))) return rc;
    const size_t bytes = per * 8 * B;
    // the state always lands in d_bin (dmd_scatt_time_eval_batched replays it)
    if ((rc = to_device(h, h->d_bin, rho_B, bytes))) return rc;
    double2 *out = is_device_ptr(ddmdt_B) ? (double2 *)ddmdt_B : h->d_bout;
    if ((rc = set_clock(h, t, 0.0, 0))) return rc;
    if ((rc = launch_batched(h, B, h->d_bin, 0.0, out))) return rc;
    if (out == h->d_bout && (rc = from_device(h, ddmdt_B, h->d_bout, bytes))) return rc;
    DMD_HIP(hipStreamSynchronize(h->st));
    h->last_B = B;
    return DMD_OK;
}

extern "C" int dmd_scatt_time_eval_batched(dmd_scatt *h, double t, int n, double *kernel_ms)
{
    if (!h || n < 1) return fail(DMD_ERR_ARG, "dmd_scatt_time_eval_batched: bad arguments");
    if (!batched_native(h) || h->last_B < 1)
        return fail(DMD_ERR_STATE, "dmd_scatt_time_eval_batched: needs a reduced-precision nb = 4 handle after dmd_scatt_eval_batched");
    int rc;
    const int B = h->last_B;
    if ((rc = set_clock(h, t, 0.0, 0))) return rc;
    if ((rc = launch_batched(h, B, h->d_bin, 0.0, h->d_bout))) return rc;  // warm
    DMD_HIP(hipStreamSynchronize(h->st));
    DMD_HIP(hipEventRecord(h->ev0, h->st));
    for (int i = 0; i < n; i++)
        if ((rc = launch_batched(h, B, h->d_bin, 0.0, h->d_bout))) return rc;
    DMD_HIP(hipEventRecord(h->ev1, h->st));
    DMD_HIP(hipEventSynchronize(h->ev1));
    float ms = 0;
    DMD_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    if (kernel_ms) *kernel_ms = ms / n;
    h->last_ms = ms / n;
    return DMD_OK;
}

extern "C" int dmd_scatt_integrate_rk4_batched(dmd_scatt *h, int B, double t0, double dt, int64_t n_steps,
                                               int report_every, double *rho_B, double *obs_B)
{
    if (!h || B < 1 || !rho_B || n_steps < 0) return fail(DMD_ERR_ARG, "dmd_scatt_integrate_rk4_batched: bad arguments");
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    const bool want_obs = report_every > 0 && h->has_obs;
    if (want_obs && !obs_B) return fail(DMD_ERR_ARG, "obs_B is null");
    const int64_t nobs = want_obs ? n_steps / report_every + 1 : 0;
    const size_t per = (size_t)h->nk * h->nbd * h->nbd * 2;
    int rc;
    if (!batched_native(h)) {  // one trajectory after another
        for (int b = 0; b < B; b++)
            if ((rc = dmd_scatt_integrate_rk4(h, t0, dt, n_steps, report_every, rho_B + b * per,
                                              want_obs ? obs_B + b * nobs : nullptr)))
                return rc;
        return DMD_OK;
    }
    if (h->nbd != h->nb) return fail(DMD_ERR_UNSUPPORTED, "batched RK4: states are nk x nb^2 (no band window)");
    if ((rc = batched_prepare(h, B)) || (rc = ensure_rkb(h, B))) return rc;
    double *d_obs = nullptr;
    if (nobs && (rc = dalloc_t(&d_obs, (size_t)nobs * B, "obs series"))) return rc;
    struct Guard { double *p; ~Guard() { dfree(p); } } guard{ d_obs };
    const size_t bytes = per * 8 * B;
    if ((rc = to_device(h, h->d_yB0, rho_B, bytes))) return rc;
    if ((rc = set_clock(h, t0, dt, 0))) return rc;
    StageArgs A = stage_args(h, dt);
    A.nkb = h->nk * (int64_t)B;
    const size_t n1 = (size_t)h->nk * h->nb2;
    auto report = [&](int64_t iobs) -> int {
        for (int b = 0; b < B; b++) {
            int r = launch_obs(h, h->d_yB0 + b * n1, d_obs, (int)(b * nobs + iobs));
            if (r) return r;
        }
        return DMD_OK;
    };
    int64_t iobs = 0;
    if (want_obs && (rc = report(iobs++))) return rc;
    const char *env = std::getenv("DMD_GRAPH");
    const bool use_graph = !(env && env[0] == '0') && n_steps > 1;
    hipGraphExec_t exec = nullptr;
    if (use_graph) {
        hipGraph_t graph = nullptr;
        DMD_HIP(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
        rc = launch_step_b(h, B, A);
        hipError_t e = hipStreamEndCapture(h->st, &graph);
        if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
        if (e != hipSuccess) return fail(DMD_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(DMD_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    for (int64_t it = 0; it < n_steps; it++) {
        if (exec) {
            hipError_t e = hipGraphLaunch(exec, h->st);
            if (e != hipSuccess) { (void)hipGraphExecDestroy(exec); return fail(DMD_ERR_HIP, hipGetErrorString(e)); }
        } else if ((rc = launch_step_b(h, B, A))) {
            return rc;
        }
        if (want_obs && (it + 1) % report_every == 0 && (rc = report(iobs++))) {
            if (exec) (void)hipGraphExecDestroy(exec);
            return rc;
        }
    }
    hipError_t e = hipStreamSynchronize(h->st);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (e != hipSuccess) return fail(DMD_ERR_HIP, hipGetErrorString(e));
    if ((rc = from_device(h, rho_B, h->d_yB0, bytes))) return rc;
    if (nobs) DMD_HIP(hipMemcpyAsync(obs_B, d_obs, (size_t)nobs * B * 8, hipMemcpyDeviceToHost, h->st));
    DMD_HIP(hipStreamSynchronize(h->st));
    return DMD_OK;
}
