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
    int semic;          // alg_semiclassical: off-diagonals of the state and of the result are zero
    double ph_tau_inv;  // alg_phenom_relax: 1/tau (0: off), bands [ph_b0, ph_b1), f_eq nk x nbd
    int ph_b0, ph_b1;
    const double *ph_feq;
};

// right-hand side element (i, j) of state kk (row k = kk mod nk)
__device__ __forceinline__ double2 rhs_elem(const StageArgs &A, int64_t k, int64_t kk, int i, int j,
                                            const double2 *yk, const double2 *__restrict__ w, double t)
{
    double2 kv = make_double2(0, 0);
    const int n = A.nbd;
    // alg_semiclassical (DensityMatrix_Dynamics.h:249,277): the state's off-diagonals are zeroed before
    // compute and the result's after, so the coherent term -i[H, diag y] leaves nothing
    if (A.semic && i != j) return kv;
    if (A.coh && !A.semic) kv = coherent_elem(A.H ? A.H + k * n * n : nullptr, A.edm + k * n, yk, n, i, j, t, A.schrod);
    int wi = i - A.b0, wj = j - A.b0;
    if (wi >= 0 && wi < A.nb && wj >= 0 && wj < A.nb) kv = cadd(kv, w[kk * A.nb * A.nb + wi * A.nb + wj]);
    // phenom_relax::evolve (scattering_model/phenomenon_relax.h:42-51): -(rho - rho_eq) / tau
    if (A.ph_tau_inv != 0.0 && i >= A.ph_b0 && i < A.ph_b1 && j >= A.ph_b0 && j < A.ph_b1) {
        double2 d = yk[i * n + j];
        if (i == j) d.x -= A.ph_feq[k * n + i];
        kv = csub(kv, make_double2(d.x * A.ph_tau_inv, d.y * A.ph_tau_inv));
    }
    return kv;
}

__device__ __forceinline__ void rotate_out(const StageArgs &A, int64_t k, int i, int j, double2 v, double t,
                                           double2 *__restrict__ rhoS)
{
    int wi = i - A.b0, wj = j - A.b0;
    if (!rhoS || wi < 0 || wi >= A.nb || wj < 0 || wj >= A.nb) return;
    if (A.expt && wi != wj) v = cmul(v, ccis(-(A.e[k * A.nb + wi] - A.e[k * A.nb + wj]) * t));
    if (A.fsub && wi == wj) v.x -= A.fsub[k * A.nb + wi];
    if (A.semic && wi != wj) v = make_double2(0.0, 0.0);
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
    A.semic = h->semiclassical;
    A.ph_tau_inv = h->ph_tau_inv;
    A.ph_b0 = h->ph_b0;
    A.ph_b1 = h->ph_b1;
    A.ph_feq = h->d_phfeq;
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

static int run_rhs_hook(dmd_scatt *h, double t, const double2 *y);

// one RK4 step on the stream (captured into a graph by the caller); with a
// right-hand-side hook (not captured) it runs before every stage at t0 + (it + a) dt
static int launch_step(dmd_scatt *h, const StageArgs &A, int64_t it = 0, double t0 = 0.0, double dt = 0.0)
{
    const int64_t n = h->nk * (int64_t)h->nbd * h->nbd;
    dim3 g(nblk(n, 256)), b(256);
    const double a_s[4] = { 0.0, 0.5, 0.5, 1.0 };
    double2 *yin[4] = { h->d_y0, h->d_ya, h->d_yb, h->d_ya };
    double2 *yout[3] = { h->d_ya, h->d_yb, h->d_ya };
    int rc;
    for (int s = 0; s < 4; s++) {
        if (h->rhs_hook && (rc = run_rhs_hook(h, t0 + ((double)it + a_s[s]) * dt, yin[s]))) return rc;
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
    if ((rc = launch_rotate(h, h->d_y0, nullptr, h->nbd, h->b0, 0.0, h->semiclassical))) return rc;
    int iobs = 0;
    if (want_obs && (rc = launch_obs(h, h->d_y0, d_obs, iobs++))) return rc;
    const char *env = std::getenv("DMD_GRAPH");
    const bool use_graph = !(env && env[0] == '0') && n_steps > 1 && !h->rhs_hook;
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
        } else if ((rc = launch_step(h, A, it, t0, dt))) {
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
    if (h->rhs_hook) return fail(DMD_ERR_UNSUPPORTED, "a right-hand-side hook acts on one trajectory's state");
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    if (h->semiclassical || h->ph_tau_inv != 0.0)
        return fail(DMD_ERR_UNSUPPORTED, "the batched integrator has no semiclassical mode or phenomenological relaxation");
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

// ---------------------------------------------------------------------------
// Adaptive and first-order integrators of the reference's dm_dynamics:
// evolve_gsl (DensityMatrix_Dynamics.h:189-225) drives GSL's rkf45 stepper
// with the standard step-size control and hmax = tstep; evolve_euler (:177-187,
// 227-234) takes rho += dt (ddmdt + ddmdt^+)/2 (DenMat.cpp:236-248).  Both
// evaluate the same right-hand side as the RK4 path (coherent + scattering
// window, compute() at :243-278) on the device.
//
// The RKF45 stepper is Fehlberg's embedded 4(5) pair (the published tableau);
// the fourth-order solution advances the state and the difference to the
// fifth-order one estimates the local error.  Step-size control follows the
// standard rule documented for GSL's odeiv2 control: D_i = eps_abs + eps_rel
// |y_i| per real component, r = max_i |err_i| / D_i; r > 1.1 rejects the step
// and retries with h S r^(-1/q), r < 0.5 grows the next step to h S
// r^(-1/(q+1)), S = 0.9, q = 4, each change limited to [1/5, 5]; the step
// never exceeds hmax and the last one ends exactly on t1.
// ---------------------------------------------------------------------------
namespace dmd {

struct Coef6 {
    double c[6];
};
struct Ptr6 {
    const double2 *p[6];
};

// kout = f(t + a h, y): the full right-hand side from the window ddm w
__global__ void k_rhs_full(StageArgs A, const double2 *__restrict__ y, const double2 *__restrict__ w, double a,
                           double2 *__restrict__ kout)
{
    const int n = A.nbd, n2 = n * n;
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= A.nk * n2) return;
    int64_t k = idx / n2;
    int ij = (int)(idx - k * n2), i = ij / n, j = ij - i * n;
    kout[idx] = rhs_elem(A, k, k, i, j, y + k * n2, w, clock_t_at(A.clk, a));
}

// out = y0 + h sum_j c_j k_j (m terms)
__global__ void k_comb(int64_t n, const double2 *__restrict__ y0, double h, int m, Coef6 c, Ptr6 k,
                       double2 *__restrict__ out)
{
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    double2 s = make_double2(0, 0);
    for (int j = 0; j < m; j++) {
        const double2 v = k.p[j][idx];
        s.x += c.c[j] * v.x;
        s.y += c.c[j] * v.y;
    }
    out[idx] = make_double2(y0[idx].x + h * s.x, y0[idx].y + h * s.y);
}

// ynew = y0 + h sum b_j k_j; err = h sum e_j k_j; per-block max of |err| / D
__global__ __launch_bounds__(256) void k_rkf_final(int64_t n, const double2 *__restrict__ y0, double h, Coef6 b,
                                                   Coef6 e, Ptr6 k, double eps_abs, double eps_rel,
                                                   double2 *__restrict__ ynew, double *__restrict__ part)
{
    __shared__ double red[256];
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double r = 0.0;
    if (idx < n) {
        double2 s = make_double2(0, 0), q = make_double2(0, 0);
        for (int j = 0; j < 6; j++) {
            const double2 v = k.p[j][idx];
            s.x += b.c[j] * v.x; s.y += b.c[j] * v.y;
            q.x += e.c[j] * v.x; q.y += e.c[j] * v.y;
        }
        const double2 y = make_double2(y0[idx].x + h * s.x, y0[idx].y + h * s.y);
        ynew[idx] = y;
        r = fmax(fabs(h * q.x) / (eps_abs + eps_rel * fabs(y.x)), fabs(h * q.y) / (eps_abs + eps_rel * fabs(y.y)));
    }
    red[threadIdx.x] = r;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_max_final(int nparts, const double *__restrict__ part, double *__restrict__ out)
{
    __shared__ double red[256];
    double r = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) r = fmax(r, part[i]);
    red[threadIdx.x] = r;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// update_dm_euler (DenMat.cpp:236-248): y_ij += dt (k_ij + conj(k_ji)) / 2
// semic: evolve_euler's compute() zeroes the state's own off-diagonals
// (DensityMatrix_Dynamics.h:229,249), so they stay zero after the step
__global__ void k_euler(int64_t nk, int nbd, double dt, const double2 *__restrict__ kv, double2 *__restrict__ y,
                        int semic)
{
    const int n2 = nbd * nbd;
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nk * n2) return;
    int64_t k = idx / n2;
    int ij = (int)(idx - k * n2), i = ij / nbd, j = ij - i * nbd;
    if (semic && i != j) {
        y[idx] = make_double2(0.0, 0.0);
        return;
    }
    const double2 a = kv[idx], b = kv[k * n2 + j * nbd + i];
    y[idx].x += dt * 0.5 * (a.x + b.x);
    y[idx].y += dt * 0.5 * (a.y - b.y);
}

// adaptive scratch: y0, ytmp, ynew and six stage derivatives (9 states), the
// error partials and the reduced ratio
static int ensure_adaptive(dmd_scatt *h)
{
    const int64_t n = h->nk * (int64_t)h->nbd * h->nbd;
    int rc;
    if ((rc = ensure_rk(h))) return rc;
    if (h->ad_cap >= n) return DMD_OK;
    dfree(h->d_ad); dfree(h->d_adpart);
    h->d_ad = nullptr; h->d_adpart = nullptr;
    h->ad_cap = 0;
    if ((rc = dalloc_t(&h->d_ad, (size_t)9 * n, "adaptive integrator states"))) return rc;
    if ((rc = dalloc_t(&h->d_adpart, (size_t)nblk(n, 256) + 1, "error partials"))) return rc;
    h->ad_cap = n;
    return DMD_OK;
}

// update_scatt_inside: the caller's hook on the stage state y at time t, with
// the stream idle; the reduced forms' F_eq follows a changed P
static int run_rhs_hook(dmd_scatt *h, double t, const double2 *y)
{
    const size_t n = (size_t)h->nk * h->nbd * h->nbd;
    if (h->h_hook.size() < 2 * n) h->h_hook.resize(2 * n);
    DMD_HIP(hipMemcpyAsync(h->h_hook.data(), y, n * 16, hipMemcpyDeviceToHost, h->st));
    DMD_HIP(hipStreamSynchronize(h->st));
    if (h->rhs_hook(h->rhs_ctx, t, h->h_hook.data()) != 0)
        return fail(DMD_ERR_STATE, "the right-hand-side hook stopped the integrator at t = " + std::to_string(t));
    return tiles4_prepare(h);
}

// k = f(t + a hstep, y) with the clock at (t, hstep)
static int full_rhs(dmd_scatt *h, const StageArgs &A, const double2 *y, double a, double2 *kout)
{
    int rc;
    if (h->rhs_hook &&
        (rc = run_rhs_hook(h, h->h_clk.t0 + ((double)h->h_clk.step + a) * h->h_clk.dt, y)))
        return rc;
    if ((rc = launch_rotate(h, y, nullptr, h->nbd, h->b0, a, h->semiclassical))) return rc;
    if ((rc = launch_contract(h, false, a, h->d_w))) return rc;
    const int64_t n = h->nk * (int64_t)h->nbd * h->nbd;
    hipLaunchKernelGGL(k_rhs_full, dim3(nblk(n, 256)), dim3(256), 0, h->st, A, y, h->d_w, a, kout);
    DMD_HIP(hipGetLastError());
    return DMD_OK;
}

// one RKF45 attempt of size hs from (t, y0): ynew and the error ratio
static int rkf45_attempt(dmd_scatt *h, double t, double hs, double eps_abs, double eps_rel, double *ratio)
{
    static const double c[6] = { 0.0, 1.0 / 4, 3.0 / 8, 12.0 / 13, 1.0, 1.0 / 2 };
    static const double a[6][5] = {
        { 0, 0, 0, 0, 0 },
        { 1.0 / 4, 0, 0, 0, 0 },
        { 3.0 / 32, 9.0 / 32, 0, 0, 0 },
        { 1932.0 / 2197, -7200.0 / 2197, 7296.0 / 2197, 0, 0 },
        { 439.0 / 216, -8.0, 3680.0 / 513, -845.0 / 4104, 0 },
        { -8.0 / 27, 2.0, -3544.0 / 2565, 1859.0 / 4104, -11.0 / 40 },
    };
    static const double b4[6] = { 25.0 / 216, 0, 1408.0 / 2565, 2197.0 / 4104, -1.0 / 5, 0 };
    static const double b5[6] = { 16.0 / 135, 0, 6656.0 / 12825, 28561.0 / 56430, -9.0 / 50, 2.0 / 55 };
    const int64_t n = h->nk * (int64_t)h->nbd * h->nbd;
    double2 *y0 = h->d_ad, *ytmp = h->d_ad + n, *ynew = h->d_ad + 2 * n, *K[6];
    for (int j = 0; j < 6; j++) K[j] = h->d_ad + (3 + j) * n;
    int rc;
    if ((rc = set_clock(h, t, hs, 0))) return rc;
    StageArgs A = stage_args(h, hs);
    const dim3 g(nblk(n, 256)), bl(256);
    Ptr6 kp;
    for (int j = 0; j < 6; j++) kp.p[j] = K[j];
    for (int s = 0; s < 6; s++) {
        const double2 *ys = y0;
        if (s > 0) {
            Coef6 cf{};
            for (int j = 0; j < s; j++) cf.c[j] = a[s][j];
            hipLaunchKernelGGL(k_comb, g, bl, 0, h->st, n, y0, hs, s, cf, kp, ytmp);
            ys = ytmp;
        }
        if ((rc = full_rhs(h, A, ys, c[s], K[s]))) return rc;
    }
    Coef6 cb{}, ce{};
    for (int j = 0; j < 6; j++) { cb.c[j] = b4[j]; ce.c[j] = b5[j] - b4[j]; }
    hipLaunchKernelGGL(k_rkf_final, g, bl, 0, h->st, n, y0, hs, cb, ce, kp, eps_abs, eps_rel, ynew, h->d_adpart);
    const int nparts = (int)nblk(n, 256);
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, h->st, nparts, h->d_adpart, h->d_adpart + nparts);
    DMD_HIP(hipGetLastError());
    DMD_HIP(hipMemcpyAsync(ratio, h->d_adpart + nparts, 8, hipMemcpyDeviceToHost, h->st));
    DMD_HIP(hipStreamSynchronize(h->st));
    return DMD_OK;
}

}  // namespace dmd

extern "C" int dmd_scatt_rkf45_apply(dmd_scatt *h, double *t, double t1, double *h_step, double hmin, double hmax,
                                     double eps_abs, double eps_rel, double *rho, dmd_ode_stats *stats)
{
    if (!h || !t || !h_step || !rho || !(hmax > 0) || hmin < 0 || hmin > hmax || !(*h_step > 0) ||
        !(eps_abs >= 0) || !(eps_rel >= 0) || (eps_abs == 0 && eps_rel == 0))
        return fail(DMD_ERR_ARG, "dmd_scatt_rkf45_apply: bad arguments");
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    int rc;
    if ((rc = ensure_adaptive(h))) return rc;
    if ((rc = tiles4_prepare(h))) return rc;
    const int64_t n = h->nk * (int64_t)h->nbd * h->nbd;
    if ((rc = to_device(h, h->d_ad, rho, (size_t)n * 16))) return rc;
    dmd_ode_stats st{};
    double tc = *t, hs = std::min(*h_step, hmax);
    constexpr double S = 0.9, Q = 4.0;
    while (tc < t1) {
        double htry = std::min(hs, t1 - tc);
        bool final_step = htry >= t1 - tc;
        double ratio = 0;
        for (;;) {
            // an error estimate that never passes would shrink h to nothing: stop instead
            if (!(htry > 0) || tc + htry == tc)
                return fail(DMD_ERR_STATE, "rkf45: the step size underflowed at t = " + std::to_string(tc));
            if ((rc = rkf45_attempt(h, tc, htry, eps_abs, eps_rel, &ratio))) return rc;
            st.rhs_evals += 6;
            if (!(ratio == ratio)) return fail(DMD_ERR_STATE, "rkf45: the error estimate is not finite");
            if (ratio > 1.1 && htry > hmin) {
                double f = S * std::pow(ratio, -1.0 / Q);
                htry = std::max(hmin, htry * std::max(0.2, f));
                final_step = false;
                st.rejected++;
                continue;
            }
            break;
        }
        // accept: y0 <- ynew
        DMD_HIP(hipMemcpyAsync(h->d_ad, h->d_ad + 2 * n, (size_t)n * 16, hipMemcpyDeviceToDevice, h->st));
        tc = final_step ? t1 : tc + htry;
        st.steps++;
        st.last_err_ratio = ratio;
        double hnext = htry;
        if (ratio < 0.5) hnext = htry * std::min(5.0, std::max(1.0, S * std::pow(std::max(ratio, 1e-300), -1.0 / (Q + 1))));
        // a step clamped to land on t1 does not shrink the suggestion
        hs = std::min(hmax, final_step ? std::max(hs, hnext) : hnext);
    }
    DMD_HIP(hipStreamSynchronize(h->st));
    if ((rc = from_device(h, rho, h->d_ad, (size_t)n * 16))) return rc;
    DMD_HIP(hipStreamSynchronize(h->st));
    *t = tc;
    *h_step = hs;
    st.h_next = hs;
    if (stats) *stats = st;
    return DMD_OK;
}

extern "C" int dmd_scatt_integrate_euler(dmd_scatt *h, double t0, double dt, int64_t n_steps, double *rho)
{
    if (!h || !rho || n_steps < 0) return fail(DMD_ERR_ARG, "dmd_scatt_integrate_euler: bad arguments");
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    int rc;
    if ((rc = ensure_adaptive(h))) return rc;
    if ((rc = tiles4_prepare(h))) return rc;
    const int64_t n = h->nk * (int64_t)h->nbd * h->nbd;
    double2 *y = h->d_ad, *kv = h->d_ad + 3 * n;
    if ((rc = to_device(h, y, rho, (size_t)n * 16))) return rc;
    for (int64_t it = 0; it < n_steps; it++) {
        const double t = t0 + (double)it * dt;
        if ((rc = set_clock(h, t, dt, 0))) return rc;
        StageArgs A = stage_args(h, dt);
        if ((rc = full_rhs(h, A, y, 0.0, kv))) return rc;
        hipLaunchKernelGGL(k_euler, dim3(nblk(n, 256)), dim3(256), 0, h->st, h->nk, h->nbd, dt, kv, y,
                           h->semiclassical);
        DMD_HIP(hipGetLastError());
    }
    DMD_HIP(hipStreamSynchronize(h->st));
    if ((rc = from_device(h, rho, y, (size_t)n * 16))) return rc;
    DMD_HIP(hipStreamSynchronize(h->st));
    return DMD_OK;
}

// dm_dynamics::compute (DensityMatrix_Dynamics.h:243-278): the whole
// right-hand side at t (coherent + scattering window), as the integrators use it
extern "C" int dmd_scatt_compute_rhs(dmd_scatt *h, double t, const double *rho, double *ddmdt)
{
    if (!h || !rho || !ddmdt) return fail(DMD_ERR_ARG, "dmd_scatt_compute_rhs: bad arguments");
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    int rc;
    if ((rc = ensure_adaptive(h))) return rc;
    if ((rc = tiles4_prepare(h))) return rc;
    const int64_t n = h->nk * (int64_t)h->nbd * h->nbd;
    if ((rc = to_device(h, h->d_ad, rho, (size_t)n * 16))) return rc;
    if ((rc = set_clock(h, t, 0.0, 0))) return rc;
    StageArgs A = stage_args(h, 0.0);
    if ((rc = full_rhs(h, A, h->d_ad, 0.0, h->d_ad + 3 * n))) return rc;
    if ((rc = from_device(h, ddmdt, h->d_ad + 3 * n, (size_t)n * 16))) return rc;
    DMD_HIP(hipStreamSynchronize(h->st));
    return DMD_OK;
}

extern "C" int dmd_scatt_set_rhs_hook(dmd_scatt *h, dmd_rhs_hook fn, void *ctx)
{
    if (!h) return fail(DMD_ERR_ARG, "null handle");
    h->rhs_hook = fn;
    h->rhs_ctx = fn ? ctx : nullptr;
    return DMD_OK;
}

// alg_semiclassical (DensityMatrix_Dynamics.h:249,277)
extern "C" int dmd_scatt_set_semiclassical(dmd_scatt *h, int on)
{
    if (!h) return fail(DMD_ERR_ARG, "null handle");
    h->semiclassical = on ? 1 : 0;
    return DMD_OK;
}

// alg_phenom_relax (phenom_relax, scattering_model/phenomenon_relax.h; added in
// compute, DensityMatrix_Dynamics.h:272-275): -(rho - diag f_eq)/tau on the
// density-matrix bands [b_start, b_end); tau <= 0 switches it off.
extern "C" int dmd_scatt_set_phenom_relax(dmd_scatt *h, double tau, int b_start, int b_end, const double *f_eq)
{
    if (!h) return fail(DMD_ERR_ARG, "null handle");
    if (!(tau > 0)) {
        h->ph_tau_inv = 0.0;
        return DMD_OK;
    }
    if (!f_eq || b_start < 0 || b_end > h->nbd || b_start >= b_end)
        return fail(DMD_ERR_ARG, "dmd_scatt_set_phenom_relax: bands must satisfy 0 <= b_start < b_end <= " +
                                     std::to_string(h->nbd) + " and f_eq must be given");
    int rc;
    const size_t n = (size_t)h->nk * h->nbd;
    if (!h->d_phfeq && (rc = dalloc_t(&h->d_phfeq, n, "phenomenological f_eq"))) return rc;
    if ((rc = to_device(h, h->d_phfeq, f_eq, n * 8))) return rc;
    DMD_HIP(hipStreamSynchronize(h->st));
    h->ph_tau_inv = 1.0 / tau;
    h->ph_b0 = b_start;
    h->ph_b1 = b_end;
    return DMD_OK;
}
