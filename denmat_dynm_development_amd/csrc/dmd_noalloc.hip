// dmd_noalloc.hip -- dmd_scatt_eval without any device allocation.
//
// dmd_scatt_eval works on per-handle device scratch: the Schroedinger-picture
// state rho_S, the row accumulators the contraction kernels add into, the
// FP32 drho operand of the reduced forms and the device clock.  The handle
// allocates them at create time or on first use.  dmd_scatt_eval_noalloc runs
// the same kernels with that scratch taken from one file-scope device array
// instead, so a caller that must not allocate (code under stream capture,
// tight serving loops, harnesses that forbid hipMalloc) can evaluate the
// operator with nothing but its own input and output buffers.
//
// The array holds at most DMD_NOALLOC_SCRATCH_BYTES (16 MiB) per GPU for the
// life of the process.  Every call on one GPU shares it: calls that overlap
// on one GPU (two host threads, or two streams) race.  Its address is taken
// once per device with hipGetSymbolAddress; host code never passes the
// array's own name to a kernel.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "dmd_internal.h"

namespace dmd {

static __device__ __align__(256) unsigned char scratch[DMD_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on device dev, looked up once per device
// with that device current.
int scratch_on(int dev, unsigned char **out)
{
    if (dev < 0 || dev >= kMaxDevices) return fail(DMD_ERR_ARG, "dmd_scatt_eval_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        DMD_HIP(hipGetDevice(&cur));
        if (cur != dev) DMD_HIP(hipSetDevice(dev));
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(DMD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[dev];
    return DMD_OK;
}

__global__ void k_zero_words(int64_t n, double *__restrict__ p)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void k_put_clock(DClock *clk, double t0)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        clk->t0 = t0;
        clk->dt = 0.0;
        clk->step = 0;
        clk->pad = 0;
    }
}

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

// The handle scratch one evaluation touches, redirected into the array for
// the duration of a call and put back afterwards.
struct Redirect {
    dmd_scatt *h;
    double2 *rhoS;
    double *acc_herm, *acc_rows, *hacc16;
    float2 *x32;
    DClock *clk;
    explicit Redirect(dmd_scatt *hh)
        : h(hh), rhoS(hh->d_rhoS), acc_herm(hh->d_acc_herm), acc_rows(hh->d_acc_rows), hacc16(hh->d_hacc16),
          x32(hh->d_x32), clk(hh->d_clk)
    {
    }
    ~Redirect()
    {
        h->d_rhoS = rhoS;
        h->d_acc_herm = acc_herm;
        h->d_acc_rows = acc_rows;
        h->d_hacc16 = hacc16;
        h->d_x32 = x32;
        h->d_clk = clk;
    }
};

}  // namespace
}  // namespace dmd

using namespace dmd;

extern "C" int dmd_scatt_eval_noalloc(dmd_scatt *h, double t, const double *rho, double *ddmdt)
{
    if (!h) return fail(DMD_ERR_ARG, "null handle");
    if (!rho || !ddmdt) return fail(DMD_ERR_ARG, "dmd_scatt_eval_noalloc: rho/ddmdt null");
    if (!is_device_ptr(rho) || !is_device_ptr(ddmdt))
        return fail(DMD_ERR_ARG, "dmd_scatt_eval_noalloc: rho and ddmdt must be device memory (no staging is allocated)");
    if ((const void *)rho == (const void *)ddmdt)
        return fail(DMD_ERR_ARG, "dmd_scatt_eval_noalloc: ddmdt must not alias rho");
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    if (h->precision != DMD_PREC_FP64 && h->feq_dirty)
        return fail(DMD_ERR_STATE,
                    "dmd_scatt_eval_noalloc: the reduced form's equilibrium sums are out of date (P or the split "
                    "occupations changed); one dmd_scatt_eval rebuilds them");
    const int kind = contract_kind(h);
    const size_t nk = (size_t)h->nk, nb2 = (size_t)h->nb2;
    // scratch layout: [clock | rho_S | accumulator | FP32 drho]
    const size_t clk_b = 256, rho_b = up256(nk * nb2 * sizeof(double2));
    size_t acc_words = 0, x32_b = 0;
    if (h->linear && kind != DMD_KERNEL_TILES) acc_words = 0;                // deterministic row kernel
    else if (kind == DMD_KERNEL_TILES && h->nb == 4) {
        acc_words = nk * 16;                                                  // packed Hermitian rows
        if (h->precision != DMD_PREC_FP64 && !h->linear) x32_b = up256(nk * 16 * sizeof(float2));
    } else if (kind == DMD_KERNEL_TILES) acc_words = nk * nb2;                // Hermitian rows, nb = 2
    else if (kind == DMD_KERNEL_SINGLE_READ) acc_words = nk * 4 * nb2;        // Pdm | dm1P rows
    const size_t acc_b = up256(acc_words * sizeof(double));
    const size_t need = clk_b + rho_b + acc_b + x32_b;
    if (need > (size_t)DMD_NOALLOC_SCRATCH_BYTES)
        return fail(DMD_ERR_ARG, "dmd_scatt_eval_noalloc: this operator needs " + std::to_string(need) +
                                     " bytes of scratch (nk = " + std::to_string(nk) + ", nb = " +
                                     std::to_string(h->nb) + "), more than the " +
                                     std::to_string((size_t)DMD_NOALLOC_SCRATCH_BYTES) +
                                     " bytes the allocation-free entry point holds; use dmd_scatt_eval");
    unsigned char *base = nullptr;
    int rc;
    if ((rc = scratch_on(h->dev, &base))) return rc;
    Redirect keep(h);
    h->d_clk = (DClock *)base;
    h->d_rhoS = (double2 *)(base + clk_b);
    double *acc = (double *)(base + clk_b + rho_b);
    h->d_acc_herm = acc;
    h->d_acc_rows = acc;
    h->d_hacc16 = acc;
    if (x32_b) h->d_x32 = (float2 *)(base + clk_b + rho_b + acc_b);
    // the kernels add into the accumulator, which the previous call on this
    // GPU may have left in any state: clear it (a fill kernel, no memset node)
    if (acc_words)
        hipLaunchKernelGGL(k_zero_words, dim3(nblk((int64_t)acc_words, 256)), dim3(256), 0, h->st,
                           (int64_t)acc_words, acc);
    hipLaunchKernelGGL(k_put_clock, dim3(1), dim3(64), 0, h->st, h->d_clk, t);
    DMD_HIP(hipGetLastError());
    if ((rc = launch_rotate(h, (const double2 *)rho, nullptr, h->nb, 0, 0.0))) return rc;
    if ((rc = launch_contract(h, false, 0.0, (double2 *)ddmdt))) return rc;
    DMD_HIP(hipStreamSynchronize(h->st));
    return DMD_OK;
}
