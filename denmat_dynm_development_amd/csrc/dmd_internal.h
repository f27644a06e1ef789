// dmd_internal.h -- shared declarations of libdmd_scatt (host + device).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "dmd_scatt.h"

namespace dmd {

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define DMD_HIP(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return ::dmd::fail(DMD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// True when p points to device memory (hipMalloc'd); host otherwise.
bool is_device_ptr(const void *p);

// Host-side synthetic pair list (sorted by (k1,k2), duplicate free).
int synth_kpairs_vec(const dmd_synth_params &p, std::vector<uint64_t> &k1, std::vector<uint64_t> &k2);
int hw_threads();

inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// rows per tile of the tile kernel (a multiple of 256)
#define DMD_TILE_R 512
// chunks of one row block a tile-kernel workgroup takes in sequence
#define DMD_TILE_GROUP 1

// Device clock: t(a) = (t0 + step*dt) + a*dt, the same expression as the RK4
// stage times of the oracle's driver.
struct DClock {
    double t0, dt;
    long long step;
    long long pad;
};
__device__ __forceinline__ double clock_t_at(const DClock *c, double a)
{
    return (c->t0 + (double)c->step * c->dt) + a * c->dt;
}

// ---------------------------------------------------------------------------
// complex128 helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.y * b.x + a.x * b.y);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b)  // a * conj(b)
{
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// acc += a * b and acc += a * conj(b) as four FMAs each (the sum is
// accumulated term by term rather than product first)
__device__ __forceinline__ void cfma(double2 &acc, double2 a, double2 b)
{
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.y, b.x, acc.y);
    acc.y = fma(a.x, b.y, acc.y);
}
__device__ __forceinline__ void cfmac(double2 &acc, double2 a, double2 b)
{
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(a.y, b.x, acc.y);
    acc.y = fma(-a.x, b.y, acc.y);
}
__device__ __forceinline__ double2 ccis(double x)
{
    double s, c;
    sincos(x, &s, &c);
    return make_double2(c, s);
}


}  // namespace dmd

using dmd::DClock;

struct dmd_scatt {
    int dev = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    int64_t nk = 0, np = 0, nrev = 0;
    int nb = 0, nb2 = 0, nb4 = 0;
    double prefac = 0;
    int precision = DMD_PREC_FP64;
    int expt = 1, schrodinger = 0, kernel = DMD_KERNEL_AUTO;
    bool identity_order = true;
    bool has_P = false, has_eq = false;
    double last_ms = 0;
    // window / driver pieces
    int nbd = 0, b0 = 0;
    // device arrays
    double *d_e = nullptr, *d_edm = nullptr;
    double2 *d_P = nullptr;
    int64_t *d_inv = nullptr, *d_rowptr = nullptr, *d_revptr = nullptr, *d_revpos = nullptr;
    int32_t *d_col = nullptr, *d_revcol = nullptr, *d_k1 = nullptr, *d_k2 = nullptr;
    int64_t *d_fwdpos = nullptr;         // forward list entry -> physical position of P
    uint32_t *d_pk = nullptr;            // per physical pair: (k mod R) + (k' mod R) * R
    int64_t *d_cbeg = nullptr;           // tile chunks: first pair, length, tile (I, J)
    int32_t *d_clen = nullptr;
    int2 *d_cij = nullptr;
    int32_t *d_gbeg = nullptr;           // workgroup -> first chunk of its group
    int64_t nchunks = 0, ngroups = 0;
    int64_t nslots = 0;                  // physical pair slots (stored pairs + tile padding)
    int tile_R = 0;
    int soa = 0;                         // P stored as SoA granules of 8 pairs (tile kernel)
    double2 *d_rhoS = nullptr, *d_rho1S = nullptr, *d_out = nullptr, *d_eq = nullptr;
    double2 *d_in = nullptr, *d_in1 = nullptr, *d_outw = nullptr;
    DClock *d_clk = nullptr;
    // driver: coherent H, observable, RK4 buffers (allocated on first use)
    double2 *d_H = nullptr, *d_s = nullptr;
    double *d_f = nullptr, *d_obs = nullptr, *d_part = nullptr;
    double *d_acc_rows = nullptr;  // nk x [Pdm | dm1P] accumulator of the single-read kernel
    double *d_acc_herm = nullptr;  // nk x [H00, H11, Re H01, Im H01] accumulator of the tile kernel
    int obs_i0 = 0, obs_i1 = 0;
    double obs_prefac = 0;
    bool has_H = false, has_obs = false;
    double2 *d_y0 = nullptr, *d_ya = nullptr, *d_yb = nullptr, *d_acc = nullptr, *d_w = nullptr;
    int64_t rk_cap = 0;  // states the RK buffers hold
    DClock h_clk{};
    std::vector<void *> extra;  // other device allocations owned by the handle
    ~dmd_scatt();
};


namespace dmd {
void dfree(void *p);
template <class T>
inline int dalloc_t(T **p, size_t count, const char *what)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(DMD_ERR_NOMEM, std::string("hipMalloc failed for ") + what + " (" +
                                       std::to_string(count * sizeof(T)) + " bytes): " + hipGetErrorString(e));
    }
    return DMD_OK;
}


int to_device(dmd_scatt *h, void *dst, const void *src, size_t bytes);
int from_device(dmd_scatt *h, void *dst, const void *src, size_t bytes);
int set_clock(dmd_scatt *h, double t0, double dt, long long step);
// rotate the window [b0, b0+nb) of an nbd x nbd state to the Schroedinger picture (h->d_rhoS)
int launch_rotate(dmd_scatt *h, const double2 *in, const double2 *in1, int nbd, int b0, double a);
// contraction + Hermitian epilogue on h->d_rhoS (and h->d_rho1S when dm1) into out (nk x nb^2)
int launch_contract(dmd_scatt *h, bool dm1, double a, double2 *out);
// DMD_KERNEL_ROWS or DMD_KERNEL_SINGLE_READ: the kernel launch_contract uses
int contract_kind(const dmd_scatt *h);
}  // namespace dmd
