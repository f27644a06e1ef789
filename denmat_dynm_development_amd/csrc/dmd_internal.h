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
// nb = 4 panel kernel: rows per panel, pairs per chunk (one workgroup) at most
#define DMD_TILE4_R 32
#define DMD_TILE4_GROUP 1
#define DMD_TILE4_CMAX 2048

// nb = 4 blocked pair record (512 complex elements, P1 then P2): offset of
// P_mat[row][col] for U elements per 16-byte unit.  Lane l = 4 (row/4) +
// col/4 of a 16-lane group holds the 4 x 4 block; its element e = 4 (row%4) +
// col%4 sits in unit e / U.
__host__ __device__ inline int blk_off(int mat, int row, int col, int U)
{
    const int l = 4 * (row >> 2) + (col >> 2), e = 4 * (row & 3) + (col & 3);
    return mat * 256 + ((e / U) * 16 + l) * U + e % U;
}

// Address of element el of the pair in slot s: pair-contiguous, or the
// structure-of-arrays granules of 8 pairs the tile kernel streams.
__host__ __device__ __forceinline__ int64_t paddr(int64_t s, int el, int ep, int soa)
{
    if (soa == 2) return s * 512 + blk_off(el >> 8, (el >> 4) & 15, el & 15, 1);  // nb = 4 blocked records
    return soa ? (s >> 3) * (8 * (int64_t)ep) + (int64_t)el * 8 + (s & 7) : s * ep + el;
}

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
// Loads of what a launch streams exactly once (P records, their pair codes,
// batched visit records): the nontemporal hint keeps that stream from
// evicting the rho rows, accumulators and partner operands that are reused
// out of L2.  k_tiles_soa at 58 M pairs: 5.19 -> 4.72 ms per launch.
typedef double dmd_dv2 __attribute__((ext_vector_type(2)));
typedef unsigned int dmd_uv4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ double2 ld_stream(const double2 *p)
{
    const dmd_dv2 v = __builtin_nontemporal_load((const dmd_dv2 *)p);
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ uint4 ld_stream_u4(const void *p)
{
    const dmd_uv4 v = __builtin_nontemporal_load((const dmd_uv4 *)p);
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint32_t ld_stream(const uint32_t *p) { return __builtin_nontemporal_load(p); }

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

// nb = 2 tile helpers (shared by the forward and the adjoint tile kernels)
__device__ __forceinline__ double2 one_minus(const double2 *r, int q)
{
    return make_double2((q % 3 == 0 ? 1.0 : 0.0) - r[q].x, -r[q].y);  // nb = 2: diagonal q = 0, 3
}
// The packed Hermitian part H = X + X^+ of a 2 x 2 matrix X (the linearized
// operator's per-pair increment, evolve_eph_linear.cpp:47-53).
__device__ __forceinline__ void herm_of(const double2 *c, double hv[4])
{
    hv[0] = 2.0 * c[0].x;
    hv[1] = 2.0 * c[3].x;
    hv[2] = c[1].x + c[2].x;
    hv[3] = c[1].y - c[2].y;
}

// Rows [base_row, base_row + 64) of one wave: lane l holds the 4 doubles of
// row base_row + l (zeros past nk); the wave regroups them so every atomic
// instruction covers 16 consecutive rows' 32-byte records (4 full lines).
__device__ __forceinline__ void flush_herm_wave(const double hv[4], int lane, int64_t base_row, int64_t nk,
                                                double *__restrict__ hacc)
{
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int src = 16 * q + (lane >> 2), d = lane & 3;
        const double a0 = __shfl(hv[0], src), a1 = __shfl(hv[1], src), a2 = __shfl(hv[2], src),
                     a3 = __shfl(hv[3], src);
        const double v = d == 0 ? a0 : (d == 1 ? a1 : (d == 2 ? a2 : a3));
        const int64_t row = base_row + src;
        if (v != 0.0 && row < nk) unsafeAtomicAdd(hacc + row * 4 + d, v);
    }
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
    int soa = 0;                         // P layout: 0 pair-contiguous, 1 SoA granules of 8 pairs
                                         // (nb = 2 tiles), 2 blocked 4x4 per lane (nb = 4 tiles)
    // nb = 4 tiles and the reduced-precision forms (DMD_PREC_C64 / _BF16 store P
    // only in the blocked layout; d_P then holds float2 or packed bf16 re/im words)
    size_t p_esize = 16;                 // bytes per stored complex element of P
    float *d_pscale = nullptr;           // BF16: per-slot power-of-two scale
    double *d_fsplit = nullptr;          // nk x nb occupations f of the split rho = f + drho
    double *d_feq = nullptr;             // nk x 64: F1 | F2 row sums at rho = diag(f)
    float2 *d_x32 = nullptr;             // nk x 16: drho_S in FP32 for the reduced forms
    double *d_hacc16 = nullptr;          // nk x 16 packed Hermitian row accumulator (nb = 4)
    bool feq_dirty = true;
    // batched trajectories (BF16, nb = 4): visit records, segments, operands
    uint16_t *d_vrec = nullptr;          // visits x 2048 BF16: M1 hi | M2 hi | M1 lo | M2 lo (bhalf: x 1024 FP16 M1 | M2)
    bool bhalf = false;                  // BF16 store: FP16 records of M x bhalf_g, FP16 hi/lo operand
    double bhalf_g = 1.0;                // power-of-two factor that puts every record entry under 2^15
    uint32_t *d_bxmax = nullptr;         // Bp: bits of max |drho| of every trajectory (bhalf)
    uint32_t *d_bqueue = nullptr;        // Bp / 32: segment queue counters of the persistent pass
    int bq_resident[6] = { 0, 0, 0, 0, 0, 0 }; // workgroups resident at once, persistent pass instances
    int32_t *d_vq = nullptr;             // partner row of each visit
    uint32_t *d_vsd = nullptr;           // slot << 1 | reverse of each visit
    bool bdirect = false;                // batched pass forms M from the stored P (no visit records)
    int4 *d_seg = nullptr;               // segments: (row, first visit lo, hi, count)
    int64_t nseg = 0;
    bool bvis_ready = false, bvis_dirty = true;
    int b_cap = 0;                       // padded trajectory count the buffers hold
    int batch_nblk = 2;                  // 32-trajectory blocks per wavefront (autotuned)
    int seg_len = 0;                     // visits per segment (autotuned)
    int tuned_Bp = 0;                    // padded B the tuning was done for
    std::vector<int64_t> h_vptr;         // first visit of every row (host)
    std::vector<int32_t> h_vq;           // partner row of every visit (host; sorted within each row)
    int seg_pblock = 0;                  // partner-block size of the segment order (0: row order; autotuned)
    uint16_t *d_xb = nullptr;            // nk x 32 x Bp BF16 hi/lo drho operand
    float *d_Fb = nullptr;               // nk x 64 x Bp row sums
    double2 *d_rhoSB = nullptr;          // Bp x nk x 16 rho_S
    double2 *d_bin = nullptr, *d_bout = nullptr;  // B x nk x 16 staging
    int64_t bio_cap = 0;
    double2 *d_yB0 = nullptr, *d_yBa = nullptr, *d_yBb = nullptr, *d_accB = nullptr, *d_wB = nullptr;
    int64_t rkb_cap = 0;
    int last_B = 0;                      // trajectories of the last dmd_scatt_eval_batched
    // adjoint scratch
    int64_t ndiag = 0;                   // stored pairs with k = k' (no reverse direction)
    int64_t *d_dpos = nullptr;           // their P positions
    int32_t *d_dk = nullptr;             // and rows
    double2 *d_adj = nullptr;            // 3 x nk x nb^2
    double *d_dotpart = nullptr;         // dot-product partials
    // compressed P (dmd_scatt_compress_P, nb = 2 tiles): per-slot nonzero
    // masks, per-granule value offsets and the nonzero values; d_P is
    // released while they are set
    uint32_t *d_spmask = nullptr, *d_spgofs = nullptr;
    double2 *d_spv = nullptr;
    int64_t sp_values = 0;  // stored values
    // linearized operator (alg_linearize): the P store holds Lscij | Lscji
    bool linear = false;
    int semiclassical = 0;  // alg_semiclassical: the integrators keep rho diagonal
    // alg_phenom_relax: -(rho - diag f_eq) / tau on density-matrix bands [ph_b0, ph_b1)
    double ph_tau_inv = 0.0;
    int ph_b0 = 0, ph_b1 = 0;
    double *d_phfeq = nullptr;  // nk x nbd
    double *d_flin = nullptr;            // f_eq (nk x nb) the operator was linearized at
    // adaptive / Euler integrators: 9 states (y0, ytmp, ynew, k1..k6) and the error partials
    double2 *d_ad = nullptr;
    double *d_adpart = nullptr;
    int64_t ad_cap = 0;
    double2 *d_rhoS = nullptr, *d_rho1S = nullptr, *d_out = nullptr, *d_eq = nullptr;
    // alg_linearize_dPee: the Hermitian row term of the e-e P's change, added
    // to every evaluation times cis((e_i - e_j) t) (dmd_scatt_set_ee_dP)
    double2 *d_dPH = nullptr;
    bool has_dP = false;
    // host callback before every integrator right-hand side (dmd_scatt_set_rhs_hook)
    dmd_rhs_hook rhs_hook = nullptr;
    void *rhs_ctx = nullptr;
    std::vector<double> h_hook;  // host copy of the stage state handed to the hook
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


// refusal of the operations that need the dense P store once P is compressed
inline int sp_guard(const dmd_scatt *h, const char *who)
{
    if (h && h->d_spv)
        return fail(DMD_ERR_STATE, std::string(who) + ": P is stored compressed (dmd_scatt_compress_P); "
                                                     "set P up again on a new handle to change it");
    return DMD_OK;
}
int to_device(dmd_scatt *h, void *dst, const void *src, size_t bytes);
int from_device(dmd_scatt *h, void *dst, const void *src, size_t bytes);
int set_clock(dmd_scatt *h, double t0, double dt, long long step);
// rotate the window [b0, b0+nb) of an nbd x nbd state to the Schroedinger picture (h->d_rhoS)
int launch_rotate(dmd_scatt *h, const double2 *in, const double2 *in1, int nbd, int b0, double a, int semic = 0);
// contraction + Hermitian epilogue on h->d_rhoS (and h->d_rho1S when dm1) into out (nk x nb^2)
int launch_contract(dmd_scatt *h, bool dm1, double a, double2 *out);
// adds the alg_linearize_dPee row term at the clock time a into out (nk x nb^2)
int launch_dP_term(dmd_scatt *h, double a, double2 *out);
// nb = 2: out = prefac/2 x the packed Hermitian rows of h->d_acc_herm, rotated to
// the interaction picture at stage a (no ddmdt_eq); clears the accumulator
int launch_assemble_herm(dmd_scatt *h, double a, double prefac, double2 *out);
// DMD_KERNEL_ROWS, _SINGLE_READ or _TILES: the kernel launch_contract uses
int contract_kind(const dmd_scatt *h);
// nb = 4 tile path (dmd_tiles4.hip)
int tiles4_place(dmd_scatt *h, const double2 *P1, const double2 *P2, int64_t n, int64_t first, double scale, int add);
int place_P_dev(dmd_scatt *h, const double2 *a, const double2 *b, int64_t n, int64_t first, double scale, int add);
std::string ldbd_path(const char *dir, const std::string &name);
std::string ldbd_P_name(int which, const char *scatt, int impurity, int hole);
std::string ldbd_sP_name(int which, const char *scatt, int hole, const char *part);
int file_bytes(const std::string &path, int64_t *bytes);
int check_size(const std::string &path, int64_t expect);
int pread_all(const std::string &path, int64_t off, int64_t n, void *buf, int nthreads);
int tiles4_fetch(dmd_scatt *h, int64_t first, int64_t n, double2 *P1, double2 *P2);
int tiles4_generate(dmd_scatt *h, const dmd_synth_params &p);
int tiles4_relayout(dmd_scatt *h, int to_blocked);
int tiles4_prepare(dmd_scatt *h);  // F_eq of the reduced forms when P or f changed (not capturable)
int launch_tiles4(dmd_scatt *h, bool dm1, double a, double2 *out);
// raw row sums [Pdm | dm1P] of the single-read contraction into d_acc_rows with
// X in place of rho_S and Y in place of 1 - rho_S (Y null: 1 - X); P must be
// pair-contiguous (the single-read layout)
int launch_raw_rows(dmd_scatt *h, const double2 *X, const double2 *Y);
int launch_linear(dmd_scatt *h, double a, double2 *out);
// batched path (dmd_batched.hip)
int batched_prepare(dmd_scatt *h, int B);  // not capturable
int launch_batched(dmd_scatt *h, int B, const double2 *rho, double a, double2 *out);
}  // namespace dmd
