// dmd_sparse.hip -- P stored by its nonzero elements, the streaming side of the
// reference's sparse modes (alg_sparseP / alg_Pin_is_sparse: evolve runs
// sparse_zgemm over per-pair element lists, src_v4.5.7/evolve_eph.cpp:69-86,
// 194-203, comm/sparse_matrix.cpp:6-40, so a zero element costs nothing).
//
// Layout (nb = 2, the tile kernel's slot order): per slot a 32-bit mask of its
// nonzero elements (bit el: P1 element el, bit 16 + el: P2 element el) and a
// 32-bit offset; the slot's nonzero values follow one another in element
// order from that offset.  A pair with n nonzeros streams 16 n + 8 bytes
// against 512 dense: about half at half the density.  The lanes of a
// wavefront take consecutive slots, whose value runs are adjacent, so what a
// launch reads is one contiguous stretch per wavefront.  k_tiles_soa<.., SP =
// true> rebuilds each slot's 32 elements in registers and runs the dense
// arithmetic, so the result is the dense kernel's on the same (zero-filled) P.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "dmd_internal.h"

namespace dmd {

// per slot: the nonzero mask of its 32 elements (SoA granule layout, EP = 32)
__global__ void k_sp_mask(int64_t nslots, const double2 *__restrict__ P, uint32_t *__restrict__ mask)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const double2 *g = P + (s >> 3) * (8 * 32) + (s & 7);
    uint32_t m = 0;
#pragma unroll
    for (int el = 0; el < 32; el++) {
        const double2 v = g[el * 8];
        if (v.x != 0.0 || v.y != 0.0) m |= 1u << el;
    }
    mask[s] = m;
}

// each slot's nonzero values, in element order, from its offset
__global__ void k_sp_pack(int64_t nslots, const double2 *__restrict__ P, const uint32_t *__restrict__ mask,
                          const uint32_t *__restrict__ ofs, double2 *__restrict__ V)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const double2 *g = P + (s >> 3) * (8 * 32) + (s & 7);
    double2 *o = V + ofs[s];
    const uint32_t m = mask[s];
    int pos = 0;
    for (int el = 0; el < 32; el++)
        if ((m >> el) & 1u) o[pos++] = g[el * 8];
}

}  // namespace dmd

using namespace dmd;

extern "C" int dmd_scatt_compress_P(dmd_scatt *h, int64_t *stored_values)
{
    if (!h) return fail(DMD_ERR_ARG, "null handle");
    if (h->d_spv) {
        if (stored_values) *stored_values = h->sp_values;
        return DMD_OK;
    }
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    if (h->linear) return fail(DMD_ERR_STATE, "dmd_scatt_compress_P: the handle holds the linearized operator");
    if (h->nb != 2 || h->precision != DMD_PREC_FP64 || h->soa != 1 || !h->tile_R)
        return fail(DMD_ERR_UNSUPPORTED, "dmd_scatt_compress_P: built for the nb = 2 FP64 tile layout (kernel AUTO/TILES)");
    const int64_t ns = h->nslots;
    if (ns % 8) return fail(DMD_ERR_STATE, "tile slots are not a whole number of granules");
    uint32_t *d_mask = nullptr, *d_ofs = nullptr;
    double2 *d_v = nullptr;
    int rc;
    if ((rc = dalloc_t(&d_mask, (size_t)ns, "sparse masks")) || (rc = dalloc_t(&d_ofs, (size_t)ns, "sparse offsets"))) {
        dfree(d_mask); dfree(d_ofs);
        return rc;
    }
    std::vector<uint32_t> mask((size_t)ns), ofs((size_t)ns);
    if (ns > 0) hipLaunchKernelGGL(k_sp_mask, dim3(nblk(ns, 256)), dim3(256), 0, h->st, ns, h->d_P, d_mask);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && ns > 0) e = hipMemcpyAsync(mask.data(), d_mask, (size_t)ns * 4, hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    uint64_t tot = 0;
    for (int64_t q = 0; q < ns; q++) {
        ofs[q] = (uint32_t)tot;
        tot += (uint64_t)__builtin_popcount(mask[q]);
    }
    if (e == hipSuccess && tot >= (1ull << 32)) {
        dfree(d_mask); dfree(d_ofs);
        return fail(DMD_ERR_UNSUPPORTED, "dmd_scatt_compress_P: 2^32 or more nonzero values");
    }
    if (e == hipSuccess && ns > 0) e = hipMemcpyAsync(d_ofs, ofs.data(), (size_t)ns * 4, hipMemcpyHostToDevice, h->st);
    if (e == hipSuccess && (rc = dalloc_t(&d_v, (size_t)std::max<uint64_t>(tot, 1), "sparse P values"))) {
        (void)hipStreamSynchronize(h->st);
        dfree(d_mask); dfree(d_ofs);
        return rc;
    }
    if (e == hipSuccess && ns > 0) {
        hipLaunchKernelGGL(k_sp_pack, dim3(nblk(ns, 256)), dim3(256), 0, h->st, ns, h->d_P, d_mask, d_ofs, d_v);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    if (e != hipSuccess) {
        dfree(d_mask); dfree(d_ofs); dfree(d_v);
        return fail(DMD_ERR_HIP, hipGetErrorString(e));
    }
    dfree(h->d_P);
    h->d_P = nullptr;
    h->d_spmask = d_mask;
    h->d_spofs = d_ofs;
    h->d_spv = d_v;
    h->sp_values = (int64_t)tot;
    if (stored_values) *stored_values = h->sp_values;
    return DMD_OK;
}
