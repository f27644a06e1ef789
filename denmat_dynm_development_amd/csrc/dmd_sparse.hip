// dmd_sparse.hip -- P stored by its nonzero elements, the streaming side of the
// reference's sparse modes (alg_sparseP / alg_Pin_is_sparse: evolve runs
// sparse_zgemm over per-pair element lists, src_v4.5.7/evolve_eph.cpp:69-86,
// 194-203, comm/sparse_matrix.cpp:6-40, so a zero element costs nothing).
//
// Layout (nb = 2, the tile kernel's slot order and 8-slot granules): per slot
// a 32-bit mask of its nonzero elements (bit el: P1 element el, bit 16 + el:
// P2 element el); per granule the offset of its values, which follow
// element-major: for el = 0 .. 31 the values of the granule's slots that have
// element el, in slot order.  A pair with n nonzeros streams 16 n + 4.5 bytes
// against 512 dense.  The 8 lanes of a granule read one element together, so
// one load instruction of a wavefront covers 8 runs of consecutive values (the
// dense granules' line pattern); k_tiles_soa<.., SP = true> finds each lane's
// place with one ballot per element, rebuilds the pair's dense P in registers
// and runs the dense GEMVs.
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

// one thread per granule: its values, element-major, from its offset
__global__ void k_sp_pack(int64_t ngran, const double2 *__restrict__ P, const uint32_t *__restrict__ mask,
                          const uint32_t *__restrict__ gofs, double2 *__restrict__ V)
{
    const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= ngran) return;
    const double2 *g = P + gi * (8 * 32);
    double2 *o = V + gofs[gi];
    uint32_t m[8];
    for (int t = 0; t < 8; t++) m[t] = mask[gi * 8 + t];
    int64_t pos = 0;
    for (int el = 0; el < 32; el++)
        for (int t = 0; t < 8; t++)
            if ((m[t] >> el) & 1u) o[pos++] = g[el * 8 + t];
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
    const int64_t ns = h->nslots, ng = ns / 8;
    if (ns % 8) return fail(DMD_ERR_STATE, "tile slots are not a whole number of granules");
    uint32_t *d_mask = nullptr, *d_gofs = nullptr;
    double2 *d_v = nullptr;
    int rc;
    if ((rc = dalloc_t(&d_mask, (size_t)std::max<int64_t>(ns, 1), "sparse masks")) ||
        (rc = dalloc_t(&d_gofs, (size_t)std::max<int64_t>(ng, 1), "sparse granule offsets"))) {
        dfree(d_mask); dfree(d_gofs);
        return rc;
    }
    std::vector<uint32_t> mask((size_t)ns), gofs((size_t)ng);
    if (ns > 0) hipLaunchKernelGGL(k_sp_mask, dim3(nblk(ns, 256)), dim3(256), 0, h->st, ns, h->d_P, d_mask);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && ns > 0) e = hipMemcpyAsync(mask.data(), d_mask, (size_t)ns * 4, hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    uint64_t tot = 0;
    for (int64_t g = 0; g < ng; g++) {
        gofs[g] = (uint32_t)tot;
        for (int t = 0; t < 8; t++) tot += (uint64_t)__builtin_popcount(mask[g * 8 + t]);
        if (tot >= (1ull << 32)) break;
    }
    if (e == hipSuccess && tot >= (1ull << 32)) {
        dfree(d_mask); dfree(d_gofs);
        return fail(DMD_ERR_UNSUPPORTED, "dmd_scatt_compress_P: 2^32 or more nonzero values");
    }
    // above 3/4 of the elements nonzero the compressed store saves little
    // memory and streams slower than the dense one: keep the dense store
    if (e == hipSuccess && 4 * tot > 3 * (uint64_t)h->np * 2 * (uint64_t)h->nb4) {
        dfree(d_mask); dfree(d_gofs);
        if (stored_values) *stored_values = (int64_t)tot;
        return DMD_OK;
    }
    if (e == hipSuccess && ng > 0) e = hipMemcpyAsync(d_gofs, gofs.data(), (size_t)ng * 4, hipMemcpyHostToDevice, h->st);
    if (e == hipSuccess && (rc = dalloc_t(&d_v, (size_t)std::max<uint64_t>(tot, 1), "sparse P values"))) {
        (void)hipStreamSynchronize(h->st);
        dfree(d_mask); dfree(d_gofs);
        return rc;
    }
    if (e == hipSuccess && ng > 0) {
        hipLaunchKernelGGL(k_sp_pack, dim3(nblk(ng, 256)), dim3(256), 0, h->st, ng, h->d_P, d_mask, d_gofs, d_v);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    if (e != hipSuccess) {
        dfree(d_mask); dfree(d_gofs); dfree(d_v);
        return fail(DMD_ERR_HIP, hipGetErrorString(e));
    }
    dfree(h->d_P);
    h->d_P = nullptr;
    h->d_spmask = d_mask;
    h->d_spgofs = d_gofs;
    h->d_spv = d_v;
    h->sp_values = (int64_t)tot;
    if (stored_values) *stored_values = h->sp_values;
    return DMD_OK;
}
