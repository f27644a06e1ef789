// dmd_model.h -- pieces shared by the device-built scattering models (the
// ionized-impurity model in dmd_eimp.hip, the e-e model in dmd_ee.hip): the
// model constants, the static screened Coulomb potential, the k-mesh map.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <vector>

#include "dmd_internal.h"

namespace dmd {

struct EimpConst {
    dmd_eimp_model m;
    double prefac_delta, prefac_exp;
};

// |wrap(ka - kb)|^2 under GGT (comm/kmap.h:6-15, matrix3.h:119-122)
__host__ __device__ inline double model_q2(const double *GGT, const double *ka, const double *kb)
{
    double v[3];
    for (int d = 0; d < 3; d++) {
        v[d] = ka[d] - kb[d];
        v[d] -= floor(0.5 + v[d]);
        if (fabs(v[d] - 0.5) < 1e-6) v[d] = -0.5;
    }
    const double *G = GGT;
    return v[0] * v[0] * G[0] + v[1] * v[1] * G[4] + v[2] * v[2] * G[8] +
           2 * (v[0] * v[1] * G[1] + v[0] * v[2] * G[2] + v[1] * v[2] * G[5]);
}

// the static v(q) of the closed-form screenings (coulomb_model::vq,
// Coulomb_Model.h:632-662)
__device__ inline double eimp_vq(const dmd_eimp_model &m, const double *ka, const double *kb)
{
    const double q2 = model_q2(m.GGT, ka, kb);
    if (m.screening == DMD_SCR_DEBYE) return m.prefac_vq / (m.qscr2_debye + q2);
    if (q2 < 1e-20) return 0.0;  // Gamma is skipped by the unscreened and Bechstedt forms
    if (m.screening == DMD_SCR_BECHSTEDT) {
        const double epsq = 1 + 1. / (m.fac0 + m.fac2 * q2 + m.fac4 * q2 * q2);
        return m.prefac_vq_bare / epsq / q2;
    }
    return m.prefac_vq / q2;
}

// kIndexMap::ikvec3 / qIndexMap::iqvec3 (comm/kmap.cpp): the mesh cell of a
// k (or q) in crystal coordinates, wrapped into [0, 1) first
inline void mesh_cell(const double *k, const int *km, int *v)
{
    for (int d = 0; d < 3; d++) {
        const double x = k[d] - std::floor(k[d]);
        v[d] = (int)std::round(km[d] * x);
        if (v[d] == km[d]) v[d] = 0;
    }
}

static __global__ void k_fill_int(int64_t n, int *__restrict__ a, int v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
// mesh cell -> smallest k index in it (INT_MAX if none): kIndexMap keeps the first k
static __global__ void k_kgrid(int64_t nk, const int *__restrict__ ik3, int m1, int m2, int *__restrict__ kgrid)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nk) atomicMin(kgrid + ((int64_t)ik3[3 * k] * m1 + ik3[3 * k + 1]) * m2 + ik3[3 * k + 2], (int)k);
}

// v(q) of every stored pair for the k-summed screenings (static RPA, lindhard),
// q_scr^2 summed on the device for each distinct q of the stored pairs
// (dmd_eimp.hip).  dU: the device copy of U.
int rpa_vq(dmd_scatt *h, const uint64_t *k1st, const uint64_t *k2nd, const double2 *dU, const double *kvec,
           const double *f, const dmd_eimp_model &m, std::vector<double> &vq);
// v(q) of every mesh cell that is a difference of two k points (static RPA):
// vq_cell[cell] = prefac_vq / (|q|^2 + q_scr^2(q)) with q the first (ik, jk)
// difference in that cell, 0 at q = 0 and for cells no difference reaches.
int rpa_vq_cells(dmd_scatt *h, const double2 *dU, const double *kvec, const double *f, const dmd_eimp_model &m,
                 std::vector<double> &vq_cell);

}  // namespace dmd
