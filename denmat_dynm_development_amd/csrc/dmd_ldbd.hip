// dmd_ldbd.hip -- ldbd_data straight into HBM: the P files streamed through
// pinned double buffers into the handle's layout, and a handle built from an
// ldbd_data folder the way the reference's electronphonon sets itself up.
#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "dmd_internal.h"

using namespace dmd;

namespace {

// Pinned host and device staging of the P stream; freed on every exit path.
struct Stream2 {
    double *hb[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    double2 *db[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t st;
    explicit Stream2(hipStream_t s) : st(s) {}
    ~Stream2()
    {
        (void)hipStreamSynchronize(st);  // no copy or placement still reads the staging
        for (int s = 0; s < 2; s++) {
            for (int j = 0; j < 2; j++) {
                if (hb[s][j]) (void)hipHostFree(hb[s][j]);
                if (db[s][j]) dfree(db[s][j]);
            }
            if (ev[s]) (void)hipEventDestroy(ev[s]);
        }
    }
};

}  // namespace

// read_ldbd_eph (ElectronPhonon.cpp:209-233) and, with impurity = iD+1, the
// e-i P of read_ldbd_imp_P (ElectronImpurity.cpp:3-35) folded in by
// add_scatt_contrib (ElectronPhonon.cpp:87-97): P (add = 0) or P += (add = 1)
// scale * the file's P1, P2.  The files are read in chunks of about 64 MiB
// per array by parallel preads into pinned buffers; chunk c+1 is read while
// chunk c is copied and placed, so host memory stays at 256 MiB whatever the
// file size.
extern "C" int dmd_scatt_load_P(dmd_scatt *h, const char *dir, const dmd_ldbd_info *info, const char *scatt,
                                int impurity, double scale, int add)
{
    if (!h || !info || impurity < 0) return fail(DMD_ERR_ARG, "dmd_scatt_load_P: bad arguments");
    const int nbe = info->b_top_eph - info->b_bot_eph;
    if (nbe != h->nb || info->nkpair != h->np)
        return fail(DMD_ERR_ARG, "dmd_scatt_load_P: the files hold " + std::to_string(info->nkpair) + " pairs of " +
                                     std::to_string(nbe) + " bands, the handle " + std::to_string(h->np) +
                                     " pairs of " + std::to_string(h->nb));
    if (h->linear && add) return fail(DMD_ERR_STATE, "the handle holds the linearized operator: load P first");
    const int hole = info->kind == 1;
    const std::string f[2] = {ldbd_path(dir, ldbd_P_name(1, scatt, impurity, hole)),
                              ldbd_path(dir, ldbd_P_name(2, scatt, impurity, hole))};
    const int64_t per = (int64_t)h->nb4 * 16;
    int rc;
    for (int j = 0; j < 2; j++)
        if ((rc = check_size(f[j], h->np * per))) return rc;
    if (h->np == 0) {
        h->has_P = true;
        return DMD_OK;
    }
    const int64_t chunk = std::min<int64_t>(h->np, std::max<int64_t>(1, (64ll << 20) / per));
    const int nt = std::max(1, std::min(hw_threads(), 16));
    Stream2 S(h->st);
    for (int s = 0; s < 2; s++) {
        for (int j = 0; j < 2; j++) {
            hipError_t e = hipHostMalloc((void **)&S.hb[s][j], (size_t)(chunk * per), hipHostMallocDefault);
            if (e != hipSuccess) return fail(DMD_ERR_NOMEM, std::string("pinned staging: ") + hipGetErrorString(e));
            if ((rc = dalloc_t(&S.db[s][j], (size_t)chunk * h->nb4, "P stream staging"))) return rc;
        }
        DMD_HIP(hipEventCreateWithFlags(&S.ev[s], hipEventDisableTiming));
    }
    const int64_t nchunks = (h->np + chunk - 1) / chunk;
    for (int64_t c = 0; c < nchunks; c++) {
        const int s = (int)(c & 1);
        const int64_t first = c * chunk, n = std::min(chunk, h->np - first);
        // the pinned buffers of slot s are free once chunk c-2 has been placed
        if (c >= 2) DMD_HIP(hipEventSynchronize(S.ev[s]));
        for (int j = 0; j < 2; j++)
            if ((rc = pread_all(f[j], first * per, n * per, S.hb[s][j], nt))) return rc;
        for (int j = 0; j < 2; j++)
            DMD_HIP(hipMemcpyAsync(S.db[s][j], S.hb[s][j], (size_t)(n * per), hipMemcpyHostToDevice, h->st));
        if ((rc = place_P_dev(h, S.db[s][0], S.db[s][1], n, first, scale, add))) return rc;
        DMD_HIP(hipEventRecord(S.ev[s], h->st));
    }
    DMD_HIP(hipStreamSynchronize(h->st));
    h->has_P = true;
    h->linear = false;
    return DMD_OK;
}

// The reference's set-up of one electronphonon from ldbd_data
// (ElectronPhonon.h:53-93: read_ldbd_size, get_brange, get_nkpair,
// read_ldbd_kpair, read_ldbd_eph; e of the e-ph window cut from e_dm,
// ElectronPhonon.h:79): reads ldbd_size.dat, ldbd_ek.bin (+ scissor on the
// conduction bands, electron.cpp:177-179), the stored pairs and P (scaled by
// scale_eph, ElectronPhonon.cpp:225,232), and sets the band window of
// evolve_driver when the e-ph window is narrower than the density matrix's.
extern "C" int dmd_scatt_create_from_ldbd(const char *dir, const char *scatt, double scale_eph, double scissor,
                                          int precision, dmd_scatt **out)
{
    if (!out) return fail(DMD_ERR_ARG, "dmd_scatt_create_from_ldbd: null out");
    *out = nullptr;
    dmd_ldbd_info info;
    int rc;
    if ((rc = dmd_ldbd_read_info(dir, &info))) return rc;
    const int64_t nk = info.nk;
    const int nbd = info.b_top_dm - info.b_bot_dm, nbe = info.b_top_eph - info.b_bot_eph;
    const int b0 = info.b_bot_eph - info.b_bot_dm;
    if (nbe <= 0) return fail(DMD_ERR_IO, "ldbd_size.dat: empty e-ph band window");
    std::vector<double> e((size_t)nk * info.nb), edm((size_t)nk * nbd), eeph((size_t)nk * nbe);
    if ((rc = dmd_ldbd_read_ek(dir, &info, scissor, e.data()))) return rc;
    for (int64_t k = 0; k < nk; k++) {
        for (int b = 0; b < nbd; b++) edm[k * nbd + b] = e[k * info.nb + info.b_bot_dm + b];
        for (int b = 0; b < nbe; b++) eeph[k * nbe + b] = edm[k * nbd + b0 + b];
    }
    std::vector<uint64_t> k1(info.nkpair), k2(info.nkpair);
    if ((rc = dmd_ldbd_read_kpairs(dir, &info, 0, info.nkpair, k1.data(), k2.data()))) return rc;
    dmd_scatt *h = nullptr;
    if ((rc = dmd_scatt_create(nk, nbe, info.nkpair, k1.data(), k2.data(), eeph.data(), info.nk_full, precision, &h)))
        return rc;
    if (nbd != nbe && (rc = dmd_scatt_set_window(h, nbd, b0, edm.data()))) {
        dmd_scatt_destroy(h);
        return rc;
    }
    if ((rc = dmd_scatt_load_P(h, dir, &info, scatt, 0, scale_eph, 0))) {
        dmd_scatt_destroy(h);
        return rc;
    }
    *out = h;
    return DMD_OK;
}
