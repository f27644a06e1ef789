// dmd_ldbd.hip -- ldbd_data straight into HBM: the P files streamed through
// pinned double buffers into the handle's layout, and a handle built from an
// ldbd_data folder the way the reference's electronphonon sets itself up.
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
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

// One sparse matrix set (sP1 or sP2): the four file names, the per-pair
// counts and their prefix sums.
struct SparseFiles {
    std::string f[4];  // _ns, _s, _i, _j
    std::vector<int32_t> ns;
    std::vector<int64_t> off;
};

// The counts of every stored pair and the sizes of the element files (the
// reference reads them without checks: sparse2D::get_nsk_fromfile and
// read_smat, comm/sparse2D.h:81-119).
int open_sparse(const char *dir, int which, const char *scatt, int hole, int64_t np, int nt, SparseFiles &S)
{
    static const char *part[4] = {"ns", "s", "i", "j"};
    for (int m = 0; m < 4; m++) S.f[m] = ldbd_path(dir, ldbd_sP_name(which, scatt, hole, part[m]));
    int rc;
    if ((rc = check_size(S.f[0], np * 4))) return rc;
    S.ns.resize((size_t)np);
    if ((rc = pread_all(S.f[0], 0, np * 4, S.ns.data(), nt))) return rc;
    S.off.assign((size_t)np + 1, 0);
    for (int64_t p = 0; p < np; p++) {
        // more than nb^4 is allowed: a repeated index keeps its last value
        if (S.ns[p] < 0)
            return fail(DMD_ERR_IO, S.f[0] + ": pair " + std::to_string(p) + " lists " + std::to_string(S.ns[p]) +
                                        " elements");
        S.off[p + 1] = S.off[p] + S.ns[p];
    }
    const int64_t tot = S.off[np];
    if ((rc = check_size(S.f[1], tot * 16))) return rc;
    if ((rc = check_size(S.f[2], tot * 4))) return rc;
    return check_size(S.f[3], tot * 4);
}

// |z| as the reference's complex::abs (comm/scalar.h:102-103)
inline double cabs_ref(double x, double y) { return std::sqrt(x * x + y * y); }

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
    if (const int sg = sp_guard(h, "dmd_scatt_load_P")) return sg;
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
    h->has_P = false;  // until every chunk is placed: a failure leaves no half-loaded P in use
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

// read_ldbd_eph's sparse branch (ElectronPhonon.cpp:235-246): the sparse
// files densified chunk by chunk on the host (value x scale, kept when
// |value x scale| > thr; a repeated index keeps its last value, as todense
// does) into pinned buffers and placed in the handle's layout while the next
// chunk is read.
extern "C" int dmd_scatt_load_sP(dmd_scatt *h, const char *dir, const dmd_ldbd_info *info, const char *scatt,
                                 double scale, double thr)
{
    if (!h || !info || !(thr >= 0)) return fail(DMD_ERR_ARG, "dmd_scatt_load_sP: bad arguments");
    if (const int sg = sp_guard(h, "dmd_scatt_load_sP")) return sg;
    const int nbe = info->b_top_eph - info->b_bot_eph;
    if (nbe != h->nb || info->nkpair != h->np)
        return fail(DMD_ERR_ARG, "dmd_scatt_load_sP: the files hold " + std::to_string(info->nkpair) + " pairs of " +
                                     std::to_string(nbe) + " bands, the handle " + std::to_string(h->np) +
                                     " pairs of " + std::to_string(h->nb));
    const int hole = info->kind == 1, nb2 = h->nb * h->nb, nb4 = h->nb4;
    const int nt = std::max(1, std::min(hw_threads(), 16));
    SparseFiles S[2];
    int rc;
    for (int j = 0; j < 2; j++)
        if ((rc = open_sparse(dir, j + 1, scatt, hole, h->np, nt, S[j]))) return rc;
    if (h->np == 0) {
        h->has_P = true;
        h->linear = false;
        return DMD_OK;
    }
    const int64_t per = (int64_t)nb4 * 16;
    const int64_t chunk = std::min<int64_t>(h->np, std::max<int64_t>(1, (64ll << 20) / per));
    Stream2 T(h->st);
    for (int s = 0; s < 2; s++) {
        for (int j = 0; j < 2; j++) {
            hipError_t e = hipHostMalloc((void **)&T.hb[s][j], (size_t)(chunk * per), hipHostMallocDefault);
            if (e != hipSuccess) return fail(DMD_ERR_NOMEM, std::string("pinned staging: ") + hipGetErrorString(e));
            if ((rc = dalloc_t(&T.db[s][j], (size_t)chunk * nb4, "P stream staging"))) return rc;
        }
        DMD_HIP(hipEventCreateWithFlags(&T.ev[s], hipEventDisableTiming));
    }
    std::vector<double2> sv;
    std::vector<int32_t> iv, jv;
    const int64_t nchunks = (h->np + chunk - 1) / chunk;
    h->has_P = false;  // until every chunk is placed: a failure leaves no half-loaded P in use
    for (int64_t c = 0; c < nchunks; c++) {
        const int s = (int)(c & 1);
        const int64_t first = c * chunk, n = std::min(chunk, h->np - first);
        if (c >= 2) DMD_HIP(hipEventSynchronize(T.ev[s]));
        for (int j = 0; j < 2; j++) {
            const SparseFiles &F = S[j];
            const int64_t a = F.off[first], m = F.off[first + n] - a;
            sv.resize((size_t)m);
            iv.resize((size_t)m);
            jv.resize((size_t)m);
            if ((rc = pread_all(F.f[1], a * 16, m * 16, sv.data(), nt))) return rc;
            if ((rc = pread_all(F.f[2], a * 4, m * 4, iv.data(), nt))) return rc;
            if ((rc = pread_all(F.f[3], a * 4, m * 4, jv.data(), nt))) return rc;
            double2 *dense = (double2 *)T.hb[s][j];
            std::vector<int64_t> bad(nt, -1);
            auto work = [&](int t) {
                for (int64_t p = first + n * t / nt; p < first + n * (t + 1) / nt; p++) {
                    double2 *d = dense + (p - first) * nb4;
                    std::memset((void *)d, 0, (size_t)per);
                    for (int64_t q = F.off[p] - a; q < F.off[p + 1] - a; q++) {
                        const int32_t r = iv[q], col = jv[q];
                        if (r < 0 || r >= nb2 || col < 0 || col >= nb2) {
                            bad[t] = p;
                            return;
                        }
                        const double x = sv[q].x * scale, y = sv[q].y * scale;
                        d[r * nb2 + col] = cabs_ref(x, y) > thr ? make_double2(x, y) : make_double2(0.0, 0.0);
                    }
                }
            };
            std::vector<std::thread> th;
            for (int t = 1; t < nt; t++) th.emplace_back(work, t);
            work(0);
            for (auto &x : th) x.join();
            for (int t = 0; t < nt; t++)
                if (bad[t] >= 0)
                    return fail(DMD_ERR_IO, F.f[2] + ", " + F.f[3] + ": pair " + std::to_string(bad[t]) +
                                                " has an index outside 0.." + std::to_string(nb2 - 1));
        }
        for (int j = 0; j < 2; j++)
            DMD_HIP(hipMemcpyAsync(T.db[s][j], T.hb[s][j], (size_t)(n * per), hipMemcpyHostToDevice, h->st));
        if ((rc = place_P_dev(h, T.db[s][0], T.db[s][1], n, first, 1.0, 0))) return rc;
        DMD_HIP(hipEventRecord(T.ev[s], h->st));
    }
    DMD_HIP(hipStreamSynchronize(h->st));
    h->has_P = true;
    h->linear = false;
    return DMD_OK;
}

// write_smat of set_sparseP (ElectronPhonon.cpp:126-128): P downloaded in
// chunks, the elements above thr written per pair in row-major order.
extern "C" int dmd_ldbd_write_sP(dmd_scatt *h, const char *dir, const dmd_ldbd_info *info, const char *scatt,
                                 double thr)
{
    {
        const int sg = sp_guard(h, "dmd_ldbd_write_sP");
        if (sg) return sg;
    }
    if (!h || !info || !(thr >= 0)) return fail(DMD_ERR_ARG, "dmd_ldbd_write_sP: bad arguments");
    if (!h->has_P) return fail(DMD_ERR_STATE, "P has not been uploaded or generated");
    if (h->linear) return fail(DMD_ERR_STATE, "the handle holds the linearized operator, not P");
    if (info->b_top_eph - info->b_bot_eph != h->nb || info->nkpair != h->np)
        return fail(DMD_ERR_ARG, "dmd_ldbd_write_sP: info does not describe the handle");
    const int hole = info->kind == 1, nb2 = h->nb * h->nb, nb4 = h->nb4;
    static const char *part[4] = {"ns", "s", "i", "j"};
    FILE *fp[2][4] = {{nullptr}};
    bool todo[2];
    std::string names[2][4];
    auto close_all = [&]() {
        bool ok = true;
        for (auto &row : fp)
            for (auto &f : row)
                if (f) {
                    ok = ok && !std::ferror(f);
                    ok = (std::fclose(f) == 0) && ok;
                    f = nullptr;
                }
        return ok;
    };
    for (int j = 0; j < 2; j++) {
        for (int m = 0; m < 4; m++) names[j][m] = ldbd_path(dir, ldbd_sP_name(j + 1, scatt, hole, part[m]));
        struct stat st;
        todo[j] = stat(names[j][1].c_str(), &st) != 0;  // write_smat: no overwriting (comm/sparse2D.h:137)
        for (int m = 0; m < 4 && todo[j]; m++)
            if (!(fp[j][m] = std::fopen(names[j][m].c_str(), "wb"))) {
                close_all();
                return fail(DMD_ERR_IO, "cannot create " + names[j][m]);
            }
    }
    if (!todo[0] && !todo[1]) return DMD_OK;
    const int64_t chunk = std::min<int64_t>(std::max<int64_t>(h->np, 1), std::max<int64_t>(1, (64ll << 20) / (nb4 * 16)));
    std::vector<double2> P[2] = {std::vector<double2>((size_t)chunk * nb4), std::vector<double2>((size_t)chunk * nb4)};
    std::vector<double2> sv((size_t)nb4);
    std::vector<int32_t> iv((size_t)nb4), jv((size_t)nb4);
    for (int64_t first = 0; first < h->np; first += chunk) {
        const int64_t n = std::min(chunk, h->np - first);
        int rc = dmd_scatt_download_P(h, first, n, (double *)P[0].data(), (double *)P[1].data());
        if (rc) {
            close_all();
            return rc;
        }
        for (int j = 0; j < 2; j++) {
            if (!todo[j]) continue;
            for (int64_t p = 0; p < n; p++) {
                const double2 *A = P[j].data() + p * nb4;
                int32_t ns = 0;
                for (int r = 0; r < nb2; r++)
                    for (int col = 0; col < nb2; col++) {
                        const double2 v = A[r * nb2 + col];
                        if (cabs_ref(v.x, v.y) > thr) {
                            sv[ns] = v;
                            iv[ns] = r;
                            jv[ns] = col;
                            ns++;
                        }
                    }
                std::fwrite(&ns, 4, 1, fp[j][0]);
                std::fwrite(sv.data(), 16, (size_t)ns, fp[j][1]);
                std::fwrite(iv.data(), 4, (size_t)ns, fp[j][2]);
                std::fwrite(jv.data(), 4, (size_t)ns, fp[j][3]);
            }
        }
    }
    if (!close_all()) return fail(DMD_ERR_IO, "writing the sparse P files in " + std::string(dir ? dir : "."));
    return DMD_OK;
}

// The reference's set-up of one electronphonon from ldbd_data
// (ElectronPhonon.h:53-93: read_ldbd_size, get_brange, get_nkpair,
// read_ldbd_kpair, read_ldbd_eph; e of the e-ph window cut from e_dm,
// ElectronPhonon.h:79): reads ldbd_size.dat, ldbd_ek.bin (+ scissor on the
// conduction bands, electron.cpp:177-179), the stored pairs and P (scaled by
// scale_eph, ElectronPhonon.cpp:225,232), and sets the band window of
// evolve_driver when the e-ph window is narrower than the density matrix's.
// P source of create_from: the dense files, the sparse files, none (P unset),
// or no scattering (no stored pairs, empty P)
enum { SRC_DENSE = 0, SRC_SPARSE = 1, SRC_NONE = 2, SRC_NO_PAIRS = 3 };

static int create_from(const char *dir, const char *scatt, double scale_eph, double scissor, int precision,
                       int source, dmd_scatt **out)
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
    const int64_t np = source == SRC_NO_PAIRS ? 0 : info.nkpair;
    std::vector<uint64_t> k1(np), k2(np);
    if (np && (rc = dmd_ldbd_read_kpairs(dir, &info, 0, np, k1.data(), k2.data()))) return rc;
    dmd_scatt *h = nullptr;
    if ((rc = dmd_scatt_create(nk, nbe, np, k1.data(), k2.data(), eeph.data(), info.nk_full, precision, &h)))
        return rc;
    if (nbd != nbe && (rc = dmd_scatt_set_window(h, nbd, b0, edm.data()))) {
        dmd_scatt_destroy(h);
        return rc;
    }
    if (source == SRC_SPARSE) rc = dmd_scatt_load_sP(h, dir, &info, scatt, scale_eph, 1e-40);
    else if (source == SRC_DENSE) rc = dmd_scatt_load_P(h, dir, &info, scatt, 0, scale_eph, 0);
    else if (source == SRC_NO_PAIRS) h->has_P = true;  // nothing to scatter
    if (rc) {
        dmd_scatt_destroy(h);
        return rc;
    }
    *out = h;
    return DMD_OK;
}

extern "C" int dmd_scatt_create_from_ldbd(const char *dir, const char *scatt, double scale_eph, double scissor,
                                          int precision, dmd_scatt **out)
{
    return create_from(dir, scatt, scale_eph, scissor, precision, SRC_DENSE, out);
}

// alg_eph_enable = 0 (P for the e-i fold-in to fill) or alg_scatt_enable = 0
// (no stored pairs)
extern "C" int dmd_scatt_create_from_ldbd_noP(const char *dir, double scissor, int precision, int no_pairs,
                                              dmd_scatt **out)
{
    return create_from(dir, "lindblad", 0.0, scissor, precision, no_pairs ? SRC_NO_PAIRS : SRC_NONE, out);
}

// alg_Pin_is_sparse: the sparse files' P, scaled and re-sparsified at 1e-40
// (ElectronPhonon.cpp:245-246)
extern "C" int dmd_scatt_create_from_ldbd_sparse(const char *dir, const char *scatt, double scale_eph, double scissor,
                                                 int precision, dmd_scatt **out)
{
    return create_from(dir, scatt, scale_eph, scissor, precision, SRC_SPARSE, out);
}
