// ldbd_io.cpp -- readers of the ldbd_data files the initializer writes and the
// dynamics reads, and the restart files: the reference's formats, read with
// size checks and error codes in place of exit().
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include "dmd_internal.h"

namespace dmd {

std::string ldbd_path(const char *dir, const std::string &name)
{
    std::string d(dir ? dir : ".");
    if (!d.empty() && d.back() != '/') d += '/';
    return d + name;
}

// File name of a P file: ldbd_P<which>_<scatt>[_D<imp>][_hole].bin
// (ElectronPhonon.cpp:216-217, ElectronImpurity.cpp:7-9).
std::string ldbd_P_name(int which, const char *scatt, int impurity, int hole)
{
    std::string s = "ldbd_P" + std::to_string(which) + "_" + std::string(scatt && *scatt ? scatt : "lindblad");
    if (impurity > 0) s += "_D" + std::to_string(impurity);
    if (hole) s += "_hole";
    return s + ".bin";
}

// sP1_<scatt>[_hole]_{ns,s,i,j}.bin, the sparse P files (ElectronPhonon.cpp:240-242)
std::string ldbd_sP_name(int which, const char *scatt, int hole, const char *part)
{
    std::string s = "sP" + std::to_string(which) + "_" + std::string(scatt && *scatt ? scatt : "lindblad");
    if (hole) s += "_hole";
    return s + "_" + part + ".bin";
}

int file_bytes(const std::string &path, int64_t *bytes)
{
    struct stat st;
    if (stat(path.c_str(), &st) != 0) return fail(DMD_ERR_IO, "cannot open " + path + ": " + std::strerror(errno));
    *bytes = (int64_t)st.st_size;
    return DMD_OK;
}

// The reference's check_file_size (src_v4.5.7/comm/myio.cpp:126-132): the
// whole file must have exactly the expected size.
int check_size(const std::string &path, int64_t expect)
{
    int64_t got;
    int rc = file_bytes(path, &got);
    if (rc) return rc;
    if (got != expect)
        return fail(DMD_ERR_IO, path + ": file size is " + std::to_string(got) + " while expected size is " +
                                    std::to_string(expect));
    return DMD_OK;
}

// pread of [off, off+n) into buf, in pieces, by up to nthreads threads.
int pread_all(const std::string &path, int64_t off, int64_t n, void *buf, int nthreads)
{
    if (n <= 0) return DMD_OK;
    int fd = open(path.c_str(), O_RDONLY);
    if (fd < 0) return fail(DMD_ERR_IO, "cannot open " + path + ": " + std::strerror(errno));
    const int64_t min_piece = 8ll << 20;
    int nt = (int)std::max<int64_t>(1, std::min<int64_t>(nthreads, n / min_piece));
    std::vector<int> err(nt, 0);
    auto work = [&](int t) {
        int64_t a = n * t / nt, b = n * (t + 1) / nt;
        char *p = (char *)buf + a;
        while (a < b) {
            ssize_t r = pread(fd, p, (size_t)std::min<int64_t>(b - a, 1ll << 30), (off_t)(off + a));
            if (r < 0 && errno == EINTR) continue;
            if (r <= 0) { err[t] = r < 0 ? errno : EIO; return; }
            a += r;
            p += r;
        }
    };
    if (nt == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; t++) th.emplace_back(work, t);
        for (auto &x : th) x.join();
    }
    close(fd);
    for (int e : err)
        if (e) return fail(DMD_ERR_IO, "read of " + path + " failed: " + std::strerror(e));
    return DMD_OK;
}

static int read_exact(const std::string &path, int64_t expect, void *buf)
{
    int rc = check_size(path, expect);
    if (rc) return rc;
    return pread_all(path, 0, expect, buf, 1);
}

// Integers (or numbers) of a text line before any '#'.
static int scan_numbers(const char *line, double *v, int vmax)
{
    std::string s(line);
    size_t h = s.find('#');
    if (h != std::string::npos) s.resize(h);
    const char *p = s.c_str();
    int n = 0;
    while (n < vmax) {
        char *end;
        double x = std::strtod(p, &end);
        if (end == p) break;
        v[n++] = x;
        p = end;
    }
    return n;
}

}  // namespace dmd

using namespace dmd;

// ldbd_size.dat, line by line as electron::read_ldbd_size (electron.cpp:100-132),
// electronphonon::get_brange / get_nkpair (ElectronPhonon.cpp:132-176),
// phonon::read_jdftx (phonon.h:49-79) and parameters::read_jdftx
// (parameters.cpp:7-23) read it; the file's layouts are written by
// electron::saveSize (lindbladInit.../electron_save.cpp:16-50).
extern "C" int dmd_ldbd_read_info(const char *dir, dmd_ldbd_info *info)
{
    if (!info) return fail(DMD_ERR_ARG, "dmd_ldbd_read_info: null info");
    std::memset(info, 0, sizeof(*info));
    const std::string path = ldbd_path(dir, "ldbd_size.dat");
    FILE *fp = std::fopen(path.c_str(), "r");
    if (!fp) return fail(DMD_ERR_IO, "cannot open " + path + ": " + std::strerror(errno));
    std::vector<std::string> lines;
    char buf[1024];
    while (lines.size() < 10 && std::fgets(buf, sizeof buf, fp)) lines.emplace_back(buf);
    std::fclose(fp);
    if (lines.size() < 8) return fail(DMD_ERR_IO, path + ": expected at least 8 lines, found " + std::to_string(lines.size()));
    const std::string &l1 = lines[0];
    if (l1.find("only for conduction") != std::string::npos) info->kind = 0;
    else if (l1.find("only for valence") != std::string::npos) info->kind = 1;
    else info->kind = 2;
    double v[16];
    int n = scan_numbers(lines[1].c_str(), v, 16);
    info->n_line2 = n;
    // conduction-only and both: nb nv bBot_dm bTop_dm bBot_eph bTop_eph nb_wannier bBot_probe [band_skipped];
    // valence-only: nb nv bBot_dm bTop_dm 0 0 bBot_eph_hole bTop_eph_hole nb_wannier bBot_probe band_skipped
    const int need = info->kind == 1 ? 8 : 6;
    if (n < need) return fail(DMD_ERR_IO, path + ": line 2 has " + std::to_string(n) + " integers, expected at least " + std::to_string(need));
    info->nb = (int)v[0];
    info->nv = (int)v[1];
    info->b_bot_dm = (int)v[2];
    info->b_top_dm = (int)v[3];
    const int e0 = info->kind == 1 ? 6 : 4;
    info->b_bot_eph = (int)v[e0];
    info->b_top_eph = (int)v[e0 + 1];
    info->nb_wannier = n > e0 + 2 ? (int)v[e0 + 2] : 0;
    info->b_skip_wannier = n > e0 + 3 ? (int)v[e0 + 3] : 0;  // bskipped_wannier (electron.cpp:108-111)
    if (info->nb <= 0 || info->nv < 0 || info->nv > info->nb)
        return fail(DMD_ERR_IO, path + ": need 0 <= nv <= nb and nb > 0");
    if (info->b_bot_dm < 0 || info->b_bot_dm > info->b_top_dm || info->b_top_dm > info->nb)
        return fail(DMD_ERR_IO, path + ": need 0 <= bStart_dm <= bEnd_dm <= nb");
    if (info->b_bot_eph < info->b_bot_dm || info->b_bot_eph > info->b_top_eph || info->b_top_eph > info->b_top_dm)
        return fail(DMD_ERR_IO, path + ": need bStart_dm <= bStart_eph <= bEnd_eph <= bEnd_dm");
    n = scan_numbers(lines[2].c_str(), v, 16);
    if (n < 2) return fail(DMD_ERR_IO, path + ": line 3 must hold nk_full nk");
    info->nk_full = v[0];
    info->nk = (int64_t)v[1];
    for (int i = 0; i < 3; i++) info->kmesh[i] = n > 2 + i ? (int)v[2 + i] : 0;
    n = scan_numbers(lines[3].c_str(), v, 16);
    if (n < (info->kind == 1 ? 2 : 1)) return fail(DMD_ERR_IO, path + ": line 4 must hold the pair count");
    info->nkpair = (int64_t)(info->kind == 1 ? v[1] : v[0]);
    n = scan_numbers(lines[4].c_str(), v, 16);
    if (n < 2) return fail(DMD_ERR_IO, path + ": line 5 must hold modeStart modeStop");
    info->mode_start = (int)v[0];
    info->mode_stop = (int)v[1];
    info->n_modes = info->mode_stop - info->mode_start;
    if (n >= 4 && v[3] - v[2] > 0) info->n_modes -= (int)(v[3] - v[2]);
    n = scan_numbers(lines[5].c_str(), v, 16);
    if (n < 1) return fail(DMD_ERR_IO, path + ": line 6 must hold T");
    info->temperature = v[0];
    n = scan_numbers(lines[6].c_str(), v, 16);
    if (n < 3) return fail(DMD_ERR_IO, path + ": line 7 must hold muMin muMax mu");
    info->mu_min = v[0];
    info->mu_max = v[1];
    info->mu = v[2];
    n = scan_numbers(lines[7].c_str(), v, 16);
    if (n < 2) return fail(DMD_ERR_IO, path + ": line 8 must hold degauss ndegauss");
    info->degauss = v[0];
    info->ndegauss = v[1];
    if (lines.size() > 8 && scan_numbers(lines[8].c_str(), v, 16) >= 6) {
        info->e_bot_probe = v[0];
        info->e_top_probe = v[1];
        info->e_bot_dm = v[2];
        info->e_top_dm = v[3];
        info->e_bot_eph = v[4];
        info->e_top_eph = v[5];
    }
    if (lines.size() > 9 && scan_numbers(lines[9].c_str(), v, 16) >= 1) info->omega_max = v[0];
    if (info->nk <= 0 || info->nkpair < 0) return fail(DMD_ERR_IO, path + ": need nk > 0 and nkpair >= 0");
    return DMD_OK;
}

// ldbd_ek.bin: nk x nb float64 (electron.cpp:172-180, saveEk electron_save.cpp:52-65);
// scissor added to bands nv.. as electron.cpp:177-179 does.
extern "C" int dmd_ldbd_read_ek(const char *dir, const dmd_ldbd_info *info, double scissor, double *e)
{
    if (!info || !e) return fail(DMD_ERR_ARG, "dmd_ldbd_read_ek: bad arguments");
    const int64_t cnt = info->nk * info->nb;
    int rc = read_exact(ldbd_path(dir, "ldbd_ek.bin"), cnt * 8, e);
    if (rc) return rc;
    if (scissor != 0)
        for (int64_t k = 0; k < info->nk; k++)
            for (int b = info->nv; b < info->nb; b++) e[k * info->nb + b] += scissor;
    return DMD_OK;
}

// ldbd_kvec.bin: nk x 3 float64 (electron.cpp:134-147).
extern "C" int dmd_ldbd_read_kvec(const char *dir, const dmd_ldbd_info *info, double *kvec)
{
    if (!info || !kvec) return fail(DMD_ERR_ARG, "dmd_ldbd_read_kvec: bad arguments");
    return read_exact(ldbd_path(dir, "ldbd_kvec.bin"), info->nk * 3 * 8, kvec);
}

// ldbd_smat.bin: nk x 3 x nb_dm^2 complex128 (electron.cpp:284-296).
extern "C" int dmd_ldbd_read_smat(const char *dir, const dmd_ldbd_info *info, double *s)
{
    if (!info || !s) return fail(DMD_ERR_ARG, "dmd_ldbd_read_smat: bad arguments");
    const int64_t nbd = info->b_top_dm - info->b_bot_dm;
    return read_exact(ldbd_path(dir, "ldbd_smat.bin"), info->nk * 3 * nbd * nbd * 16, s);
}

// ldbd_lmat.bin: nk x 3 x nb_dm^2 complex128, the orbital angular momentum
// matrices of needL (electron.cpp:315-335).
extern "C" int dmd_ldbd_read_lmat(const char *dir, const dmd_ldbd_info *info, double *l)
{
    if (!info || !l) return fail(DMD_ERR_ARG, "dmd_ldbd_read_lmat: bad arguments");
    const int64_t nbd = info->b_top_dm - info->b_bot_dm;
    return read_exact(ldbd_path(dir, "ldbd_lmat.bin"), info->nk * 3 * nbd * nbd * 16, l);
}

// ldbd_Bso.bin: nk x 3 float64, the internal spin-orbit field of
// alg_read_Bso (electron.cpp:159-170).
extern "C" int dmd_ldbd_read_Bso(const char *dir, const dmd_ldbd_info *info, double *bso)
{
    if (!info || !bso) return fail(DMD_ERR_ARG, "dmd_ldbd_read_Bso: bad arguments");
    return read_exact(ldbd_path(dir, "ldbd_Bso.bin"), info->nk * 3 * 8, bso);
}

// ldbd_HEzmat.bin: nk x nb_dm^2 complex128, the electric-field Hamiltonian
// of scale_Ez (electron.cpp:44-61).
extern "C" int dmd_ldbd_read_HEzmat(const char *dir, const dmd_ldbd_info *info, double *hez)
{
    if (!info || !hez) return fail(DMD_ERR_ARG, "dmd_ldbd_read_HEzmat: bad arguments");
    const int64_t nbd = info->b_top_dm - info->b_bot_dm;
    return read_exact(ldbd_path(dir, "ldbd_HEzmat.bin"), info->nk * nbd * nbd * 16, hez);
}

// ldbd_Umat.bin (electron.cpp:388-410): the full nk x nb_wannier^2 matrices,
// of which the e-ph columns b_bot_eph + b_skip_wannier .. + nb_eph are kept,
// or nk x nb_wannier x nb_eph directly.
extern "C" int dmd_ldbd_read_Umat(const char *dir, const dmd_ldbd_info *info, double *U)
{
    if (!info || !U) return fail(DMD_ERR_ARG, "dmd_ldbd_read_Umat: bad arguments");
    const int64_t nw = info->nb_wannier, nbe = info->b_top_eph - info->b_bot_eph, nk = info->nk;
    if (nw <= 0) return fail(DMD_ERR_IO, "ldbd_size.dat gives no nb_wannier for ldbd_Umat.bin");
    const std::string path = ldbd_path(dir, "ldbd_Umat.bin");
    int64_t bytes = -1;
    int rc = file_bytes(path, &bytes);
    if (rc) return rc;
    if (bytes == nk * nw * nw * 16) {
        const int64_t c0 = info->b_bot_eph + info->b_skip_wannier;
        if (c0 < 0 || c0 + nbe > nw)
            return fail(DMD_ERR_IO, path + ": e-ph columns [" + std::to_string(c0) + ", " + std::to_string(c0 + nbe) +
                                        ") outside nb_wannier = " + std::to_string(nw));
        std::vector<double> full((size_t)(nk * nw * nw * 2));
        if ((rc = read_exact(path, bytes, full.data()))) return rc;
        for (int64_t k = 0; k < nk; k++)
            for (int64_t w = 0; w < nw; w++)
                for (int64_t b = 0; b < nbe; b++) {
                    const size_t src = (size_t)(((k * nw + w) * nw + c0 + b) * 2), dst = (size_t)(((k * nw + w) * nbe + b) * 2);
                    U[dst] = full[src];
                    U[dst + 1] = full[src + 1];
                }
        return DMD_OK;
    }
    return read_exact(path, nk * nw * nbe * 16, U);
}

// ldbd_imsig.bin: nk x nb_eph float64 (electron.cpp:181-198).
extern "C" int dmd_ldbd_read_imsig(const char *dir, const dmd_ldbd_info *info, double *imsig)
{
    if (!info || !imsig) return fail(DMD_ERR_ARG, "dmd_ldbd_read_imsig: bad arguments");
    const int64_t nbe = info->b_top_eph - info->b_bot_eph;
    return read_exact(ldbd_path(dir, "ldbd_imsig.bin"), info->nk * nbe * 8, imsig);
}

// ldbd_R.dat: dim, then the three lattice vectors as rows (lattice.h:48-56).
extern "C" int dmd_ldbd_read_R(const char *dir, int *dim, double *R)
{
    if (!dim || !R) return fail(DMD_ERR_ARG, "dmd_ldbd_read_R: bad arguments");
    const std::string path = ldbd_path(dir, "ldbd_R.dat");
    FILE *fp = std::fopen(path.c_str(), "r");
    if (!fp) return fail(DMD_ERR_IO, "cannot open " + path + ": " + std::strerror(errno));
    int ok = std::fscanf(fp, "%d", dim) == 1;
    for (int i = 0; ok && i < 9; i++) ok = std::fscanf(fp, "%lf", &R[i]) == 1;
    std::fclose(fp);
    if (!ok) return fail(DMD_ERR_IO, path + ": expected dim and 9 numbers");
    return DMD_OK;
}

// ldbd_kpair_k1st/k2nd[_hole].bin: nkpair size_t each (ElectronPhonon.cpp:182-204);
// stored pairs [first, first+count).
extern "C" int dmd_ldbd_read_kpairs(const char *dir, const dmd_ldbd_info *info, int64_t first, int64_t count,
                                    uint64_t *k1st, uint64_t *k2nd)
{
    if (!info || first < 0 || count < 0 || first + count > info->nkpair || (count && (!k1st || !k2nd)))
        return fail(DMD_ERR_ARG, "dmd_ldbd_read_kpairs: bad arguments");
    const char *sfx = info->kind == 1 ? "_hole" : "";
    const std::string f1 = ldbd_path(dir, std::string("ldbd_kpair_k1st") + sfx + ".bin");
    const std::string f2 = ldbd_path(dir, std::string("ldbd_kpair_k2nd") + sfx + ".bin");
    int rc;
    if ((rc = check_size(f1, info->nkpair * 8)) || (rc = check_size(f2, info->nkpair * 8))) return rc;
    if ((rc = pread_all(f1, first * 8, count * 8, k1st, hw_threads()))) return rc;
    if ((rc = pread_all(f2, first * 8, count * 8, k2nd, hw_threads()))) return rc;
    for (int64_t i = 0; i < count; i++)
        if (k1st[i] >= (uint64_t)info->nk || k2nd[i] >= (uint64_t)info->nk)
            return fail(DMD_ERR_IO, "stored pair " + std::to_string(first + i) + " names a k beyond nk");
    return DMD_OK;
}

// ldbd_P1/P2_<scatt>[_D<imp>][_hole].bin: nkpair x nb_eph^4 complex128, pairs
// [first, first+count) on the host (ElectronPhonon.cpp:209-233,
// ElectronImpurity.cpp:3-35).  No scaling.
extern "C" int dmd_ldbd_read_P(const char *dir, const dmd_ldbd_info *info, const char *scatt, int impurity,
                               int64_t first, int64_t count, double *P1, double *P2)
{
    if (!info || first < 0 || count < 0 || first + count > info->nkpair || impurity < 0 || (count && (!P1 || !P2)))
        return fail(DMD_ERR_ARG, "dmd_ldbd_read_P: bad arguments");
    const int64_t nbe = info->b_top_eph - info->b_bot_eph;
    const int64_t per = nbe * nbe * nbe * nbe * 16;
    const int hole = info->kind == 1;
    const std::string f1 = ldbd_path(dir, ldbd_P_name(1, scatt, impurity, hole));
    const std::string f2 = ldbd_path(dir, ldbd_P_name(2, scatt, impurity, hole));
    int rc;
    if ((rc = check_size(f1, info->nkpair * per)) || (rc = check_size(f2, info->nkpair * per))) return rc;
    if ((rc = pread_all(f1, first * per, count * per, P1, hw_threads()))) return rc;
    return pread_all(f2, first * per, count * per, P2, hw_threads());
}

// restart/denmat_restart.bin (nk x nb^2 complex128) and restart/time_restart.dat
// ("%14.7le"), as singdenmat_k::read_dm_restart / write_dm_tofile
// (DenMat.cpp:301-334) and parameters.cpp:299-310 use them.  dir is the
// restart folder; dm or t may be NULL to skip that file.
extern "C" int dmd_restart_read(const char *dir, int64_t nk, int nb, double *dm, double *t)
{
    if ((dm && (nk <= 0 || nb <= 0)) || (!dm && !t)) return fail(DMD_ERR_ARG, "dmd_restart_read: bad arguments");
    int rc = dm ? read_exact(ldbd_path(dir, "denmat_restart.bin"), nk * nb * nb * 16, dm) : DMD_OK;
    if (rc || !t) return rc;
    const std::string path = ldbd_path(dir, "time_restart.dat");
    FILE *fp = std::fopen(path.c_str(), "r");
    if (!fp) return fail(DMD_ERR_IO, "restart needs " + path);
    int ok = std::fscanf(fp, "%le", t) == 1;
    std::fclose(fp);
    return ok ? DMD_OK : fail(DMD_ERR_IO, path + ": no time value");
}

extern "C" int dmd_restart_write(const char *dir, int64_t nk, int nb, const double *dm, double t)
{
    if (nk <= 0 || nb <= 0 || !dm) return fail(DMD_ERR_ARG, "dmd_restart_write: bad arguments");
    const std::string pb = ldbd_path(dir, "denmat_restart.bin"), pt = ldbd_path(dir, "time_restart.dat");
    FILE *fb = std::fopen(pb.c_str(), "wb");
    if (!fb) return fail(DMD_ERR_IO, "cannot open " + pb + ": " + std::strerror(errno));
    const size_t cnt = (size_t)nk * nb * nb * 2;
    const bool okb = std::fwrite(dm, 8, cnt, fb) == cnt;
    std::fclose(fb);
    if (!okb) return fail(DMD_ERR_IO, "write of " + pb + " failed");
    FILE *ft = std::fopen(pt.c_str(), "w");
    if (!ft) return fail(DMD_ERR_IO, "cannot open " + pt + ": " + std::strerror(errno));
    std::fprintf(ft, "%14.7le", t);
    std::fclose(ft);
    return DMD_OK;
}
