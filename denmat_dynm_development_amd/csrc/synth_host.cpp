// synth_host.cpp -- host side of the synthetic ldbd_data generator and the
// ldbd_data writer (reference file layouts).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "dmd_internal.h"
#include "synth.h"

namespace dmd {

int hw_threads()
{
    unsigned n = std::thread::hardware_concurrency();
    if (n == 0) n = 1;
    return (int)std::min(n, 64u);
}

template <class F>
static void parallel_for(int64_t n, int nthreads, F f)
{
    if (nthreads <= 1 || n < 2) { f(0, n); return; }
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) {
        int64_t a = n * t / nthreads, b = n * (t + 1) / nthreads;
        th.emplace_back([=, &f]() { f(a, b); });
    }
    for (auto &x : th) x.join();
}

// Partner count of row k: fixed, or power law pdf ~ c^-alpha on [cmin, cmax].
static int64_t partner_count(const dmd_synth_params &p, int64_t k)
{
    int64_t c;
    if (p.partners_max <= 0) {
        c = p.partners;
    } else {
        double u = uni(key4(p.seed, ST_COUNT, (uint64_t)k, 0), 0);
        double a = p.powerlaw_alpha, lo = (double)p.partners_min, hi = (double)p.partners_max;
        double x;
        if (std::fabs(a - 1.0) < 1e-12) x = lo * std::exp(u * std::log(hi / lo));
        else x = std::pow(std::pow(lo, 1 - a) + u * (std::pow(hi, 1 - a) - std::pow(lo, 1 - a)), 1.0 / (1 - a));
        c = (int64_t)std::floor(x);
    }
    return std::max<int64_t>(0, std::min<int64_t>(c, p.nk - 1));
}
static inline int64_t partner_of(const dmd_synth_params &p, uint64_t key, int64_t k, int64_t j)
{
    int64_t m = (int64_t)(uni(key, (uint64_t)j) * (double)(p.nk - 1));
    if (m >= p.nk - 1) m = p.nk - 2;
    return m >= k ? m + 1 : m;
}

// Each k draws its partner count of partners k' != k uniformly; the pair is
// stored as (min, max) and duplicates are removed.  Output sorted by (k1, k2).
int synth_kpairs_vec(const dmd_synth_params &p, std::vector<uint64_t> &k1, std::vector<uint64_t> &k2)
{
    const int64_t nk = p.nk;
    if (nk < 2) { k1.clear(); k2.clear(); return DMD_OK; }
    const int nt = hw_threads();
    std::vector<std::vector<int64_t>> hist(nt, std::vector<int64_t>(nk, 0));
    parallel_for(nt, nt, [&](int64_t t0, int64_t t1) {
        for (int64_t t = t0; t < t1; t++) {
            int64_t a = nk * t / nt, b = nk * (t + 1) / nt;
            for (int64_t k = a; k < b; k++) {
                uint64_t key = key4(p.seed, ST_PAIRS, (uint64_t)k, 0);
                int64_t c = partner_count(p, k);
                for (int64_t j = 0; j < c; j++) {
                    int64_t q = partner_of(p, key, k, j);
                    hist[t][std::min(k, q)]++;
                }
            }
        }
    });
    std::vector<int64_t> off(nk + 1, 0);
    std::vector<std::vector<int64_t>> cur(nt, std::vector<int64_t>(nk));
    for (int64_t m = 0; m < nk; m++) {
        int64_t s = off[m];
        for (int t = 0; t < nt; t++) { cur[t][m] = s; s += hist[t][m]; }
        off[m + 1] = s;
    }
    hist.clear();
    std::vector<uint32_t> buf((size_t)off[nk]);
    parallel_for(nt, nt, [&](int64_t t0, int64_t t1) {
        for (int64_t t = t0; t < t1; t++) {
            int64_t a = nk * t / nt, b = nk * (t + 1) / nt;
            for (int64_t k = a; k < b; k++) {
                uint64_t key = key4(p.seed, ST_PAIRS, (uint64_t)k, 0);
                int64_t c = partner_count(p, k);
                for (int64_t j = 0; j < c; j++) {
                    int64_t q = partner_of(p, key, k, j);
                    int64_t m = std::min(k, q);
                    buf[cur[t][m]++] = (uint32_t)std::max(k, q);
                }
            }
        }
    });
    cur.clear();
    std::vector<int64_t> cnt(nk, 0);
    parallel_for(nk, nt, [&](int64_t a, int64_t b) {
        for (int64_t m = a; m < b; m++) {
            auto first = buf.begin() + off[m], last = buf.begin() + off[m + 1];
            std::sort(first, last);
            cnt[m] = std::unique(first, last) - first;
        }
    });
    int64_t tot = 0;
    std::vector<int64_t> o2(nk + 1, 0);
    for (int64_t m = 0; m < nk; m++) { o2[m] = tot; tot += cnt[m]; }
    o2[nk] = tot;
    k1.resize(tot);
    k2.resize(tot);
    parallel_for(nk, nt, [&](int64_t a, int64_t b) {
        for (int64_t m = a; m < b; m++)
            for (int64_t i = 0; i < cnt[m]; i++) {
                k1[o2[m] + i] = (uint64_t)m;
                k2[o2[m] + i] = buf[off[m] + i];
            }
    });
    return DMD_OK;
}

}  // namespace dmd

using namespace dmd;

static bool valid_params(const dmd_synth_params *p)
{
    return p && p->nb >= 1 && p->nb <= DMD_SYNTH_MAX_NB && p->nk >= 1 && p->n_modes >= 0 && p->degauss > 0 &&
           p->temperature > 0 && p->nk < (int64_t)1 << 31;
}

extern "C" int dmd_synth_preset(int which, dmd_synth_params *p)
{
    if (!p) return fail(DMD_ERR_ARG, "dmd_synth_preset: null params");
    const double eV = 1.0 / 27.21138505, meV = eV * 1e-3;
    const double Kelvin = 1.3806488e-23 / 4.35974434e-18;
    std::memset(p, 0, sizeof(*p));
    p->n_modes = 6; p->e_max = 0.1 * eV; p->split_max = 5 * meV; p->gap_min = 20 * meV; p->gap_max = 60 * meV;
    p->omega_min = 5 * meV; p->omega_max = 50 * meV; p->sigma_g = 1e-3; p->degauss = 5 * meV;
    p->temperature = 300 * Kelvin; p->powerlaw_alpha = 1.0;
    // impurity density 1e18 cm^-3 in bohr^-3 times a GaAs-like fcc cell
    // (a/2 = 5.34145 bohr, as the reference example's ldbd_R.dat)
    const double bohr2cm = 5.291772109038e-9, half_a = 5.34145;
    const double vcell = 2.0 * half_a * half_a * half_a;
    switch (which) {
    case 0: p->seed = 7; p->nb = 2; p->nk = 64; p->partners = 8; break;
    case 1: p->seed = 1; p->nb = 2; p->nk = 4096; p->partners = 250; break;
    case 2: p->seed = 2; p->nb = 2; p->nk = 65536; p->partners = 900; break;
    case 3:
        p->seed = 3; p->nb = 4; p->nk = 20000; p->partners = 500;
        p->sigma_imp = 5e-4; p->imp_fraction = 1e18 * bohr2cm * bohr2cm * bohr2cm * vcell;
        break;
    case 4:
        p->seed = 4; p->nb = 4; p->nk = 32768; p->partners_min = 10; p->partners_max = 5000;
        p->powerlaw_alpha = 1.15;
        break;
    case 5:  // six branches plus an e-i term: the seven parameters of the adjoint workload
        p->seed = 5; p->nb = 2; p->nk = 16384; p->partners = 488;
        p->sigma_imp = 5e-4; p->imp_fraction = 1e18 * bohr2cm * bohr2cm * bohr2cm * vcell;
        break;
    default: return fail(DMD_ERR_ARG, "dmd_synth_preset: unknown preset " + std::to_string(which));
    }
    return DMD_OK;
}

extern "C" int dmd_synth_energies(const dmd_synth_params *p, double *e)
{
    if (!valid_params(p) || !e) return fail(DMD_ERR_ARG, "dmd_synth_energies: bad arguments");
    for (int64_t k = 0; k < p->nk; k++) synth_energies_k(*p, k, e + k * p->nb);
    return DMD_OK;
}

extern "C" int dmd_synth_kpairs(const dmd_synth_params *p, uint64_t *k1st, uint64_t *k2nd, int64_t *n_pairs)
{
    if (!valid_params(p) || !n_pairs) return fail(DMD_ERR_ARG, "dmd_synth_kpairs: bad arguments");
    std::vector<uint64_t> a, b;
    int rc = synth_kpairs_vec(*p, a, b);
    if (rc) return rc;
    if (k1st && k2nd) {
        if (*n_pairs < (int64_t)a.size()) return fail(DMD_ERR_ARG, "dmd_synth_kpairs: output too small");
        std::memcpy(k1st, a.data(), a.size() * 8);
        std::memcpy(k2nd, b.data(), b.size() * 8);
    }
    *n_pairs = (int64_t)a.size();
    return DMD_OK;
}

extern "C" int dmd_synth_P(const dmd_synth_params *p, const double *e, int64_t n, const uint64_t *k1,
                           const uint64_t *k2, double *P1, double *P2, int nthreads)
{
    if (!valid_params(p) || !e || n < 0 || (n > 0 && (!k1 || !k2 || !P1 || !P2)))
        return fail(DMD_ERR_ARG, "dmd_synth_P: bad arguments");
    const int nb = p->nb;
    const size_t nb4 = (size_t)nb * nb * nb * nb;
    if (nthreads <= 0) nthreads = hw_threads();
    parallel_for(n, nthreads, [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b; i++)
            synth_pair_P(*p, e + k1[i] * nb, e + k2[i] * nb, (int64_t)k1[i], (int64_t)k2[i], P1 + 2 * nb4 * i,
                         P2 + 2 * nb4 * i);
    });
    return DMD_OK;
}

extern "C" int dmd_synth_couplings(const dmd_synth_params *p, uint64_t k, uint64_t kp, double *omega, double *gm,
                                   double *gp)
{
    if (!valid_params(p) || !omega || !gm || !gp) return fail(DMD_ERR_ARG, "dmd_synth_couplings: bad arguments");
    const int nb = p->nb;
    for (int a = 0; a < p->n_modes; a++) {
        omega[a] = synth_omega(*p, a);
        for (int b1 = 0; b1 < nb; b1++)
            for (int b2 = 0; b2 < nb; b2++) {
                size_t o = 2 * ((size_t)(a * nb + b1) * nb + b2);
                synth_g(*p, (int64_t)k, (int64_t)kp, a, b1, b2, gm + o, gm + o + 1);
                double r, i;
                synth_g(*p, (int64_t)kp, (int64_t)k, a, b2, b1, &r, &i);
                gp[o] = r; gp[o + 1] = -i;
            }
    }
    return DMD_OK;
}

// Random Hermitian matrices, one per k and direction (ldbd_smat.bin layout:
// nk x 3 x nb^2 complex).  Off-diagonals CN(0, 1/4), diagonals +-1/2 spread.
extern "C" int dmd_synth_spin(const dmd_synth_params *p, double *s3)
{
    if (!valid_params(p) || !s3) return fail(DMD_ERR_ARG, "dmd_synth_spin: bad arguments");
    const int nb = p->nb;
    for (int64_t k = 0; k < p->nk; k++)
        for (int c = 0; c < 3; c++) {
            uint64_t key = key4(p->seed, ST_SPIN, (uint64_t)k, (uint64_t)c);
            double *m = s3 + 2 * ((size_t)(k * 3 + c) * nb * nb);
            for (int i = 0; i < nb; i++) {
                double r, im;
                cnormal(key, (uint64_t)(i * nb + i), 0.5, &r, &im);
                m[2 * (i * nb + i)] = r; m[2 * (i * nb + i) + 1] = 0.0;
                for (int j = i + 1; j < nb; j++) {
                    cnormal(key, (uint64_t)(i * nb + j), 0.25, &r, &im);
                    m[2 * (i * nb + j)] = r; m[2 * (i * nb + j) + 1] = im;
                    m[2 * (j * nb + i)] = r; m[2 * (j * nb + i) + 1] = -im;
                }
            }
        }
    return DMD_OK;
}

static int write_bin(const std::string &path, const void *data, size_t bytes)
{
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) return fail(DMD_ERR_IO, "cannot open " + path);
    size_t w = bytes ? std::fwrite(data, 1, bytes, fp) : 0;
    std::fclose(fp);
    if (w != bytes) return fail(DMD_ERR_IO, "short write " + path);
    return DMD_OK;
}

// ldbd_data writer: layouts of the initializer's saveSize / saveEk / P and
// k-pair writers (lindbladInit_for-DMD-4.5.6/electron_save.cpp:16-65,
// ElectronPhonon.h:252-284), conduction-electron-only form.
extern "C" int dmd_ldbd_write_synth(const char *dir, const dmd_synth_params *p, int nthreads)
{
    if (!dir || !valid_params(p)) return fail(DMD_ERR_ARG, "dmd_ldbd_write_synth: bad arguments");
    const int nb = p->nb;
    const int64_t nk = p->nk;
    std::string d(dir);
    if (!d.empty() && d.back() != '/') d += '/';
    std::vector<double> e((size_t)nk * nb), kv((size_t)nk * 3), s3((size_t)nk * 3 * nb * nb * 2);
    dmd_synth_energies(p, e.data());
    for (int64_t k = 0; k < nk; k++)
        for (int c = 0; c < 3; c++) kv[k * 3 + c] = uni(key4(p->seed, ST_KVEC, (uint64_t)k, 0), (uint64_t)c);
    dmd_synth_spin(p, s3.data());
    std::vector<uint64_t> k1, k2;
    int rc = synth_kpairs_vec(*p, k1, k2);
    if (rc) return rc;
    const size_t np = k1.size(), nb4 = (size_t)nb * nb * nb * nb;
    std::vector<double> P1(np * nb4 * 2), P2(np * nb4 * 2);
    rc = dmd_synth_P(p, e.data(), (int64_t)np, k1.data(), k2.data(), P1.data(), P2.data(), nthreads);
    if (rc) return rc;
    double emin = 1e300, emax = -1e300, wmax = 0;
    for (double x : e) { emin = std::min(emin, x); emax = std::max(emax, x); }
    for (int a = 0; a < p->n_modes; a++) wmax = std::max(wmax, synth_omega(*p, a));
    FILE *fp = std::fopen((d + "ldbd_size.dat").c_str(), "w");
    if (!fp) return fail(DMD_ERR_IO, "cannot open " + d + "ldbd_size.dat");
    std::fprintf(fp, "There are scatterings only for conduction electrons\n");
    std::fprintf(fp, "%d %d %d %d %d %d %d %d %d # nb nv bBot_dm bTop_dm bBot_eph_elec bTop_eph_elec nb_wannier bBot_probe band_dft_skipped\n",
                 nb, 0, 0, nb, 0, nb, nb, 0, 0);
    std::fprintf(fp, "%21.14le %lu %d %d %d # nk_full nk kmesh\n", (double)nk, (unsigned long)nk, 0, 0, 0);
    std::fprintf(fp, "%lu # nkpair_elec\n", (unsigned long)np);
    std::fprintf(fp, "%d %d %d %d# modeStart modeStp modeSkipStart modeSkipStop\n", 0, p->n_modes, 0, 0);
    std::fprintf(fp, "%21.14le # T\n", p->temperature);
    const double mu = -0.1 / 27.21138505;
    std::fprintf(fp, "%21.14le %21.14le %21.14le # muMin, muMax mu (given carrier density)\n", mu, mu, mu);
    std::fprintf(fp, "%lg %lg # degauss\n", p->degauss, 1000.0);
    std::fprintf(fp, "%14.7le %14.7le %14.7le %14.7le %14.7le %14.7le # EBot_probe, ETop_probe, EBot_dm, ETop_dm, EBot_eph, ETop_eph\n",
                 emin, emax, emin, emax, emin, emax);
    std::fprintf(fp, "%14.7le # omegaMax", wmax);
    std::fclose(fp);
    // ldbd_R.dat (lattice.h:48-56): a GaAs-like fcc cell, a = 10.6829 bohr
    fp = std::fopen((d + "ldbd_R.dat").c_str(), "w");
    if (!fp) return fail(DMD_ERR_IO, "cannot open " + d + "ldbd_R.dat");
    const double h = 5.34145;
    std::fprintf(fp, "3\n%15.7le%15.7le%15.7le\n%15.7le%15.7le%15.7le\n%15.7le%15.7le%15.7le\n", 0.0, h, h, h, 0.0, h, h,
                 h, 0.0);
    std::fclose(fp);
    if ((rc = write_bin(d + "ldbd_ek.bin", e.data(), e.size() * 8))) return rc;
    if ((rc = write_bin(d + "ldbd_kvec.bin", kv.data(), kv.size() * 8))) return rc;
    if ((rc = write_bin(d + "ldbd_smat.bin", s3.data(), s3.size() * 8))) return rc;
    if ((rc = write_bin(d + "ldbd_kpair_k1st.bin", k1.data(), np * 8))) return rc;
    if ((rc = write_bin(d + "ldbd_kpair_k2nd.bin", k2.data(), np * 8))) return rc;
    if ((rc = write_bin(d + "ldbd_P1_lindblad.bin", P1.data(), P1.size() * 8))) return rc;
    if ((rc = write_bin(d + "ldbd_P2_lindblad.bin", P2.data(), P2.size() * 8))) return rc;
    return DMD_OK;
}
