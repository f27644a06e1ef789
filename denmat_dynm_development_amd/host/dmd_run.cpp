// dmd_run.cpp -- the param.in-driven real-time run of the reference
// (src_v4.5.7/main.cpp:39-97 with dm_dynamics, DensityMatrix_Dynamics.h:31-77,
// 177-345) for the Lindblad scattering operator of libdmd_scatt:
//
//   ldbd_data/ + param.in  ->  s{x,y,z}_{elec,hole}_tot.out, tau_s*_tot.out, restart/
//
// run in the folder that holds param.in and ldbd_data/, as the reference is.
// What it reads: lattice (ldbd_R.dat), band energies and window, T and mu
// (ldbd_size.dat, or mu from carrier_density as electron::set_mu_and_n),
// spin matrices, stored pairs and P (x scale_eph), ab initio e-i P of every
// impurity_density<i> with impMode ab_neutral (x |n_i cell_size| x scale_ei)
// or the model ionized-impurity P (impMode model_ionized, scrMode medium,
// static RPA / unscreened / debye / Bechstedt screening) generated on the GPU,
// and the model e-e P at equilibrium (eeMode Pee_fixed_at_eq);
// the initial state is the Bpert first-order state (electron::
// compute_dm_Bpert_1st) or the restart files; alg_ddmdteq subtracts the
// equilibrium rate at the carrier-conserving f_eq (singdenmat_k::set_dm_eq);
// B (Bx, By, Bz) adds the coherent H_BS = B . s (with needL, alg_read_Bso and
// scale_Ez: (gs/2 B + Bso_k) . s + B/2 . l + scale_Ez H_Ez, electron.cpp:19-61,
// and alg_set_scv_zero; gfac_normal_dist draws g_k per k, electron.cpp:7-31);
// eeMode Pee_update rebuilds the e-e P from the state every
// freq_update_ee_model steps, or inside every right-hand side when it is
// negative, and alg_linearize_dPee adds its linearized change; alg_linearize replaces the
// operator by its linearization at that f_eq (set_Lsc).  The evolution is the
// on-device integrator alg_ode_method picks: rkf45 (the reference's default,
// adaptive with ode_hstart / ode_hmin / ode_hmax / ode_epsabs), euler, or
// fixed-step rk4 with step min(tstep, ode_hmax); reports every freq_measure
// steps of tstep, and tau_s*_tot.out every freq_compute_tau steps.
// alg_sparseP zeroes |P| <= alg_thr_sparseP and writes the sparse sP files;
// alg_Pin_is_sparse reads P from them.  alg_semiclassical, alg_phenom_relax,
// the valley k-pair filters and the scattering switches act as in the
// reference.  Features whose physics this operator does not carry (lindhard
// or electron-gas screening, partial_ionized impurities, pump pulses) stop the
// run with a message rather than being ignored.
//
//   dmd_run [--dry-run] [folder]
#include <algorithm>
#include <array>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "dmd_scatt.h"
#include "electronphonon_mi355x.h"
#include "param_in.h"
#include "units.h"

namespace {

using namespace dmdrun;

struct cplx {
    double x, y;
};

struct Settings {
    bool restart = false, tau_only = false, sparseP = false, Pin_is_sparse = false;
    bool scatt_enable = true, eph_enable = true, only_eimp = false, semiclassical = false, phenom = false;
    double tau_phenom = 0;
    int bStart_tau = 0, bEnd_tau = 0;
    bool excess = false;  // carrier_density_means_excess_density
    bool read_Bso = false, needL = false, set_scv_zero = false;
    double scale_Ez = 0;
    bool gfac_normal = false;  // gfac_normal_dist: k-resolved random g factors
    double gfac_mean = 2, gfac_sigma = 0, gfac_cap = 0;
    double thr_sparseP = 1e-40;
    std::string scatt = "lindblad", picture = "interaction";
    bool ddmdteq = false, expt = true, linearize = false, linearize_dPee = false;
    double scale_eph = 1, scale_ei = 1;
    double t0 = 0, tend = 0, tstep = 0, hmax = 0, hstart = 0, hmin = 0, epsabs = 1e-8;
    std::string ode = "rkf45";
    int freq_measure = 1, freq_compute_tau = 10;
    double mu = 0, carrier_density = 0, scissor = 0, temperature = 0;
    double Bpert[3] = {0, 0, 0}, B[3] = {0, 0, 0};
    std::vector<double> ni;  // impurity densities in bohr^-dim
    std::vector<bool> ni_model;  // impMode model_ionized (else ab_neutral)
    std::vector<double> ni_Z;    // Z_impurity
    bool eimp_db = false;        // detailBalance of the model e-i term
    std::string eeMode = "none";
    int freq_update_ee = 0;
    bool only_ee = false, ee_exchange = false;
    double scale_ee = 1, degauss_ee = 0;
    double degauss = 0, ndegauss = 0, degauss_eimp = 0;
    // screening (coulombParam, parameters.cpp:192-206)
    std::string scrMode = "none", scrFormula = "RPA", scr_dynamic = "static";
    double eps_background = 1, meff = 1;
    bool update_screening = true;
    // k-pair filter of evolve (evolve_eph.cpp:52-58)
    bool only_intra = false, only_inter = false;
    std::vector<std::array<double, 3>> vpos;  // valley positions (crystal coordinates)
    std::vector<std::pair<int, int>> vforbid;  // forbidden valley transitions
};

// electron.h:119-124
double fermi(double T, double mu, double e)
{
    const double x = (e - mu) / T;
    if (x < -46) return 1;
    if (x > 46) return 0;
    return 1. / (std::exp(x) + 1);
}

// random_normal_array(gfack, nk, mean, sigma, cap) of electron::set_gfac
// (electron.cpp:7-18, comm/myarray.cpp:299-321, comm/Random.cpp:24-45): a
// 64-bit Mersenne Twister seeded with n, normal draws mean + sigma z (each
// redrawn while |sigma z| > cap when cap > 0), then shifted to the exact mean
// and scaled to the exact population sigma about it.  The engine and the
// distribution are the standard library's, as the reference's are.
std::vector<double> gfac_normal_array(long n, double mean, double sigma, double cap)
{
    std::vector<double> a((size_t)n, 0.0);
    if (n == 0) return a;
    if (n == 1) { a[0] = mean; return a; }
    std::mt19937_64 gen((uint64_t)n);
    std::normal_distribution<double> z;
    for (;;) {
        for (long i = 0; i < n; i++) {
            double r = sigma * z(gen);
            if (cap > 0.0)
                while (std::fabs(r) > cap) r = sigma * z(gen);
            a[i] = mean + r;
        }
        if (sigma == 0) return a;
        double sum = 0;
        for (long i = 0; i < n; i++) sum += a[i];
        const double mean1 = sum / n;
        for (long i = 0; i < n; i++) a[i] = a[i] + mean - mean1;
        double ss = 0;
        for (long i = 0; i < n; i++) ss += (a[i] - mean) * (a[i] - mean);
        const double sigma1 = std::sqrt(ss / n);
        if (sigma1 == 0) continue;
        const double fac = sigma / sigma1;
        for (long i = 0; i < n; i++) a[i] = mean + fac * (a[i] - mean);
        return a;
    }
}

void refuse(const ParamIn &p, const std::string &key, double off_value, const std::string &why)
{
    if (p.has(key) && p.get(key, off_value) != off_value)
        throw std::runtime_error(key + " = " + p.str(key, "") + ": " + why + " is not carried by this operator");
}

// update_scatt_inside's e-e rebuild from a stage state (host copy, nk x nbd^2):
// compute() zeroes the off-diagonal elements first under alg_semiclassical
// (DensityMatrix_Dynamics.h:249), then reset_scatt runs on the state and f_eq
struct InsideRebuild {
    electronphonon_mi355x<cplx> *eph;
    std::vector<cplx> dm;
    double *feq;
    bool semiclassical;
    long nk;
    int nbd;
    std::string err;
};
int inside_rebuild(void *ctx, double t, const double *rho)
{
    InsideRebuild &c = *(InsideRebuild *)ctx;
    try {
        const size_t n = (size_t)c.nk * c.nbd * c.nbd;
        c.dm.assign((const cplx *)rho, (const cplx *)rho + n);
        if (c.semiclassical)
            for (long k = 0; k < c.nk; k++)
                for (int i = 0; i < c.nbd; i++)
                    for (int j = 0; j < c.nbd; j++)
                        if (i != j) c.dm[(size_t)k * c.nbd * c.nbd + i * c.nbd + j] = {0, 0};
        cplx *rows[1] = {c.dm.data()};
        double *frows[1] = {c.feq};
        c.eph->reset_scatt(false, true, rows, nullptr, t, frows);
    } catch (const std::exception &e) {
        c.err = e.what();
        std::fprintf(stderr, "dmd_run: rebuild inside the right-hand side: %s\n", e.what());
        return 1;
    }
    return 0;
}

Settings read_settings(const ParamIn &p, const dmd_ldbd_info &info, int dim, const std::string &base_dir)
{
    using namespace units;
    Settings s;
    const std::string code = p.str("code", "jdftx");
    if (code != "jdftx") throw std::runtime_error("code = " + code + ": only ldbd_data input (code = jdftx) is read");
    if (p.str("material_model", "none") != "none") throw std::runtime_error("material_model: model materials are not read");
    s.restart = p.get("restart", 0) != 0;
    s.tau_only = p.get("compute_tau_only", 0) != 0;  // parameters.cpp:73
    s.scatt_enable = p.get("alg_scatt_enable", 1) != 0;  // parameters.cpp:78-79,86
    s.eph_enable = p.get("alg_eph_enable", 1) != 0;
    s.only_eimp = p.get("alg_only_eimp", 0) != 0;
    const double scale_scatt = p.get("scale_scatt", 1.);
    s.scale_eph = p.get("scale_eph", scale_scatt);
    s.scale_ei = p.get("scale_ei", scale_scatt);
    s.scale_ee = p.get("scale_ee", scale_scatt);
    s.ddmdteq = p.get("alg_ddmdteq", 0) != 0;
    s.expt = p.get("alg_expt", 1) != 0;
    s.scatt = p.str("alg_scatt", "lindblad");
    s.picture = p.str("alg_picture", "interaction");
    if (s.picture != "interaction" && s.picture != "schrodinger")
        throw std::runtime_error("alg_picture must be interaction or schrodinger");
    refuse(p, "alg_summode", 1, "the A-matrix (non-summode) form");
    s.semiclassical = p.get("alg_semiclassical", 0) != 0;  // parameters.cpp:101
    s.linearize = p.get("alg_linearize", 0) != 0;
    s.linearize_dPee = p.get("alg_linearize_dPee", 0) != 0;
    // parameters.cpp:118-120; sparse P: the small elements are zeroed, and (nb = 2,
    // FP64) P is then held by its nonzero elements
    s.Pin_is_sparse = p.get("alg_Pin_is_sparse", 0) != 0;
    s.sparseP = p.get("alg_sparseP", 0) != 0;
    s.thr_sparseP = p.get("alg_thr_sparseP", 1e-40);
    if (!(s.thr_sparseP >= 0)) throw std::runtime_error("alg_thr_sparseP must be >= 0");
    if (s.Pin_is_sparse && s.linearize)  // ElectronPhonon_linear.cpp:140-143
        throw std::runtime_error("linearization for sparse P matrices is not implemented");
    s.phenom = p.get("alg_phenom_relax", 0) != 0;  // parameters.cpp:122-128
    if (s.phenom) {
        s.tau_phenom = p.get("tau_phenom", 1e15, 1000 * fs());
        s.bStart_tau = (int)p.get("bStart_tau", 0);
        s.bEnd_tau = (int)p.get("bEnd_tau", 0);
        if (s.bEnd_tau > info.b_top_dm - info.b_bot_dm) throw std::runtime_error("bEnd_tau > nb_dm");
        if (s.bStart_tau < 0) throw std::runtime_error("bStart_tau < 0");
    }
    s.only_ee = p.get("alg_only_ee", 0) != 0;
    s.only_intra = p.get("alg_only_intravalley", 0) != 0;  // parameters.cpp:88-89
    s.only_inter = p.get("alg_only_intervalley", 0) != 0;
    if (s.only_intra && s.only_inter)
        throw std::runtime_error("only_intravalley and only_intervalley cannot be true at the same time");
    for (int iv = 1;; iv++) {  // valley1, valley2, ... (parameters.cpp:330-350)
        std::istringstream in(p.str("valley" + std::to_string(iv), "1 1 1"));
        std::array<double, 3> v{1, 1, 1};
        in >> v[0] >> v[1] >> v[2];
        if (std::fabs(v[0]) >= 1 || std::fabs(v[1]) >= 1 || std::fabs(v[2]) >= 1) break;
        s.vpos.push_back(v);
    }
    if (s.vpos.size() == 1)
        throw std::runtime_error("if you want to analyse valley dynamics, at least you need provide positions of two valleys");
    if (s.vpos.size() >= 2) {  // parameters.cpp:25-44: one "i j" per line
        const std::string f = p.str("file_forbid_vtrans", "NONE");
        if (f != "NONE") {
            std::ifstream in(f[0] == '/' ? f : base_dir + "/" + f);
            if (!in) throw std::runtime_error("cannot open file_forbid_vtrans " + f);
            std::string line;
            while (std::getline(in, line)) {
                std::istringstream ls(line);
                int i, j;
                if (!(ls >> i >> j)) continue;
                if (i < 0 || j < 0 || i >= (int)s.vpos.size() || j >= (int)s.vpos.size())
                    throw std::runtime_error("file_forbid_vtrans: valley index outside 0.." + std::to_string(s.vpos.size() - 1));
                s.vforbid.push_back({i, j});
            }
        }
    }
    // coherent-Hamiltonian inputs (electron.cpp:19-61, parameters.cpp:115,186,295)
    s.read_Bso = p.get("alg_read_Bso", 0) != 0;
    s.needL = p.get("needL", 0) != 0;
    s.scale_Ez = p.get("scale_Ez", 0.);
    s.set_scv_zero = p.get("alg_set_scv_zero", 0) != 0;
    // k-resolved random g factors (parameters.cpp:179-183,524-525; electron.cpp:7-31)
    s.gfac_normal = p.get("gfac_normal_dist", 0) != 0;
    if (p.get("gfac_k_resolved", 1) == 0)
        throw std::runtime_error("currently, g factor should not have band dependence (gfac_k_resolved = 0)");
    s.gfac_mean = p.get("gfac_mean", 2.);
    s.gfac_sigma = p.get("gfac_sigma", 0.);
    s.gfac_cap = p.get("gfac_cap", 0.);
    s.excess = p.get("carrier_density_means_excess_density", 0) != 0;  // parameters.cpp:170
    // e-e (parameters.cpp:226-231,479-492): the P of set_eph at the equilibrium
    // occupations, kept fixed (Pee_fixed_at_eq), or rebuilt from the state every
    // freq_update_ee_model steps (Pee_update; DensityMatrix_Dynamics.h:281-320)
    s.eeMode = p.str("eeMode", "none");
    if (s.eeMode != "none" && s.eeMode != "Pee_fixed_at_eq" && s.eeMode != "Pee_update")
        throw std::runtime_error("eeMode must be none or Pee_fixed_at_eq now");
    s.ee_exchange = p.get("ee_antisymmetry", 0) != 0;
    s.degauss_ee = p.get("degauss_ee", p.get("degauss", info.degauss / eV, eV) / eV, eV);
    s.freq_update_ee = p.get("freq_update_ee_model", 0);
    const double laserA = p.get("laserA", p.get("pumpA0", 0.));
    if (laserA != 0) throw std::runtime_error("laserA / pumpA0: pump pulses are not carried by this operator");
    // alg_eph_sepr_eh / need_elec / need_hole choose which file set the
    // initializer wrote (ElectronPhonon.cpp:132-158); the files say it too
    // (parameters.cpp:92-99,454-455); main.cpp:65 builds one electronphonon with
    // isHole = !need_elec
    const bool sepr_eh = p.get("alg_eph_sepr_eh", 1) != 0;
    const bool need_elec = !sepr_eh || p.get("alg_eph_need_elec", 1) != 0;
    const bool need_hole = !sepr_eh || p.get("alg_eph_need_hole", 0) != 0;
    if (sepr_eh && !need_elec && !need_hole)
        throw std::runtime_error("if alg_eph_sepr_eh, either alg_eph_need_elec or alg_eph_need_hole");
    if (!need_elec != (info.kind == 1))
        throw std::runtime_error("alg_eph_need_elec / alg_eph_need_hole do not match the ldbd_data file set "
                                 "(first line of ldbd_size.dat)");
    // parameters.cpp:360-367; rk4 (fixed step min(tstep, ode_hmax)) is this program's addition
    s.ode = p.str("alg_ode_method", "rkf45");
    if (s.ode != "rkf45" && s.ode != "euler" && s.ode != "rk4")
        throw std::runtime_error("alg_ode_method must be rkf45, euler or rk4");
    s.temperature = info.temperature;  // code = jdftx: T of ldbd_size.dat (parameters.cpp:7-23)
    s.scissor = p.get("scissor", 0, eV);
    s.mu = p.get("mu", info.mu / eV, eV);
    s.carrier_density = p.get("carrier_density", 0, std::pow(bohr2cm, dim));
    for (int i = 0; i < 3; i++) {
        static const char *bp[3] = {"Bxpert", "Bypert", "Bzpert"}, *bb[3] = {"Bx", "By", "Bz"};
        s.Bpert[i] = p.get(bp[i], 0., Tesla2au);
        s.B[i] = p.get(bb[i], 0., Tesla2au);
    }
    s.degauss = p.get("degauss", info.degauss / eV, eV);  // parameters.cpp:131-140
    s.ndegauss = p.get("ndegauss", info.ndegauss);
    s.degauss_eimp = p.get("degauss_eimp", s.degauss / eV, eV);
    s.eimp_db = p.get("detailBalance", 0) != 0;
    s.scrMode = p.str("scrMode", "none");
    s.scrFormula = p.str("scrFormula", "RPA");
    s.scr_dynamic = p.str("dynamic_screening", "static");
    s.eps_background = p.get("epsilon_background", 1);
    s.meff = p.get("meff_screening", 1);
    s.update_screening = p.get("update_screening", 1) != 0;
    // coulombParam::check_params (Scatt_Param.h:21-49)
    if (s.scrMode != "none" && s.scrMode != "medium") throw std::runtime_error("scrMode must be none or medium now");
    if (s.scrFormula != "unscreened" && s.scrFormula != "debye" && s.scrFormula != "Bechstedt" &&
        s.scrFormula != "heg" && s.scrFormula != "lindhard" && s.scrFormula != "RPA")
        throw std::runtime_error("scrFormula must be unscreened, debye, Bechstedt, heg, lindhard or RPA now");
    if (s.scr_dynamic != "static" &&
        (s.scrFormula == "unscreened" || s.scrFormula == "debye" || s.scrFormula == "Bechstedt"))
        throw std::runtime_error("debye and Bechstedt models are inn static limit");
    if (s.scr_dynamic != "static" && s.scr_dynamic != "ppa" && s.scr_dynamic != "real-axis" && s.scr_dynamic != "model")
        throw std::runtime_error("dynamic must be static or ppa or real-axis or model now");
    if (s.meff <= 0) throw std::runtime_error("meff must > 0");
    if (s.eps_background < 1) throw std::runtime_error("epsilon_background < 1 is not right");
    for (int iD = 0;; iD++) {  // parameters.cpp:210-224,466-478
        const std::string sfx = iD == 0 ? "" : std::to_string(iD + 1);
        const double n = p.get("impurity_density" + sfx, 0, std::pow(bohr2cm, dim));
        if (n == 0) break;
        const std::string mode = p.str("impMode" + sfx, "model_ionized");
        if (mode != "model_ionized" && mode != "ab_neutral")
            throw std::runtime_error("impMode must be model_ionized or ab_neutral now");
        const double Z = p.get("Z_impurity" + sfx, 1);
        if (mode == "model_ionized") {
            if (s.scrMode == "none")
                throw std::runtime_error("scrMode should not be none if considering model electron-ionized-impurity scattering");
            // lindhard: coulomb_model fills vq_RPA for RPA only (Coulomb_Model.h:531), so vq reads an
            // empty table; heg: frequency integrals of the electron gas
            if (s.scrFormula == "lindhard" || s.scrFormula == "heg")
                throw std::runtime_error("scrFormula = " + s.scrFormula +
                                         ": the model e-i term is built for the static RPA, unscreened, debye and "
                                         "Bechstedt screening");
            // occ_of_impurity_level reads lng, which parameters.cpp never fills (Scatt_Param.h:88-93)
            // the impurity level's occupation reads lng[iD] (Scatt_Param.h:90), a vector
            // parameters.cpp never fills (its push_back is commented out, :219)
            refuse(p, "partial_ionized" + sfx, 0,
                   "partially ionized impurities (the reference's occupation of the impurity level reads an "
                   "unfilled lng vector, Scatt_Param.h:90, parameters.cpp:219)");
            if (n < 0) throw std::runtime_error("eip.ni_ionized[iD] must be postive");  // ElecImp_Model.h:33
        }
        if (Z <= 0) throw std::runtime_error("for ionized impurity, Z is defined as a positive value in this code");
        (void)p.get("E_impurity" + sfx, 0.);  // read by the reference for partial ionization only
        (void)p.get("g_impurity" + sfx, 2.);
        s.ni.push_back(n);
        s.ni_model.push_back(mode == "model_ionized");
        s.ni_Z.push_back(Z);
    }
    (void)p.get("freq_update_eimp_model", 0);  // updates only during a pump pulse (DensityMatrix_Dynamics.h:301-318)
    // parameters.cpp:459-464,485-486,493-497
    if (s.only_eimp && s.ni.empty()) throw std::runtime_error("alg_only_eimp is only possible if impurity_density is non-zero");
    const bool ee = s.eeMode != "none";
    if ((!s.ni.empty() || ee || s.eph_enable) && !s.scatt_enable)
        throw std::runtime_error("alg_scatt_enable must be true if any scattering mechanism is considered");
    if (!s.eph_enable && s.ni.empty() && !ee && s.scatt_enable)
        throw std::runtime_error("alg_scatt_enable must be false if no scattering mechanism is considered");
    if (s.scrMode == "none" && ee)
        throw std::runtime_error("scrMode should not be none if considering electron-electron scattering");
    if (s.eph_enable && (s.only_eimp || s.only_ee))
        throw std::runtime_error("if only e-imp or only e-e, e-ph should not be enabled. Set alg_eph_enable = 0");
    if (s.only_ee && !ee) throw std::runtime_error("alg_only_ee is only possible if eeMode is not none");
    if (s.only_eimp && ee) throw std::runtime_error("if you want only e-i scattering, please set eeMode to be none");
    if (s.linearize && s.linearize_dPee)  // parameters.cpp:436-443
        throw std::runtime_error("alg_linearize_dPee and alg.linearize cannot be both true");
    if (s.linearize_dPee && (s.Pin_is_sparse || s.sparseP))
        throw std::runtime_error("linearization does not support sparse matrices");
    if (s.linearize_dPee && !s.ddmdteq) throw std::runtime_error("linearize_dPee must be used ddmdteq");
    if (!ee && s.linearize_dPee)
        throw std::runtime_error("alg_linearize_dPee is only active when e-e scattering is active");
    // set_eph's e-e fold-in (ElectronPhonon.cpp:12) calls calc_P's dP branch with
    // no state, which reads dm1[ik3] (ElectronPhonon.cpp:84, ElecElec_Model.h:216):
    // the reference stops there under this flag.  What its code defines past that
    // point is run when a rebuild from the state replaces that P before the first
    // step: Pee_update with freq_update_ee_model > 0 (the rebuild at it = 1).
    if (s.linearize_dPee && s.eeMode != "Pee_update")
        throw std::runtime_error("alg_linearize_dPee with eeMode = " + s.eeMode +
                                 ": the reference's only e-e build is set_eph's, which passes calc_P's dP branch "
                                 "no density matrix (ElectronPhonon.cpp:12,84; ElecElec_Model.h:216); use "
                                 "eeMode = Pee_update");
    if (!s.linearize && s.eeMode == "Pee_update" && s.freq_update_ee == 0)  // parameters.cpp:483-484
        throw std::runtime_error("want to update Pee but set freq_update_ee_model as 0");
    if (ee && (s.scrFormula == "lindhard" || s.scrFormula == "heg" || s.scr_dynamic != "static"))
        throw std::runtime_error("eeMode: the model e-e term is built for the static RPA, unscreened, debye and "
                                 "Bechstedt screening");
    bool any_ab = false;
    for (size_t iD = 0; iD < s.ni.size(); iD++) {
        if (!s.ni_model[iD]) any_ab = true;
        if (!s.ni_model[iD] && s.sparseP) throw std::runtime_error("ab initio neutral e-imp does not support sparse matrices");
        if (s.carrier_density * s.ni[iD] < 0)
            throw std::runtime_error("currently carrier_denstiy and impurity_density must have the same sign");
    }
    if (s.Pin_is_sparse && !s.eph_enable && !any_ab)
        throw std::runtime_error("alg_Pin_is_sparse implies ab initio e-ph or e-i is enabled");
    s.t0 = p.get("t0", 0., fs());
    s.tend = p.get("tend", 0., fs());
    s.tstep = p.get("tstep", 1., fs());
    s.hmax = p.get("ode_hmax", s.tstep / fs(), fs());
    s.hstart = p.get("ode_hstart", 1e-3, fs());
    s.hmin = p.get("ode_hmin", 0, fs());
    s.epsabs = p.get("ode_epsabs", 1e-8);
    if (s.hstart <= 0 || s.hmin < 0 || s.hmin > s.hmax || s.epsabs <= 0)
        throw std::runtime_error("ode_hstart > 0, 0 <= ode_hmin <= ode_hmax and ode_epsabs > 0 are required");
    s.freq_measure = (int)p.get("freq_measure", 1);
    s.freq_compute_tau = (int)p.get("freq_compute_tau", p.get("freq_measure_ene", 10));  // parameters.cpp:324-326
    if (s.tstep <= 0 || s.hmax <= 0 || s.freq_measure < 1 || s.freq_compute_tau < 1)
        throw std::runtime_error("tstep, ode_hmax > 0, freq_measure, freq_compute_tau >= 1");
    return s;
}

// electron::find_mu by bisection (electron.cpp:537-576) on the carriers of
// the conduction (or, for a negative density, valence) bands: carrier_bvk =
// density * nk_full * cell_size, compute_nfree (electron.cpp:582-600).
double find_mu_bisection(double carrier_bvk, double T, double emin, double emax, const std::vector<double> &e,
                         long nk, int nb, int nv, bool excess = false)
{
    // excess: electrons minus holes over every band (electron.cpp:547-549)
    const bool hole = carrier_bvk < 0;
    const int i0 = excess ? 0 : (hole ? 0 : nv), i1 = excess ? nb : (hole ? nv : nb);
    double lo = emin, hi = emax, mu = 0.5 * (lo + hi);
    for (int it = 0; it < 200; it++) {
        mu = 0.5 * (lo + hi);
        double n = 0;
        for (long k = 0; k < nk; k++)
            for (int i = i0; i < i1; i++) {
                const double f = fermi(T, mu, e[(size_t)k * nb + i]);
                n += i < nv ? f - 1. : f;
            }
        const double err = n - carrier_bvk;
        if (std::fabs(err) < 1e-12) return mu;
        if (n > carrier_bvk) hi = mu;
        else lo = mu;
    }
    return mu;
}

// carrier-conserving chemical potential of bands [b0, b1) of the density
// matrix for singdenmat_k::set_dm_eq (DenMat.cpp:97-149): the root of
// sum f(e; mu) = sum Re dm_ii (holes: sum (f - 1) = sum (Re dm_ii - 1)).
double mu_conserving(const std::vector<cplx> &dm, const std::vector<double> &edm, long nk, int nbd, int b0, int b1,
                     bool hole, double T)
{
    double target = 0, emin = 1e300, emax = -1e300;
    for (long k = 0; k < nk; k++)
        for (int i = b0; i < b1; i++) {
            target += hole ? dm[(size_t)k * nbd * nbd + i * nbd + i].x - 1. : dm[(size_t)k * nbd * nbd + i * nbd + i].x;
            emin = std::min(emin, edm[(size_t)k * nbd + i]);
            emax = std::max(emax, edm[(size_t)k * nbd + i]);
        }
    double lo = emin - 60 * T, hi = emax + 60 * T;
    for (int it = 0; it < 200; it++) {  // to the resolution of double: drho = rho - f_eq needs every digit
        const double mu = 0.5 * (lo + hi);
        if (mu <= lo || mu >= hi) break;
        double n = 0;
        for (long k = 0; k < nk; k++)
            for (int i = b0; i < b1; i++) {
                const double f = fermi(T, mu, edm[(size_t)k * nbd + i]);
                n += hole ? f - 1. : f;
            }
        if (n > target) hi = mu;
        else lo = mu;
    }
    return 0.5 * (lo + hi);
}

// GGT = G G^T with G = 2 pi R^-1, R's rows the lattice vectors (lattice.h:40-41)
void reciprocal_metric(const double R[9], double GGT[9])
{
    const double a = R[0], b = R[1], c = R[2], d = R[3], e = R[4], f = R[5], g = R[6], h = R[7], i = R[8];
    const double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
    const double inv[3][3] = {{(e * i - f * h) / det, (c * h - b * i) / det, (b * f - c * e) / det},
                              {(f * g - d * i) / det, (a * i - c * g) / det, (c * d - a * f) / det},
                              {(d * h - e * g) / det, (b * g - a * h) / det, (a * e - b * d) / det}};
    for (int r = 0; r < 3; r++)
        for (int q = 0; q < 3; q++) {
            double acc = 0;
            for (int m = 0; m < 3; m++) acc += inv[r][m] * inv[q][m];
            GGT[r * 3 + q] = 4 * M_PI * M_PI * acc;
        }
}

// The k-pair filter of evolve (evolve_eph.cpp:52-58, lattice.h:82-139):
// lengths under the reciprocal metric GGT = G G^T, G = 2 pi R^-1
// (lattice.h:40-41; matrix3.h:119-122), of differences wrapped into
// [-0.5, 0.5) (comm/kmap.h:6-15); K = (1/3, 1/3, 0) and K' = -K (lattice.h:43-44).
struct KpairFilter {
    double M[3][3];
    const Settings &s;
    KpairFilter(const double R[9], const Settings &set) : s(set)
    {
        double G[9];
        reciprocal_metric(R, G);
        for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++) M[r][q] = G[r * 3 + q];
    }
    double lsq(const double *k, const double *c) const
    {
        double v[3];
        for (int d = 0; d < 3; d++) {
            v[d] = k[d] - c[d];
            v[d] -= std::floor(0.5 + v[d]);
            if (std::fabs(v[d] - 0.5) < 1e-6) v[d] = -0.5;
        }
        return v[0] * v[0] * M[0][0] + v[1] * v[1] * M[1][1] + v[2] * v[2] * M[2][2] +
               2 * (v[0] * v[1] * M[0][1] + v[0] * v[2] * M[0][2] + v[1] * v[2] * M[1][2]);
    }
    bool isK(const double *k) const
    {
        static const double K[3] = {1. / 3, 1. / 3, 0}, Kp[3] = {-1. / 3, -1. / 3, 0};
        return lsq(k, K) < lsq(k, Kp);
    }
    int valley(const double *k) const
    {
        if (s.vpos.size() < 2) return -1;
        int iv = 0;
        double best = lsq(k, s.vpos[0].data());
        for (size_t q = 1; q < s.vpos.size(); q++) {
            const double l = lsq(k, s.vpos[q].data());
            if (l < best) { best = l; iv = (int)q; }
        }
        return iv;
    }
    bool active() const { return s.only_intra || s.only_inter || (s.vpos.size() >= 2 && !s.vforbid.empty()); }
    bool allowed(const double *k1, const double *k2) const
    {
        const bool intra = isK(k1) == isK(k2);
        if (intra && s.only_inter) return false;
        if (!intra && s.only_intra) return false;
        const int v1 = valley(k1), v2 = valley(k2);
        if (v1 >= 0 && v2 >= 0)
            for (const auto &fb : s.vforbid)
                if ((fb.first == v1 && fb.second == v2) || (fb.first == v2 && fb.second == v1)) return false;
        return true;
    }
};

struct Report {
    std::string name;  // "sx_elec" ...
    int comp, i0, i1;
    FILE *fp = nullptr;
    bool orbital = false;  // lx/ly/lz: ldbd_lmat.bin in place of the spin matrices
};

void mkdir_p(const std::string &d)
{
    if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) throw std::runtime_error("cannot create " + d);
}

int run(const std::string &folder, bool dry)
{
    using namespace units;
    const std::string base = folder.empty() ? std::string(".") : folder;
    const std::string ldbd = base + "/ldbd_data";
    const ParamIn p = ParamIn::read(base + "/param.in");
    dmd_ldbd_info info;
    check(dmd_ldbd_read_info(ldbd.c_str(), &info), "ldbd_size.dat");
    int dim;
    double R[9];
    check(dmd_ldbd_read_R(ldbd.c_str(), &dim, R), "ldbd_R.dat");
    dim = (int)p.get("dim", dim);
    // lattice.h:24-37
    const double volume = std::fabs(R[0] * (R[4] * R[8] - R[5] * R[7]) - R[1] * (R[3] * R[8] - R[5] * R[6]) +
                                    R[2] * (R[3] * R[7] - R[4] * R[6]));
    double cell = volume;
    if (dim == 2) cell = volume / R[8];
    if (dim == 1) cell = R[8];
    if (dim == 0) cell = 1;
    const Settings s = read_settings(p, info, dim, base);
    const long nk = (long)info.nk;
    const int nbf = info.nb, nbd = info.b_top_dm - info.b_bot_dm, nvd = std::max(info.nv - info.b_bot_dm, 0);
    const double T = s.temperature;
    std::vector<double> e((size_t)nk * nbf);
    check(dmd_ldbd_read_ek(ldbd.c_str(), &info, s.scissor, e.data()), "ldbd_ek.bin");
    std::vector<double> edm((size_t)nk * nbd);
    for (long k = 0; k < nk; k++)
        for (int b = 0; b < nbd; b++) edm[(size_t)k * nbd + b] = e[(size_t)k * nbf + info.b_bot_dm + b];
    // electron::set_mu_and_n (electron.h:141-160) with emin/emax of electron.h:94-105
    double mu = s.mu;
    if (s.carrier_density != 0 || s.excess) {
        double evmax = -1e300, emin = 1e300, emax = -1e300, ecmin = 1e300;
        for (long k = 0; k < nk; k++)
            for (int b = 0; b < nbd; b++) {
                const double x = edm[(size_t)k * nbd + b];
                if (b >= nvd) { emax = std::max(emax, x); ecmin = std::min(ecmin, x); }
                else { evmax = std::max(evmax, x); emin = std::min(emin, x); }
            }
        if (nbd == nvd) emax = evmax + 1;
        if (nvd == 0) emin = ecmin - 1;
        mu = find_mu_bisection(s.carrier_density * info.nk_full * cell, T, emin, emax, e, nk, nbf, info.nv, s.excess);
    }
    std::vector<double> fdm((size_t)nk * nbd);
    for (size_t i = 0; i < fdm.size(); i++) fdm[i] = fermi(T, mu, edm[i]);
    const int n_sub = (int)std::ceil(s.tstep / s.hmax - 1e-9);
    const double h_rk = s.tstep / n_sub;
    const double prefac_obs = 1. / (info.nk_full * cell * std::pow(bohr2cm, dim));  // observable.h:53-55
    std::printf("dmd_run: nk = %ld nb_dm = %d (nv_dm = %d) e-ph bands [%d, %d) of the density matrix, %ld stored pairs\n",
                nk, nbd, nvd, info.b_bot_eph - info.b_bot_dm, info.b_top_eph - info.b_bot_dm, (long)info.nkpair);
    std::printf("T = %.10le mu = %.10le cell_size = %.10le dim = %d\n", T, mu, cell, dim);
    std::printf("t0 = %.10le tend = %.10le tstep = %.10le freq_measure = %d alg_ode_method = %s\n", s.t0, s.tend, s.tstep,
                s.freq_measure, s.ode.c_str());
    std::printf("freq_compute_tau = %d compute_tau_only = %d\n", s.freq_compute_tau, s.tau_only ? 1 : 0);
    if (s.ode == "rk4") std::printf("RK4 step %.10le x %d per tstep\n", h_rk, n_sub);
    if (s.ode == "rkf45")
        std::printf("rkf45: ode_hstart = %.10le ode_hmin = %.10le ode_hmax = %.10le ode_epsabs = %lg\n", s.hstart, s.hmin,
                    s.hmax, s.epsabs);
    std::printf("scale_eph = %lg scale_ei = %lg alg_ddmdteq = %d alg_linearize = %d alg_linearize_dPee = %d "
                "alg_expt = %d alg_picture = %s alg_scatt = %s\n", s.scale_eph, s.scale_ei, (int)s.ddmdteq,
                (int)s.linearize, (int)s.linearize_dPee, (int)s.expt, s.picture.c_str(), s.scatt.c_str());
    std::printf("Bpert = %.10le %.10le %.10le B = %.10le %.10le %.10le impurities = %zu restart = %d\n", s.Bpert[0],
                s.Bpert[1], s.Bpert[2], s.B[0], s.B[1], s.B[2], s.ni.size(), (int)s.restart);
    std::printf("alg_sparseP = %d alg_Pin_is_sparse = %d alg_thr_sparseP = %lg\n", (int)s.sparseP, (int)s.Pin_is_sparse,
                s.thr_sparseP);
    std::printf("alg_scatt_enable = %d alg_eph_enable = %d alg_only_eimp = %d alg_semiclassical = %d\n",
                (int)s.scatt_enable, (int)s.eph_enable, (int)s.only_eimp, (int)s.semiclassical);
    {
        int n_model = 0;
        for (bool mdl : s.ni_model) n_model += mdl;
        std::printf("scrMode = %s scrFormula = %s epsilon_background = %lg model_impurities = %d detailBalance = %d "
                    "degauss_eimp = %.10le\n", s.scrMode.c_str(), s.scrFormula.c_str(), s.eps_background, n_model,
                    (int)s.eimp_db, s.degauss_eimp);
        std::printf("eeMode = %s alg_only_ee = %d scale_ee = %lg degauss_ee = %.10le ee_antisymmetry = %d\n",
                    s.eeMode.c_str(), (int)s.only_ee, s.scale_ee, s.degauss_ee, (int)s.ee_exchange);
    }
    std::printf("alg_read_Bso = %d needL = %d scale_Ez = %lg alg_set_scv_zero = %d\n", (int)s.read_Bso, (int)s.needL,
                s.scale_Ez, (int)s.set_scv_zero);
    std::printf("gfac_normal_dist = %d gfac_mean = %lg gfac_sigma = %lg gfac_cap = %lg\n", (int)s.gfac_normal,
                s.gfac_mean, s.gfac_sigma, s.gfac_cap);
    std::printf("alg_phenom_relax = %d tau_phenom = %.10le bStart_tau = %d bEnd_tau = %d "
                "carrier_density_means_excess_density = %d\n", (int)s.phenom, s.tau_phenom, s.bStart_tau, s.bEnd_tau,
                (int)s.excess);
    std::printf("alg_only_intravalley = %d alg_only_intervalley = %d valleys = %zu forbidden_transitions = %zu\n",
                (int)s.only_intra, (int)s.only_inter, s.vpos.size(), s.vforbid.size());
    for (const auto &k : p.unread()) std::printf("note: param.in key \"%s\" is not read by this runner\n", k.c_str());
    if (dry) return 0;

    // spin matrices (ldbd_smat.bin) and the initial state
    std::vector<cplx> smat((size_t)nk * 3 * nbd * nbd), lmat;
    check(dmd_ldbd_read_smat(ldbd.c_str(), &info, (double *)smat.data()), "ldbd_smat.bin");
    if (s.needL) {  // electron.h:87
        lmat.resize(smat.size());
        check(dmd_ldbd_read_lmat(ldbd.c_str(), &info, (double *)lmat.data()), "ldbd_lmat.bin");
    }
    const size_t nbd2 = (size_t)nbd * nbd;
    if (s.set_scv_zero)  // alg_set_scv_zero: no valence-conduction elements (electron.cpp:290-297,321-328)
        for (std::vector<cplx> *m : {&smat, &lmat})
            for (size_t kc = 0; kc < m->size() / nbd2; kc++)
                for (int i = 0; i < nvd; i++)
                    for (int j = nvd; j < nbd; j++) {
                        (*m)[kc * nbd2 + i * nbd + j] = {0, 0};
                        (*m)[kc * nbd2 + j * nbd + i] = {0, 0};
                    }
    std::vector<cplx> dm((size_t)nk * nbd2);
    double t = s.t0;
    const bool interaction = s.picture == "interaction";
    if (s.restart) {
        check(dmd_restart_read((base + "/restart").c_str(), nk, nbd, (double *)dm.data(), &t), "restart files");
    } else {
        // electron::compute_dm_Bpert_1st (electron.cpp:62-98), then init_dm's
        // Hermitian average (DenMat.cpp:64-79)
        std::vector<cplx> d0((size_t)nk * nbd2);
        for (long k = 0; k < nk; k++) {
            const cplx *sk = &smat[(size_t)k * 3 * nbd2];
            for (int i = 0; i < nbd; i++)
                for (int j = i; j < nbd; j++) {
                    cplx H1 = {0, 0};
                    for (int c = 0; c < 3; c++) {
                        H1.x += s.Bpert[c] * sk[c * nbd2 + i * nbd + j].x;
                        H1.y += s.Bpert[c] * sk[c * nbd2 + i * nbd + j].y;
                    }
                    if (s.needL) {  // H1 = (gs Bpert.s + Bpert.l) / 2 (electron.cpp:74-76)
                        const cplx *lk = &lmat[(size_t)k * 3 * nbd2];
                        cplx L1 = {0, 0};
                        for (int c = 0; c < 3; c++) {
                            L1.x += s.Bpert[c] * lk[c * nbd2 + i * nbd + j].x;
                            L1.y += s.Bpert[c] * lk[c * nbd2 + i * nbd + j].y;
                        }
                        H1 = {0.5 * (2.0023193043625635 * H1.x + L1.x), 0.5 * (2.0023193043625635 * H1.y + L1.y)};
                    }
                    const double ei = edm[(size_t)k * nbd + i], ej = edm[(size_t)k * nbd + j];
                    const double fi = fdm[(size_t)k * nbd + i], fj = fdm[(size_t)k * nbd + j];
                    double dfde;
                    if (std::fabs(ei - ej) < 1e-8) {
                        const double favg = 0.5 * (fi + fj);
                        dfde = favg * (favg - 1.) / T;
                    } else {
                        dfde = (fi - fj) / (ei - ej);
                    }
                    cplx v = {dfde * H1.x, dfde * H1.y};
                    if (i != j && interaction) {
                        const double a = (ei - ej) * s.t0, c = std::cos(a), sn = std::sin(a);
                        v = {v.x * c - v.y * sn, v.x * sn + v.y * c};
                    }
                    d0[(size_t)k * nbd2 + i * nbd + j] = v;
                    if (i < j) d0[(size_t)k * nbd2 + j * nbd + i] = {v.x, -v.y};
                }
            for (int i = 0; i < nbd; i++) d0[(size_t)k * nbd2 + i * nbd + i].x += fdm[(size_t)k * nbd + i];
        }
        for (long k = 0; k < nk; k++)
            for (int i = 0; i < nbd; i++)
                for (int j = 0; j < nbd; j++) {
                    const cplx a = d0[(size_t)k * nbd2 + i * nbd + j], b = d0[(size_t)k * nbd2 + j * nbd + i];
                    dm[(size_t)k * nbd2 + i * nbd + j] = {0.5 * (a.x + b.x), 0.5 * (a.y - b.y)};
                }
    }

    // dm_eq (DensityMatrix_Dynamics.h:57-67): the carrier-conserving f_eq,
    // set only when scattering is on (dm_eq stays 0 otherwise)
    std::vector<double> feq((size_t)nk * nbd, 0.0);
    double mue = mu, muh = mu;
    if ((s.ddmdteq || s.linearize || s.phenom) && s.scatt_enable) {
        mue = nbd > nvd ? mu_conserving(dm, edm, nk, nbd, nvd, nbd, false, T) : mu;
        muh = nvd > 0 ? mu_conserving(dm, edm, nk, nbd, 0, nvd, true, T) : mu;
        for (long k = 0; k < nk; k++)
            for (int i = 0; i < nbd; i++) feq[(size_t)k * nbd + i] = fermi(T, i >= nvd ? mue : muh, edm[(size_t)k * nbd + i]);
    }

    // the scattering operator: e-ph P, ab initio e-i P, flags, ddmdt_eq, H_BS
    using Teph = electronphonon_mi355x<cplx>;
    const Teph::p_source src = !s.scatt_enable ? Teph::P_NO_SCATTERING
                               : !s.eph_enable ? Teph::P_NONE
                               : s.Pin_is_sparse ? Teph::P_SPARSE : Teph::P_DENSE;
    Teph eph(ldbd, s.scatt, s.scale_eph, s.scissor, DMD_PREC_FP64, src, s.thr_sparseP);
    eph.Pin_is_sparse = s.Pin_is_sparse;  // e-i folded into sparse input is re-sparsified
    // coulomb_model (Coulomb_Model.h:412-425,533-575), built when scrMode = medium:
    // 3D only; the free-carrier density n_free = |n_e| + |n_h| at mu over every
    // band (electron.h:153-157, electron.cpp:578-596), or, when the linearized
    // operator is built with update_screening, from f_eq over the e-ph bands
    // (ElectronPhonon_linear.cpp:11, coulomb_model::init)
    double n_free = 0;
    if (s.scrMode == "medium" && s.scatt_enable) {
        if (dim < 3) throw std::runtime_error("debye screening model for lowD is not implemented");
        if (s.linearize && s.update_screening) {
            for (long k = 0; k < nk; k++)
                for (int b = info.b_bot_eph - info.b_bot_dm; b < info.b_top_eph - info.b_bot_dm; b++)
                    n_free += b < nvd ? 1 - feq[(size_t)k * nbd + b] : feq[(size_t)k * nbd + b];
            n_free /= info.nk_full * cell;
        } else {
            double ne = 0, nh = 0;
            for (long k = 0; k < nk; k++)
                for (int b = 0; b < nbf; b++) {
                    const double f = fermi(T, mu, e[(size_t)k * nbf + b]);
                    if (b >= info.nv) ne += f;
                    else nh += f - 1.;
                }
            n_free = std::fabs(ne / (info.nk_full * cell)) + std::fabs(nh / (info.nk_full * cell));
        }
        if ((s.scrFormula == "debye" || s.scrFormula == "Bechstedt" || s.scrFormula == "heg") && n_free <= 0)
            throw std::runtime_error("nfreetot must be postive");
        std::printf("screening: %s, epsilon_background = %lg, nfree = %lg cm-3\n", s.scrFormula.c_str(), s.eps_background,
                    n_free / std::pow(bohr2cm, 3));
    }
    std::vector<uint64_t> mk1, mk2;
    std::vector<cplx> mU;
    std::vector<double> mkv;
    const int b0e = info.b_bot_eph - info.b_bot_dm, nbe = info.b_top_eph - info.b_bot_eph;
    auto model_inputs = [&](const char *who) {  // electron::read_ldbd_Umat (electron.cpp:388-410), kvec, stored pairs
        if (!mU.empty()) return;
        struct stat us;
        if (stat((ldbd + "/ldbd_Umat.bin").c_str(), &us) != 0)
            throw std::runtime_error(std::string(who) + " needs ldbd_data/ldbd_Umat.bin (the overlaps <k|k'>)");
        mU.resize((size_t)nk * info.nb_wannier * nbe);
        check(dmd_ldbd_read_Umat(ldbd.c_str(), &info, (double *)mU.data()), "ldbd_Umat.bin");
        mkv.resize((size_t)nk * 3);
        check(dmd_ldbd_read_kvec(ldbd.c_str(), &info, mkv.data()), "ldbd_kvec.bin");
        mk1.resize(info.nkpair);
        mk2.resize(info.nkpair);
        check(dmd_ldbd_read_kpairs(ldbd.c_str(), &info, 0, info.nkpair, mk1.data(), mk2.data()), "k pairs");
    };
    // coulomb_model::init_model (Coulomb_Model.h:554-575) and vq's static forms
    auto screened = [&]() {
        dmd_eimp_model m{};
        m.screening = s.scrFormula == "debye" ? DMD_SCR_DEBYE : s.scrFormula == "Bechstedt" ? DMD_SCR_BECHSTEDT
                      : s.scrFormula == "RPA" ? DMD_SCR_RPA : DMD_SCR_UNSCREENED;
        for (int d = 0; d < 3; d++) m.kmesh[d] = info.kmesh[d];
        m.nk_full = info.nk_full;
        m.lindblad = s.scatt == "lindblad";
        m.nb_wannier = info.nb_wannier;
        m.prefac_vq = 4 * M_PI / s.eps_background / cell;
        m.prefac_vq_bare = 4 * M_PI / cell;
        const double kF = std::pow(3 * M_PI * M_PI * n_free, 1. / 3.), EF = kF * kF / 2. / s.meff;
        const double qTF2 = 6 * M_PI * n_free / s.eps_background / EF;
        m.qscr2_debye = 4 * M_PI * n_free / s.eps_background / T;
        m.fac0 = 1. / (s.eps_background - 1);
        m.fac2 = 1. / qTF2;
        m.fac4 = 3. / 4. / (kF * kF) / qTF2;
        m.ethr = s.degauss * s.ndegauss;
        m.omegaL = 1e-4 * T;
        m.temperature = T;
        reciprocal_metric(R, m.GGT);
        return m;
    };
    // occupations of the e-ph bands: at mu (coulomb_model's f, Coulomb_Model.h:428), or f_eq
    // for the linearized operator's e-e term (ElectronPhonon_linear.cpp:19)
    auto eph_cols = [&](const std::vector<double> &fsrc) {
        std::vector<double> o((size_t)nk * nbe);
        for (long k = 0; k < nk; k++)
            for (int b = 0; b < nbe; b++) o[(size_t)k * nbe + b] = fsrc[(size_t)k * nbd + b0e + b];
        return o;
    };
    // Pee_update: reset_scatt rebuilds P from the e-ph (and ab initio e-i) part,
    // the model e-i terms (screening from the state under update_screening) and
    // the e-e term at the state; that base is kept before the model terms go in
    const bool pee_update = s.eeMode == "Pee_update" && !s.linearize && !s.only_eimp && s.scatt_enable;
    for (size_t iD = 0; iD < s.ni.size() && !s.only_ee; iD++)
        if (!s.ni_model[iD]) eph.add_eimp_ab_neutral((int)iD, std::fabs(s.ni[iD] * cell), s.scale_ei);
    // the e-e term's static RPA v(q) for every k-point difference, at f of mu
    // (init_RPA keeps coulomb_model's f, Coulomb_Model.h:587-590)
    std::vector<double> ee_vq_cell;
    if (s.eeMode != "none" && !s.only_eimp && s.scatt_enable && s.scrFormula == "RPA") {
        model_inputs("eeMode");
        dmd_eimp_model m = screened();
        const std::vector<double> fmu = eph_cols(fdm);
        ee_vq_cell.resize((size_t)info.kmesh[0] * info.kmesh[1] * info.kmesh[2]);
        check(dmd_scatt_ee_vq_cells(eph.handle(), (const double *)mU.data(), mkv.data(), fmu.data(), &m,
                                    ee_vq_cell.data()),
              "RPA v(q) table");
    }
    if (pee_update) {
        eph.linearize_dPee = s.linearize_dPee;
        eph.keep_base(true, ee_vq_cell);
        if (s.update_screening && (s.scrFormula == "debye" || s.scrFormula == "Bechstedt")) {
            // coulomb_model::init(dm): n_free from the window's occupations
            // (bands below nv as holes), then init_model (Coulomb_Model.h:533-575)
            const int nvw = nvd - b0e;
            eph.rescreen = [&, nvw](const std::vector<double> &occ, dmd_eimp_model &m) {
                double n = 0;
                for (long k = 0; k < nk; k++)
                    for (int b = 0; b < nbe; b++) n += b < nvw ? 1 - occ[(size_t)k * nbe + b] : occ[(size_t)k * nbe + b];
                n /= info.nk_full * cell;
                if (n <= 0) throw std::runtime_error("nfreetot must be postive");
                const double kF = std::pow(3 * M_PI * M_PI * n, 1. / 3.), EF = kF * kF / 2. / s.meff;
                const double qTF2 = 6 * M_PI * n / s.eps_background / EF;
                m.qscr2_debye = 4 * M_PI * n / s.eps_background / T;
                m.fac0 = 1. / (s.eps_background - 1);
                m.fac2 = 1. / qTF2;
                m.fac4 = 3. / 4. / (kF * kF) / qTF2;
            };
        }
    }
    for (size_t iD = 0; iD < s.ni.size() && !s.only_ee; iD++) {
        if (!s.ni_model[iD]) continue;
        model_inputs("impMode = model_ionized");
        // elecimp_model (ElecImp_Model.h:22-40)
        dmd_eimp_model m = screened();
        m.detail_balance = s.eimp_db;
        m.prefac_A = s.ni_Z[iD] * std::sqrt(s.ni[iD] * cell);
        m.degauss = s.degauss_eimp;
        std::printf("model e-i %zu: Z = %lg n_i = %lg cm-3 qscr2_debye = %lg degauss = %lg detailBalance = %d\n", iD + 1,
                    s.ni_Z[iD], s.ni[iD] / std::pow(bohr2cm, 3), m.qscr2_debye, m.degauss, (int)m.detail_balance);
        // RPA sums at f of mu (init_RPA leaves f when the screening is updated, Coulomb_Model.h:587-590)
        std::vector<double> feph;
        if (m.screening == DMD_SCR_RPA) feph = eph_cols(fdm);
        eph.add_eimp_model((int)iD, mk1.data(), mk2.data(), mU.data(), mkv.data(), feph.empty() ? nullptr : feph.data(), m,
                           s.scale_ei);
    }
    if (s.eeMode != "none" && !s.only_eimp && s.scatt_enable) {  // add_scatt_contrib("ee") (ElectronPhonon.cpp:11)
        model_inputs("eeMode");
        for (int d = 0; d < 3; d++)
            if (info.kmesh[d] <= 0) throw std::runtime_error("eeMode needs the k mesh of ldbd_size.dat (line 3)");
        dmd_eimp_model m = screened();
        m.degauss = s.degauss_ee;
        m.exchange = s.ee_exchange;
        // under alg_linearize_dPee the e-e P is taken at f_eq (calc_A1f with f_eq,
        // ElecElec_Model.h:213-215), as every rebuild takes it
        const std::vector<double> fee = eph_cols(s.linearize || s.linearize_dPee ? feq : fdm);
        std::printf("model e-e (%s): degauss_ee = %lg scale_ee = %lg\n", s.eeMode.c_str(), m.degauss, s.scale_ee);
        eph.add_ee_model(mk1.data(), mk2.data(), mU.data(), mkv.data(), fee.data(), m, s.scale_ee, ee_vq_cell);
    }
    // set_eph ends with set_sparseP(true) (ElectronPhonon.cpp:12); under
    // alg_linearize set_eph returns before it (:6)
    if (s.sparseP && !s.linearize && !s.Pin_is_sparse) {
        eph.set_sparseP(true);
        std::printf("sparse P: |P| <= %lg zeroed, sP files in %s\n", s.thr_sparseP, ldbd.c_str());
    }
    // the k-pair filter of evolve, after the sP files took the unfiltered P;
    // evolve_linear has none (evolve_eph_linear.cpp)
    const KpairFilter filt(R, s);
    if (filt.active() && s.scatt_enable && !s.linearize) {
        std::vector<double> kv((size_t)nk * 3);
        check(dmd_ldbd_read_kvec(ldbd.c_str(), &info, kv.data()), "ldbd_kvec.bin");
        std::vector<uint64_t> k1(info.nkpair), k2(info.nkpair);
        check(dmd_ldbd_read_kpairs(ldbd.c_str(), &info, 0, info.nkpair, k1.data(), k2.data()), "k pairs");
        std::vector<uint8_t> keep(info.nkpair);
        int64_t dropped = 0;
        for (int64_t q = 0; q < info.nkpair; q++) {
            keep[q] = filt.allowed(&kv[k1[q] * 3], &kv[k2[q] * 3]) ? 1 : 0;
            dropped += !keep[q];
        }
        check(dmd_scatt_mask_pairs(eph.handle(), keep.data()), "k-pair filter");
        std::printf("k-pair filter: %lld of %lld stored pairs dropped\n", (long long)dropped, (long long)info.nkpair);
    }
    eph.set_algorithm(s.expt, !interaction);
    eph.set_semiclassical(s.semiclassical);
    if ((s.ddmdteq || s.linearize || s.phenom) && s.scatt_enable) {
        double *rows[1] = {feq.data()};
        if (s.linearize) {
            eph.set_Lsc(rows);  // ddmdt_eq of the linearized operator is zero: evolve_linear at f_eq
            // add_Lsc_contrib("ee") beyond the P1/P2 form: fac2 under ee_antisymmetry,
            // and P3..P6 unless eeMode = Pee_fixed_at_eq (ElectronPhonon_linear.cpp:61-95)
            if (s.eeMode != "none" && !s.only_eimp) {
                const std::vector<double> fw = eph_cols(feq);
                eph.add_Lsc_ee(fw.data(), s.eeMode != "Pee_fixed_at_eq", 1.0, ee_vq_cell);
            }
            std::printf("linearized at mu_e = %.12le mu_h = %.12le\n", mue, muh);
        } else if (s.ddmdteq && s.linearize_dPee) {
            // DensityMatrix_Dynamics.h:60-62: neither reset_scatt at dm_eq nor
            // compute_ddmdt_eq under alg_linearize_dPee (they run again only
            // for a pump's e-i update, :282-288): ddmdt_eq stays zero
            std::printf("ddmdt_eq not computed (alg_linearize_dPee): zero\n");
        } else if (s.ddmdteq) {
            eph.compute_ddmdt_eq(rows);
            std::printf("ddmdt_eq at mu_e = %.12le mu_h = %.12le\n", mue, muh);
        }
    }
    if (s.phenom && s.bEnd_tau > s.bStart_tau) {  // phenom_relax over [bStart_tau, bEnd_tau); an empty range does nothing
        double *rows[1] = {feq.data()};
        eph.set_phenom_relax(s.tau_phenom, s.bStart_tau, s.bEnd_tau, rows);
        std::printf("phenomenological relaxation: tau = %.10le on bands [%d, %d)\n", s.tau_phenom, s.bStart_tau, s.bEnd_tau);
    }
    // the sparse modes stream P by its nonzero elements (nb = 2); Pee_update
    // rebuilds P in place and keeps it dense
    if ((s.sparseP || s.Pin_is_sparse) && !s.linearize && !pee_update && s.scatt_enable && eph.compress_P())
        std::printf("sparse P: stored by its nonzero elements\n");
    dmd_scatt *h = eph.handle();
    // freq_update_ee_model < 0: the e-e P rebuilt from every right-hand side's
    // stage state (update_scatt_inside in compute, DensityMatrix_Dynamics.h:262,
    // 281-289, 307-310), through the integrators' hook
    InsideRebuild inside{&eph, {}, feq.data(), s.semiclassical, nk, nbd, {}};
    if (pee_update && s.freq_update_ee < 0) {
        check(dmd_scatt_set_rhs_hook(h, inside_rebuild, &inside), "right-hand-side hook");
        std::printf("the e-e P is rebuilt inside every right-hand side (freq_update_ee_model < 0)\n");
    }
    // H_coh = H_BS + H_Ez (main.cpp:51-52, DenMat.cpp:7-14): H_BS = Btot . s with
    // Btot = B, or gs/2 B + Bso_k under alg_read_Bso, plus B/2 . l under needL
    // (electron.cpp:19-42); H_Ez = scale_Ez x ldbd_HEzmat.bin when the file exists (:44-61).
    // Under gfac_normal_dist gs becomes g_k and the orbital term is not added
    // (electron.cpp:25-29).
    const bool has_BS = std::sqrt(s.B[0] * s.B[0] + s.B[1] * s.B[1] + s.B[2] * s.B[2]) > 1e-10 || s.read_Bso;
    std::vector<double> bso;
    if (s.read_Bso) {
        bso.resize((size_t)nk * 3);
        check(dmd_ldbd_read_Bso(ldbd.c_str(), &info, bso.data()), "ldbd_Bso.bin");
    }
    std::vector<double> gfack;
    if (has_BS && s.gfac_normal) {
        gfack = gfac_normal_array(nk, s.gfac_mean, s.gfac_sigma, s.gfac_cap);
        const std::string fn = base + "/gfack_normal.out";
        FILE *gf = std::fopen(fn.c_str(), "w");
        if (!gf) throw std::runtime_error("cannot open " + fn);
        std::fprintf(gf, "gfack:\n");
        for (long k = 0; k < nk; k++) std::fprintf(gf, "%14.10lf\n", gfack[k]);
        std::fprintf(gf, "\n");
        std::fclose(gf);
    }
    std::vector<cplx> hez;
    if (std::fabs(s.scale_Ez) > 1e-10) {
        struct stat hs;
        if (stat((ldbd + "/ldbd_HEzmat.bin").c_str(), &hs) == 0) {
            hez.resize((size_t)nk * nbd2);
            check(dmd_ldbd_read_HEzmat(ldbd.c_str(), &info, (double *)hez.data()), "ldbd_HEzmat.bin");
        }
    }
    if (has_BS || !hez.empty()) {
        std::vector<cplx> H((size_t)nk * nbd2, cplx{0, 0});
        for (long k = 0; k < nk; k++) {
            double Bt[3];
            const double gk = s.gfac_normal ? gfack[k] : 2.0023193043625635;
            for (int c = 0; c < 3; c++)
                Bt[c] = s.read_Bso ? 0.5 * gk * s.B[c] + bso[(size_t)k * 3 + c] : s.B[c];
            for (size_t q = 0; q < nbd2; q++) {
                cplx v = {0, 0};
                if (has_BS)
                    for (int c = 0; c < 3; c++) {
                        v.x += Bt[c] * smat[((size_t)k * 3 + c) * nbd2 + q].x;
                        v.y += Bt[c] * smat[((size_t)k * 3 + c) * nbd2 + q].y;
                        if (s.needL && !s.gfac_normal) {
                            v.x += 0.5 * s.B[c] * lmat[((size_t)k * 3 + c) * nbd2 + q].x;
                            v.y += 0.5 * s.B[c] * lmat[((size_t)k * 3 + c) * nbd2 + q].y;
                        }
                    }
                if (!hez.empty()) {
                    v.x += s.scale_Ez * hez[(size_t)k * nbd2 + q].x;
                    v.y += s.scale_Ez * hez[(size_t)k * nbd2 + q].y;
                }
                H[(size_t)k * nbd2 + q] = v;
            }
        }
        check(dmd_scatt_set_coherent(h, (const double *)H.data()), "coherent H");
    }

    // reports (DensityMatrix_Dynamics.h:322-345, observable.h:131-403)
    std::vector<Report> reps;
    const char *cn[3] = {"sx", "sy", "sz"};
    for (int c = 0; c < 3; c++) {
        if (nbd > nvd) reps.push_back({std::string(cn[c]) + "_elec", c, nvd, nbd});
        if (nvd > 0) reps.push_back({std::string(cn[c]) + "_hole", c, 0, nvd});
    }
    const char *ln[3] = {"lx", "ly", "lz"};
    for (int c = 0; s.needL && c < 3; c++) {  // DensityMatrix_Dynamics.h:333-335
        if (nbd > nvd) reps.push_back({std::string(ln[c]) + "_elec", c, nvd, nbd, nullptr, true});
        if (nvd > 0) reps.push_back({std::string(ln[c]) + "_hole", c, 0, nvd, nullptr, true});
    }
    std::vector<cplx> sc((size_t)nk * nbd2);
    for (auto &r : reps) {
        const std::string fn = base + "/" + r.name + "_tot.out";
        struct stat st;
        const bool exists = stat(fn.c_str(), &st) == 0;
        r.fp = std::fopen(fn.c_str(), "a");
        if (!r.fp) throw std::runtime_error("cannot open " + fn);
        if (!exists) std::fprintf(r.fp, r.orbital ? "#time(au), l(t), l(t) without exp(iwt)\n"
                                                  : "#time(au), s(t), s(t) without exp(iwt)\n");
    }
    mkdir_p(base + "/restart");
    auto report = [&](int it) {
        if (it % s.freq_measure != 0) return;
        if (it > 0) check(dmd_restart_write((base + "/restart").c_str(), nk, nbd, (const double *)dm.data(), t), "restart");
        for (auto &r : reps) {
            const std::vector<cplx> &om = r.orbital ? lmat : smat;
            for (long k = 0; k < nk; k++)
                std::memcpy(&sc[(size_t)k * nbd2], &om[((size_t)k * 3 + r.comp) * nbd2], nbd2 * sizeof(cplx));
            check(dmd_scatt_set_observable(h, (const double *)sc.data(), fdm.data(), r.i0, r.i1, prefac_obs), "observable");
            double v = 0, amp = 0;
            check(dmd_scatt_measure(h, t, (const double *)dm.data(), &v), "measure");
            check(dmd_scatt_measure(h, 0.0, (const double *)dm.data(), &amp), "measure");
            if (std::fabs(t) < 1e-30) std::fprintf(r.fp, " 0.00000000000000e+01 %21.14le %21.14le\n", v, amp);
            else std::fprintf(r.fp, "%21.14le %21.14le %21.14le\n", t, v, amp);
            std::fflush(r.fp);
        }
    };
    // instantaneous relaxation times (report_tau, DensityMatrix_Dynamics.h:346-354;
    // observable.h:254-274,392-402): tau = -S / dS/dt from the right-hand side
    // at t, without and with the precession term of the interaction picture,
    // and by a forward difference over dt_tiny = min(ode_hstart, tstep / 10)
    std::vector<Report> taus;
    const double Bp = std::sqrt(s.Bpert[0] * s.Bpert[0] + s.Bpert[1] * s.Bpert[1] + s.Bpert[2] * s.Bpert[2]);
    for (auto &r : reps) {
        if (r.orbital || (Bp > 1e-10 && s.Bpert[r.comp] == 0)) continue;  // spin only
        taus.push_back({"tau_" + r.name, r.comp, r.i0, r.i1});
        const std::string fn = base + "/" + taus.back().name + "_tot.out";
        struct stat st2;
        const bool exists = stat(fn.c_str(), &st2) == 0;
        taus.back().fp = std::fopen(fn.c_str(), "a");
        if (!taus.back().fp) throw std::runtime_error("cannot open " + fn);
        if (!exists) std::fprintf(taus.back().fp, "#time(au), tau by FD, tau without precession, tau\n");
    }
    std::vector<cplx> ddm((size_t)nk * nbd2), dmp((size_t)nk * nbd2), s2((size_t)nk * nbd2);
    const std::vector<double> fzero((size_t)nk * nbd, 0.0);
    const double dt_tiny = std::min(s.hstart, s.tstep / 10);
    auto report_tau = [&](int it) {
        if (it > 0 && (it - 1) % s.freq_compute_tau != 0) return;
        if (taus.empty()) return;
        check(dmd_scatt_compute_rhs(h, t, (const double *)dm.data(), (double *)ddm.data()), "compute");
        for (size_t q = 0; q < dm.size(); q++) dmp[q] = {dm[q].x + ddm[q].x * dt_tiny, dm[q].y + ddm[q].y * dt_tiny};
        for (auto &r : taus) {
            for (long k = 0; k < nk; k++) {
                std::memcpy(&sc[(size_t)k * nbd2], &smat[((size_t)k * 3 + r.comp) * nbd2], nbd2 * sizeof(cplx));
                for (int i = 0; i < nbd; i++)
                    for (int b2 = 0; b2 < nbd; b2++) {  // 2i e_i s_ib: Re(2i e_i c) = -2 Im(e_i c)
                        const cplx v = sc[(size_t)k * nbd2 + i * nbd + b2];
                        const double ei = edm[(size_t)k * nbd + i];
                        s2[(size_t)k * nbd2 + i * nbd + b2] = {-2 * ei * v.y, 2 * ei * v.x};
                    }
            }
            double tot = 0, dot = 0, term2 = 0, tplus = 0;
            check(dmd_scatt_set_observable(h, (const double *)sc.data(), fdm.data(), r.i0, r.i1, prefac_obs), "observable");
            check(dmd_scatt_measure(h, t, (const double *)dm.data(), &tot), "measure");
            check(dmd_scatt_measure(h, t + dt_tiny, (const double *)dmp.data(), &tplus), "measure");
            check(dmd_scatt_set_observable(h, (const double *)sc.data(), fzero.data(), r.i0, r.i1, prefac_obs), "observable");
            check(dmd_scatt_measure(h, t, (const double *)ddm.data(), &dot), "measure");
            if (interaction) {
                check(dmd_scatt_set_observable(h, (const double *)s2.data(), fdm.data(), r.i0, r.i1, prefac_obs),
                      "observable");
                check(dmd_scatt_measure(h, t, (const double *)dm.data(), &term2), "measure");
            }
            const double tau_fd = -tot * dt_tiny / (tplus - tot);
            const double psu = ps();
            if (interaction) {
                if (std::fabs(t) < 1e-30)
                    std::fprintf(r.fp, " 0.0000000000e+01 fs: %14.7le %14.7le %14.7le ps\n", tau_fd / psu, -tot / dot / psu,
                                 -tot / (dot + term2) / psu);
                else
                    std::fprintf(r.fp, "%17.10le fs: %14.7le %14.7le %14.7le ps\n", t / fs(), tau_fd / psu, -tot / dot / psu,
                                 -tot / (dot + term2) / psu);
            } else {
                if (std::fabs(t) < 1e-30) std::fprintf(r.fp, " 0.0000000000e+01 fs: %14.7le ps\n", -tot / dot / psu);
                else std::fprintf(r.fp, "%17.10le fs: %14.7le ps\n", t / fs(), -tot / dot / psu);
            }
            std::fflush(r.fp);
        }
    };
    if (!s.restart) report(0);
    if (s.tau_only) {  // DensityMatrix_Dynamics.h:75 and main.cpp:79: taus at t0, no evolution
        report_tau(0);
        for (auto &r : reps) std::fclose(r.fp);
        for (auto &r : taus) std::fclose(r.fp);
        return 0;
    }
    // the time loop of evolve_gsl / evolve_euler (DensityMatrix_Dynamics.h:177-225)
    double ti = t, hstep = s.hstart;
    dmd_ode_stats st{}, tot{};
    const std::string exit_file = base + "/EXIT_DMD";
    for (int it = 1; t < s.tend; it++) {
        struct stat xs;
        if (stat(exit_file.c_str(), &xs) == 0) {  // "touch EXIT_DMD" stops the run (DensityMatrix_Dynamics.h:244)
            std::remove(exit_file.c_str());
            std::printf("dmd_run: EXIT_DMD found at t = %.10le, clean exit\n", t);
            for (auto &r : reps) std::fclose(r.fp);
            for (auto &r : taus) std::fclose(r.fp);
            return 0;
        }
        // update_scatt_outside (DensityMatrix_Dynamics.h:282-292,316-318): the e-e P
        // rebuilt from the state at it = 1, 1 + freq, ...; evolve_euler_one_step
        // rebuilds before report_tau, evolve_gsl after it
        auto update_ee = [&]() {
            if (!pee_update || s.freq_update_ee <= 0 || (it - 1) % s.freq_update_ee != 0) return;
            cplx *rows_dm[1] = {dm.data()};
            double *rows_feq[1] = {feq.data()};
            eph.reset_scatt(false, true, rows_dm, nullptr, t, rows_feq);
        };
        if (s.ode == "euler") update_ee();
        report_tau(it);
        if (s.ode != "euler") update_ee();
        ti += s.tstep;
        if (s.ode == "rkf45") {
            check(dmd_scatt_rkf45_apply(h, &t, ti, &hstep, s.hmin, s.hmax, s.epsabs, 0.0, (double *)dm.data(), &st),
                  "rkf45");
            tot.steps += st.steps;
            tot.rejected += st.rejected;
            tot.rhs_evals += st.rhs_evals;
        } else if (s.ode == "euler") {
            check(dmd_scatt_integrate_euler(h, t, s.tstep, 1, (double *)dm.data()), "euler");
            tot.rhs_evals += 1;
        } else {
            check(dmd_scatt_integrate_rk4(h, t, h_rk, n_sub, 0, (double *)dm.data(), nullptr), "RK4");
            tot.rhs_evals += 4 * n_sub;
        }
        t = ti;
        report(it);
    }
    std::printf("dmd_run: %lld right-hand-side evaluations", (long long)tot.rhs_evals);
    if (s.ode == "rkf45") std::printf(" (%lld steps, %lld rejected)", (long long)tot.steps, (long long)tot.rejected);
    std::printf("\n");
    for (auto &r : reps) std::fclose(r.fp);
    for (auto &r : taus) std::fclose(r.fp);
    std::printf("dmd_run: reached t = %.10le\n", t);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    bool dry = false;
    std::string folder;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--dry-run") dry = true;
        else if (a == "-h" || a == "--help") {
            std::printf("usage: dmd_run [--dry-run] [folder with param.in and ldbd_data/]\n");
            return 0;
        } else folder = a;
    }
    try {
        return run(folder, dry);
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "dmd_run: %s\n", ex.what());
        return 1;
    }
}
