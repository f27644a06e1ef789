// electronphonon_mi355x.h -- a Teph for the reference's dm_dynamics
// (src_v4.5.7/DensityMatrix_Dynamics.h:15-25, instantiated in main.cpp:73-74):
// the members the driver touches (DensityMatrix_Dynamics.h:59-63,70,266-319;
// signatures ElectronPhonon.h:150-165) forwarded to the C-ABI of libdmd_scatt.
// complex_t is the caller's complex type, two doubles (src_v4.5.7/comm/
// scalar.h:67-69); the arrays are the caller's row-pointer tables whose rows
// sit in one contiguous pool (src_v4.5.7/comm/myarray.cpp:145-155), so row 0
// is the flat nk x nb^2 buffer the C-ABI takes.  Errors throw
// std::runtime_error with the library's message where the reference exits.
#pragma once
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "dmd_scatt.h"

namespace dmdrun {

inline void check(int rc, const std::string &what)
{
    if (rc != DMD_OK) throw std::runtime_error(what + ": " + dmd_scatt_last_error());
}

// dm_dynamics reads eimp[iD]->eimp_model (DensityMatrix_Dynamics.h:306-314); P
// here is fixed after set-up, so no defect carries a model to update.
struct impurity_slot {
    void *eimp_model = nullptr;
};

template <class complex_t>
class electronphonon_mi355x {
    static_assert(sizeof(complex_t) == 16, "complex must be two doubles");

public:
    // e-ph window relative to the density matrix's first band, as the
    // reference's bStart/bEnd (ElectronPhonon.cpp:132-158)
    int bStart = 0, bEnd = 0, nb_expand = 0, nb = 0;
    long nk_glob = 0, nkpair_glob = 0;
    double prefac_eph = 0;
    std::vector<impurity_slot> eimp_slots;
    std::vector<impurity_slot *> eimp;
    void *ee_model = nullptr;
    // where the e-ph P comes from: the ldbd_P files, the sparse sP files
    // (alg_Pin_is_sparse, ElectronPhonon.cpp:235-246), none (alg_eph_enable =
    // 0: the e-i fold-in fills P, :211), or no scattering at all
    // (alg_scatt_enable = 0, ElectronPhonon.h:72)
    enum p_source { P_DENSE = 0, P_SPARSE = 1, P_NONE = 2, P_NO_SCATTERING = 3 };
    // alg.Pin_is_sparse and alg.thr_sparseP (parameters.cpp:118-120)
    bool Pin_is_sparse = false;
    double thr_sparseP = 1e-40;

    // electronphonon's set-up from ldbd_data (ElectronPhonon.h:53-93 and
    // set_eph, ElectronPhonon.cpp:3-13): P x scale_eph streamed into HBM.
    explicit electronphonon_mi355x(const std::string &ldbd_dir, const std::string &scatt = "lindblad",
                                   double scale_eph = 1.0, double scissor = 0.0, int precision = DMD_PREC_FP64,
                                   p_source source = P_DENSE, double thr_sparseP_ = 1e-40)
        : Pin_is_sparse(source == P_SPARSE), thr_sparseP(thr_sparseP_), dir_(ldbd_dir), scatt_(scatt),
          P_set_(source != P_NONE)
    {
        check(dmd_ldbd_read_info(dir_.c_str(), &info_), "ldbd_size.dat");
        int rc;
        if (source == P_SPARSE)
            rc = dmd_scatt_create_from_ldbd_sparse(dir_.c_str(), scatt_.c_str(), scale_eph, scissor, precision, &h_);
        else if (source == P_DENSE)
            rc = dmd_scatt_create_from_ldbd(dir_.c_str(), scatt_.c_str(), scale_eph, scissor, precision, &h_);
        else
            rc = dmd_scatt_create_from_ldbd_noP(dir_.c_str(), scissor, precision, source == P_NO_SCATTERING, &h_);
        check(rc, "scattering operator from " + dir_);
        bStart = info_.b_bot_eph - info_.b_bot_dm;
        bEnd = info_.b_top_eph - info_.b_bot_dm;
        nb_expand = info_.b_top_dm - info_.b_bot_dm;
        nb = bEnd - bStart;
        nk_glob = (long)info_.nk;
        nkpair_glob = (long)info_.nkpair;
        prefac_eph = 2 * M_PI / info_.nk_full;
    }
    ~electronphonon_mi355x()
    {
        if (h_) dmd_scatt_destroy(h_);
    }
    electronphonon_mi355x(const electronphonon_mi355x &) = delete;
    electronphonon_mi355x &operator=(const electronphonon_mi355x &) = delete;

    // add_scatt_contrib("eimp", iD) of an ab initio defect (ElectronPhonon.cpp:
    // 44-113 with read_ldbd_imp_P, ElectronImpurity.cpp:3-48): P += fraction *
    // scale_ei * P_D<iD+1>, fraction = |n_i cell_size| (ElectronImpurity.h:17-18);
    // under Pin_is_sparse the sum is re-sparsified at thr_sparseP (:91-94).
    void add_eimp_ab_neutral(int iD, double fraction, double scale_ei)
    {
        // without e-ph P the first defect sets P (P = 0 + P_ei, as alg_eph_enable = 0 leaves it)
        check(dmd_scatt_load_P(h_, dir_.c_str(), &info_, scatt_.c_str(), iD + 1, fraction * scale_ei, P_set_ ? 1 : 0),
              "e-i P of defect " + std::to_string(iD + 1));
        P_set_ = true;
        if (Pin_is_sparse) check(dmd_scatt_threshold_P(h_, thr_sparseP), "sparsification");
        if ((int)eimp_slots.size() <= iD) eimp_slots.resize(iD + 1);
        eimp.clear();
        for (auto &s : eimp_slots) eimp.push_back(&s);
    }

    // add_scatt_contrib("eimp", iD) with impMode = model_ionized
    // (ElectronPhonon.cpp:44-113, elecimp_model::calc_P, ElecImp_Model.h:185-213):
    // P += scale_ei * P_imp, P_imp generated on the device from the overlaps
    // U_k^H U_k' (U: nk x nb_wannier x nb of this window), the k points and the
    // screened Coulomb model m; re-sparsified under Pin_is_sparse as above.
    // f: nk x nb occupations of this window at mu, for the RPA / lindhard screening (else null)
    void add_eimp_model(int iD, const uint64_t *k1st, const uint64_t *k2nd, const complex_t *U, const double *kvec,
                        const double *f, const dmd_eimp_model &m, double scale_ei)
    {
        check(dmd_scatt_add_eimp_model(h_, k1st, k2nd, (const double *)U, kvec, f, &m, scale_ei, P_set_ ? 1 : 0),
              "model e-i P of defect " + std::to_string(iD + 1));
        P_set_ = true;
        if (Pin_is_sparse) check(dmd_scatt_threshold_P(h_, thr_sparseP), "sparsification");
        if ((int)eimp_slots.size() <= iD) eimp_slots.resize(iD + 1);
        eimp.clear();
        for (auto &s : eimp_slots) eimp.push_back(&s);
    }

    // add_scatt_contrib("ee") (ElectronPhonon.cpp:11,44-113) with eeMode =
    // Pee_fixed_at_eq: P += scale_ee * P_ee(f), the model e-e P at the
    // occupations f (nk x nb of this window) generated on the device; alg_only_ee
    // (no e-ph P set) makes it P itself
    void add_ee_model(const uint64_t *k1st, const uint64_t *k2nd, const complex_t *U, const double *kvec,
                      const double *f, const dmd_eimp_model &m, double scale_ee)
    {
        check(dmd_scatt_add_ee_model(h_, k1st, k2nd, (const double *)U, kvec, f, &m, scale_ee, P_set_ ? 1 : 0),
              "model e-e P");
        P_set_ = true;
        if (Pin_is_sparse) check(dmd_scatt_threshold_P(h_, thr_sparseP), "sparsification");
    }

    // set_sparseP(true) at the end of set_eph (ElectronPhonon.cpp:12,115-129)
    // under alg_sparseP: P with |P| <= thr_sparseP zeroed and, with
    // write_files, the sparse sP files written next to ldbd_data's P files
    // (an existing _s file is kept).
    void set_sparseP(bool write_files = true)
    {
        if (Pin_is_sparse) return;
        check(dmd_scatt_threshold_P(h_, thr_sparseP), "sparsification");
        if (write_files) check(dmd_ldbd_write_sP(h_, dir_.c_str(), &info_, scatt_.c_str(), thr_sparseP), "sP files");
    }

    // alg.semiclassical (parameters.cpp:101): the driver-side right-hand side
    // keeps rho diagonal (DensityMatrix_Dynamics.h:249,277)
    void set_semiclassical(bool on) { check(dmd_scatt_set_semiclassical(h_, on ? 1 : 0), "semiclassical"); }

    // phenom_relax (scattering_model/phenomenon_relax.h:14-51): -(rho - dm_eq)/tau
    // on density-matrix bands [bStart_tau, bEnd_tau); f_eq_expand is nk x nb_expand
    void set_phenom_relax(double tau, int bStart_tau, int bEnd_tau, double **f_eq_expand)
    {
        check(dmd_scatt_set_phenom_relax(h_, tau, bStart_tau, bEnd_tau, f_eq_expand ? f_eq_expand[0] : nullptr),
              "phenomenological relaxation");
    }

    // alg.expt and alg.picture (parameters.cpp:104,107)
    void set_algorithm(bool expt, bool schrodinger)
    {
        check(dmd_scatt_set_flags(h_, expt ? 1 : 0, schrodinger ? 1 : 0, DMD_KERNEL_AUTO), "flags");
    }

    // evolve_eph.cpp:21-32 (compute_eq is the library's own business: the
    // stored ddmdt_eq is subtracted whenever one is set)
    void evolve_driver(double t, complex_t **dm_expand, complex_t **dm1_expand, complex_t **ddmdt_expand,
                       bool compute_eq = false)
    {
        (void)compute_eq;
        check(dmd_scatt_evolve_driver(h_, t, (const double *)dm_expand[0],
                                      dm1_expand ? (const double *)dm1_expand[0] : nullptr, (double *)ddmdt_expand[0]),
              "evolve_driver");
    }

    // evolve_eph.cpp:34-166 on the e-ph window arrays (nk x nb^2)
    void evolve(double t, complex_t **dm, complex_t **dm1, complex_t **ddmdt, bool compute_eq = false)
    {
        (void)compute_eq;
        check(dmd_scatt_eval(h_, t, (const double *)dm[0], dm1 ? (const double *)dm1[0] : nullptr, (double *)ddmdt[0]),
              "evolve");
    }

    // evolve_eph.cpp:3-19: f0_expand is nk x nb_expand (the density matrix's bands)
    void compute_ddmdt_eq(double **f0_expand)
    {
        std::vector<double> f((size_t)nk_glob * nb);
        for (long k = 0; k < nk_glob; k++)
            for (int i = 0; i < nb; i++) f[(size_t)k * nb + i] = f0_expand[0][(size_t)k * nb_expand + bStart + i];
        check(dmd_scatt_compute_ddmdt_eq(h_, f.data()), "compute_ddmdt_eq");
    }

    // set_Lsc (ElectronPhonon_linear.cpp:3-36): alg_linearize at the nk x
    // nb_expand occupations f_eq_expand (cut to the e-ph window as :4 does)
    void set_Lsc(double **f_eq_expand)
    {
        std::vector<double> f((size_t)nk_glob * nb);
        for (long k = 0; k < nk_glob; k++)
            for (int i = 0; i < nb; i++) f[(size_t)k * nb + i] = f_eq_expand[0][(size_t)k * nb_expand + bStart + i];
        check(dmd_scatt_set_Lsc(h_, f.data()), "set_Lsc");
    }

    // ElectronPhonon.cpp:15-42: P is rebuilt only for model e-i and e-e
    // scattering, neither of which this operator carries
    void reset_scatt(bool, bool, complex_t **, complex_t **, double, double **) {}

    dmd_scatt *handle() const { return h_; }
    const dmd_ldbd_info &info() const { return info_; }

private:
    std::string dir_, scatt_;
    bool P_set_ = true;
    dmd_ldbd_info info_{};
    dmd_scatt *h_ = nullptr;
};

}  // namespace dmdrun
