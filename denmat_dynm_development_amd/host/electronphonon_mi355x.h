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
#include <functional>
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
          scissor_(scissor), precision_(precision), P_set_(source != P_NONE)
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
        if (base_) dmd_scatt_destroy(base_);
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
        keep_inputs(k1st, k2nd, U, kvec, f, m.nb_wannier);
        eimp_models_.push_back({iD, m, scale_ei});
        if (Pin_is_sparse) check(dmd_scatt_threshold_P(h_, thr_sparseP), "sparsification");
        if ((int)eimp_slots.size() <= iD) eimp_slots.resize(iD + 1);
        eimp.clear();
        for (auto &s : eimp_slots) eimp.push_back(&s);
    }

    // add_scatt_contrib("ee") (ElectronPhonon.cpp:11,44-113) with eeMode =
    // Pee_fixed_at_eq: P += scale_ee * P_ee(f), the model e-e P at the
    // occupations f (nk x nb of this window) generated on the device; alg_only_ee
    // (no e-ph P set) makes it P itself
    // vq_cell: the static RPA v(q) of every k-mesh cell (dmd_scatt_ee_vq_cells);
    // empty: the closed-form screening, or the RPA summed for the stored pairs' q
    void add_ee_model(const uint64_t *k1st, const uint64_t *k2nd, const complex_t *U, const double *kvec,
                      const double *f, const dmd_eimp_model &m, double scale_ee, const std::vector<double> &vq_cell = {})
    {
        if (vq_cell.empty())
            check(dmd_scatt_add_ee_model(h_, k1st, k2nd, (const double *)U, kvec, f, &m, scale_ee, P_set_ ? 1 : 0),
                  "model e-e P");
        else
            check(dmd_scatt_add_ee_model_state(h_, k1st, k2nd, (const double *)U, kvec, f, nullptr, nullptr,
                                               vq_cell.data(), &m, scale_ee, P_set_ ? 1 : 0),
                  "model e-e P");
        P_set_ = true;
        if (Pin_is_sparse) check(dmd_scatt_threshold_P(h_, thr_sparseP), "sparsification");
        keep_inputs(k1st, k2nd, U, kvec, nullptr, m.nb_wannier);
        ee_f_.assign(f, f + (size_t)nk_glob * nb);
        has_ee_ = true;
        ee_m_ = m;
        scale_ee_ = scale_ee;
    }

    // The Pee_update set-up: the P of the parts reset_scatt does not rebuild
    // (the e-ph P and ab initio e-i, set_ephmat and read_ldbd_imp_P,
    // ElectronPhonon.cpp:34-38) is kept in a second handle before the model
    // terms are added.  update_ee: rebuild the e-e P from the state on
    // reset_scatt (eeMode = Pee_update); vq_cell: the static RPA v(q) per mesh
    // cell for the e-e term (empty for the closed-form screenings).
    void keep_base(bool update_ee, const std::vector<double> &vq_cell = {})
    {
        if (base_) return;
        std::vector<uint64_t> k1(info_.nkpair), k2(info_.nkpair);
        check(dmd_ldbd_read_kpairs(dir_.c_str(), &info_, 0, info_.nkpair, k1.data(), k2.data()), "k pairs");
        std::vector<double> ek((size_t)info_.nk * info_.nb), e((size_t)info_.nk * nb);
        check(dmd_ldbd_read_ek(dir_.c_str(), &info_, scissor_, ek.data()), "ldbd_ek.bin");
        const int b0 = info_.b_bot_eph;
        for (long k = 0; k < nk_glob; k++)
            for (int b = 0; b < nb; b++) e[(size_t)k * nb + b] = ek[(size_t)k * info_.nb + b0 + b];
        e_window_ = e;
        check(dmd_scatt_create((int64_t)nk_glob, nb, (int64_t)info_.nkpair, k1.data(), k2.data(), e.data(), info_.nk_full,
                               precision_, &base_),
              "base P handle");
        check(dmd_scatt_copy_P(base_, h_, 1.0, 0), "base P copy");
        update_ee_ = update_ee;
        ee_vq_cell_ = vq_cell;
    }
    // coulomb_model::init(dm) under update_screening (Coulomb_Model.h:533-552):
    // given the occupations of this window (nk x nb, the state's diagonal) it
    // re-derives a closed-form screening of model m; unset: the screening stays.
    std::function<void(const std::vector<double> &occ, dmd_eimp_model &m)> rescreen;

    // The linearized operator's e-e terms beyond the P1/P2 form
    // (add_Lsc_contrib("ee"), ElectronPhonon_linear.cpp:38-104,183-253), after
    // set_Lsc: the fac2 doubling under ee_antisymmetry and, with dP_terms (eeMode
    // other than Pee_fixed_at_eq), P3/P4 and P5/P6 at f_eq (nk x nb of this window).
    // coef34 = 1 keeps the reference's sign of the P3/P4 terms.
    void add_Lsc_ee(const double *f_eq, bool dP_terms, double coef34 = 1.0, const std::vector<double> &vq_cell = {})
    {
        if (!has_ee_) return;
        check(dmd_scatt_add_Lsc_ee(h_, mk1_.data(), mk2_.data(), (const double *)mU_.data(), mkv_.data(), f_eq,
                                   vq_cell.empty() ? nullptr : vq_cell.data(), &ee_m_, scale_ee_,
                                   ee_m_.exchange ? 1.0 : 0.0, coef34, dP_terms ? 1 : 0),
              "linearized e-e terms");
    }

    // set_sparseP(true) at the end of set_eph (ElectronPhonon.cpp:12,115-129)
    // under alg_sparseP: P with |P| <= thr_sparseP zeroed and, with
    // write_files, the sparse sP files written next to ldbd_data's P files
    // (an existing _s file is kept).
    void set_sparseP(bool write_files = true)
    {
        if (Pin_is_sparse) return;
        sparse_on_ = true;
        check(dmd_scatt_threshold_P(h_, thr_sparseP), "sparsification");
        if (write_files) check(dmd_ldbd_write_sP(h_, dir_.c_str(), &info_, scatt_.c_str(), thr_sparseP), "sP files");
    }

    // The sparse modes' per-pair element lists (evolve_eph.cpp:69-86): P kept by
    // its nonzero elements so the bytes held and streamed fall with the density
    // (nb = 2 FP64 tile layout; false where the store stays dense)
    bool compress_P()
    {
        if (nb != 2) return false;
        int64_t np = 0, before = 0, after = 0;
        double ms = 0;
        check(dmd_scatt_info(h_, &np, &before, &ms), "info");
        const int rc = dmd_scatt_compress_P(h_, nullptr);
        if (rc == DMD_ERR_UNSUPPORTED) return false;  // another layout or precision: the store stays dense
        check(rc, "compressed P");
        check(dmd_scatt_info(h_, &np, &after, &ms), "info");
        return after != before;
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
        interaction_ = !schrodinger;
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

    // reset_scatt (ElectronPhonon.cpp:15-42): with a base kept (keep_base) and
    // either flag set, P = base + every model e-i term (screening re-derived
    // from the state when rescreen is set) + the e-e term at the state
    // (Pee_update) or at f, and the re-sparsification of alg_sparseP.  The
    // e-e term takes the window of dm (and dm1; null: 1 - dm) turned from the
    // interaction to the Schroedinger picture as add_scatt_contrib("ee") does
    // (:56-65): the upper triangle times cis((e_j - e_i) t), the lower its
    // conjugate, dm1 = -dm off the diagonal.  Under linearize_dPee the e-e P
    // is taken at f_eq_expand (required; cut to the window as :28-31 does) and
    // its change dP towards the state is built, whose row term every later
    // evolve adds (calc_P with dP, ElectronPhonon.cpp:84; evolve_eph.cpp:107-148).
    void reset_scatt(bool reset_eimp, bool reset_ee, complex_t **dm_expand, complex_t **dm1_expand, double t,
                     double **f_eq_expand)
    {
        if ((!reset_eimp && !reset_ee) || !base_ || (eimp_models_.empty() && !has_ee_)) return;
        if (linearize_dPee && has_ee_ && !f_eq_expand)
            throw std::runtime_error("reset_scatt: alg_linearize_dPee needs f_eq");
        const long nk = nk_glob;
        const int nb2 = nb * nb, nbx = nb_expand;
        std::vector<double> dm((size_t)nk * nb2 * 2), dm1((size_t)nk * nb2 * 2), occ((size_t)nk * nb);
        const double *src = (const double *)dm_expand[0], *src1 = dm1_expand ? (const double *)dm1_expand[0] : nullptr;
        for (long k = 0; k < nk; k++)
            for (int i = 0; i < nb; i++)
                for (int j = 0; j < nb; j++) {
                    const size_t o = ((size_t)k * nbx + bStart + i) * nbx + bStart + j, q = (size_t)k * nb2 + i * nb + j;
                    dm[2 * q] = src[2 * o];
                    dm[2 * q + 1] = src[2 * o + 1];
                    dm1[2 * q] = src1 ? src1[2 * o] : (i == j ? 1.0 : 0.0) - src[2 * o];
                    dm1[2 * q + 1] = src1 ? src1[2 * o + 1] : -src[2 * o + 1];
                    if (i == j) occ[(size_t)k * nb + i] = src[2 * o];
                }
        if (interaction_)
            for (long k = 0; k < nk; k++)
                for (int i = 0; i < nb; i++)
                    for (int j = i + 1; j < nb; j++) {
                        const size_t q = (size_t)k * nb2 + i * nb + j, r = (size_t)k * nb2 + j * nb + i;
                        const double a = (e_window_[(size_t)k * nb + j] - e_window_[(size_t)k * nb + i]) * t;
                        const double c = std::cos(a), sn = std::sin(a), x = dm[2 * q], y = dm[2 * q + 1];
                        dm[2 * q] = x * c - y * sn;
                        dm[2 * q + 1] = x * sn + y * c;
                        dm1[2 * q] = -dm[2 * q];
                        dm1[2 * q + 1] = -dm[2 * q + 1];
                        dm[2 * r] = dm[2 * q];
                        dm[2 * r + 1] = -dm[2 * q + 1];
                        dm1[2 * r] = dm1[2 * q];
                        dm1[2 * r + 1] = -dm1[2 * q + 1];
                    }
        check(dmd_scatt_copy_P(h_, base_, 1.0, 0), "P reset");
        for (auto &em : eimp_models_) {
            dmd_eimp_model m = em.m;
            if (rescreen) rescreen(occ, m);
            check(dmd_scatt_add_eimp_model(h_, mk1_.data(), mk2_.data(), (const double *)mU_.data(), mkv_.data(),
                                           mf_.empty() ? nullptr : mf_.data(), &m, em.scale, 1),
                  "model e-i P of defect " + std::to_string(em.iD + 1));
        }
        if (has_ee_ && linearize_dPee) {
            dmd_eimp_model m = ee_m_;
            if (rescreen) rescreen(occ, m);
            std::vector<double> feq((size_t)nk * nb);
            for (long k = 0; k < nk; k++)
                for (int i = 0; i < nb; i++) feq[(size_t)k * nb + i] = f_eq_expand[0][(size_t)k * nbx + bStart + i];
            const double *vc = ee_vq_cell_.empty() ? nullptr : ee_vq_cell_.data();
            check(dmd_scatt_add_ee_model_state(h_, mk1_.data(), mk2_.data(), (const double *)mU_.data(), mkv_.data(),
                                               feq.data(), nullptr, nullptr, vc, &m, scale_ee_, 1),
                  "model e-e P at f_eq");
            check(dmd_scatt_set_ee_dP(h_, mk1_.data(), mk2_.data(), (const double *)mU_.data(), mkv_.data(), feq.data(),
                                      dm.data(), dm1.data(), vc, &m),
                  "linearize_dPee: dP of the model e-e P");
        } else if (has_ee_) {
            dmd_eimp_model m = ee_m_;
            if (rescreen) rescreen(occ, m);
            const bool at_state = update_ee_ && reset_ee;
            check(dmd_scatt_add_ee_model_state(h_, mk1_.data(), mk2_.data(), (const double *)mU_.data(), mkv_.data(),
                                               ee_f_.data(), at_state ? dm.data() : nullptr, at_state ? dm1.data() : nullptr,
                                               ee_vq_cell_.empty() ? nullptr : ee_vq_cell_.data(), &m, scale_ee_, 1),
                  "model e-e P");
        }
        if (Pin_is_sparse || sparse_on_) check(dmd_scatt_threshold_P(h_, thr_sparseP), "sparsification");
    }

    // alg_linearize_dPee (set before the first reset_scatt)
    bool linearize_dPee = false;

    dmd_scatt *handle() const { return h_; }
    const dmd_ldbd_info &info() const { return info_; }

private:
    struct eimp_entry {
        int iD;
        dmd_eimp_model m;
        double scale;
    };
    // the model terms' inputs, kept for reset_scatt (pairs, U, kvec, f at mu)
    void keep_inputs(const uint64_t *k1st, const uint64_t *k2nd, const complex_t *U, const double *kvec,
                     const double *f, int nw)
    {
        if (f && mf_.empty()) mf_.assign(f, f + (size_t)nk_glob * nb);
        if (!mk1_.empty() || info_.nkpair == 0) return;
        mk1_.assign(k1st, k1st + info_.nkpair);
        mk2_.assign(k2nd, k2nd + info_.nkpair);
        mU_.assign(U, U + (size_t)nk_glob * nw * nb);
        mkv_.assign(kvec, kvec + (size_t)nk_glob * 3);
    }
    std::string dir_, scatt_;
    double scissor_ = 0.0;
    int precision_ = DMD_PREC_FP64;
    bool P_set_ = true;
    bool interaction_ = true, sparse_on_ = false;
    dmd_ldbd_info info_{};
    dmd_scatt *h_ = nullptr;
    dmd_scatt *base_ = nullptr;
    std::vector<uint64_t> mk1_, mk2_;
    std::vector<complex_t> mU_;
    std::vector<double> mkv_, mf_, ee_f_, e_window_, ee_vq_cell_;
    std::vector<eimp_entry> eimp_models_;
    bool has_ee_ = false, update_ee_ = false;
    dmd_eimp_model ee_m_{};
    double scale_ee_ = 1.0;
};

}  // namespace dmdrun
