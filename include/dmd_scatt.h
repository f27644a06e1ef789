/*
 * dmd_scatt.h -- C-ABI of the MI355X Lindblad scattering operator for DMD.
 *
 * Drop-in for the reference's scattering term dRho/dt|scatt
 * (electronphonon::evolve_driver / evolve, src_v4.5.7/evolve_eph.cpp:21-166,
 * and the initializer's per-pair ElectronPhonon::compute_ddm,
 * src_FeynWann_forDMD/lindbladInit_for-DMD-4.5.6/ElectronPhonon.h:855-862),
 * with the e-i term of the same form folded into P
 * (src_v4.5.7/ElectronPhonon.cpp:44-113).
 *
 * Conventions
 *   complex arrays  interleaved (re, im) float64, i.e. the reference's
 *                   struct complex {double x, y} (src_v4.5.7/comm/scalar.h:67-69);
 *   2-D arrays      flat row-major pools, dm[0] of the reference's row-pointer
 *                   tables (src_v4.5.7/comm/myarray.cpp:145-155);
 *   k-pairs         stored pairs k1st[i] <= k2nd[i] as in
 *                   ldbd_kpair_k1st/k2nd.bin (ElectronPhonon.cpp:182-204);
 *   P1, P2          n_pairs x nb^4 complex, element [(i1*nb+i2)*nb^2 + i3*nb+i4]
 *                   (ldbd_P1/P2_lindblad.bin, ElectronPhonon.cpp:209-233);
 *   time            Hartree atomic units (src_v4.5.7/comm/Units.h:62).
 * Pointers to state arrays (rho, ddmdt, ...) may be host or device memory;
 * the library detects which.  Every call returns 0 on success and a negative
 * DMD_ERR_* code otherwise; dmd_scatt_last_error() gives the message (thread
 * local).  One host thread per handle; calls are ordered on the handle's
 * HIP stream and return after the results are in the caller's buffers.
 */
#ifndef DMD_SCATT_H
#define DMD_SCATT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMD_OK 0
#define DMD_ERR_ARG (-1)
#define DMD_ERR_HIP (-2)
#define DMD_ERR_STATE (-3)
#define DMD_ERR_UNSUPPORTED (-4)
#define DMD_ERR_IO (-5)
#define DMD_ERR_NOMEM (-6)

/* precision of the streamed P */
#define DMD_PREC_FP64 0 /* complex128 P, FP64 arithmetic: the drop-in           */
#define DMD_PREC_C64 1  /* complex64 P, FP32 products, FP64 row sums and state (nb = 4) */
#define DMD_PREC_BF16 2 /* bf16 re/im P with a per-pair FP32 scale (nb = 4)           */

/* contraction kernel */
#define DMD_KERNEL_AUTO 0        /* fastest available for the shape               */
#define DMD_KERNEL_ROWS 1        /* deterministic: every row gathers its own pairs  */
#define DMD_KERNEL_SINGLE_READ 2 /* P read once; the reverse direction is summed  */
                                 /* with FP64 atomics (order not fixed), nb <= 2   */
#define DMD_KERNEL_TILES 3       /* P read once; both directions summed in LDS per  */
                                 /* (k, k') row-block tile, one atomic flush per   */
                                 /* tile chunk (order not fixed), nb = 2           */

typedef struct dmd_scatt dmd_scatt;

/* Synthetic ldbd_data description (energies in Hartree).  See dmd_synth_*. */
typedef struct {
    uint64_t seed;
    int nb;               /* bands in the e-ph window                          */
    int n_modes;          /* phonon branches                                   */
    int64_t nk;           /* k points                                          */
    int64_t partners;     /* partners drawn per k (uniform mode)               */
    int64_t partners_min; /* power-law mode bounds; partners_max = 0: uniform   */
    int64_t partners_max;
    double powerlaw_alpha;/* pdf ~ c^-alpha on [partners_min, partners_max]    */
    double e_max;         /* band-edge energies E_k ~ U(0, e_max) above CBM     */
    double split_max;     /* Kramers splitting ~ U(0, split_max)               */
    double gap_min, gap_max; /* second Kramers pair offset (nb = 4)            */
    double omega_min, omega_max; /* phonon energies per branch ~ U(.,.)         */
    double sigma_g;       /* e-ph coupling std dev per component                */
    double degauss;       /* gaussian smearing of the energy delta             */
    double temperature;   /* k_B T                                             */
    double sigma_imp;     /* e-i coupling std dev per component                */
    double imp_fraction;  /* |n_i V_cell| * scale_ei; 0 disables the e-i term   */
    uint32_t branch_mask; /* 0: every term; else bit a keeps phonon branch a and
                             bit 31 the e-i term (per-component P for gradients) */
} dmd_synth_params;

/* ---- errors --------------------------------------------------------------- */
const char *dmd_scatt_last_error(void);
int dmd_scatt_version(void);

/* ---- operator lifetime ------------------------------------------------------ */
/* Replaces electronphonon's constructor + alloc_ephmat + read_ldbd_kpair
 * (src_v4.5.7/ElectronPhonon.h:53-93, ElectronPhonon.cpp:182-204,255-304):
 * takes the stored k-pair lists and band energies e (nk x nb, the e-ph window
 * of e_dm, ElectronPhonon.h:84), builds the CSR structure-of-arrays layout
 * grouped by k and allocates HBM.  prefac = 2*pi/nk_full (ElectronPhonon.h:61). */
int dmd_scatt_create(int64_t nk, int nb, int64_t n_pairs, const uint64_t *k1st, const uint64_t *k2nd,
                     const double *e, double nk_full, int precision, dmd_scatt **out);
int dmd_scatt_destroy(dmd_scatt *h);

/* P = scale * (P1, P2), in ldbd_P file order; replaces read_ldbd_eph
 * (ElectronPhonon.cpp:209-233, scale = scale_eph). */
int dmd_scatt_upload_P(dmd_scatt *h, const double *P1, const double *P2, double scale);
/* P += scale * (P1, P2); the e-i fold-in add_scatt_contrib("eimp")
 * (ElectronPhonon.cpp:87-97; ab_neutral scale = |n_i V_cell| * scale_ei,
 * ElectronImpurity.cpp:38-39 and ElectronImpurity.h:17-18). */
int dmd_scatt_add_P(dmd_scatt *h, const double *P1, const double *P2, double scale);
/* Model ionized-impurity scattering (impMode = model_ionized): the screened
 * Coulomb matrix element of elecimp_model (scattering_model/ElecImp_Model.h:
 * 67-137) with the static screening of coulomb_model::vq (Coulomb_Model.h:
 * 632-662): the closed forms in |q| (unscreened, debye, Bechstedt) and the
 * static RPA / lindhard sums over the k mesh (calc_qscr2_static_RPA,
 * :690-744, and calc_vq_RPA, :820-828). */
typedef struct {
    int screening;        /* DMD_SCR_UNSCREENED, _DEBYE, _BECHSTEDT, _RPA or _LINDHARD   */
    int lindblad;         /* 1: alg_scatt = lindblad (A1 = A2), 0: conventional          */
    int detail_balance;   /* detailBalance: emission/absorption at +-omegaL              */
    int nb_wannier;       /* rows of each U_k                                            */
    double prefac_A;      /* Z sqrt(n_ionized V_cell)                                   */
    double prefac_vq;     /* 4 pi / (epsilon_background V_cell)                         */
    double prefac_vq_bare;/* 4 pi / V_cell                                               */
    double qscr2_debye;   /* 4 pi n_free / (epsilon_background T)                       */
    double fac0, fac2, fac4; /* Bechstedt: 1/(eps-1), 1/q_TF^2, 3/(4 k_F^2 q_TF^2)        */
    double degauss;       /* Gaussian width of the delta (degauss_eimp)                  */
    double ethr;          /* |de| cut-off, degauss * ndegauss of the e-ph set           */
    double omegaL;        /* 1e-4 T (detail_balance only)                                */
    double temperature;
    double GGT[9];        /* reciprocal metric (2 pi)^2 R^-1 R^-T                        */
    int kmesh[3];         /* k mesh of the k points (RPA and lindhard)                   */
    double nk_full;       /* k points of the full mesh (RPA and lindhard)                */
    int exchange;         /* e-e only: the exchange term (ee_antisymmetry)               */
} dmd_eimp_model;
#define DMD_SCR_UNSCREENED 0
#define DMD_SCR_DEBYE 1
#define DMD_SCR_BECHSTEDT 2
#define DMD_SCR_RPA 3
#define DMD_SCR_LINDHARD 4
/* P += scale * P_imp (add = 1) or P = scale * P_imp (add = 0) with P_imp of
 * every stored pair generated on the device: elecimp_model::calc_P
 * (ElecImp_Model.h:185-213) called from add_scatt_contrib("eimp")
 * (ElectronPhonon.cpp:44-113; scale = scale_ei).  k1st/k2nd are the stored
 * pairs in file order (host); U is nk x nb_wannier x nb complex128, the
 * e-ph band columns of ldbd_Umat.bin (dmd_ldbd_read_Umat); kvec nk x 3.
 * The overlap <k|k'> = U_k^H U_k', the screened v(q), the Gaussian delta
 * and the products A1 x conj(A2) are formed per element in one pass.  f
 * (nk x nb occupations of the e-ph bands at mu) enters the RPA and lindhard
 * screening only and may be null otherwise. */
int dmd_scatt_add_eimp_model(dmd_scatt *h, const uint64_t *k1st, const uint64_t *k2nd, const double *U,
                             const double *kvec, const double *f, const dmd_eimp_model *m, double scale, int add);
/* P += scale * P_ee (add = 1) or P = scale * P_ee (add = 0): the model e-e P
 * at the equilibrium occupations f (nk x nb, e-ph bands at mu) that set_eph
 * folds in through add_scatt_contrib("ee") (ElectronPhonon.cpp:10-11,44-113;
 * elecelec_model::calc_P with dm = null, ElecElec_Model.h:185-262), kept fixed
 * as eeMode = Pee_fixed_at_eq does; m->exchange adds the exchange term of
 * ee_antisymmetry (closed-form screenings only).
 * m: the screening as for dmd_scatt_add_eimp_model (unscreened, debye,
 * Bechstedt or RPA; prefac_A unused), degauss = degauss_ee, ethr, lindblad,
 * nb_wannier, kmesh (required: k4 = k1 + k3 - k2 is looked up on it), nk_full;
 * scale = scale_ee. */
int dmd_scatt_add_ee_model(dmd_scatt *h, const uint64_t *k1st, const uint64_t *k2nd, const double *U,
                           const double *kvec, const double *f, const dmd_eimp_model *m, double scale, int add);
/* The e-e rebuild of reset_scatt for eeMode = Pee_update (ElectronPhonon.cpp:
 * 15-42,44-113 with elecelec_model::calc_P at the density matrix,
 * ElecElec_Model.h:185-262): P (+)= scale * P_ee(dm, dm1), with dm, dm1 the
 * e-ph window of the state and of 1 - state in the Schroedinger picture
 * (nk x nb^2, Hermitian, read from the upper triangle as the reference's
 * zhemm does); dm = dm1 = NULL: at the occupations f (as
 * dmd_scatt_add_ee_model).  vq_cell: v(q) per k-mesh cell from
 * dmd_scatt_ee_vq_cells, required for the static RPA (unused otherwise). */
int dmd_scatt_add_ee_model_state(dmd_scatt *h, const uint64_t *k1st, const uint64_t *k2nd, const double *U,
                                 const double *kvec, const double *f, const double *dm, const double *dm1,
                                 const double *vq_cell, const dmd_eimp_model *m, double scale, int add);
/* alg_linearize_dPee.  The reference's reset_scatt builds, with the e-e P at
 * f_eq, the change dP1ee, dP2ee of that P towards the state (ElectronPhonon.cpp:
 * 28-31,84; elecelec_model::calc_P with dP, ElecElec_Model.h:185-262, whose
 * calc_A1rho_dP, :271-281, takes f_eq from the state's rows: ddm[i][j] =
 * dm[i][j] - f_eq[i] for every j, read Hermitian from the upper triangle), and
 * evolve adds (evolve_eph.cpp:107-131,138,144-148)
 *   ddmdt_contrib[ij] += ((1 - f_eq_i) dPdm_ij - dm1dP_ij f_eq_j) cis((e_i - e_j) t),
 *   dPdm_k[i12] = sum_k' sum_i3 dP1[(i12),(i3 i3)] f_eq_k'[i3] (and the conjugate
 *   reverse terms for k < k'), dm1dP_k likewise from dP2 with 1 - f_eq,
 * unscaled by scale_ee, and not when computing ddmdt_eq.
 * dmd_scatt_add_ee_dP: P (+)= scale * dP of every stored pair (the dP1ee,
 * dP2ee tensors themselves).  f_eq: nk x nb occupations of the window; dm, dm1:
 * the state and 1 - state in the Schroedinger picture (nk x nb^2); vq_cell,
 * m: as for dmd_scatt_add_ee_model_state. */
int dmd_scatt_add_ee_dP(dmd_scatt *h, const uint64_t *k1st, const uint64_t *k2nd, const double *U, const double *kvec,
                        const double *f_eq, const double *dm, const double *dm1, const double *vq_cell,
                        const dmd_eimp_model *m, double scale, int add);
/* dmd_scatt_set_ee_dP: builds dP at the state and keeps its row term on the
 * handle; every later evaluation (dmd_scatt_eval, evolve_driver, the
 * integrators) adds prefac/2 (D + D^+)_ij cis((e_i - e_j) t) with D the row term
 * above; compute_ddmdt_eq and compute_ddm do not.  P is left as it is.  dm =
 * dm1 = NULL clears the term.  The batched forms and the adjoint refuse a handle
 * holding it.  dmd_scatt_get_ee_dP: that Hermitian row term prefac/2 (D + D^+)
 * (nk x nb^2 complex). */
int dmd_scatt_set_ee_dP(dmd_scatt *h, const uint64_t *k1st, const uint64_t *k2nd, const double *U, const double *kvec,
                        const double *f_eq, const double *dm, const double *dm1, const double *vq_cell,
                        const dmd_eimp_model *m);
int dmd_scatt_get_ee_dP(dmd_scatt *h, double *H);
/* The static RPA v(q) of every k-mesh cell that is a difference of two k
 * points (coulomb_model's q list over all (ik, jk), Coulomb_Model.h:690-744,
 * 820-828), 0 at q = 0 and elsewhere: vq_cell holds kmesh[0]*kmesh[1]*kmesh[2]
 * doubles, cell ((x m1 + y) m2 + z).  f: occupations at mu (nk x nb). */
int dmd_scatt_ee_vq_cells(dmd_scatt *h, const double *U, const double *kvec, const double *f, const dmd_eimp_model *m,
                          double *vq_cell);
/* The e-e terms of add_Lsc_contrib("ee") beyond the P1/P2 form set_Lsc holds
 * (ElectronPhonon_linear.cpp:38-104,183-253), added into the linearized
 * operator of a handle after dmd_scatt_set_Lsc at f_eq: extra x the Lscij /
 * Lscji of P_ee(f_eq) (extra = 1 under ee_antisymmetry, whose fac2 doubles
 * them), and with dP_terms (eeMode other than Pee_fixed_at_eq) the P3/P4
 * (ElecElec_Model_linear.cpp:3-57, none with exchange) and P5/P6 (:59-111)
 * terms, all times scale (scale_ee).  coef34 = 1 subtracts the P3/P4 terms as
 * the reference does; coef34 = -1 adds them, which makes the operator the
 * exact Jacobian of the e-e right-hand side at f_eq. */
int dmd_scatt_add_Lsc_ee(dmd_scatt *h, const uint64_t *k1st, const uint64_t *k2nd, const double *U,
                         const double *kvec, const double *f_eq, const double *vq_cell, const dmd_eimp_model *m,
                         double scale, double extra, double coef34, int dP_terms);
/* dst's P = scale x src's stored P (add = 1: added), pair by pair in file
 * order, converted on the device between precisions and layouts (a reduced
 * store's values into an FP64 handle, or FP64 P quantised into a reduced
 * store).  Both handles hold the same stored pairs on one device.  No
 * reference counterpart (the reference keeps one complex128 P). */
int dmd_scatt_copy_P(dmd_scatt *dst, dmd_scatt *src, double scale, int add);
/* P of stored pairs [first, first+count) in file order, as complex128. */
int dmd_scatt_download_P(dmd_scatt *h, int64_t first, int64_t count, double *P1, double *P2);
/* Fills P on the device from the synthetic description (the handle's e must
 * be the one dmd_synth_energies gives for p); no host traffic. */
int dmd_scatt_generate_P(dmd_scatt *h, const dmd_synth_params *p);

/* alg.expt (interaction-picture phases on P, evolve_eph.cpp:91-98),
 * alg.picture == "schrodinger" (ddmdt_eq subtracted without phase,
 * evolve_eph.cpp:158; -i[e, rho] joins the integrator's right-hand side,
 * DenMat.cpp:16-24,45-46) and the contraction kernel (DMD_KERNEL_*).
 * Defaults: expt 1, interaction picture, DMD_KERNEL_AUTO. */
int dmd_scatt_set_flags(dmd_scatt *h, int expt, int schrodinger, int kernel);
/* ddmdt_eq subtracted after the assembly (alg.ddmdteq, evolve_eph.cpp:154-162);
 * NULL clears it.  nk x nb^2 complex. */
int dmd_scatt_set_ddmdt_eq(dmd_scatt *h, const double *ddmdt_eq);
/* compute_ddmdt_eq (evolve_eph.cpp:3-19): ddmdt_eq = evolve(t=0, diag f,
 * diag 1-f) and stores it.  f: nk x nb occupations of the window. */
int dmd_scatt_compute_ddmdt_eq(dmd_scatt *h, const double *f);
/* Reduced-precision forms (DMD_PREC_C64 / _BF16): the diagonal occupations f
 * (nk x nb) around which rho is split, rho = diag(f) + drho.  The rows sums at
 * diag(f) are formed once in FP64 from the stored P, and only drho enters the
 * FP32 products, as the reference's ddmdt_eq idea separates the equilibrium
 * part (evolve_eph.cpp:3-19,154-162).  NULL: f = 0.  No effect for FP64. */
int dmd_scatt_set_split(dmd_scatt *h, const double *f);

/* Linearized operator (alg_linearize): electronphonon::set_Lsc
 * (ElectronPhonon_linear.cpp:3-36, e-ph part :147-207) at the occupations
 * f_eq (nk x nb, the e-ph window).  The stored P (e-i P already folded in)
 * is replaced, in place, by Lscij | Lscji per stored pair; Lscii of row k is
 * folded into its diagonal pair (k, k), and rows without one drop it, as the
 * reference does.  From then on eval, evolve_driver, integrate_rk4 and
 * time_eval compute evolve_linear (evolve_eph_linear.cpp:3-53): ddmdt_k =
 * prefac/2 sum (X + X^+), X = Lsc (rho_k' - f_eq_k'), with compute_Pt's phase
 * when expt is set and no ddmdt_eq.  Uploading, generating or loading the
 * whole P again returns the handle to the P form.  FP64 only. */
int dmd_scatt_set_Lsc(dmd_scatt *h, const double *f_eq);

/* ---- the hot path --------------------------------------------------------- */
/* electronphonon::evolve (evolve_eph.cpp:34-166): ddmdt (nk x nb^2, fully
 * overwritten) from rho (nk x nb^2) at time t.  one_minus_rho may be NULL,
 * then 1-rho is formed as set_oneminusdm does (DenMat.cpp:249-261). */
int dmd_scatt_eval(dmd_scatt *h, double t, const double *rho, const double *one_minus_rho, double *ddmdt);
/* Allocation-free form of dmd_scatt_eval.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP stream
 * capture, tight serving loops, harnesses that forbid hipMalloc).  rho and
 * ddmdt must be device memory (nk x nb^2 complex, ddmdt fully overwritten,
 * 1 - rho formed from rho); the rotated state, the row accumulators and the
 * device clock live in one static device array of DMD_NOALLOC_SCRATCH_BYTES
 * per GPU instead of the handle's buffers, and the call makes no hipMalloc
 * or hipFree.  Operators whose scratch exceeds that size (about nk * nb^2 *
 * 24 bytes plus 256; nk up to 65536 at nb = 2 and 32768 at nb = 4 fit) get
 * DMD_ERR_ARG.  A reduced-precision handle needs one dmd_scatt_eval after its
 * P or split occupations change (DMD_ERR_STATE otherwise).  Every call on one
 * GPU shares the array: calls that overlap on one GPU race. */
#define DMD_NOALLOC_SCRATCH_BYTES (16u << 20)
int dmd_scatt_eval_noalloc(dmd_scatt *h, double t, const double *rho, double *ddmdt);
/* evolve (evolve_eph.cpp:34-166) for B independent states: rho_B, ddmdt_B
 * are B x nk x nb^2 (one_minus_rho formed internally; no band window).  On
 * a DMD_PREC_C64 or DMD_PREC_BF16 nb = 4 handle the B states share one pass
 * over P on the matrix cores (C64: P and drho as two BF16 terms each; BF16:
 * the folded P entries as one FP16 word, drho as two FP16 terms, each scaled
 * by a power of two; FP32 sums, FP64 equilibrium part and assembly); on
 * other handles it evaluates the states one after another.  No reference
 * counterpart for B > 1. */
int dmd_scatt_eval_batched(dmd_scatt *h, double t, int B, const double *rho_B, double *ddmdt_B);
/* The initializer's one-step form over all stored pairs
 * (lindbladInit.../ElectronPhonon.h:774-862 without the file reads): ddm +=
 * sum of per-pair compute_ddm, no phases.  Same stream as eval at t=0 with
 * expt off; ddm accumulated. */
int dmd_scatt_compute_ddm(dmd_scatt *h, const double *rho, const double *one_minus_rho, double *ddm);

/* ---- driver-side pieces (DenMat, observables, integrator) -------------------- */
/* Band window of evolve_driver (evolve_eph.cpp:21-32): the dynamics state is
 * nb_dm x nb_dm per k, the scattering acts on rows/cols [b0, b0+nb).  e_dm is
 * nk x nb_dm.  Default: nb_dm = nb, b0 = 0, e_dm = e. */
int dmd_scatt_set_window(dmd_scatt *h, int nb_dm, int b0, const double *e_dm);
/* evolve_driver itself: cut, evolve, expand (nk x nb_dm^2 arrays). */
int dmd_scatt_evolve_driver(dmd_scatt *h, double t, const double *dm_expand, const double *dm1_expand,
                            double *ddmdt_expand);
/* Coherent Hamiltonian H (nk x nb_dm^2, the H_BS(+H_Ez) of init_Hcoh,
 * DenMat.cpp:7-28); -i[H(t), rho] joins the right-hand side (DenMat.cpp:29-55).
 * NULL removes it. */
int dmd_scatt_set_coherent(dmd_scatt *h, const double *H);
/* Observable of ob_1dmk (observable.h:301-366): S(t) = prefac * sum_k
 * sum_{i in [i0,i1)} sum_b Re(s_t[i,b] (rho[b,i] - f_i delta_bi)).
 * s: nk x nb_dm^2 (one component of ldbd_smat.bin), f: nk x nb_dm. */
int dmd_scatt_set_observable(dmd_scatt *h, const double *s, const double *f, int i0, int i1, double prefac);
int dmd_scatt_measure(dmd_scatt *h, double t, const double *rho, double *value);
/* On-device fixed-step RK4 over compute() (DensityMatrix_Dynamics.h:243-278:
 * coherent + scattering), with re-Hermitization after every step, replacing
 * the GSL loop of evolve_gsl (:189-225).  rho_inout: nk x nb_dm^2 at t0, state
 * at t0 + n_steps*dt on return.  If report_every > 0 and an observable is set,
 * obs_out receives S at steps 0, report_every, 2*report_every, ...
 * (n_steps/report_every + 1 values). */
int dmd_scatt_integrate_rk4(dmd_scatt *h, double t0, double dt, int64_t n_steps, int report_every,
                            double *rho_inout, double *obs_out);
/* B trajectories (rho_inout: B x nk x nb_dm^2, obs_out: B x (n_steps /
 * report_every + 1)).  C64 / BF16 nb = 4 handles: one pass over P per RK
 * stage for all B (a captured HIP graph per step; no band window); other
 * handles: dmd_scatt_integrate_rk4 per trajectory. */
int dmd_scatt_integrate_rk4_batched(dmd_scatt *h, int B, double t0, double dt, int64_t n_steps,
                                    int report_every, double *rho_inout, double *obs_out);

/* dm_dynamics::compute (DensityMatrix_Dynamics.h:243-278): the whole
 * right-hand side d rho / dt at t (coherent term and scattering window, as
 * the integrators evaluate it); rho, ddmdt: nk x nb_dm^2, host or device. */
int dmd_scatt_compute_rhs(dmd_scatt *h, double t, const double *rho, double *ddmdt);
/* Statistics of an adaptive integration call. */
typedef struct {
    int64_t steps;          /* accepted steps                                   */
    int64_t rejected;       /* rejected attempts                                */
    int64_t rhs_evals;      /* right-hand-side evaluations                      */
    double h_next;          /* suggested next step (also returned in *h_step)   */
    double last_err_ratio;  /* max |err_i| / D_i of the last accepted step      */
} dmd_ode_stats;
/* gsl_odeiv2_driver_apply of evolve_gsl (DensityMatrix_Dynamics.h:189-225:
 * rkf45 stepper, standard control with a_y = 1, a_dydt = 0): integrates the
 * same right-hand side as dmd_scatt_integrate_rk4 from *t to t1 with
 * Fehlberg's embedded 4(5) pair on the device, D_i = eps_abs + eps_rel |y_i|
 * per real component, steps in [hmin, hmax] (ode_hmin, ode_hmax = tstep,
 * parameters.cpp:360-367), the last one ending exactly on t1.  *h_step carries
 * the controller's suggestion between calls as the GSL driver's state does
 * (start with ode_hstart); on return *t = t1.  rho_inout: nk x nb_dm^2, host
 * or device.  GSL's own rkf45 trajectories are not reproduced bit for bit
 * (its source is not part of the reference). */
int dmd_scatt_rkf45_apply(dmd_scatt *h, double *t, double t1, double *h_step, double hmin, double hmax,
                          double eps_abs, double eps_rel, double *rho_inout, dmd_ode_stats *stats);
/* A host callback run before every right-hand side the integrators
 * (dmd_scatt_integrate_rk4, dmd_scatt_rkf45_apply, dmd_scatt_integrate_euler)
 * and dmd_scatt_compute_rhs evaluate: the place where compute() runs
 * update_scatt_inside (DensityMatrix_Dynamics.h:243-262,281-289; eeMode =
 * Pee_update with freq_update_ee_model < 0 rebuilds P from every stage state).
 * t: the stage time; rho: a host copy of the stage state (nk x nb_dm^2
 * complex, interaction picture), valid for the duration of the call; the
 * stream is idle.  The callback may change the handle's P (dmd_scatt_copy_P,
 * dmd_scatt_add_*, dmd_scatt_set_ee_dP) and returns 0, or nonzero to stop the
 * integrator with DMD_ERR_STATE.  With a hook, RK4 runs its stages without a
 * captured graph.  fn = NULL removes it.  No counterpart as a function: the
 * reference calls update_scatt_inside from compute. */
typedef int (*dmd_rhs_hook)(void *ctx, double t, const double *rho);
int dmd_scatt_set_rhs_hook(dmd_scatt *h, dmd_rhs_hook fn, void *ctx);
/* evolve_euler (DensityMatrix_Dynamics.h:177-187,227-234): n_steps of
 * rho += dt (ddmdt + ddmdt^+)/2 (update_dm_euler, DenMat.cpp:236-248) from t0. */
int dmd_scatt_integrate_euler(dmd_scatt *h, double t0, double dt, int64_t n_steps, double *rho_inout);

/* ---- adjoint ------------------------------------------------------------------ */
/* Vector-Jacobian product of evolve (evolve_eph.cpp:34-166) at the state rho
 * (Hermitian, nk x nb^2) and time t: g = (d ddmdt / d rho)^T v under the real
 * pairing Re sum conj(a) b, projected on Hermitian matrices.  FP64, nb <= 2,
 * scattering term only.  nb = 2 runs one pass over P on the tile layout (the
 * handle switches to the tile kernel), nb = 1 two passes of the single-read
 * kernel (the handle switches to it).  No reference counterpart (the
 * reference has no derivatives). */
int dmd_scatt_vjp(dmd_scatt *h, double t, const double *rho, const double *v, double *g);
/* P = sum_c weights[c] P_c over component handles holding FP64 P of the same
 * pairs (e.g. per-branch and impurity parts; scale_eph generalised,
 * ElectronPhonon.cpp:225,232 and :87-97).  No reference counterpart. */
int dmd_scatt_combine_P(dmd_scatt *h, int ncomp, dmd_scatt *const *comps, const double *weights);
/* Checkpointed discrete adjoint of dmd_scatt_integrate_rk4 (scattering term
 * alone, FP64, nb <= 2): given dL/dS at the n_steps/report_every + 1 reports
 * of the observable, returns dL/dtheta_c = sum over stages of Re <Kbar,
 * f_c(Y)> for the component handles (the derivative with respect to the
 * weight of P_c) and dL/drho0.  The run is cut into segments of
 * checkpoint_every steps: every segment's first state is kept, and the stage
 * states of a segment are recomputed from it for the reverse sweep (none are
 * recomputed for checkpoint_every >= n_steps; the stage states then take
 * 4 n_steps nk nb^2 x 16 bytes of device memory).  nb = 2 on tiles folds the
 * parameter cotangents into one per-pair sum dotted with every P_c at the
 * end (components must hold P of the handle's pairs, FP64, same phase flag);
 * other layouts evaluate every component at every stage.  No reference
 * counterpart. */
int dmd_scatt_adjoint_rk4(dmd_scatt *h, double t0, double dt, int64_t n_steps, int report_every,
                          int checkpoint_every, const double *rho0, const double *dL_dobs, int ncomp,
                          dmd_scatt *const *comps, double *dL_dtheta, double *dL_drho0);

/* ---- introspection ---------------------------------------------------------- */
/* Bytes of P streamed per eval and stored pairs; the last kernel time in ms
 * measured with HIP events on the handle's stream (0 if none). */
int dmd_scatt_info(dmd_scatt *h, int64_t *n_pairs, int64_t *p_bytes, double *last_kernel_ms);
/* Times n evals of the contraction at time t on the current device state
 * (rho from the last eval); returns the mean ms per eval of the dominant
 * kernel and of the whole eval, measured with HIP events on the handle's stream. */
int dmd_scatt_time_eval(dmd_scatt *h, double t, int n, double *kernel_ms, double *eval_ms);
/* Times n batched evaluations (X operand, matrix-core pass, assembly) of the
 * states of the last dmd_scatt_eval_batched (C64 / BF16 nb = 4 handles); no
 * reference counterpart. */
int dmd_scatt_time_eval_batched(dmd_scatt *h, double t, int n, double *kernel_ms);

/* ---- synthetic ldbd_data ------------------------------------------------------- */
/* Presets: 1..5 = the five benchmark workloads (see the package docs); 0 = a
 * tiny case.  Returns DMD_ERR_ARG for an unknown preset. */
int dmd_synth_preset(int which, dmd_synth_params *p);
int dmd_synth_energies(const dmd_synth_params *p, double *e /* nk x nb */);
/* Stored pairs: pass k1st = k2nd = NULL to get the count in *n_pairs. */
int dmd_synth_kpairs(const dmd_synth_params *p, uint64_t *k1st, uint64_t *k2nd, int64_t *n_pairs);
/* P1/P2 of the given pairs on the host (nthreads <= 0: all cores). */
int dmd_synth_P(const dmd_synth_params *p, const double *e, int64_t n_pairs, const uint64_t *k1st,
                const uint64_t *k2nd, double *P1, double *P2, int nthreads);
/* Couplings of one stored pair: omega (n_modes), gm, gp (n_modes x nb^2). */
int dmd_synth_couplings(const dmd_synth_params *p, uint64_t k, uint64_t kp, double *omega, double *gm,
                        double *gp);
/* Random Hermitian spin matrices, nk x 3 x nb^2 (ldbd_smat.bin layout). */
int dmd_synth_spin(const dmd_synth_params *p, double *s3);

/* ---- ldbd_data files (reference formats) ---------------------------------------- */
/* Writes ldbd_size.dat, ldbd_R.dat, ldbd_ek.bin, ldbd_kvec.bin, ldbd_smat.bin,
 * ldbd_kpair_k1st/k2nd.bin and ldbd_P1/P2_lindblad.bin into dir
 * (layouts of lindbladInit.../electron_save.cpp:16-65 and ElectronPhonon.h:252-284). */
int dmd_ldbd_write_synth(const char *dir, const dmd_synth_params *p, int nthreads);

/* ---- ldbd_data reader ------------------------------------------------------------ */
/* ldbd_size.dat as the reference reads it: electron::read_ldbd_size
 * (src_v4.5.7/electron.cpp:100-132), electronphonon::get_brange / get_nkpair
 * (ElectronPhonon.cpp:132-176), phonon::read_jdftx (phonon.h:49-79) and
 * parameters::read_jdftx (parameters.cpp:7-23); the layouts are those of
 * electron::saveSize (lindbladInit.../electron_save.cpp:16-50).  Band
 * indices count from the first band of ldbd_ek.bin.  A valence-only file set
 * (kind 1) names its e-ph window, pair count and pair/P files "_hole". */
typedef struct {
    int kind;                 /* first line: 0 conduction only, 1 valence only, 2 both */
    int n_line2;              /* integers on line 2 (8 or 9; 11 for kind 1)              */
    int nb, nv;               /* bands of ldbd_ek.bin; valence bands among them          */
    int b_bot_dm, b_top_dm;   /* density-matrix bands [b_bot_dm, b_top_dm)               */
    int b_bot_eph, b_top_eph; /* e-ph bands [b_bot_eph, b_top_eph) of this file set      */
    int nb_wannier;
    double nk_full;           /* k points of the full mesh (prefac = 2 pi / nk_full)     */
    int64_t nk;               /* k points kept                                           */
    int kmesh[3];
    int64_t nkpair;           /* stored pairs (nkpair_hole for kind 1)                    */
    int mode_start, mode_stop, n_modes;
    double temperature;       /* Hartree                                                 */
    double mu_min, mu_max, mu;
    double degauss, ndegauss;
    double e_bot_probe, e_top_probe, e_bot_dm, e_top_dm, e_bot_eph, e_top_eph;
    double omega_max;         /* 0 when the file has no tenth line                        */
    int b_skip_wannier;       /* line 2 after nb_wannier: bands below the Wannier set     */
} dmd_ldbd_info;
int dmd_ldbd_read_info(const char *dir, dmd_ldbd_info *info);
/* ldbd_ek.bin, nk x nb float64; scissor is added to bands nv.. (electron.cpp:172-180). */
int dmd_ldbd_read_ek(const char *dir, const dmd_ldbd_info *info, double scissor, double *e);
/* ldbd_kvec.bin, nk x 3 float64 (electron.cpp:134-147). */
int dmd_ldbd_read_kvec(const char *dir, const dmd_ldbd_info *info, double *kvec);
/* ldbd_smat.bin, nk x 3 x nb_dm^2 complex128 (electron.cpp:284-296). */
int dmd_ldbd_read_smat(const char *dir, const dmd_ldbd_info *info, double *s);
/* ldbd_lmat.bin (needL, electron.cpp:315-335): nk x 3 x nb_dm^2 complex. */
int dmd_ldbd_read_lmat(const char *dir, const dmd_ldbd_info *info, double *l);
/* ldbd_Bso.bin (alg_read_Bso, electron.cpp:159-170): nk x 3 float64. */
int dmd_ldbd_read_Bso(const char *dir, const dmd_ldbd_info *info, double *bso);
/* ldbd_HEzmat.bin (scale_Ez, electron.cpp:44-61): nk x nb_dm^2 complex. */
int dmd_ldbd_read_HEzmat(const char *dir, const dmd_ldbd_info *info, double *hez);
/* ldbd_Umat.bin (electron.cpp:388-410) as nk x nb_wannier x nb_eph complex:
 * the full nk x nb_wannier^2 file gives columns b_bot_eph + b_skip_wannier
 * onward, a file of that reduced size is read as it is. */
int dmd_ldbd_read_Umat(const char *dir, const dmd_ldbd_info *info, double *U);
/* ldbd_imsig.bin, nk x nb_eph float64 (electron.cpp:181-198). */
int dmd_ldbd_read_imsig(const char *dir, const dmd_ldbd_info *info, double *imsig);
/* ldbd_R.dat: dim and the lattice vectors as rows of R[9] (lattice.h:48-56). */
int dmd_ldbd_read_R(const char *dir, int *dim, double *R);
/* Stored pairs [first, first+count) of ldbd_kpair_k1st/k2nd[_hole].bin
 * (ElectronPhonon.cpp:182-204). */
int dmd_ldbd_read_kpairs(const char *dir, const dmd_ldbd_info *info, int64_t first, int64_t count, uint64_t *k1st,
                         uint64_t *k2nd);
/* P1, P2 of pairs [first, first+count) from ldbd_P{1,2}_<scatt>[_D<impurity>][_hole].bin
 * (scatt "lindblad" or "conventional"; impurity 0: e-ph, i > 0: defect i),
 * unscaled, on the host (ElectronPhonon.cpp:209-233, ElectronImpurity.cpp:3-35). */
int dmd_ldbd_read_P(const char *dir, const dmd_ldbd_info *info, const char *scatt, int impurity, int64_t first,
                    int64_t count, double *P1, double *P2);
/* read_ldbd_eph (ElectronPhonon.cpp:209-233; scale = scale_eph) and the e-i
 * fold-in of read_ldbd_imp_P (ElectronImpurity.cpp:3-35, ElectronPhonon.cpp:
 * 87-97; impurity = iD+1, scale = fraction * scale_ei, add = 1) straight into
 * the handle: the files are streamed in 64 MiB chunks through pinned buffers,
 * so no host copy of P is made. */
int dmd_scatt_load_P(dmd_scatt *h, const char *dir, const dmd_ldbd_info *info, const char *scatt, int impurity,
                     double scale, int add);
/* electronphonon's set-up from an ldbd_data folder (ElectronPhonon.h:53-93):
 * sizes, energies (e-ph window of e_dm, ElectronPhonon.h:79; scissor as
 * electron.cpp:177-179), stored pairs and P x scale_eph, and the band window
 * of evolve_driver when the e-ph bands are fewer than the density matrix's. */
int dmd_scatt_create_from_ldbd(const char *dir, const char *scatt, double scale_eph, double scissor, int precision,
                               dmd_scatt **out);
/* P of stored pairs [first, first+count) (file order), set (add = 0) or added
 * (add = 1), scaled; the range form of dmd_scatt_upload_P / dmd_scatt_add_P for
 * callers that read P in pieces, as the reference's ranks read their share
 * (ElectronPhonon.cpp:222-223).  Pairs never given keep P = 0. */
int dmd_scatt_upload_P_range(dmd_scatt *h, int64_t first, int64_t count, const double *P1, const double *P2,
                             double scale, int add);

/* alg_semiclassical (dm_dynamics::compute, DensityMatrix_Dynamics.h:249,277):
 * the integrators' right-hand side (integrate_rk4, rkf45_apply,
 * integrate_euler, compute_rhs) zeroes the state's off-diagonal elements
 * before the scattering term and the result's after, so only the
 * populations evolve and the coherent term drops out.  Not for the batched
 * integrator or the adjoint.  dmd_scatt_eval (electronphonon::evolve) is
 * unaffected, as in the reference. */
int dmd_scatt_set_semiclassical(dmd_scatt *h, int on);

/* alg_phenom_relax (phenom_relax, scattering_model/phenomenon_relax.h:5-51,
 * added by compute, DensityMatrix_Dynamics.h:272-275): the integrators'
 * right-hand side gains -(rho - diag f_eq) / tau on the density-matrix bands
 * [b_start, b_end) (bStart_tau, bEnd_tau); f_eq is nk x nb_dm (dm_eq of
 * set_dm_eq).  tau <= 0 switches it off.  Not for the batched integrator or
 * the adjoint; dmd_scatt_eval is unaffected. */
int dmd_scatt_set_phenom_relax(dmd_scatt *h, double tau, int b_start, int b_end, const double *f_eq);

/* The k-pair filter of evolve (evolve_eph.cpp:52-58: alg_only_intravalley,
 * alg_only_intervalley and forbidden valley transitions, lattice.h:82-139):
 * the P of every stored pair with keep[p] = 0 (file order, n_pairs bytes)
 * becomes 0, so neither direction of that pair scatters.  The pairs keep
 * their slots.  Not on a linearized handle (evolve_linear has no filter). */
int dmd_scatt_mask_pairs(dmd_scatt *h, const uint8_t *keep);

/* ---- sparse P (alg_sparseP, alg_Pin_is_sparse) ------------------------------------ */
/* Sparse files sP{1,2}_<scatt>[_hole]_{ns,s,i,j}.bin (sparse2D::read_smat,
 * comm/sparse2D.h:98-119; sparse_mat(FILE*, ...), comm/sparse_matrix.h:30-36):
 * per stored pair an int count ns in _ns, and that many complex128 values in
 * _s, int row indices i1*nb+i2 in _i and int column indices i3*nb+i4 in _j.
 * The device keeps P dense in its usual layout (nb <= 4: a pair's record is
 * 256 B to 8 KiB), so every kernel serves the sparse modes unchanged. */
/* read_ldbd_eph's sparse branch (ElectronPhonon.cpp:235-246): P from the
 * sparse files, each value times scale and kept when |scale*s| > thr (the
 * re-sparsification at :245-246 keeps |.| > 1e-40); an index pair given twice
 * keeps the last value (sparse_mat::todense, comm/sparse_matrix.h:76-80).
 * Sizes and indices are checked against the handle. */
int dmd_scatt_load_sP(dmd_scatt *h, const char *dir, const dmd_ldbd_info *info, const char *scatt, double scale,
                      double thr);
/* Stores P by its nonzero elements (the reference's sparse modes alg_sparseP /
 * alg_Pin_is_sparse evolve over per-pair element lists, evolve_eph.cpp:69-86,
 * 194-203, comm/sparse_matrix.cpp:6-40): per pair a 32-bit nonzero mask and
 * its nonzero values, per 8 pairs a 32-bit offset, 16 n + 4.5 bytes for n
 * nonzeros, so the bytes held and streamed per evaluation fall with the
 * density; the dense P is released.  nb = 2 FP64 on the tile kernel.  With
 * more than 3/4 of the elements nonzero the dense store is kept (DMD_OK).
 * Evaluation and the integrators run as before; operations that change or
 * read back P (upload, add, download, threshold, masks, set_Lsc, the
 * adjoint, kernel changes) return DMD_ERR_STATE afterwards.  stored_values
 * (optional): the nonzero values. */
int dmd_scatt_compress_P(dmd_scatt *h, int64_t *stored_values);
/* The sparsification of alg_sparseP and of the e-i fold-in under
 * alg_Pin_is_sparse (set_sparseP, ElectronPhonon.cpp:115-129;
 * sparse_plus_dense, comm/sparse_matrix.cpp:42-47; sparse_mat::sparse keeps
 * |A| > thr, comm/sparse_matrix.h:57-70): every stored element of P1 and P2
 * with |P| <= thr becomes 0.  Not on a linearized handle. */
int dmd_scatt_threshold_P(dmd_scatt *h, double thr);
/* write_smat of set_sparseP (ElectronPhonon.cpp:126-128; comm/sparse2D.h:
 * 136-151, sparse_mat::write_to_files, comm/sparse_matrix.h:37-42): the
 * handle's P1 and P2 as sparse files, elements with |P| > thr in row-major
 * order; a matrix whose _s file exists is not written again.  All four files
 * carry the <scatt>[_hole] suffix the reader looks for (ElectronPhonon.cpp:241). */
int dmd_ldbd_write_sP(dmd_scatt *h, const char *dir, const dmd_ldbd_info *info, const char *scatt, double thr);
/* The handle of an ldbd_data folder without the e-ph P: alg_eph_enable = 0
 * (ElectronPhonon.cpp:211, P left zero for the e-i fold-in to fill; P counts
 * as unset until dmd_scatt_load_P / _load_sP / _upload_P gives it), or, with
 * no_pairs = 1, alg_scatt_enable = 0 (ElectronPhonon.h:72: no scattering; the
 * handle has no stored pairs and its right-hand side is the coherent term). */
int dmd_scatt_create_from_ldbd_noP(const char *dir, double scissor, int precision, int no_pairs, dmd_scatt **out);
/* dmd_scatt_create_from_ldbd with P from the sparse files (alg_Pin_is_sparse),
 * scaled by scale_eph and kept when |scale_eph*s| > 1e-40. */
int dmd_scatt_create_from_ldbd_sparse(const char *dir, const char *scatt, double scale_eph, double scissor,
                                      int precision, dmd_scatt **out);

/* ---- restart files ---------------------------------------------------------------- */
/* denmat_restart.bin (nk x nb^2 complex128) and time_restart.dat ("%14.7le")
 * in the folder dir, as singdenmat_k::read_dm_restart / write_dm_tofile
 * (DenMat.cpp:301-334) and parameters.cpp:299-310 use them; on read, dm or t
 * may be NULL to skip that file. */
int dmd_restart_read(const char *dir, int64_t nk, int nb, double *dm, double *t);
int dmd_restart_write(const char *dir, int64_t nk, int nb, const double *dm, double t);

#ifdef __cplusplus
}
#endif
#endif
