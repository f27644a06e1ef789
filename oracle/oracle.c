/*
 * oracle.c -- CPU restatement of the DMD Lindblad scattering hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke()
 * entry and the cpu_baseline leg of bench.py load it (as oracle/liboracle.so)
 * and compare the HIP library against it.  Nothing in the product library
 * links, loads or calls it, and it is never the thing that is measured as the
 * GPU result.
 *
 * Every routine follows one reference routine (paths relative to the reference
 * root) in loop order and operation order, with the reference's interleaved
 * complex128 layout (struct {double x, y}, src_v4.5.7/comm/scalar.h:67-106) and
 * flat row-major 2-D pools (src_v4.5.7/comm/myarray.cpp:145-155).
 *
 * Pinning (what checks each routine; the tests hold them):
 *   oracle_compute_Pt      t=0 identity, factorised-phase identity, periodicity.
 *   oracle_evolve          brute-force Eq. 9 over an explicit both-direction
 *                          pair list (tests/test_oracle.py), Hermiticity,
 *                          particle conservation, rho=0 / rho=I zeros,
 *                          interaction vs Schroedinger picture, per-pair
 *                          initializer form (oracle_compute_ddm_pairs).
 *   oracle_compute_ddm_pairs  equality with oracle_evolve at t=0 (linearity).
 *   oracle_evolve_sparse   equality with oracle_evolve on the densified P
 *                          (zeros where no element is listed) at t = 0 and
 *                          with phases, for any element order.
 *   oracle_compute_P       Gram-form positivity, the P1(k,k')/P2(k',k)
 *                          conjugate identity, ImSigma from the P diagonal
 *                          against a direct golden-rule sum; with
 *                          oracle_evolve, the initializer's golden-rule T1
 *                          (T1_rate_dRho) equals the one-step Lindblad T1
 *                          for any drho (tests/test_oracle_t1_rate.py).
 *   oracle_eimp_vq /       the reference's own GaAs ldbd_vq.out: every |q|^2
 *   oracle_eimp_q2         it lists is a wrapped k - k' length of that k mesh
 *                          (smallest nonzero 2.815e-05), and v = prefac_vq /
 *                          (q^2 + q_scr^2) with its q_scr^2 column reproduces
 *                          its v column; the Debye length and the Thomas-Fermi
 *                          inputs against its header; Bechstedt limits (bare
 *                          4 pi / V q^2 at large q, background-screened at
 *                          small q).
 *   oracle_eimp_rpa_vq     the reference's GaAs run's q_scr^2 at q = 0 from its own
 *                          bands, k points and mu (there the overlaps are the
 *                          identity); the nondegenerate limit at q = 0 is the
 *                          Debye q^2 times sum f(1-f) / sum f; every q's sum
 *                          against a brute-force double sum over (k, k') with
 *                          k - k' in the q cell, on meshes with missing points.
 *   oracle_eimp_model_P    the P diagonal is the golden-rule |M|^2 times a
 *                          normalised Gaussian delta for both scattering forms
 *                          (and the emission/absorption pair with detailed
 *                          balance), P2(k,k') = conj(P1(k',k)), the reshaped
 *                          P1[(i1 i3),(i2 i4)] is Hermitian PSD of rank 1
 *                          (2 with detailed balance).
 *   oracle_ee_model_P      (with dP, alg_linearize_dPee) dP is the exact
 *     dP part              derivative (central differences of a quadratic) of
 *                          P at the state along the row-subtracted directions
 *                          calc_A1rho_dP forms; for nb = 1, where no row
 *                          subtraction enters, the derivative along rho - f_eq.
 *   oracle_ee_dP_term      evolve of dP at diag(f_eq), diag(1 - f_eq) without
 *                          phases, times cis((e_i - e_j) t) elementwise.
 *   oracle_ee_model_P      the P diagonal is the e-e golden rule
 *                          (1/nk_full) sum_k3 |M|^2 (1-f3) f4 delta(de) with a
 *                          normalised Gaussian, for both scattering forms;
 *                          P2(k1,k2) = conj(P1(k2,k1)); the Lindblad P1
 *                          reshaped to [(i1 i3),(i2 i4)] is Hermitian PSD.
 *   oracle_coherent        two-level precession closed form, tracelessness.
 *   oracle_rk4             1+z+z^2/2+z^3/6+z^4/24 amplification, O(h^4).
 *   oracle_spin            brute force, zero at rho=f, linearity, and the
 *                          reference's own GaAs restart/sz_elec_tot.out pair.
 *   oracle_dm_Bpert_1st    closed form of the first-order perturbation,
 *                          degenerate limit, Hermiticity.
 *   oracle_set_Lsc +       the Jacobian of oracle_evolve at diag(f_eq) (exact
 *   oracle_evolve_linear   central differences of a quadratic map) when every
 *                          k has its diagonal pair; Lscii dropped for rows
 *                          without one, as the reference does; linearity.
 *   oracle_mt19937_64      the C++ standard's known answer (10000th output
 *                          of the default seed 5489 is 9981545732273789042).
 *   oracle_normal_draws    bit for bit against the host C++ library's
 *                          std::normal_distribution over std::mt19937_64.
 *   oracle_gfac_normal     (g - mean)/sigma is the raw draws of a generator
 *                          seeded with n standardised by their own mean and
 *                          population sigma, with the draws beyond the cap
 *                          skipped; n = 0, 1 and sigma = 0.
 * The bit-level rounding of the reference's CBLAS zgemm/zhemm (GSL 2.6 or MKL)
 * and libm sincos is not reproduced bit for bit ("parity unpinned" at the last
 * ulp); summation orders follow the reference's loops so results agree to a
 * few ulp.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double x, y; } zc;

static inline zc zmk(double x, double y) { zc r; r.x = x; r.y = y; return r; }
/* complex product as in src_v4.5.7/comm/scalar.h:89 */
static inline zc zmul(zc a, zc b) { return zmk(a.x * b.x - a.y * b.y, a.y * b.x + a.x * b.y); }
static inline zc zadd(zc a, zc b) { return zmk(a.x + b.x, a.y + b.y); }
static inline zc zsub(zc a, zc b) { return zmk(a.x - b.x, a.y - b.y); }
static inline zc zconj(zc a) { return zmk(a.x, -a.y); }
static inline zc zscale(zc a, double s) { return zmk(a.x * s, a.y * s); }
/* cis(x) = cos x + i sin x, src_v4.5.7/comm/scalar.h:122-125 */
static inline zc zcis(double x) { return zmk(cos(x), sin(x)); }

#define Z(p, i) (((const zc *)(p))[i])
#define ZW(p, i) (((zc *)(p))[i])

/* ------------------------------------------------------------------------ */
/* compute_Pt: interaction-picture P, src_v4.5.7/evolve_eph.cpp:169-187.     */
/* P1t[(i1 i2),(i3 i4)] = P1 * cis((ek[i1] - ekp[i3] - ek[i2] + ekp[i4]) t)  */
/* P2t (minus)          = P2 * cis((ek[i3] - ekp[i1] - ek[i4] + ekp[i2]) t)  */
/* ------------------------------------------------------------------------ */
void oracle_compute_Pt(int nb, double t, const double *ek, const double *ekp,
                       const double *P, double *Pt, int minus)
{
    for (int i1 = 0; i1 < nb; i1++) {
        int n1 = i1 * nb;
        for (int i2 = 0; i2 < nb; i2++) {
            int n12 = (n1 + i2) * nb;
            for (int i3 = 0; i3 < nb; i3++) {
                int n123 = (n12 + i3) * nb;
                for (int i4 = 0; i4 < nb; i4++) {
                    double arg = !minus ? (ek[i1] - ekp[i3] - ek[i2] + ekp[i4]) * t
                                        : (ek[i3] - ekp[i1] - ek[i4] + ekp[i2]) * t;
                    ZW(Pt, n123 + i4) = zmul(Z(P, n123 + i4), zcis(arg));
                }
            }
        }
    }
}

/* C(m x 1) += A(m x m) * b, the zgemm_interface call at evolve_eph.cpp:99
 * (src_v4.5.7/comm/mymatrix.cpp:82-86).  Row-major; the k loop is the outer
 * accumulation order of a reference CBLAS zgemm (C += A[i,k]*B[k,j]). */
static void gemv_Ab(zc *C, const zc *A, const zc *b, int m)
{
    for (int i = 0; i < m; i++)
        for (int k = 0; k < m; k++)
            C[i] = zadd(C[i], zmul(A[i * m + k], b[k]));
}
/* C(1 x m) += a(1 x m) * B(m x m), the call at evolve_eph.cpp:100. */
static void gemv_aB(zc *C, const zc *a, const zc *B, int m)
{
    for (int k = 0; k < m; k++) {
        zc ak = a[k];
        for (int j = 0; j < m; j++)
            C[j] = zadd(C[j], zmul(ak, B[k * m + j]));
    }
}

/* Element of a Hermitian matrix stored in its upper triangle, as CBLAS zhemm
 * with CblasUpper reads it: A[i][j] for i<=j, conj(A[j][i]) below, and only
 * the real part on the diagonal. */
static inline zc herm_upper(const zc *A, int n, int i, int j)
{
    if (i < j) return A[i * n + j];
    if (i > j) return zconj(A[j * n + i]);
    return zmk(A[i * n + i].x, 0.0);
}

/* zhemm_interface(C, left, A, B, n, alpha, beta), src_v4.5.7/comm/mymatrix.cpp:73-76:
 * left:  C = alpha*A*B + beta*C;  right: C = alpha*B*A + beta*C; A Hermitian
 * (upper triangle), row-major n x n. */
static void zhemm_upper(zc *C, int left, const zc *A, const zc *B, int n, zc alpha, zc beta)
{
    zc tmp[64];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            zc s = zmk(0, 0);
            for (int k = 0; k < n; k++) {
                if (left) s = zadd(s, zmul(herm_upper(A, n, i, k), B[k * n + j]));
                else      s = zadd(s, zmul(B[i * n + k], herm_upper(A, n, k, j)));
            }
            tmp[i * n + j] = s;
        }
    /* BLAS semantics: with beta == 0, C is not read on input */
    int beta0 = (beta.x == 0.0 && beta.y == 0.0);
    for (int i = 0; i < n * n; i++)
        C[i] = beta0 ? zmul(alpha, tmp[i]) : zadd(zmul(alpha, tmp[i]), zmul(beta, C[i]));
}

/* The per-pair accumulation body of electronphonon::evolve for pairs
 * [p0, p1): src_v4.5.7/evolve_eph.cpp:49-106 (dense summode branch, the one the
 * reference's default flags select: alg.summode=1, alg.sparseP=0). */
static void evolve_pairs(int64_t p0, int64_t p1, int nb, const uint64_t *k1, const uint64_t *k2,
                         const double *P1, const double *P2, const double *e,
                         const double *dm, const double *dm1, double t, int expt,
                         zc *Pdm, zc *dm1P)
{
    int nb2 = nb * nb, nb4 = nb2 * nb2;
    zc *P1t = (zc *)malloc(sizeof(zc) * nb4), *P2t = (zc *)malloc(sizeof(zc) * nb4);
    zc *P1n = (zc *)malloc(sizeof(zc) * nb4), *P2n = (zc *)malloc(sizeof(zc) * nb4);
    for (int64_t ip = p0; ip < p1; ip++) {
        int64_t ik = (int64_t)k1[ip], ikp = (int64_t)k2[ip];
        const double *p1 = P1 + (size_t)ip * nb4 * 2, *p2 = P2 + (size_t)ip * nb4 * 2;
        if (!expt) {                                   /* evolve_eph.cpp:91-94 */
            memcpy(P1t, p1, sizeof(zc) * nb4);
            memcpy(P2t, p2, sizeof(zc) * nb4);
        } else {                                       /* evolve_eph.cpp:95-98 */
            oracle_compute_Pt(nb, t, e + ik * nb, e + ikp * nb, p1, (double *)P1t, 0);
            oracle_compute_Pt(nb, t, e + ik * nb, e + ikp * nb, p2, (double *)P2t, 1);
        }
        gemv_Ab(Pdm + ik * nb2, P1t, (const zc *)dm + ikp * nb2, nb2);    /* :99  */
        gemv_aB(dm1P + ik * nb2, (const zc *)dm1 + ikp * nb2, P2t, nb2);  /* :100 */
        if (ik < ikp) {                                                    /* :101 */
            for (int i = 0; i < nb4; i++) { P2n[i] = zconj(P1t[i]); P1n[i] = zconj(P2t[i]); }
            gemv_Ab(Pdm + ikp * nb2, P1n, (const zc *)dm + ik * nb2, nb2);   /* :104 */
            gemv_aB(dm1P + ikp * nb2, (const zc *)dm1 + ik * nb2, P2n, nb2); /* :105 */
        }
    }
    free(P1t); free(P2t); free(P1n); free(P2n);
}

/* ------------------------------------------------------------------------ */
/* electronphonon::evolve, src_v4.5.7/evolve_eph.cpp:34-166.                 */
/*   Pdm_k  = sum_k' P1t_kk' . vec(rho_k')      (+ reverse from stored pairs) */
/*   dm1P_k = sum_k' vec(1-rho_k')^T . P2t_kk'                                */
/*   C_k = (1-rho_k) mat(Pdm_k) - mat(dm1P_k) rho_k   (zhemm, :140-143)       */
/*   ddmdt_k = prefac * 0.5 * (C_k + C_k^dagger)      (:149-151)              */
/*   optional ddmdteq subtraction (:154-162).                                 */
/* nthreads <= 1 reproduces the reference loop order on one rank.  With      */
/* nthreads > 1 the stored pairs are block-partitioned over threads like     */
/* MPI ranks (src_v4.5.7/comm/mymp.cpp:8-21), each thread keeps private      */
/* Pdm/dm1P and the partials are summed in thread order (the allreduce at    */
/* :137 without its per-row collectives).                                    */
/* ------------------------------------------------------------------------ */
static void assemble(int64_t nk, int nb, const double *e, const double *dm, const double *dm1, const zc *Pdm,
                     const zc *dm1P, double t, double prefac, const double *ddmdt_eq, int schrodinger, double *ddmdt);

int oracle_evolve(int64_t nk, int nb, int64_t npairs, const uint64_t *k1, const uint64_t *k2,
                  const double *P1, const double *P2, const double *e,
                  const double *dm, const double *dm1, double t, double prefac, int expt,
                  const double *ddmdt_eq, int schrodinger, int nthreads, double *ddmdt)
{
    int nb2 = nb * nb;
    if (nb < 1 || nb > 8 || nk < 0 || npairs < 0) return -1;
    for (int64_t ip = 0; ip < npairs; ip++)
        if (k1[ip] > k2[ip] || (int64_t)k2[ip] >= nk) return -2;
    size_t nrow = (size_t)nk * nb2;
    zc *Pdm = (zc *)calloc(nrow, sizeof(zc)), *dm1P = (zc *)calloc(nrow, sizeof(zc));
    if (!Pdm || !dm1P) { free(Pdm); free(dm1P); return -3; }
    if (nthreads <= 1) {
        evolve_pairs(0, npairs, nb, k1, k2, P1, P2, e, dm, dm1, t, expt, Pdm, dm1P);
    } else {
        zc **pPdm = (zc **)calloc(nthreads, sizeof(zc *)), **pdm1P = (zc **)calloc(nthreads, sizeof(zc *));
        int fail = 0;
        for (int it = 0; it < nthreads; it++) {
            pPdm[it] = (zc *)calloc(nrow, sizeof(zc));
            pdm1P[it] = (zc *)calloc(nrow, sizeof(zc));
            if (!pPdm[it] || !pdm1P[it]) fail = 1;
        }
        if (!fail) {
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
            for (int it = 0; it < nthreads; it++) {
                int64_t p0 = npairs * it / nthreads, p1 = npairs * (it + 1) / nthreads;
                evolve_pairs(p0, p1, nb, k1, k2, P1, P2, e, dm, dm1, t, expt, pPdm[it], pdm1P[it]);
            }
#pragma omp parallel for num_threads(nthreads) schedule(static)
            for (int64_t i = 0; i < (int64_t)nrow; i++)
                for (int it = 0; it < nthreads; it++) {
                    Pdm[i] = zadd(Pdm[i], pPdm[it][i]);
                    dm1P[i] = zadd(dm1P[i], pdm1P[it][i]);
                }
        }
        for (int it = 0; it < nthreads; it++) { free(pPdm[it]); free(pdm1P[it]); }
        free(pPdm); free(pdm1P);
        if (fail) { free(Pdm); free(dm1P); return -3; }
    }
    assemble(nk, nb, e, dm, dm1, Pdm, dm1P, t, prefac, ddmdt_eq, schrodinger, ddmdt);
    free(Pdm); free(dm1P);
    return 0;
}

/* final assembly, evolve_eph.cpp:140-152, and the ddmdteq subtraction, :154-162 */
static void assemble(int64_t nk, int nb, const double *e, const double *dm, const double *dm1, const zc *Pdm,
                     const zc *dm1P, double t, double prefac, const double *ddmdt_eq, int schrodinger, double *ddmdt)
{
    const int nb2 = nb * nb;
    const zc c1 = zmk(1, 0), cm1 = zmk(-1, 0), c0 = zmk(0, 0);
    for (int64_t ik = 0; ik < nk; ik++) {
        zc contrib[64];
        for (int i = 0; i < nb2; i++) contrib[i] = c0;
        zhemm_upper(contrib, 1, (const zc *)dm1 + ik * nb2, Pdm + ik * nb2, nb, c1, c0);
        zhemm_upper(contrib, 0, (const zc *)dm + ik * nb2, dm1P + ik * nb2, nb, cm1, c1);
        for (int i = 0; i < nb; i++)
            for (int j = 0; j < nb; j++)
                ZW(ddmdt, ik * nb2 + i * nb + j) =
                    zscale(zadd(contrib[i * nb + j], zconj(contrib[j * nb + i])), prefac * 0.5);
    }
    if (ddmdt_eq) {                                    /* evolve_eph.cpp:154-162 */
        for (int64_t ik = 0; ik < nk; ik++)
            for (int i = 0; i < nb; i++)
                for (int j = 0; j < nb; j++) {
                    size_t o = ik * nb2 + i * nb + j;
                    if (i == j || schrodinger)
                        ZW(ddmdt, o) = zsub(Z(ddmdt, o), Z(ddmdt_eq, o));
                    else
                        ZW(ddmdt, o) = zsub(Z(ddmdt, o), zmul(Z(ddmdt_eq, o), zcis((e[ik * nb + i] - e[ik * nb + j]) * t)));
                }
    }
}

/* ------------------------------------------------------------------------ */
/* The sparse branch of electronphonon::evolve (alg_sparseP /               */
/* alg_Pin_is_sparse), src_v4.5.7/evolve_eph.cpp:69-86: per stored pair the */
/* COO lists of P1 and P2 (ns values, row indices i1*nb+i2, column indices  */
/* i3*nb+i4, in the order of the files sP{1,2}_*_{ns,s,i,j}.bin), phases by */
/* compute_sPt (:194-203), products by sparse_zgemm (src_v4.5.7/comm/       */
/* sparse_matrix.cpp:6-40): left, n = 1: Pdm[i] += s * dm[j]; right, m = 1: */
/* dm1P[j] += dm1[i] * s; the reverse direction with the conjugated lists   */
/* of the other matrix (:81-85).  Then the same assembly as oracle_evolve.  */
/* ------------------------------------------------------------------------ */
int oracle_evolve_sparse(int64_t nk, int nb, int64_t npairs, const uint64_t *k1, const uint64_t *k2,
                         const int32_t *ns1, const double *s1, const int32_t *i1, const int32_t *j1,
                         const int32_t *ns2, const double *s2, const int32_t *i2, const int32_t *j2,
                         const double *e, const double *dm, const double *dm1, double t, double prefac, int expt,
                         double *ddmdt)
{
    const int nb2 = nb * nb;
    if (nb < 1 || nb > 8 || nk < 0 || npairs < 0) return -1;
    for (int64_t ip = 0; ip < npairs; ip++)
        if (k1[ip] > k2[ip] || (int64_t)k2[ip] >= nk || ns1[ip] < 0 || ns2[ip] < 0) return -2;
    size_t nrow = (size_t)nk * nb2;
    zc *Pdm = (zc *)calloc(nrow, sizeof(zc)), *dm1P = (zc *)calloc(nrow, sizeof(zc));
    zc *t1 = (zc *)malloc(sizeof(zc) * (size_t)nb2 * nb2), *t2 = (zc *)malloc(sizeof(zc) * (size_t)nb2 * nb2);
    if (!Pdm || !dm1P || !t1 || !t2) { free(Pdm); free(dm1P); free(t1); free(t2); return -3; }
    const zc *S1 = (const zc *)s1, *S2 = (const zc *)s2, *D = (const zc *)dm, *D1 = (const zc *)dm1;
    int64_t o1 = 0, o2 = 0;
    for (int64_t ip = 0; ip < npairs; ip++) {
        const int64_t ik = (int64_t)k1[ip], ikp = (int64_t)k2[ip];
        const double *ek = e + ik * nb, *ekp = e + ikp * nb;
        const int n1 = ns1[ip], n2 = ns2[ip];
        if (n1 > nb2 * nb2 || n2 > nb2 * nb2) { free(Pdm); free(dm1P); free(t1); free(t2); return -2; }
        for (int q = 0; q < n1; q++) {   /* compute_sPt(..., sP1, smat1_time, false) */
            const int r = i1[o1 + q], c = j1[o1 + q];
            if (r < 0 || r >= nb2 || c < 0 || c >= nb2) { free(Pdm); free(dm1P); free(t1); free(t2); return -2; }
            const int a = r / nb, b = r % nb, k = c / nb, l = c % nb;
            t1[q] = expt ? zmul(S1[o1 + q], zcis((ek[a] - ekp[k] - ek[b] + ekp[l]) * t)) : S1[o1 + q];
        }
        for (int q = 0; q < n2; q++) {   /* compute_sPt(..., sP2, smat2_time, true) */
            const int r = i2[o2 + q], c = j2[o2 + q];
            if (r < 0 || r >= nb2 || c < 0 || c >= nb2) { free(Pdm); free(dm1P); free(t1); free(t2); return -2; }
            const int a = r / nb, b = r % nb, k = c / nb, l = c % nb;
            t2[q] = expt ? zmul(S2[o2 + q], zcis((ek[k] - ekp[a] - ek[l] + ekp[b]) * t)) : S2[o2 + q];
        }
        for (int q = 0; q < n1; q++) {   /* :78 */
            const size_t o = ik * nb2 + i1[o1 + q];
            Pdm[o] = zadd(Pdm[o], zmul(t1[q], D[ikp * nb2 + j1[o1 + q]]));
        }
        for (int q = 0; q < n2; q++) {   /* :79 */
            const size_t o = ik * nb2 + j2[o2 + q];
            dm1P[o] = zadd(dm1P[o], zmul(D1[ikp * nb2 + i2[o2 + q]], t2[q]));
        }
        if (ik < ikp) {                  /* :80-85: sm1_next = conj(smat2_time), sm2_next = conj(smat1_time) */
            for (int q = 0; q < n2; q++) {
                const size_t o = ikp * nb2 + i2[o2 + q];
                Pdm[o] = zadd(Pdm[o], zmul(zconj(t2[q]), D[ik * nb2 + j2[o2 + q]]));
            }
            for (int q = 0; q < n1; q++) {
                const size_t o = ikp * nb2 + j1[o1 + q];
                dm1P[o] = zadd(dm1P[o], zmul(D1[ik * nb2 + i1[o1 + q]], zconj(t1[q])));
            }
        }
        o1 += n1;
        o2 += n2;
    }
    assemble(nk, nb, e, dm, dm1, Pdm, dm1P, t, prefac, NULL, 0, ddmdt);
    free(Pdm); free(dm1P); free(t1); free(t2);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Linearized operator (alg_linearize), e-ph part.                           */
/* electronphonon::set_Lsc, src_v4.5.7/ElectronPhonon_linear.cpp:3-36, with  */
/* add_Lscii_from_P1_P2 (:147-181) and add_Lscij_from_P1_P2 (:183-207, fac2  */
/* off) over the stored pairs in file order, then Lscii[k] added to both     */
/* matrices of every diagonal pair (k, k) (:25-34).  f_eq: nk x nb.          */
/* ------------------------------------------------------------------------ */
int oracle_set_Lsc(int64_t nk, int nb, int64_t npairs, const uint64_t *k1, const uint64_t *k2,
                   const double *P1, const double *P2, const double *f_eq, double *Lij, double *Lji)
{
    const int nb2 = nb * nb, nb4 = nb2 * nb2;
    if (nb < 1 || nb > 8 || nk < 0 || npairs < 0) return -1;
    for (int64_t ip = 0; ip < npairs; ip++)
        if (k1[ip] > k2[ip] || (int64_t)k2[ip] >= nk) return -2;
    zc *Lii = (zc *)calloc((size_t)nk * nb4 + 1, sizeof(zc));
    if (!Lii) return -3;
    memset(Lij, 0, sizeof(zc) * (size_t)npairs * nb4);
    memset(Lji, 0, sizeof(zc) * (size_t)npairs * nb4);
    for (int64_t ip = 0; ip < npairs; ip++) {
        const int64_t ik = (int64_t)k1[ip], jk = (int64_t)k2[ip];
        const zc *p1 = (const zc *)P1 + (size_t)ip * nb4, *p2 = (const zc *)P2 + (size_t)ip * nb4;
        const double *fi = f_eq + ik * nb, *fj = f_eq + jk * nb;
        zc *li = Lii + (size_t)ik * nb4, *lj = Lii + (size_t)jk * nb4;
        for (int i1 = 0; i1 < nb; i1++)
            for (int i2 = 0; i2 < nb; i2++) {
                const int n12 = (i1 * nb + i2) * nb2;
                /* -delta_1a P1_{b2,33} f^jk_3 (and the conj(P2) mirror on jk) */
                for (int ib = 0; ib < nb; ib++) {
                    const int nb2b = (ib * nb + i2) * nb2, i12ab = n12 + i1 * nb + ib;
                    for (int i3 = 0; i3 < nb; i3++) {
                        const int ib233 = nb2b + i3 * nb + i3;
                        li[i12ab] = zsub(li[i12ab], zscale(p1[ib233], fj[i3]));
                        if (jk > ik) lj[i12ab] = zsub(lj[i12ab], zscale(zconj(p2[ib233]), fi[i3]));
                    }
                }
                /* -(1 - f^jk_3) P2_{33,1a} delta_2b (and the conj(P1) mirror) */
                for (int ia = 0; ia < nb; ia++) {
                    const int i1a = i1 * nb + ia, i12ab = n12 + ia * nb + i2;
                    for (int i3 = 0; i3 < nb; i3++) {
                        const int i331a = (i3 * nb + i3) * nb2 + i1a;
                        li[i12ab] = zsub(li[i12ab], zscale(p2[i331a], 1 - fj[i3]));
                        if (jk > ik) lj[i12ab] = zsub(lj[i12ab], zscale(zconj(p1[i331a]), 1 - fi[i3]));
                    }
                }
            }
        zc *a = (zc *)Lij + (size_t)ip * nb4, *b = (zc *)Lji + (size_t)ip * nb4;
        for (int i1 = 0; i1 < nb; i1++)
            for (int i2 = 0; i2 < nb; i2++) {
                const int i12 = i1 * nb + i2, n12 = i12 * nb2;
                for (int ia = 0; ia < nb; ia++)
                    for (int ib = 0; ib < nb; ib++) {
                        const int iab = ia * nb + ib, i12ab = n12 + iab, iab12 = iab * nb2 + i12;
                        a[i12ab] = zadd(a[i12ab], zadd(zscale(p1[i12ab], 1 - fi[i1]), zscale(p2[iab12], fi[i2])));
                        b[i12ab] = zadd(b[i12ab], zadd(zscale(zconj(p2[i12ab]), 1 - fj[i1]),
                                                       zscale(zconj(p1[iab12]), fj[i2])));
                    }
            }
    }
    for (int64_t ip = 0; ip < npairs; ip++)
        if (k1[ip] == k2[ip]) {
            const zc *l = Lii + (size_t)k1[ip] * nb4;
            zc *a = (zc *)Lij + (size_t)ip * nb4, *b = (zc *)Lji + (size_t)ip * nb4;
            for (int i = 0; i < nb4; i++) { a[i] = zadd(a[i], l[i]); b[i] = zadd(b[i], l[i]); }
        }
    free(Lii);
    return 0;
}

/* The alg_linearize_dPee part of electronphonon::evolve (compute_eq false),  */
/* src_v4.5.7/evolve_eph.cpp:107-131 (row sums), 138, 144-151 (row term and  */
/* the Hermitian assembly): for every stored pair (k, k')                     */
/*   dPdm[k][i12]  += sum_i3 dP1[(i12),(i3 i3)] f_eq[k'][i3]                  */
/*   dm1dP[k][i12] += sum_i3 f1_eq[k'][i3] dP2[(i3 i3),(i12)]                 */
/* and for k < k' the same from P2_next = conj(dP1), P1_next = conj(dP2) into */
/* row k'; then contrib[ij] = (f1_eq[k][i] dPdm[ij] - dm1dP[ij] f_eq[k][j])   */
/* cis((e_i - e_j) t) and out[ij] += prefac/2 (contrib[ij] + conj(contrib[ji])).  */
/* f1_eq = 1 - f_eq (ElectronPhonon.cpp:28-31).  out: nk x nb^2, added to.    */
int oracle_ee_dP_term(int64_t nk, int nb, int64_t npairs, const uint64_t *k1, const uint64_t *k2, const double *dP1,
                      const double *dP2, const double *e, const double *f_eq, double t, double prefac, double *out)
{
    int nb2 = nb * nb, nb4 = nb2 * nb2;
    zc *dPdm = (zc *)calloc((size_t)nk * nb2, sizeof(zc)), *dm1dP = (zc *)calloc((size_t)nk * nb2, sizeof(zc));
    zc *P1n = (zc *)malloc(sizeof(zc) * nb4), *P2n = (zc *)malloc(sizeof(zc) * nb4), *c = (zc *)malloc(sizeof(zc) * nb2);
    if (!dPdm || !dm1dP || !P1n || !P2n || !c) { free(dPdm); free(dm1dP); free(P1n); free(P2n); free(c); return 1; }
    for (int64_t p = 0; p < npairs; p++) {
        int64_t k = (int64_t)k1[p], kp = (int64_t)k2[p];
        const zc *D1 = (const zc *)dP1 + p * nb4, *D2 = (const zc *)dP2 + p * nb4;
        for (int i1 = 0; i1 < nb; i1++)
            for (int i2 = 0; i2 < nb; i2++) {
                int i12 = i1 * nb + i2, n12 = i12 * nb2;
                for (int i3 = 0; i3 < nb; i3++) {
                    dPdm[k * nb2 + i12] = zadd(dPdm[k * nb2 + i12], zscale(D1[n12 + i3 * nb + i3], f_eq[kp * nb + i3]));
                    dm1dP[k * nb2 + i12] = zadd(dm1dP[k * nb2 + i12],
                                                zscale(D2[(i3 * nb + i3) * nb2 + i12], 1 - f_eq[kp * nb + i3]));
                }
            }
        if (k < kp) {
            for (int q = 0; q < nb4; q++) { P2n[q] = zconj(D1[q]); P1n[q] = zconj(D2[q]); }
            for (int i1 = 0; i1 < nb; i1++)
                for (int i2 = 0; i2 < nb; i2++) {
                    int i12 = i1 * nb + i2, n12 = i12 * nb2;
                    for (int i3 = 0; i3 < nb; i3++) {
                        dPdm[kp * nb2 + i12] = zadd(dPdm[kp * nb2 + i12], zscale(P1n[n12 + i3 * nb + i3], f_eq[k * nb + i3]));
                        dm1dP[kp * nb2 + i12] = zadd(dm1dP[kp * nb2 + i12],
                                                     zscale(P2n[(i3 * nb + i3) * nb2 + i12], 1 - f_eq[k * nb + i3]));
                    }
                }
        }
    }
    for (int64_t k = 0; k < nk; k++) {
        for (int i = 0; i < nb; i++)
            for (int j = 0; j < nb; j++) {
                zc v = zsub(zscale(dPdm[k * nb2 + i * nb + j], 1 - f_eq[k * nb + i]),
                            zscale(dm1dP[k * nb2 + i * nb + j], f_eq[k * nb + j]));
                c[i * nb + j] = zmul(v, zcis((e[k * nb + i] - e[k * nb + j]) * t));
            }
        for (int i = 0; i < nb; i++)
            for (int j = 0; j < nb; j++) {
                size_t o = (size_t)k * nb2 + i * nb + j;
                ZW(out, o) = zadd(Z(out, o), zscale(zadd(c[i * nb + j], zconj(c[j * nb + i])), 0.5 * prefac));
            }
    }
    free(dPdm); free(dm1dP); free(P1n); free(P2n); free(c);
    return 0;
}

/* electronphonon::evolve_linear, src_v4.5.7/evolve_eph_linear.cpp:3-45, with
 * compute_ddmdt (:47-53): drho = dm - diag(f_eq); per stored pair
 * X = Lsct_ij drho_k' added to row k as prefac/2 (X + X^+), and for k < k'
 * Y = Lsct_ji drho_k added to row k'.  Lsct carries the P1-type phase of
 * compute_Pt (:55-73) when expt is set. */
int oracle_evolve_linear(int64_t nk, int nb, int64_t npairs, const uint64_t *k1, const uint64_t *k2,
                         const double *Lij, const double *Lji, const double *e, const double *dm,
                         const double *f_eq, double t, double prefac, int expt, double *ddmdt)
{
    const int nb2 = nb * nb, nb4 = nb2 * nb2;
    if (nb < 1 || nb > 8 || nk < 0 || npairs < 0) return -1;
    for (int64_t ip = 0; ip < npairs; ip++)
        if (k1[ip] > k2[ip] || (int64_t)k2[ip] >= nk) return -2;
    zc *d = (zc *)malloc(sizeof(zc) * ((size_t)nk * nb2 + 1)), *Lt = (zc *)malloc(sizeof(zc) * nb4);
    if (!d || !Lt) { free(d); free(Lt); return -3; }
    memcpy(d, dm, sizeof(zc) * (size_t)nk * nb2);
    for (int64_t ik = 0; ik < nk; ik++)
        for (int i = 0; i < nb; i++) d[ik * nb2 + i * nb + i].x -= f_eq[ik * nb + i];
    memset(ddmdt, 0, sizeof(zc) * (size_t)nk * nb2);
    zc contrib[64];
    for (int64_t ip = 0; ip < npairs; ip++) {
        const int64_t ik = (int64_t)k1[ip], ikp = (int64_t)k2[ip];
        for (int dir = 0; dir < 2; dir++) {
            if (dir == 1 && !(ik < ikp)) break;
            const int64_t row = dir ? ikp : ik, col = dir ? ik : ikp;
            const double *L = (dir ? Lji : Lij) + (size_t)ip * nb4 * 2;
            if (expt) oracle_compute_Pt(nb, t, e + row * nb, e + col * nb, L, (double *)Lt, 0);
            else memcpy(Lt, L, sizeof(zc) * nb4);
            for (int i = 0; i < nb2; i++) contrib[i] = zmk(0, 0);
            gemv_Ab(contrib, Lt, d + col * nb2, nb2);
            zc *out = (zc *)ddmdt + row * nb2;
            for (int i = 0; i < nb; i++)
                for (int j = 0; j < nb; j++)
                    out[i * nb + j] = zadd(out[i * nb + j],
                                           zscale(zadd(contrib[i * nb + j], zconj(contrib[j * nb + i])), prefac * 0.5));
        }
    }
    free(d); free(Lt);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Initializer one-step form: ElectronPhonon::compute_ddm and its caller's   */
/* pair loop, lindbladInit_for-DMD-4.5.6/ElectronPhonon.h:755-773 (term1_P,  */
/* term2_P), :815-837 (reverse with P1_next=conj(P2), P2_next=conj(P1)) and  */
/* :855-862 (per-pair Hermitian assembly).  The e-i twin                     */
/* (ElectronImpurity.h:663-771) is the same with P scaled by the defect      */
/* fraction.  ddm is accumulated (+=), as in the reference.                  */
/* ------------------------------------------------------------------------ */
static void compute_ddm_one(int nb, const zc *dmk, const zc *dmkp, const zc *dm1k, const zc *dm1kp,
                            const zc *p1, const zc *p2, double prefac, zc *ddmk)
{
    int nb2 = nb * nb;
    zc maux1[64] = {{0, 0}}, maux2[64] = {{0, 0}}, contrib[64];
    const zc c1 = zmk(1, 0), c0 = zmk(0, 0);
    for (int i = 0; i < nb2; i++) contrib[i] = c0;
    /* term1_P: + (1-dm_k) . mat(P1 . vec dm_k') */
    for (int i = 0; i < nb2; i++) maux1[i] = c0;
    gemv_Ab(maux1, p1, dmkp, nb2);
    zhemm_upper(maux2, 1, dm1k, maux1, nb, c1, c0);
    for (int i = 0; i < nb2; i++) contrib[i] = zadd(contrib[i], maux2[i]);
    /* term2_P: - mat(vec(1-dm_k')^T . P2) . dm_k */
    for (int i = 0; i < nb2; i++) maux1[i] = c0;
    gemv_aB(maux1, dm1kp, p2, nb2);
    zhemm_upper(maux2, 0, dmk, maux1, nb, c1, c0);
    for (int i = 0; i < nb2; i++) contrib[i] = zsub(contrib[i], maux2[i]);
    for (int i = 0; i < nb; i++)
        for (int j = 0; j < nb; j++)
            ddmk[i * nb + j] = zadd(ddmk[i * nb + j],
                                    zscale(zadd(contrib[i * nb + j], zconj(contrib[j * nb + i])), prefac * 0.5));
}

int oracle_compute_ddm_pairs(int64_t nk, int nb, int64_t npairs, const uint64_t *k1, const uint64_t *k2,
                             const double *P1, const double *P2, const double *dm, const double *dm1,
                             double prefac, double *ddm)
{
    int nb2 = nb * nb, nb4 = nb2 * nb2;
    if (nb < 1 || nb > 8) return -1;
    zc *P1n = (zc *)malloc(sizeof(zc) * nb4), *P2n = (zc *)malloc(sizeof(zc) * nb4);
    for (int64_t ip = 0; ip < npairs; ip++) {
        int64_t ik = (int64_t)k1[ip], ikp = (int64_t)k2[ip];
        if (ik > ikp || ikp >= nk) { free(P1n); free(P2n); return -2; }
        const zc *p1 = (const zc *)P1 + (size_t)ip * nb4, *p2 = (const zc *)P2 + (size_t)ip * nb4;
        compute_ddm_one(nb, (const zc *)dm + ik * nb2, (const zc *)dm + ikp * nb2,
                        (const zc *)dm1 + ik * nb2, (const zc *)dm1 + ikp * nb2, p1, p2, prefac,
                        (zc *)ddm + ik * nb2);
        if (ik < ikp) {
            for (int i = 0; i < nb4; i++) { P1n[i] = zconj(p2[i]); P2n[i] = zconj(p1[i]); }
            compute_ddm_one(nb, (const zc *)dm + ikp * nb2, (const zc *)dm + ik * nb2,
                            (const zc *)dm1 + ikp * nb2, (const zc *)dm1 + ik * nb2, P1n, P2n, prefac,
                            (zc *)ddm + ikp * nb2);
        }
    }
    free(P1n); free(P2n);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* P build of the initializer, lindbladInit_for-DMD-4.5.6/ElectronPhonon.h:  */
/* 432-575 (lindblad=true, gaussian smearing, no energy-window cut, optional */
/* detailed balance).  Inputs per branch a: omega[a] (= wqp = wqm), gm[a]    */
/* (nb x nb, g^-_kk') and gp[a] (nb x nb, g^+_kk' = dagger(g^-_k'k)), all    */
/* interleaved complex.  P1/P2 are overwritten.                              */
/* ------------------------------------------------------------------------ */
static double bose(double x) { return 1.0 / (exp(x) - 1.0); }

int oracle_compute_P(int nb, const double *Ek, const double *Ekp, int nm, const double *omega,
                     const double *gm, const double *gp, double T, double sigma, int detail_balance,
                     double *P1, double *P2)
{
    int nb2 = nb * nb, nb4 = nb2 * nb2;
    if (nb < 1 || nb > 8) return -1;
    zc App[64], Amm[64], Amp[64], Apm[64], A2pp[64], A2mm[64], A2mp[64], A2pm[64];
    for (int i = 0; i < nb4; i++) { ZW(P1, i) = zmk(0, 0); ZW(P2, i) = zmk(0, 0); }
    double sigma2 = sigma * sigma;
    double prefac_exp = -0.5 / sigma2;
    double prefac_delta = 1. / (sigma * sqrt(2. * M_PI));
    double prefac_sqrtdelta = sqrt(prefac_delta);
    for (int a = 0; a < nm; a++) {
        double wqm = omega[a], wqp = omega[a];
        double nqm = bose(fmax(1e-3, wqm / T)), nqp = bose(fmax(1e-3, wqp / T));
        const zc *gpa = (const zc *)gp + a * nb2, *gma = (const zc *)gm + a * nb2;
        for (int b1 = 0; b1 < nb; b1++)
            for (int b2 = 0; b2 < nb; b2++) {
                double dE = Ek[b1] - Ekp[b2];
                /* phonon emission, :475-487 and :497-503 */
                double deltaplus = exp(prefac_exp * pow(dE + wqp, 2));
                zc G2p = zscale(gpa[b1 * nb + b2], prefac_sqrtdelta * sqrt(deltaplus));
                double dEbyT = dE / T;
                double nq1_DB = (detail_balance && -dEbyT < 46) ? (exp(-dEbyT) * nqp) : nqp + 1;
                App[b1 * nb + b2] = zscale(G2p, sqrt(nq1_DB));
                Apm[b1 * nb + b2] = zscale(G2p, sqrt(nqp));
                A2pp[b1 * nb + b2] = zscale(G2p, sqrt(nq1_DB));
                A2pm[b1 * nb + b2] = zscale(G2p, sqrt(nqp));
                /* phonon absorption, :507-519 and :535-540 */
                double deltaminus = exp(prefac_exp * pow(dE - wqm, 2));
                zc G2m = zscale(gma[b1 * nb + b2], prefac_sqrtdelta * sqrt(deltaminus));
                nq1_DB = (detail_balance && dEbyT < 46) ? (exp(dEbyT) * nqm) : nqm + 1;
                Amp[b1 * nb + b2] = zscale(G2m, sqrt(nq1_DB));
                Amm[b1 * nb + b2] = zscale(G2m, sqrt(nqm));
                A2mp[b1 * nb + b2] = zscale(G2m, sqrt(nq1_DB));
                A2mm[b1 * nb + b2] = zscale(G2m, sqrt(nqm));
            }
        /* :558-571 */
        for (int i1 = 0; i1 < nb; i1++)
            for (int i2 = 0; i2 < nb; i2++) {
                int n12 = (i1 * nb + i2) * nb2;
                for (int i3 = 0; i3 < nb; i3++) {
                    int i13 = i1 * nb + i3, i31 = i3 * nb + i1;
                    for (int i4 = 0; i4 < nb; i4++) {
                        zc t1 = zadd(zmul(App[i13], zconj(A2pp[i2 * nb + i4])), zmul(Amm[i13], zconj(A2mm[i2 * nb + i4])));
                        zc t2 = zadd(zmul(Amp[i31], zconj(A2mp[i4 * nb + i2])), zmul(Apm[i31], zconj(A2pm[i4 * nb + i2])));
                        ZW(P1, n12 + i3 * nb + i4) = zadd(Z(P1, n12 + i3 * nb + i4), t1);
                        ZW(P2, n12 + i3 * nb + i4) = zadd(Z(P2, n12 + i3 * nb + i4), t2);
                    }
                }
            }
    }
    return 0;
}

/* e-i P in the model form without detailed balance,                        */
/* src_v4.5.7/scattering_model/ElecImp_Model.h:185-214:                     */
/* P1[(i1 i2),(i3 i4)] = A1[i1,i3] conj(A2[i2,i4]);  P2 = A1[i3,i1] conj(A2[i4,i2]). */
int oracle_compute_P_eimp(int nb, const double *A1, const double *A2, double *P1, double *P2)
{
    int nb2 = nb * nb;
    for (int i1 = 0; i1 < nb; i1++)
        for (int i2 = 0; i2 < nb; i2++) {
            int n12 = (i1 * nb + i2) * nb2;
            for (int i3 = 0; i3 < nb; i3++) {
                int i13 = i1 * nb + i3, i31 = i3 * nb + i1;
                for (int i4 = 0; i4 < nb; i4++) {
                    ZW(P1, n12 + i3 * nb + i4) = zmul(Z(A1, i13), zconj(Z(A2, i2 * nb + i4)));
                    ZW(P2, n12 + i3 * nb + i4) = zmul(Z(A1, i31), zconj(Z(A2, i4 * nb + i2)));
                }
            }
        }
    return 0;
}

/* Static screened Coulomb potential, coulomb_model::vq                      */
/* (src_v4.5.7/scattering_model/Coulomb_Model.h:632-662) for the formulas     */
/* that depend on |q| alone, q wrapped into [-0.5, 0.5) (comm/kmap.h:6-15)    */
/* and measured with GGT (comm/matrix3.h:119-122).                            */
/* screening: 0 unscreened, 1 debye, 2 Bechstedt.                            */
double oracle_eimp_q2(const double *GGT, const double *ka, const double *kb)
{
    double v[3];
    for (int d = 0; d < 3; d++) {
        v[d] = ka[d] - kb[d];
        v[d] -= floor(0.5 + v[d]);
        if (fabs(v[d] - 0.5) < 1e-6) v[d] = -0.5;
    }
    return v[0] * v[0] * GGT[0] + v[1] * v[1] * GGT[4] + v[2] * v[2] * GGT[8] +
           2 * (v[0] * v[1] * GGT[1] + v[0] * v[2] * GGT[2] + v[1] * v[2] * GGT[5]);
}
double oracle_eimp_vq(int screening, double q2, double prefac_vq, double prefac_vq_bare, double qscr2_debye,
                      double fac0, double fac2, double fac4)
{
    if (screening == 1) return prefac_vq / (qscr2_debye + q2);
    if (q2 < 1e-20) return 0.0;
    if (screening == 2) {
        double epsq = 1 + 1. / (fac0 + fac2 * q2 + fac4 * q2 * q2);
        return prefac_vq_bare / epsq / q2;
    }
    return prefac_vq / q2;
}

/* Model ionized-impurity P of stored pairs, elecimp_model::calc_P with       */
/* calc_ovlp, calc_eimp and calc_A (ElecImp_Model.h:67-137,185-213):          */
/* ovlp = U_k^H U_k' (U: nk x nw x nb), eimp = prefac_A v(q) ovlp,            */
/* A2 = eimp sqrt-Gaussian (lindblad) or Gaussian (conventional) of            */
/* de = e_k,b1 - e_k',b2 inside |de| < ethr, A1 = A2 or eimp; with detailed   */
/* balance the +-omegaL pair and the exp(-+de/2T) factors.  P1, P2 are        */
/* npairs x nb^4, overwritten.                                                 */
/* kIndexMap::ikvec3 / qIndexMap::iqvec3 (src_v4.5.7/comm/kmap.cpp): mesh    */
/* cell of a point in crystal coordinates, wrapped into [0, 1) first.         */
static int64_t mesh_code(const double *k, const int *km)
{
    int v[3];
    for (int d = 0; d < 3; d++) {
        double x = k[d] - floor(k[d]);
        v[d] = (int)round(km[d] * x);
        if (v[d] == km[d]) v[d] = 0;
    }
    return ((int64_t)v[0] * km[1] + v[1]) * km[2] + v[2];
}

/* Static RPA screening and v(q) of every stored pair, coulomb_model:        */
/* qIndexMap::build over all (ik, jk) (kmap.h:43-55), kIndexMap (first k of  */
/* a cell kept), calc_qscr2_static_RPA (Coulomb_Model.h:690-744),            */
/* calc_vq_RPA's static table (:820-828) and vq (:670-675).  f: nk x nb at mu. */
int oracle_eimp_rpa_vq(int64_t nk, int nb, int nw, const double *kvec, const int *kmesh, const double *U,
                       const double *e, const double *f, double T, const double *GGT, double prefac_vq,
                       double nk_full, int64_t npairs, const uint64_t *k1, const uint64_t *k2, double *vq_pair,
                       double *qscr2_out, int64_t *nq_out, double *vq_cell_out)
{
    int64_t cells = (int64_t)kmesh[0] * kmesh[1] * kmesh[2];
    int64_t *kmap = (int64_t *)malloc(sizeof(int64_t) * cells), *qmap = (int64_t *)malloc(sizeof(int64_t) * cells);
    double *qvec = (double *)malloc(sizeof(double) * 3 * cells);
    for (int64_t c = 0; c < cells; c++) kmap[c] = qmap[c] = -1;
    for (int64_t ik = 0; ik < nk; ik++) {
        int64_t c = mesh_code(kvec + 3 * ik, kmesh);
        if (kmap[c] < 0) kmap[c] = ik;
    }
    int64_t nq = 0;
    for (int64_t ik = 0; ik < nk; ik++)
        for (int64_t jk = 0; jk < nk; jk++) {
            double q[3];
            for (int d = 0; d < 3; d++) {
                q[d] = kvec[3 * ik + d] - kvec[3 * jk + d];
                q[d] -= floor(0.5 + q[d]);
                if (fabs(q[d] - 0.5) < 1e-6) q[d] = -0.5;
            }
            int64_t c = mesh_code(q, kmesh);
            if (qmap[c] < 0) {
                qmap[c] = nq;
                memcpy(qvec + 3 * nq, q, sizeof q);
                nq++;
            }
        }
    double *qscr2 = (double *)calloc(nq, sizeof(double)), *vq = (double *)malloc(sizeof(double) * nq);
    for (int64_t iq = 0; iq < nq; iq++) {
        double s = 0;
        for (int64_t ik = 0; ik < nk; ik++) {
            double kj[3];
            for (int d = 0; d < 3; d++) kj[d] = kvec[3 * ik + d] - qvec[3 * iq + d];
            int64_t jk = kmap[mesh_code(kj, kmesh)];
            if (jk < 0) continue;
            const zc *Ui = (const zc *)U + ik * nw * nb, *Uj = (const zc *)U + jk * nw * nb;
            for (int b1 = 0; b1 < nb; b1++)
                for (int b2 = 0; b2 < nb; b2++) {
                    double de = e[ik * nb + b1] - e[jk * nb + b2], dfde;
                    if (fabs(de) < 1e-8) {
                        double favg = 0.5 * (f[ik * nb + b1] + f[jk * nb + b2]);
                        dfde = (1 - favg) * favg / T;
                    } else {
                        dfde = (f[jk * nb + b2] - f[ik * nb + b1]) / de;
                    }
                    zc o = zmk(0, 0);
                    for (int w = 0; w < nw; w++) o = zadd(o, zmul(zconj(Ui[w * nb + b1]), Uj[w * nb + b2]));
                    s += dfde * (o.x * o.x + o.y * o.y);
                }
        }
        qscr2[iq] = prefac_vq / nk_full * s;
        vq[iq] = prefac_vq / (oracle_eimp_q2(GGT, qvec + 3 * iq, (const double[3]){0, 0, 0}) + qscr2[iq]);
    }
    for (int64_t p = 0; p < npairs; p++) {
        double q[3];
        for (int d = 0; d < 3; d++) q[d] = kvec[3 * k1[p] + d] - kvec[3 * k2[p] + d];
        double q2 = oracle_eimp_q2(GGT, kvec + 3 * k1[p], kvec + 3 * k2[p]);
        vq_pair[p] = q2 < 1e-20 ? 0.0 : vq[qmap[mesh_code(q, kmesh)]];
    }
    if (qscr2_out) memcpy(qscr2_out, qscr2, sizeof(double) * nq);
    if (nq_out) *nq_out = nq;
    /* v(q) of every mesh cell that is a difference of two k points (0 at q = 0,
     * as vq skips the Gamma point, :670-675); other cells 0 */
    if (vq_cell_out)
        for (int64_t c = 0; c < cells; c++) {
            vq_cell_out[c] = 0.0;
            if (qmap[c] < 0) continue;
            double q2 = oracle_eimp_q2(GGT, qvec + 3 * qmap[c], (const double[3]){0, 0, 0});
            vq_cell_out[c] = q2 < 1e-20 ? 0.0 : vq[qmap[c]];
        }
    free(kmap); free(qmap); free(qvec); free(qscr2); free(vq);
    return 0;
}

int oracle_eimp_model_P(int64_t nk, int nb, int nw, int64_t npairs, const uint64_t *k1, const uint64_t *k2,
                        const double *U, const double *kvec, const double *e, const double *GGT, int screening,
                        int lindblad, int detail_balance, double prefac_A, double prefac_vq, double prefac_vq_bare,
                        double qscr2_debye, double fac0, double fac2, double fac4, double degauss, double ethr,
                        double omegaL, double T, const double *vq_pair, double *P1, double *P2)
{
    int nb2 = nb * nb, nb4 = nb2 * nb2;
    double pgauss = 1. / (degauss * sqrt(2. * M_PI));
    double pdelta = lindblad ? sqrt(pgauss) : pgauss;
    if (detail_balance) pdelta *= 0.5;
    double pexp = lindblad ? -0.25 / (degauss * degauss) : -0.5 / (degauss * degauss);
    for (int64_t p = 0; p < npairs; p++)
        if (k1[p] >= (uint64_t)nk || k2[p] >= (uint64_t)nk) return -1;
    zc *ovlp = (zc *)malloc(sizeof(zc) * nb2), *eimp = (zc *)malloc(sizeof(zc) * nb2);
    zc *A[8];
    for (int a = 0; a < 8; a++) A[a] = (zc *)calloc(nb2, sizeof(zc));
    for (int64_t p = 0; p < npairs; p++) {
        int64_t ik = (int64_t)k1[p], jk = (int64_t)k2[p];
        const zc *Ui = (const zc *)U + ik * nw * nb, *Uj = (const zc *)U + jk * nw * nb;
        for (int b1 = 0; b1 < nb; b1++)
            for (int b2 = 0; b2 < nb; b2++) {
                zc o = zmk(0, 0);
                for (int w = 0; w < nw; w++) o = zadd(o, zmul(zconj(Ui[w * nb + b1]), Uj[w * nb + b2]));
                ovlp[b1 * nb + b2] = o;
            }
        double vq = vq_pair ? vq_pair[p]
                            : oracle_eimp_vq(screening, oracle_eimp_q2(GGT, kvec + 3 * ik, kvec + 3 * jk), prefac_vq,
                                             prefac_vq_bare, qscr2_debye, fac0, fac2, fac4);
        for (int q = 0; q < nb2; q++) eimp[q] = zmul(zmk(prefac_A * vq, 0), ovlp[q]);
        for (int b1 = 0; b1 < nb; b1++)
            for (int b2 = 0; b2 < nb; b2++) {
                int q = b1 * nb + b2;
                double de = e[ik * nb + b1] - e[jk * nb + b2];
                if (!detail_balance) {
                    A[1][q] = zmk(0, 0);
                    if (fabs(de) < ethr) A[1][q] = zscale(eimp[q], pdelta * exp(pexp * pow(de, 2)));
                    A[0][q] = lindblad ? A[1][q] : eimp[q];
                } else {
                    zc G2p = zmk(0, 0), G2m = zmk(0, 0);
                    if (fabs(de + omegaL) < ethr) G2p = zscale(eimp[q], pdelta * exp(pexp * pow(de + omegaL, 2)));
                    zc G1p = lindblad ? G2p : eimp[q];
                    double dEbyT = de / T;
                    double fDB = (-dEbyT < 46) ? exp(-dEbyT / 2) : 1;
                    A[0][q] = zscale(G1p, fDB); A[1][q] = G1p; A[2][q] = zscale(G2p, fDB); A[3][q] = G2p;
                    if (fabs(de - omegaL) < ethr) G2m = zscale(eimp[q], pdelta * exp(pexp * pow(de - omegaL, 2)));
                    zc G1m = lindblad ? G2m : eimp[q];
                    fDB = (dEbyT < 46) ? exp(dEbyT / 2) : 1;
                    A[4][q] = zscale(G1m, fDB); A[5][q] = G1m; A[6][q] = zscale(G2m, fDB); A[7][q] = G2m;
                }
            }
        zc *Q1 = (zc *)P1 + p * nb4, *Q2 = (zc *)P2 + p * nb4;
        for (int i1 = 0; i1 < nb; i1++)
            for (int i2 = 0; i2 < nb; i2++) {
                int n12 = (i1 * nb + i2) * nb2;
                for (int i3 = 0; i3 < nb; i3++) {
                    int i13 = i1 * nb + i3, i31 = i3 * nb + i1;
                    for (int i4 = 0; i4 < nb; i4++) {
                        int i24 = i2 * nb + i4, i42 = i4 * nb + i2;
                        if (!detail_balance) {
                            Q1[n12 + i3 * nb + i4] = zmul(A[0][i13], zconj(A[1][i24]));
                            Q2[n12 + i3 * nb + i4] = zmul(A[0][i31], zconj(A[1][i42]));
                        } else {  /* pp + mm, mp + pm */
                            Q1[n12 + i3 * nb + i4] = zadd(zmul(A[0][i13], zconj(A[2][i24])), zmul(A[5][i13], zconj(A[7][i24])));
                            Q2[n12 + i3 * nb + i4] = zadd(zmul(A[4][i31], zconj(A[6][i42])), zmul(A[1][i31], zconj(A[3][i42])));
                        }
                    }
                }
            }
    }
    free(ovlp); free(eimp);
    for (int a = 0; a < 8; a++) free(A[a]);
    return 0;
}

/* v(k_a - k_b) of the e-e model: the per-cell table (static RPA, the q list  */
/* of every k-point difference) when given, else the closed-form screening   */
/* (coulomb_model::vq, Coulomb_Model.h:632-662).                              */
static double ee_vq(int64_t a, int64_t b, const double *kvec, const int *kmesh, const double *GGT,
                    const double *vq_cell, int screening, double prefac_vq, double prefac_vq_bare,
                    double qscr2_debye, double fac0, double fac2, double fac4)
{
    double q2 = oracle_eimp_q2(GGT, kvec + 3 * a, kvec + 3 * b);
    if (vq_cell) {
        if (q2 < 1e-20) return 0.0;
        double q[3];
        for (int d = 0; d < 3; d++) q[d] = kvec[3 * a + d] - kvec[3 * b + d];
        return vq_cell[mesh_code(q, kmesh)];
    }
    return oracle_eimp_vq(screening, q2, prefac_vq, prefac_vq_bare, qscr2_debye, fac0, fac2, fac4);
}

/* calc_A1rho_dP (ElecElec_Model.h:271-281): with ddml[i][j] = dmleft[i][j] -  */
/* fl[i] and ddmr[i][j] = dmright[i][j] - fr[i] (the reference subtracts the   */
/* row's occupation from every element of the row, and its zhemm reads the   */
/* upper triangle), out = fl o (A ddmr) + (ddml A) o fr, with fl[i] scaling   */
/* rows and fr[j] columns.  fl/fr are f (or 1 - f when *_om is set).          */
static void a1rho_dP(zc *out, const zc *A, const zc *dmleft, const double *fl0, int fl_om, const zc *dmright,
                     const double *fr0, int fr_om, int nb)
{
    zc ddml[64], ddmr[64], t[64];
    double fl[8], fr[8];
    for (int i = 0; i < nb; i++) {
        fl[i] = fl_om ? 1 - fl0[i] : fl0[i];
        fr[i] = fr_om ? 1 - fr0[i] : fr0[i];
    }
    for (int i = 0; i < nb; i++)
        for (int j = 0; j < nb; j++) {
            ddml[i * nb + j] = zsub(dmleft[i * nb + j], zmk(fl[i], 0));
            ddmr[i * nb + j] = zsub(dmright[i * nb + j], zmk(fr[i], 0));
        }
    const zc c1 = zmk(1, 0), c0 = zmk(0, 0);
    zhemm_upper(t, 0, ddmr, A, nb, c1, c0);
    for (int i = 0; i < nb; i++)
        for (int j = 0; j < nb; j++) out[i * nb + j] = zscale(t[i * nb + j], fl[i]);
    zhemm_upper(t, 1, ddml, A, nb, c1, c0);
    for (int i = 0; i < nb; i++)
        for (int j = 0; j < nb; j++) out[i * nb + j] = zadd(out[i * nb + j], zscale(t[i * nb + j], fr[j]));
}

/* Model e-e P at the equilibrium occupations, elecelec_model::calc_P with  */
/* dm = null and no exchange (scattering_model/ElecElec_Model.h:69-112,      */
/* 144-184,185-262,285-296): for each stored pair (k1, k2), k1 != k2, and    */
/* every k3 whose k4 = k1 + k3 - k2 is a kept k point (first k of a mesh      */
/* cell), mee = v(k1-k2) <1|2><3|4>, delta = exp(pexp de^2) inside ethr,      */
/* A2 = delta mee, A1 = A2 (lindblad) or mee;                                 */
/* P1[(i1 i2),(i3 i4)] += sum_xy (1-f3x) A1[(i1 i3),(x y)] f4y conj(A2[(i2 i4),(x y)]), */
/* P2[(i1 i2),(i3 i4)] += sum_xy f3x A1[(i3 i1),(x y)] (1-f4y) conj(A2[(i4 i2),(x y)]), */
/* times prefac_gauss / nk_full.  vq_pair: v(q) of each stored pair.          */
int oracle_ee_model_P(int64_t nk, int nb, int nw, int64_t npairs, const uint64_t *k1, const uint64_t *k2,
                      const double *U, const double *kvec, const int *kmesh, const double *e, const double *f,
                      const double *vq_pair, int lindblad, double degauss, double ethr, double nk_full, double *P1,
                      double *P2, int antisym, const double *GGT, int screening, double prefac_vq, double prefac_vq_bare,
                      double qscr2_debye, double fac0, double fac2, double fac4, const double *vq_cell,
                      const double *dm, const double *dm1, double *dP1, double *dP2)
{
    int nb2 = nb * nb, nb4 = nb2 * nb2;
    int64_t cells = (int64_t)kmesh[0] * kmesh[1] * kmesh[2];
    int64_t *kmap = (int64_t *)malloc(sizeof(int64_t) * cells);
    for (int64_t c = 0; c < cells; c++) kmap[c] = -1;
    for (int64_t ik = 0; ik < nk; ik++) {
        int64_t c = mesh_code(kvec + 3 * ik, kmesh);
        if (kmap[c] < 0) kmap[c] = ik;
    }
    double pgauss = 1. / (degauss * sqrt(2. * M_PI));
    double pexp = lindblad ? -0.25 / (degauss * degauss) : -0.5 / (degauss * degauss);
    zc *o12 = (zc *)malloc(sizeof(zc) * nb2), *o34 = (zc *)malloc(sizeof(zc) * nb2);
    zc *A1 = (zc *)malloc(sizeof(zc) * nb4), *A2 = (zc *)malloc(sizeof(zc) * nb4);
    zc *Ar = (zc *)malloc(sizeof(zc) * nb2), *Arb = (zc *)malloc(sizeof(zc) * nb2);
    zc *o14 = (zc *)malloc(sizeof(zc) * nb2), *o32 = (zc *)malloc(sizeof(zc) * nb2);
    double *dl = (double *)malloc(sizeof(double) * nb4);
    const zc c1 = zmk(1, 0), c0 = zmk(0, 0);
    zc *Ard = (zc *)malloc(sizeof(zc) * nb2), *Arbd = (zc *)malloc(sizeof(zc) * nb2);
    for (int64_t p = 0; p < npairs; p++) {
        int64_t a = (int64_t)k1[p], b = (int64_t)k2[p];
        zc *Q1 = (zc *)P1 + p * nb4, *Q2 = (zc *)P2 + p * nb4;
        zc *D1 = dP1 ? (zc *)dP1 + p * nb4 : NULL, *D2 = dP2 ? (zc *)dP2 + p * nb4 : NULL;
        for (int q = 0; q < nb4; q++) Q1[q] = Q2[q] = zmk(0, 0);
        if (D1)
            for (int q = 0; q < nb4; q++) D1[q] = D2[q] = zmk(0, 0);
        if (a == b && !antisym) continue;
        const zc *Ua = (const zc *)U + a * nw * nb, *Ub = (const zc *)U + b * nw * nb;
        for (int i = 0; i < nb; i++)
            for (int j = 0; j < nb; j++) {
                zc o = zmk(0, 0);
                for (int w = 0; w < nw; w++) o = zadd(o, zmul(zconj(Ua[w * nb + i]), Ub[w * nb + j]));
                o12[i * nb + j] = zmul(o, zmk(vq_pair[p], 0));
            }
        for (int64_t c3 = 0; c3 < nk; c3++) {
            if (antisym && a == b && b == c3) continue;
            double k4v[3];
            for (int d = 0; d < 3; d++) k4v[d] = kvec[3 * a + d] + kvec[3 * c3 + d] - kvec[3 * b + d];
            int64_t c4 = kmap[mesh_code(k4v, kmesh)];
            if (c4 < 0) continue;
            int conserved = 0;
            for (int i1 = 0; i1 < nb; i1++)
                for (int i2 = 0; i2 < nb; i2++)
                    for (int i3 = 0; i3 < nb; i3++)
                        for (int i4 = 0; i4 < nb; i4++) {
                            double de = e[a * nb + i1] - e[b * nb + i2] + e[c3 * nb + i3] - e[c4 * nb + i4];
                            int q = ((i1 * nb + i2) * nb + i3) * nb + i4;
                            dl[q] = 0;
                            if (fabs(de) < ethr) { dl[q] = exp(pexp * pow(de, 2)); conserved = 1; }
                        }
            if (!conserved) continue;
            const zc *U3 = (const zc *)U + c3 * nw * nb, *U4 = (const zc *)U + c4 * nw * nb;
            for (int i = 0; i < nb; i++)
                for (int j = 0; j < nb; j++) {
                    zc o = zmk(0, 0);
                    for (int w = 0; w < nw; w++) o = zadd(o, zmul(zconj(U3[w * nb + i]), U4[w * nb + j]));
                    o34[i * nb + j] = o;
                }
            if (antisym) {  /* calc_ovlps' <1|v|4> and <3|2> (ElecElec_Model.h:69-75) */
                double v14 = ee_vq(a, c4, kvec, kmesh, GGT, vq_cell, screening, prefac_vq, prefac_vq_bare,
                                   qscr2_debye, fac0, fac2, fac4);
                const zc *U2 = (const zc *)U + b * nw * nb;
                for (int i = 0; i < nb; i++)
                    for (int j = 0; j < nb; j++) {
                        zc o = zmk(0, 0), r = zmk(0, 0);
                        for (int w = 0; w < nw; w++) {
                            o = zadd(o, zmul(zconj(Ua[w * nb + i]), U4[w * nb + j]));
                            r = zadd(r, zmul(zconj(U3[w * nb + i]), U2[w * nb + j]));
                        }
                        o14[i * nb + j] = zmul(o, zmk(v14, 0));
                        o32[i * nb + j] = r;
                    }
            }
            for (int q = 0; q < nb4; q++) {
                zc mee = zmul(o12[q / nb2], o34[q % nb2]);
                if (antisym) {  /* mee = (mee - mee_ex) / 2, mee_ex = <1|v|4><3|2> (:114-149) */
                    int i1 = q / (nb * nb2), i2 = (q / nb2) % nb, i3 = (q / nb) % nb, i4 = q % nb;
                    zc ex = zmul(o14[i1 * nb + i4], o32[i3 * nb + i2]);
                    mee = zadd(zmul(zmk(-0.5, 0), ex), zmul(zmk(0.5, 0), mee));
                }
                A2[q] = zscale(mee, dl[q]);
                A1[q] = lindblad ? A2[q] : mee;
            }
            const double *f3 = f + c3 * nb, *f4 = f + c4 * nb;
            const int with_rho = dm != NULL && dP1 == NULL;
            for (int i1 = 0; i1 < nb; i1++)
                for (int i3 = 0; i3 < nb; i3++) {
                    if (with_rho) {  /* calc_A1rho (:266-270): dm1_k3 A1 dm_k4, dm_k3 A1 dm1_k4 */
                        zc t[64];
                        zhemm_upper(t, 0, (const zc *)dm + c4 * nb2, A1 + (i1 * nb + i3) * nb2, nb, c1, c0);
                        zhemm_upper(Ar, 1, (const zc *)dm1 + c3 * nb2, t, nb, c1, c0);
                        zhemm_upper(t, 0, (const zc *)dm1 + c4 * nb2, A1 + (i3 * nb + i1) * nb2, nb, c1, c0);
                        zhemm_upper(Arb, 1, (const zc *)dm + c3 * nb2, t, nb, c1, c0);
                    } else {  /* calc_A1f / calc_A1fbar (:282-293) */
                        for (int x = 0; x < nb; x++)
                            for (int y = 0; y < nb; y++) {
                                Ar[x * nb + y] = zscale(zscale(A1[(i1 * nb + i3) * nb2 + x * nb + y], 1 - f3[x]), f4[y]);
                                Arb[x * nb + y] = zscale(zscale(A1[(i3 * nb + i1) * nb2 + x * nb + y], f3[x]), 1 - f4[y]);
                            }
                    }
                    if (dP1) {  /* calc_A1rho_dP (:271-281) */
                        a1rho_dP(Ard, A1 + (i1 * nb + i3) * nb2, (const zc *)dm1 + c3 * nb2, f3, 1, (const zc *)dm + c4 * nb2,
                                 f4, 0, nb);
                        a1rho_dP(Arbd, A1 + (i3 * nb + i1) * nb2, (const zc *)dm + c3 * nb2, f3, 0, (const zc *)dm1 + c4 * nb2,
                                 f4, 1, nb);
                    }
                    for (int i2 = 0; i2 < nb; i2++)
                        for (int i4 = 0; i4 < nb; i4++) {
                            int q = ((i1 * nb + i2) * nb + i3) * nb + i4, n24 = (i2 * nb + i4) * nb2, n42 = (i4 * nb + i2) * nb2;
                            for (int xy = 0; xy < nb2; xy++) {
                                Q1[q] = zadd(Q1[q], zmul(Ar[xy], zconj(A2[n24 + xy])));
                                Q2[q] = zadd(Q2[q], zmul(Arb[xy], zconj(A2[n42 + xy])));
                                if (dP1) {
                                    D1[q] = zadd(D1[q], zmul(Ard[xy], zconj(A2[n24 + xy])));
                                    D2[q] = zadd(D2[q], zmul(Arbd[xy], zconj(A2[n42 + xy])));
                                }
                            }
                        }
                }
        }
        double pf = (antisym ? 2 : 1) * pgauss / nk_full;
        for (int q = 0; q < nb4; q++) { Q1[q] = zscale(Q1[q], pf); Q2[q] = zscale(Q2[q], pf); }
        if (dP1)
            for (int q = 0; q < nb4; q++) { D1[q] = zscale(D1[q], pf); D2[q] = zscale(D2[q], pf); }
    }
    free(o14); free(o32); free(Ard); free(Arbd);
    free(kmap); free(o12); free(o34); free(A1); free(A2); free(Ar); free(Arb); free(dl);
    return 0;
}

/* The e-e model's A tensor for the k quadruple (ka, kb, kc, kd) of calc_A   */
/* (ElecElec_Model.h:144-166): mee = v(ka - kb) <a|b><c|d> (exchange: (mee -  */
/* v(ka - kd) <a|d><c|b>) / 2), delta = exp(pexp de^2) inside ethr with       */
/* de = e_a - e_b + e_c - e_d, A2 = delta mee, A1 = A2 (lindblad) or mee.      */
/* Returns 0 when no band combination conserves energy (A untouched).        */
static int ee_calc_A(int64_t ka, int64_t kb, int64_t kc, int64_t kd, int nb, int nw, const double *U,
                     const double *kvec, const int *kmesh, const double *e, int lindblad, double pexp, double ethr,
                     int antisym, const double *GGT, const double *vq_cell, int screening, double prefac_vq,
                     double prefac_vq_bare, double qscr2_debye, double fac0, double fac2, double fac4, zc *A1, zc *A2)
{
    const int nb2 = nb * nb, nb4 = nb2 * nb2;
    double dl[4096];
    int conserved = 0;
    for (int q = 0; q < nb4; q++) {
        int i1 = q / (nb * nb2), i2 = (q / nb2) % nb, i3 = (q / nb) % nb, i4 = q % nb;
        double de = e[ka * nb + i1] - e[kb * nb + i2] + e[kc * nb + i3] - e[kd * nb + i4];
        dl[q] = 0;
        if (fabs(de) < ethr) { dl[q] = exp(pexp * pow(de, 2)); conserved = 1; }
    }
    if (!conserved) return 0;
    zc oab[64], ocd[64], oad[64], ocb[64];
    const zc *Ua = (const zc *)U + ka * nw * nb, *Ub = (const zc *)U + kb * nw * nb;
    const zc *Uc = (const zc *)U + kc * nw * nb, *Ud = (const zc *)U + kd * nw * nb;
    double vab = ee_vq(ka, kb, kvec, kmesh, GGT, vq_cell, screening, prefac_vq, prefac_vq_bare, qscr2_debye, fac0,
                       fac2, fac4);
    double vad = antisym ? ee_vq(ka, kd, kvec, kmesh, GGT, vq_cell, screening, prefac_vq, prefac_vq_bare, qscr2_debye,
                                 fac0, fac2, fac4)
                         : 0.0;
    for (int i = 0; i < nb; i++)
        for (int j = 0; j < nb; j++) {
            zc x = zmk(0, 0), y = zmk(0, 0), z = zmk(0, 0), w = zmk(0, 0);
            for (int r = 0; r < nw; r++) {
                x = zadd(x, zmul(zconj(Ua[r * nb + i]), Ub[r * nb + j]));
                y = zadd(y, zmul(zconj(Uc[r * nb + i]), Ud[r * nb + j]));
                if (antisym) {
                    z = zadd(z, zmul(zconj(Ua[r * nb + i]), Ud[r * nb + j]));
                    w = zadd(w, zmul(zconj(Uc[r * nb + i]), Ub[r * nb + j]));
                }
            }
            oab[i * nb + j] = zscale(x, vab);
            ocd[i * nb + j] = y;
            oad[i * nb + j] = zscale(z, vad);
            ocb[i * nb + j] = w;
        }
    for (int q = 0; q < nb4; q++) {
        int i1 = q / (nb * nb2), i2 = (q / nb2) % nb, i3 = (q / nb) % nb, i4 = q % nb;
        zc mee = zmul(oab[i1 * nb + i2], ocd[i3 * nb + i4]);
        if (antisym) {
            zc ex = zmul(oad[i1 * nb + i4], ocb[i3 * nb + i2]);
            mee = zadd(zmul(zmk(-0.5, 0), ex), zmul(zmk(0.5, 0), mee));
        }
        A2[q] = zscale(mee, dl[q]);
        A1[q] = lindblad ? A2[q] : mee;
    }
    return 1;
}

/* The linearised e-e P3..P6 of every stored pair (k1, ka) at the            */
/* occupations f, elecelec_model::calc_P3P4 and calc_P5P6                    */
/* (scattering_model/ElecElec_Model_linear.cpp:3-111):                        */
/*  P3[(i1 i2),(ia ib)] = pf sum_k3 sum_43 A1[(i1 i4),(i3 ia)] (1-f3) f4       */
/*                              conj(A2[(i2 i4),(i3 ib)]),  k4 = k1 + k3 - ka,  */
/*  P4 the same with f3 (1-f4), A from calc_A(k1, k4, k3, ka);                 */
/*  P5[(i1 i2),(ia ib)] = pf sum_k3 sum_34 A1[(i1 i3),(ib i4)] f3 f4           */
/*                              conj(A2[(i2 i3),(ia i4)]),  k4 = k1 + ka - k3,  */
/*  P6 the same with (1-f3)(1-f4), A from calc_A(k1, k3, ka, k4);              */
/* pf = prefac_gauss / nk_full (x 2 with exchange), and the reference's k3     */
/* exclusions.  Any of P3..P6 may be NULL only together with its partner.    */
int oracle_ee_model_P3P6(int64_t nk, int nb, int nw, int64_t npairs, const uint64_t *k1, const uint64_t *k2,
                         const double *U, const double *kvec, const int *kmesh, const double *e, const double *f,
                         int lindblad, double degauss, double ethr, double nk_full, int antisym, const double *GGT,
                         const double *vq_cell, int screening, double prefac_vq, double prefac_vq_bare,
                         double qscr2_debye, double fac0, double fac2, double fac4, double *P3, double *P4,
                         double *P5, double *P6)
{
    int nb2 = nb * nb, nb4 = nb2 * nb2;
    if (nb < 1 || nb > 4) return -1;
    int64_t cells = (int64_t)kmesh[0] * kmesh[1] * kmesh[2];
    int64_t *kmap = (int64_t *)malloc(sizeof(int64_t) * cells);
    for (int64_t c = 0; c < cells; c++) kmap[c] = -1;
    for (int64_t ik = 0; ik < nk; ik++) {
        int64_t c = mesh_code(kvec + 3 * ik, kmesh);
        if (kmap[c] < 0) kmap[c] = ik;
    }
    double pgauss = 1. / (degauss * sqrt(2. * M_PI));
    double pexp = lindblad ? -0.25 / (degauss * degauss) : -0.5 / (degauss * degauss);
    double pf = (antisym ? 2 : 1) * pgauss / nk_full;
    zc A1[256], A2[256], Ar[16], Arb[16];
    for (int64_t p = 0; p < npairs; p++) {
        int64_t a = (int64_t)k1[p], b = (int64_t)k2[p];
        zc *Q3 = P3 ? (zc *)P3 + p * nb4 : NULL, *Q4 = P4 ? (zc *)P4 + p * nb4 : NULL;
        zc *Q5 = P5 ? (zc *)P5 + p * nb4 : NULL, *Q6 = P6 ? (zc *)P6 + p * nb4 : NULL;
        if (Q3) for (int q = 0; q < nb4; q++) Q3[q] = Q4[q] = zmk(0, 0);
        if (Q5) for (int q = 0; q < nb4; q++) Q5[q] = Q6[q] = zmk(0, 0);
        for (int64_t c3 = 0; c3 < nk && Q3; c3++) {  /* calc_P3P4(ik1 = a, ika = b) */
            if (b == c3 && !antisym) continue;
            if (antisym && b == c3 && a == b) continue;
            double k4v[3];
            for (int d = 0; d < 3; d++) k4v[d] = kvec[3 * a + d] + kvec[3 * c3 + d] - kvec[3 * b + d];
            int64_t c4 = kmap[mesh_code(k4v, kmesh)];
            if (c4 < 0) continue;
            if (!ee_calc_A(a, c4, c3, b, nb, nw, U, kvec, kmesh, e, lindblad, pexp, ethr, antisym, GGT, vq_cell,
                           screening, prefac_vq, prefac_vq_bare, qscr2_debye, fac0, fac2, fac4, A1, A2))
                continue;
            const double *f3 = f + c3 * nb, *f4 = f + c4 * nb;
            for (int i1 = 0; i1 < nb; i1++)
                for (int ia = 0; ia < nb; ia++) {
                    for (int i4 = 0; i4 < nb; i4++)
                        for (int i3 = 0; i3 < nb; i3++) {
                            zc x = A1[(i1 * nb + i4) * nb2 + i3 * nb + ia];
                            Ar[i4 * nb + i3] = zscale(zscale(x, 1 - f3[i3]), f4[i4]);
                            Arb[i4 * nb + i3] = zscale(zscale(x, f3[i3]), 1 - f4[i4]);
                        }
                    for (int i2 = 0; i2 < nb; i2++)
                        for (int ib = 0; ib < nb; ib++) {
                            int q = (i1 * nb + i2) * nb2 + ia * nb + ib;
                            for (int i4 = 0; i4 < nb; i4++)
                                for (int i3 = 0; i3 < nb; i3++) {
                                    zc y = zconj(A2[(i2 * nb + i4) * nb2 + i3 * nb + ib]);
                                    Q3[q] = zadd(Q3[q], zmul(Ar[i4 * nb + i3], y));
                                    Q4[q] = zadd(Q4[q], zmul(Arb[i4 * nb + i3], y));
                                }
                        }
                }
        }
        for (int64_t c3 = 0; c3 < nk && Q5; c3++) {  /* calc_P5P6(ik1 = a, ika = b) */
            if (a == c3 && !antisym) continue;
            if (antisym && a == c3 && b == c3) continue;
            double k4v[3];
            for (int d = 0; d < 3; d++) k4v[d] = kvec[3 * a + d] + kvec[3 * b + d] - kvec[3 * c3 + d];
            int64_t c4 = kmap[mesh_code(k4v, kmesh)];
            if (c4 < 0) continue;
            if (!ee_calc_A(a, c3, b, c4, nb, nw, U, kvec, kmesh, e, lindblad, pexp, ethr, antisym, GGT, vq_cell,
                           screening, prefac_vq, prefac_vq_bare, qscr2_debye, fac0, fac2, fac4, A1, A2))
                continue;
            const double *f3 = f + c3 * nb, *f4 = f + c4 * nb;
            for (int i1 = 0; i1 < nb; i1++)
                for (int ib = 0; ib < nb; ib++) {
                    for (int i3 = 0; i3 < nb; i3++)
                        for (int i4 = 0; i4 < nb; i4++) {
                            zc x = A1[(i1 * nb + i3) * nb2 + ib * nb + i4];
                            Ar[i3 * nb + i4] = zscale(zscale(x, f3[i3]), f4[i4]);
                            Arb[i3 * nb + i4] = zscale(zscale(x, 1 - f3[i3]), 1 - f4[i4]);
                        }
                    for (int i2 = 0; i2 < nb; i2++)
                        for (int ia = 0; ia < nb; ia++) {
                            int q = (i1 * nb + i2) * nb2 + ia * nb + ib;
                            for (int i3 = 0; i3 < nb; i3++)
                                for (int i4 = 0; i4 < nb; i4++) {
                                    zc y = zconj(A2[(i2 * nb + i3) * nb2 + ia * nb + i4]);
                                    Q5[q] = zadd(Q5[q], zmul(Ar[i3 * nb + i4], y));
                                    Q6[q] = zadd(Q6[q], zmul(Arb[i3 * nb + i4], y));
                                }
                        }
                }
        }
        if (Q3) for (int q = 0; q < nb4; q++) { Q3[q] = zscale(Q3[q], pf); Q4[q] = zscale(Q4[q], pf); }
        if (Q5) for (int q = 0; q < nb4; q++) { Q5[q] = zscale(Q5[q], pf); Q6[q] = zscale(Q6[q], pf); }
    }
    free(kmap);
    return 0;
}

/* The e-e terms of the linearised operator beyond set_Lsc's P1/P2 form,      */
/* added into (Lscij, Lscji) of every stored pair (file order) at f_eq:       */
/*  - extra x add_Lscij_from_P1_P2 of (Pee1, Pee2) (ElectronPhonon_linear.cpp:  */
/*    183-207): the reference's fac2 = ee_antisymmetry doubles the e-e         */
/*    Lscij/Lscji (not Lscii), i.e. extra = 1 with exchange, else 0;           */
/*  - add_Lscij_from_P3_P4 (:209-231) and add_Lscij_from_P5_P6 (:233-253)      */
/*    when P3..P6 are given (eeMode other than Pee_fixed_at_eq).               */
int oracle_add_Lsc_ee(int64_t nk, int nb, int64_t npairs, const uint64_t *k1, const uint64_t *k2, const double *f_eq,
                      double extra, const double *Pee1, const double *Pee2, const double *P3, const double *P4,
                      const double *P5, const double *P6, double *Lij, double *Lji)
{
    const int nb2 = nb * nb, nb4 = nb2 * nb2;
    if (nb < 1 || nb > 8 || nk < 0 || npairs < 0) return -1;
    for (int64_t ip = 0; ip < npairs; ip++) {
        const int64_t ik = (int64_t)k1[ip], jk = (int64_t)k2[ip];
        if (ik >= nk || jk >= nk) return -2;
        const double *fi = f_eq + ik * nb, *fj = f_eq + jk * nb;
        zc *a = (zc *)Lij + (size_t)ip * nb4, *b = (zc *)Lji + (size_t)ip * nb4;
        for (int i1 = 0; i1 < nb; i1++)
            for (int i2 = 0; i2 < nb; i2++) {
                const int i12 = i1 * nb + i2, n12 = i12 * nb2, i21 = i2 * nb + i1;
                for (int ia = 0; ia < nb; ia++)
                    for (int ib = 0; ib < nb; ib++) {
                        const int iab = ia * nb + ib, i12ab = n12 + iab, iab12 = iab * nb2 + i12;
                        if (extra != 0 && Pee1) {
                            const zc *p1 = (const zc *)Pee1 + (size_t)ip * nb4, *p2 = (const zc *)Pee2 + (size_t)ip * nb4;
                            zc x = zadd(zscale(p1[i12ab], 1 - fi[i1]), zscale(p2[iab12], fi[i2]));
                            zc y = zadd(zscale(zconj(p2[i12ab]), 1 - fj[i1]), zscale(zconj(p1[iab12]), fj[i2]));
                            a[i12ab] = zadd(a[i12ab], zscale(x, extra));
                            b[i12ab] = zadd(b[i12ab], zscale(y, extra));
                        }
                        if (P3) {
                            const zc *p3 = (const zc *)P3 + (size_t)ip * nb4, *p4 = (const zc *)P4 + (size_t)ip * nb4;
                            a[i12ab] = zsub(a[i12ab], zadd(zscale(p3[i12ab], 1 - fi[i1]), zscale(p4[i12ab], fi[i2])));
                            b[i12ab] = zsub(b[i12ab], zadd(zscale(zconj(p4[iab12]), 1 - fj[i1]),
                                                           zscale(zconj(p3[iab12]), fj[i2])));
                        }
                        if (P5) {
                            const zc *p5 = (const zc *)P5 + (size_t)ip * nb4, *p6 = (const zc *)P6 + (size_t)ip * nb4;
                            const int iba21 = (ib * nb + ia) * nb2 + i21;
                            a[i12ab] = zsub(a[i12ab], zadd(zscale(p5[i12ab], 1 - fi[i1]), zscale(p6[i12ab], fi[i2])));
                            b[i12ab] = zsub(b[i12ab], zadd(zscale(p5[iba21], 1 - fj[i1]), zscale(p6[iba21], fj[i2])));
                        }
                    }
            }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Coherent term, singdenmat_k::evolve_coh + compute_Hcoht,                  */
/* src_v4.5.7/DenMat.cpp:29-55: ddmdt = -i (H(t) dm - dm H(t)),              */
/* H(t)_ij = H_ij cis((e_i - e_j) t) off the diagonal in the interaction      */
/* picture.  Output overwritten.                                             */
/* ------------------------------------------------------------------------ */
int oracle_coherent(int64_t nk, int nb, const double *H, const double *e, const double *dm,
                    double t, int interaction, double *out)
{
    int nb2 = nb * nb;
    const zc c1 = zmk(1, 0), cm1 = zmk(-1, 0), c0 = zmk(0, 0), cmi = zmk(0, -1);
    zc Ht[64];
    for (int64_t ik = 0; ik < nk; ik++) {
        const zc *Hk = (const zc *)H + ik * nb2;
        const double *ek = e + ik * nb;
        for (int i = 0; i < nb; i++)
            for (int j = 0; j < nb; j++)
                Ht[i * nb + j] = (i == j || !interaction) ? Hk[i * nb + j] : zmul(Hk[i * nb + j], zcis((ek[i] - ek[j]) * t));
        zc *o = (zc *)out + ik * nb2;
        for (int i = 0; i < nb2; i++) o[i] = c0;
        zhemm_upper(o, 0, Ht, (const zc *)dm + ik * nb2, nb, c1, c0);   /* dm*H      */
        zhemm_upper(o, 1, Ht, (const zc *)dm + ik * nb2, nb, c1, cm1);  /* H*dm - .. */
        for (int i = 0; i < nb2; i++) o[i] = zmul(o[i], cmi);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Spin-type observable, ob_1dmk::measure_brange for sx/sy/sz with diff=true, */
/* src_v4.5.7/observable.h:301-336 and totals :356-366:                      */
/*   S = prefac * sum_k sum_{i in [i0,i1)} sum_b Re(s_t[i,b] * dmk_bi),       */
/*   s_t[i,b] = s[i,b] cis((e_i - e_b) t) (i != b, interaction picture),      */
/*   dmk_bi = dm[b,i] - f_i delta_bi.                                         */
/* s is one component (nk x nb x nb); returns S, and S_amp (no phase) via     */
/* *amp when non-null (the third column of sz_elec_tot.out).                  */
/* ------------------------------------------------------------------------ */
double oracle_spin(int64_t nk, int nb, int i0, int i1, const double *s, const double *e,
                   const double *f, const double *dm, double t, int interaction, double prefac, double *amp)
{
    int nb2 = nb * nb;
    double tot = 0, tot_amp = 0;
    for (int64_t ik = 0; ik < nk; ik++) {
        for (int i = i0; i < i1; i++) {
            double ob = 0, ob_amp = 0;
            for (int b = 0; b < nb; b++) {
                zc ob_kib = Z(s, ik * nb2 + i * nb + b);
                zc ob_kib_t = (i == b || !interaction) ? ob_kib : zmul(ob_kib, zcis((e[ik * nb + i] - e[ik * nb + b]) * t));
                zc dm_kbi = Z(dm, ik * nb2 + b * nb + i);
                if (i == b) dm_kbi.x -= f[ik * nb + i];
                ob += zmul(ob_kib_t, dm_kbi).x;
                ob_amp += zmul(ob_kib, dm_kbi).x;
            }
            tot += ob; tot_amp += ob_amp;
        }
    }
    if (amp) *amp = tot_amp * prefac;
    return tot * prefac;
}

/* ------------------------------------------------------------------------ */
/* Initial state rho0 = f + (df/de) H1, electron::compute_dm_Bpert_1st,       */
/* src_v4.5.7/electron.cpp:62-98 (degthr 1e-8, degenerate limit             */
/* favg(favg-1)/T, interaction-picture phase at t0).  s holds 3 components   */
/* per k (nk x 3 x nb x nb), as ldbd_smat.bin.                               */
/* ------------------------------------------------------------------------ */
int oracle_dm_Bpert_1st(int64_t nk, int nb, const double *e, const double *f, const double *s,
                        const double *Bpert, double T, double t0, int interaction, double *dm)
{
    int nb2 = nb * nb;
    const double degthr = 1e-8;
    for (int64_t ik = 0; ik < nk; ik++) {
        zc *d = (zc *)dm + ik * nb2;
        for (int i = 0; i < nb2; i++) d[i] = zmk(0, 0);
        for (int i = 0; i < nb; i++)
            for (int j = i; j < nb; j++) {
                zc H1 = zmk(0, 0);
                for (int c = 0; c < 3; c++)
                    H1 = zadd(H1, zscale(Z(s, (ik * 3 + c) * nb2 + i * nb + j), Bpert[c]));
                double ei = e[ik * nb + i], ej = e[ik * nb + j];
                double fi = f[ik * nb + i], fj = f[ik * nb + j];
                double dfde;
                if (fabs(ei - ej) < degthr) {
                    double favg = 0.5 * (fi + fj);
                    dfde = favg * (favg - 1.) / T;
                } else
                    dfde = (fi - fj) / (ei - ej);
                d[i * nb + j] = zscale(H1, dfde);
                if (i != j && interaction) d[i * nb + j] = zmul(d[i * nb + j], zcis((ei - ej) * t0));
                if (i < j) d[j * nb + i] = zconj(d[i * nb + j]);
            }
        for (int i = 0; i < nb; i++) d[i * nb + i].x += f[ik * nb + i];
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Right-hand side of dm_dynamics::compute for the configuration this path  */
/* covers, src_v4.5.7/DensityMatrix_Dynamics.h:243-278: 1-rho                */
/* (set_oneminusdm, DenMat.cpp:249-261), optional coherent term, the         */
/* scattering term through evolve_driver's band window (evolve_eph.cpp:21-32: */
/* rows/cols [b0, b0+nb) of the nb_dm x nb_dm state), summed.               */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t nk; int nb_dm, b0, nb;
    int64_t npairs; const uint64_t *k1, *k2; const double *P1, *P2;
    const double *e_dm;     /* nk x nb_dm */
    const double *H;        /* nk x nb_dm^2 or NULL */
    const double *ddmdt_eq; /* nk x nb^2 (e-ph window) or NULL */
    double prefac; int expt;
    int nthreads;
} oracle_system;

static int oracle_rhs(const oracle_system *S, double t, const double *dm, double *ddmdt,
                      double *w_dm, double *w_dm1, double *w_ddm, double *w_e, double *w_coh)
{
    int64_t nk = S->nk; int nbd = S->nb_dm, nb = S->nb, b0 = S->b0;
    int nbd2 = nbd * nbd, nb2 = nb * nb;
    memset(ddmdt, 0, sizeof(double) * 2 * nk * nbd2);
    if (S->H) {
        oracle_coherent(nk, nbd, S->H, S->e_dm, dm, t, 1, w_coh);
        for (int64_t i = 0; i < 2 * nk * nbd2; i++) ddmdt[i] += w_coh[i];
    }
    for (int64_t ik = 0; ik < nk; ik++) {
        for (int i = 0; i < nb; i++) {
            w_e[ik * nb + i] = S->e_dm[ik * nbd + b0 + i];
            for (int j = 0; j < nb; j++) {
                zc v = Z(dm, ik * nbd2 + (b0 + i) * nbd + b0 + j);
                ZW(w_dm, ik * nb2 + i * nb + j) = v;
                ZW(w_dm1, ik * nb2 + i * nb + j) = (i == j) ? zmk(1.0 - v.x, -v.y) : zmk(-v.x, -v.y);
            }
        }
    }
    int rc = oracle_evolve(nk, nb, S->npairs, S->k1, S->k2, S->P1, S->P2, w_e, w_dm, w_dm1, t,
                           S->prefac, S->expt, S->ddmdt_eq, 0, S->nthreads, w_ddm);
    if (rc) return rc;
    for (int64_t ik = 0; ik < nk; ik++)
        for (int i = 0; i < nb; i++)
            for (int j = 0; j < nb; j++) {
                size_t o = ik * nbd2 + (b0 + i) * nbd + b0 + j;
                ZW(ddmdt, o) = zadd(Z(ddmdt, o), Z(w_ddm, ik * nb2 + i * nb + j));
            }
    return 0;
}

/* Classical fourth-order Runge-Kutta on the compute() right-hand side, with */
/* the state re-Hermitized after every step as update_dm_euler does for the  */
/* reference's Euler step (DenMat.cpp:236-248).  The reference has no fixed- */
/* step RK4 (its integrators are GSL rkf45 and Euler,                        */
/* DensityMatrix_Dynamics.h:177-234); this is the textbook method.           */
/* Observable (if s_obs) recorded at step 0 and every report_every steps.    */
int oracle_rk4(int64_t nk, int nb_dm, int b0, int nb, int64_t npairs, const uint64_t *k1, const uint64_t *k2,
               const double *P1, const double *P2, const double *e_dm, const double *H, const double *ddmdt_eq,
               double prefac, int expt, double t0, double dt, int64_t nsteps, int report_every,
               const double *s_obs, const double *f_dm, int i0, int i1, double obs_prefac,
               int nthreads, double *dm, double *obs_out)
{
    oracle_system S = { nk, nb_dm, b0, nb, npairs, k1, k2, P1, P2, e_dm, H, ddmdt_eq, prefac, expt, nthreads };
    size_t n = (size_t)2 * nk * nb_dm * nb_dm, nw = (size_t)2 * nk * nb * nb;
    double *kk = (double *)malloc(sizeof(double) * n), *acc = (double *)malloc(sizeof(double) * n);
    double *y = (double *)malloc(sizeof(double) * n), *coh = (double *)malloc(sizeof(double) * n);
    double *wdm = (double *)malloc(sizeof(double) * nw), *wdm1 = (double *)malloc(sizeof(double) * nw);
    double *wddm = (double *)malloc(sizeof(double) * nw), *we = (double *)malloc(sizeof(double) * nk * nb);
    int rc = 0, iobs = 0;
    double t = t0;
    if (s_obs && report_every > 0) obs_out[iobs++] = oracle_spin(nk, nb_dm, i0, i1, s_obs, e_dm, f_dm, dm, t, 1, obs_prefac, NULL);
    static const double a[4] = { 0.0, 0.5, 0.5, 1.0 }, w[4] = { 1.0, 2.0, 2.0, 1.0 };
    for (int64_t it = 0; it < nsteps && !rc; it++) {
        memcpy(y, dm, sizeof(double) * n);
        memset(acc, 0, sizeof(double) * n);
        for (int st = 0; st < 4 && !rc; st++) {
            if (st > 0)
                for (size_t i = 0; i < n; i++) y[i] = dm[i] + a[st] * dt * kk[i];
            rc = oracle_rhs(&S, t + a[st] * dt, y, kk, wdm, wdm1, wddm, we, coh);
            for (size_t i = 0; i < n; i++) acc[i] += w[st] * kk[i];
        }
        for (size_t i = 0; i < n; i++) dm[i] += dt / 6.0 * acc[i];
        /* re-Hermitize: dm = (dm + dm^dagger)/2 per k */
        for (int64_t ik = 0; ik < nk; ik++) {
            zc *d = (zc *)dm + ik * nb_dm * nb_dm;
            for (int i = 0; i < nb_dm; i++) {
                d[i * nb_dm + i].y = 0.0;
                for (int j = i + 1; j < nb_dm; j++) {
                    zc u = d[i * nb_dm + j], l = d[j * nb_dm + i];
                    zc h = zmk(0.5 * (u.x + l.x), 0.5 * (u.y - l.y));
                    d[i * nb_dm + j] = h; d[j * nb_dm + i] = zconj(h);
                }
            }
        }
        t = t0 + (double)(it + 1) * dt;
        if (s_obs && report_every > 0 && (it + 1) % report_every == 0)
            obs_out[iobs++] = oracle_spin(nk, nb_dm, i0, i1, s_obs, e_dm, f_dm, dm, t, 1, obs_prefac, NULL);
    }
    free(kk); free(acc); free(y); free(coh); free(wdm); free(wdm1); free(wddm); free(we);
    return rc;
}

/* Number of OpenMP threads this build can use (1 without OpenMP). */
int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* k-resolved random g factors, gfac_normal_dist (src_v4.5.7/electron.cpp:7-18,  */
/* comm/myarray.cpp:299-321, comm/Random.cpp:24-45).  The reference draws from   */
/* a std::mt19937_64 seeded with n and a std::normal_distribution<double>; both */
/* are restated here from their published definitions: the 64-bit Mersenne     */
/* Twister of the C++ standard ([rand.eng.mers], [rand.predef]) and the         */
/* Marsaglia polar method as libstdc++ implements normal_distribution, with     */
/* uniforms in [0, 1) from generate_canonical<double, 53> (one 64-bit draw      */
/* divided by 2^64).                                                            */
/* ------------------------------------------------------------------------ */
typedef struct { uint64_t mt[312]; int i; } omt64;

static void omt64_seed(omt64 *g, uint64_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < 312; i++)
        g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->i = 312;
}

static uint64_t omt64_next(omt64 *g)
{
    if (g->i >= 312) {
        const uint64_t up = ~0ULL << 31, lo = (1ULL << 31) - 1;
        for (int i = 0; i < 312; i++) {
            uint64_t x = (g->mt[i] & up) | (g->mt[(i + 1) % 312] & lo);
            uint64_t xa = x >> 1;
            if (x & 1) xa ^= 0xB5026F5AA96619E9ULL;
            g->mt[i] = g->mt[(i + 156) % 312] ^ xa;
        }
        g->i = 0;
    }
    uint64_t y = g->mt[g->i++];
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    y ^= y >> 43;
    return y;
}

/* Raw engine outputs after seeding (for the standard's known answer). */
void oracle_mt19937_64(uint64_t seed, int64_t n, uint64_t *out)
{
    omt64 g;
    omt64_seed(&g, seed);
    for (int64_t i = 0; i < n; i++) out[i] = omt64_next(&g);
}

typedef struct { omt64 g; double saved; int has_saved; } onorm;

static double onorm_uniform(onorm *s)
{
    double u = (double)omt64_next(&s->g) / 18446744073709551616.0;
    return u >= 1.0 ? nextafter(1.0, 0.0) : u;
}

static double onorm_next(onorm *s)
{
    if (s->has_saved) { s->has_saved = 0; return s->saved; }
    double x, y, r2;
    do {
        x = 2.0 * onorm_uniform(s) - 1.0;
        y = 2.0 * onorm_uniform(s) - 1.0;
        r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    const double mult = sqrt(-2.0 * log(r2) / r2);
    s->saved = x * mult;
    s->has_saved = 1;
    return y * mult;
}

/* n standard normal draws from a generator seeded with seed. */
void oracle_normal_draws(uint64_t seed, int64_t n, double *out)
{
    onorm s;
    omt64_seed(&s.g, seed);
    s.has_saved = 0;
    for (int64_t i = 0; i < n; i++) out[i] = onorm_next(&s);
}

/* random_normal_array(a, n, mean, sigma, cap) with no weights: n draws of
 * mean + sigma z (each z redrawn while |sigma z| > cap when cap > 0), then
 * shifted to the exact mean and scaled to the exact sigma (population form,
 * about the target mean); a draw whose sigma comes out zero is repeated from
 * the running generator. */
void oracle_gfac_normal(int64_t n, double mean, double sigma, double cap, double *a)
{
    if (n == 0) return;
    if (n == 1) { a[0] = mean; return; }
    onorm s;
    omt64_seed(&s.g, (uint64_t)n);
    s.has_saved = 0;
    for (;;) {
        for (int64_t i = 0; i < n; i++) {
            double r = sigma * onorm_next(&s);
            if (cap > 0.0)
                while (fabs(r) > cap) r = sigma * onorm_next(&s);
            a[i] = mean + r;
        }
        if (sigma == 0) return;
        double sum = 0;
        for (int64_t i = 0; i < n; i++) sum += a[i];
        const double mean1 = sum / (double)n;
        for (int64_t i = 0; i < n; i++) a[i] = a[i] + mean - mean1;
        double ss = 0;
        for (int64_t i = 0; i < n; i++) ss += (a[i] - mean) * (a[i] - mean);
        const double sigma1 = sqrt(ss / (double)n);
        if (sigma1 == 0) continue;
        const double fac = sigma / sigma1;
        for (int64_t i = 0; i < n; i++) a[i] = mean + fac * (a[i] - mean);
        return;
    }
}
