// synth.h -- counter-based synthetic ldbd_data generator (energies, k-pairs,
// e-ph and e-i couplings, P1/P2).  The same inline functions run on the host
// (writing ldbd_data files, small cases) and inside the device fill kernel
// (full-size workloads generated straight into HBM), so every pair's P is a
// pure function of (seed, k, k'), independent of thread or launch layout.
//
// The P formula follows the reference initializer's compute_P
// (src_FeynWann_forDMD/lindbladInit_for-DMD-4.5.6/ElectronPhonon.h:432-575,
// lindblad form, gaussian delta, no energy-window cut) and the e-i model form
// (src_v4.5.7/scattering_model/ElecImp_Model.h:185-214).
#pragma once
#include <stdint.h>
#include <math.h>
#include "dmd_scatt.h"

#if defined(__HIPCC__)
#define DMD_HD __host__ __device__
#else
#define DMD_HD
#endif

namespace dmd {

// splitmix64 finaliser; keyed mixing gives a counter-based stream.
DMD_HD inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
DMD_HD inline uint64_t key4(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b)
{
    return mix64(mix64(mix64(seed ^ (stream * 0xD6E8FEB86659FD93ull)) ^ a) + b);
}
// uniform in the open interval (0,1) from the top 53 bits
DMD_HD inline double u01(uint64_t x) { return ((double)(x >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
DMD_HD inline double uni(uint64_t key, uint64_t ctr) { return u01(mix64(key + ctr)); }

// Box-Muller: one complex Gaussian with standard deviation sigma per component.
DMD_HD inline void cnormal(uint64_t key, uint64_t ctr, double sigma, double *re, double *im)
{
    double u1 = u01(mix64(key + 2 * ctr)), u2 = u01(mix64(key + 2 * ctr + 1));
    double r = sigma * sqrt(-2.0 * log(u1));
    double ang = 6.283185307179586476925286766559 * u2;
    *re = r * cos(ang);
    *im = r * sin(ang);
}

enum : uint64_t {
    ST_ENERGY = 1, ST_PAIRS = 2, ST_COUNT = 3, ST_GEPH = 4, ST_OMEGA = 5, ST_GIMP = 6, ST_SPIN = 7, ST_KVEC = 8
};

// Band energies of one k (Hartree, CBM at 0): Kramers pair E, E+split; for
// nb=4 a second pair gap above.  Extra bands (nb>4) continue in pairs.
DMD_HD inline void synth_energies_k(const dmd_synth_params &p, int64_t k, double *ek)
{
    uint64_t key = key4(p.seed, ST_ENERGY, (uint64_t)k, 0);
    double E = p.e_max * uni(key, 0);
    for (int b = 0; b < p.nb; b += 2) {
        double base = E;
        if (b >= 2) base = ek[b - 2] + p.gap_min + (p.gap_max - p.gap_min) * uni(key, 10 + b);
        double split = p.split_max * uni(key, 1 + b);
        ek[b] = base;
        if (b + 1 < p.nb) ek[b + 1] = base + split;
    }
}

// Phonon frequency of branch a (fixed per branch), Hartree.
DMD_HD inline double synth_omega(const dmd_synth_params &p, int a)
{
    return p.omega_min + (p.omega_max - p.omega_min) * uni(key4(p.seed, ST_OMEGA, (uint64_t)a, 0), 0);
}

// g^-_{kk'} of branch a, element (b1,b2): complex Gaussian keyed by the ordered pair.
DMD_HD inline void synth_g(const dmd_synth_params &p, int64_t k, int64_t kp, int a, int b1, int b2,
                           double *re, double *im)
{
    uint64_t key = key4(p.seed, ST_GEPH, (uint64_t)k, (uint64_t)kp);
    cnormal(key, (uint64_t)((a * p.nb + b1) * p.nb + b2), p.sigma_g, re, im);
}
DMD_HD inline void synth_gimp(const dmd_synth_params &p, int64_t k, int64_t kp, int b1, int b2,
                              double *re, double *im)
{
    uint64_t key = key4(p.seed, ST_GIMP, (uint64_t)k, (uint64_t)kp);
    cnormal(key, (uint64_t)(b1 * p.nb + b2), p.sigma_imp, re, im);
}

DMD_HD inline double synth_bose(double x) { return 1.0 / (exp(x) - 1.0); }

// P1/P2 of stored pair (k <= kp), written as interleaved complex into
// P1/P2 (nb^4 each).  Accumulation order per branch follows compute_P:
// P += App (x) conj(A2pp) + Amm (x) conj(A2mm) (ElectronPhonon.h:566-570).
// The e-i term (when imp_fraction != 0) is added after the branches with
// weight imp_fraction, the fold-in of add_scatt_contrib
// (src_v4.5.7/ElectronPhonon.cpp:87-97).  A non-zero branch_mask keeps only
// the branches (bits 0..) and the e-i term (bit 31) it names, so P is the sum
// of its components.  Needs nb <= DMD_SYNTH_MAX_NB.
#define DMD_SYNTH_MAX_NB 4
template <int NB>
DMD_HD inline void synth_pair_P_nb(const dmd_synth_params &p, const double *Ek, const double *Ekp,
                                   int64_t k, int64_t kp, double *P1, double *P2)
{
    constexpr int nb = NB, nb2 = nb * nb, nb4 = nb2 * nb2;
    for (int i = 0; i < 2 * nb4; i++) { P1[i] = 0.0; P2[i] = 0.0; }
    const double sigma2 = p.degauss * p.degauss;
    const double prefac_exp = -0.5 / sigma2;
    const double prefac_sqrtdelta = sqrt(1.0 / (p.degauss * sqrt(2.0 * 3.141592653589793238462643383279)));
    double Ar[NB * NB][4], Ai[NB * NB][4];
    // Ar/Ai[.][0]=App=A2pp, [1]=Amm=A2mm, [2]=Amp=A2mp, [3]=Apm=A2pm
    for (int a = 0; a < p.n_modes; a++) {
        if (p.branch_mask && !((p.branch_mask >> a) & 1u)) continue;
        const double w = synth_omega(p, a);
        const double nq = synth_bose(fmax(1e-3, w / p.temperature));
        const double sq1 = sqrt(nq + 1.0), sq0 = sqrt(nq);
        for (int b1 = 0; b1 < nb; b1++)
            for (int b2 = 0; b2 < nb; b2++) {
                const double dE = Ek[b1] - Ekp[b2];
                // g^+_{kk'} = dagger(g^-_{k'k}): element (b1,b2) = conj(g_{k'k}(b2,b1))
                double gpr, gpi, gmr, gmi;
                synth_g(p, kp, k, a, b2, b1, &gpr, &gpi); gpi = -gpi;
                synth_g(p, k, kp, a, b1, b2, &gmr, &gmi);
                const double dp = dE + w, dm = dE - w;
                const double sp = prefac_sqrtdelta * sqrt(exp(prefac_exp * (dp * dp)));
                const double sm = prefac_sqrtdelta * sqrt(exp(prefac_exp * (dm * dm)));
                const int ij = b1 * nb + b2;
                Ar[ij][0] = gpr * sp * sq1; Ai[ij][0] = gpi * sp * sq1;
                Ar[ij][3] = gpr * sp * sq0; Ai[ij][3] = gpi * sp * sq0;
                Ar[ij][2] = gmr * sm * sq1; Ai[ij][2] = gmi * sm * sq1;
                Ar[ij][1] = gmr * sm * sq0; Ai[ij][1] = gmi * sm * sq0;
            }
        for (int i1 = 0; i1 < nb; i1++)
            for (int i2 = 0; i2 < nb; i2++)
                for (int i3 = 0; i3 < nb; i3++)
                    for (int i4 = 0; i4 < nb; i4++) {
                        const int o = 2 * (((i1 * nb + i2) * nb + i3) * nb + i4);
                        const int i13 = i1 * nb + i3, i24 = i2 * nb + i4, i31 = i3 * nb + i1, i42 = i4 * nb + i2;
                        // x * conj(y) = (xr yr + xi yi) + i (xi yr - xr yi)
                        double t1r = (Ar[i13][0] * Ar[i24][0] + Ai[i13][0] * Ai[i24][0])
                                   + (Ar[i13][1] * Ar[i24][1] + Ai[i13][1] * Ai[i24][1]);
                        double t1i = (Ai[i13][0] * Ar[i24][0] - Ar[i13][0] * Ai[i24][0])
                                   + (Ai[i13][1] * Ar[i24][1] - Ar[i13][1] * Ai[i24][1]);
                        double t2r = (Ar[i31][2] * Ar[i42][2] + Ai[i31][2] * Ai[i42][2])
                                   + (Ar[i31][3] * Ar[i42][3] + Ai[i31][3] * Ai[i42][3]);
                        double t2i = (Ai[i31][2] * Ar[i42][2] - Ar[i31][2] * Ai[i42][2])
                                   + (Ai[i31][3] * Ar[i42][3] - Ar[i31][3] * Ai[i42][3]);
                        P1[o] += t1r; P1[o + 1] += t1i;
                        P2[o] += t2r; P2[o + 1] += t2i;
                    }
    }
    if (p.imp_fraction != 0.0 && (!p.branch_mask || ((p.branch_mask >> 31) & 1u))) {
        // A = g_imp sqrt(delta(E_k - E_k')), P1 = A(x)conj(A), P2 = A^T(x)conj(A^T)
        for (int b1 = 0; b1 < nb; b1++)
            for (int b2 = 0; b2 < nb; b2++) {
                const double dE = Ek[b1] - Ekp[b2];
                double gr, gi;
                synth_gimp(p, k, kp, b1, b2, &gr, &gi);
                const double s = prefac_sqrtdelta * sqrt(exp(prefac_exp * dE * dE));
                Ar[b1 * nb + b2][0] = gr * s; Ai[b1 * nb + b2][0] = gi * s;
            }
        const double fr = p.imp_fraction;
        for (int i1 = 0; i1 < nb; i1++)
            for (int i2 = 0; i2 < nb; i2++)
                for (int i3 = 0; i3 < nb; i3++)
                    for (int i4 = 0; i4 < nb; i4++) {
                        const int o = 2 * (((i1 * nb + i2) * nb + i3) * nb + i4);
                        const int i13 = i1 * nb + i3, i24 = i2 * nb + i4, i31 = i3 * nb + i1, i42 = i4 * nb + i2;
                        double t1r = Ar[i13][0] * Ar[i24][0] + Ai[i13][0] * Ai[i24][0];
                        double t1i = Ai[i13][0] * Ar[i24][0] - Ar[i13][0] * Ai[i24][0];
                        double t2r = Ar[i31][0] * Ar[i42][0] + Ai[i31][0] * Ai[i42][0];
                        double t2i = Ai[i31][0] * Ar[i42][0] - Ar[i31][0] * Ai[i42][0];
                        P1[o] += fr * t1r; P1[o + 1] += fr * t1i;
                        P2[o] += fr * t2r; P2[o + 1] += fr * t2i;
                    }
    }
}

// Runtime-nb front end (host).
DMD_HD inline void synth_pair_P(const dmd_synth_params &p, const double *Ek, const double *Ekp,
                                int64_t k, int64_t kp, double *P1, double *P2)
{
    switch (p.nb) {
    case 1: synth_pair_P_nb<1>(p, Ek, Ekp, k, kp, P1, P2); break;
    case 2: synth_pair_P_nb<2>(p, Ek, Ekp, k, kp, P1, P2); break;
    case 3: synth_pair_P_nb<3>(p, Ek, Ekp, k, kp, P1, P2); break;
    default: synth_pair_P_nb<4>(p, Ek, Ekp, k, kp, P1, P2); break;
    }
}

}  // namespace dmd
