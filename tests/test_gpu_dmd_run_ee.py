"""dmd_run with the e-e model beyond the fixed-at-equilibrium P, against the
oracle's run of the same physics:

- eeMode = Pee_update with Debye screening and a model impurity: every
  freq_update_ee_model steps reset_scatt rebuilds P = P_eph + scale_ei P_imp +
  scale_ee P_ee(rho) with the screening re-derived from the state's
  occupations (update_screening, coulomb_model::init(dm)) and P_ee at the state
  in the Schroedinger picture (ElectronPhonon.cpp:15-42,56-65).
- The key set of the reference's linearized e-ph + e-i + e-e example
  (alg_linearize = 1, eeMode = Pee_update, impMode = model_ionized, scrFormula
  at its default, the static RPA): set_Lsc of P_eph + P_imp + P_ee at f_eq
  with the P3..P6 terms in the reference's sign (ElectronPhonon_linear.cpp:
  3-104), RPA q_scr^2 at the occupations of mu.
- alg_linearize_dPee with Pee_update: the e-e P at f_eq and the row term of
  its change dP towards the state added to every right-hand side
  (evolve_eph.cpp:107-148), ddmdt_eq zero."""
import os
import subprocess

import numpy as np
import pytest

from test_gpu_dmd_run import BOHR2CM, FS, RUN, TESLA, _fermi, _mu_conserving

pytestmark = pytest.mark.gpu


def _folder(tmp_path, seed):
    from denmat_dynm_development_amd import ldbd, scatt
    p = scatt.synth_params(0)
    d = tmp_path / "run"
    L = d / "ldbd_data"
    L.mkdir(parents=True)
    scatt.write_ldbd(str(L), p)
    lines = (L / "ldbd_size.dat").read_text().split("\n")
    f2 = lines[1].split("#")[0].split()
    f2[6], f2[7] = "5", "1"  # nb_wannier, bands skipped below the Wannier set
    lines[1] = " ".join(f2) + " # nb nv bBot_dm bTop_dm bBot_eph bTop_eph nb_wannier bskipped_wannier band_dft_skipped"
    lines[2] = " 6.40000000000000e+01 64 4 4 4 # nk_full nk kmesh"
    (L / "ldbd_size.dat").write_text("\n".join(lines))
    g = np.array([(a, b, c) for a in range(4) for b in range(4) for c in range(4)], np.float64) / 4
    g.tofile(L / "ldbd_kvec.bin")
    info = ldbd.read_info(str(L))
    rng = np.random.default_rng(seed)
    Ufull = (rng.normal(size=(info.nk, 5, 5)) + 1j * rng.normal(size=(info.nk, 5, 5))) / np.sqrt(5)
    Ufull.tofile(L / "ldbd_Umat.bin")
    U = ldbd.read_Umat(str(L), info)
    e = ldbd.read_ek(str(L), info)
    kv = ldbd.read_kvec(str(L), info)
    k1, k2 = ldbd.read_kpairs(str(L), info)
    P1, P2 = ldbd.read_P(str(L), info)
    s3 = ldbd.read_smat(str(L), info)
    _, R = ldbd.read_R(str(L))
    return d, info, U, e, kv, k1, k2, P1, P2, s3, R


def _rk4_run(g, rho, nsteps, dt, spin, before_step=None):
    """RK4 with the runner's Hermitization after each step; S per step."""
    nk, nb2 = rho.shape
    nb = int(round(np.sqrt(nb2)))
    ref = [spin(rho, 0.0)]
    y = rho.copy()
    for i in range(nsteps):
        t = i * dt
        if before_step:
            before_step(i + 1, t, y)
        a1 = g(t, y)
        a2 = g(t + dt / 2, y + dt / 2 * a1)
        a3 = g(t + dt / 2, y + dt / 2 * a2)
        a4 = g(t + dt, y + dt * a3)
        y = y + dt / 6 * (a1 + 2 * a2 + 2 * a3 + a4)
        yy = y.reshape(nk, nb, nb)
        y = (0.5 * (yy + yy.conj().transpose(0, 2, 1))).reshape(nk, -1)
        ref.append(spin(y, (i + 1) * dt))
    return np.array(ref)


def test_dmd_run_pee_update_rebuilds_P_from_the_state(tmp_path):
    from denmat_dynm_development_amd import scatt
    from oracle import oracle
    print("case dmd_run Pee_update", flush=True)
    d, info, U, e, kv, k1, k2, P1, P2, s3, R = _folder(tmp_path, 41)
    nk, nb = e.shape
    V = abs(np.linalg.det(R))
    T = info.temperature
    f = _fermi(T, info.mu, e)
    eps = 4.0
    km = tuple(info.kmesh)
    sig, ethr = info.degauss, info.degauss * info.ndegauss

    def model(n_free, n_i=1.0):
        return scatt.eimp_model("debye", n_free, T, V, scatt.reciprocal_metric(R), 5, n_ionized=n_i,
                                epsilon_background=eps, degauss=sig, ethr=ethr, kmesh=km,
                                nk_full=info.nk_full).as_dict()

    def vq_of(m):
        return np.array([oracle.eimp_vq(m, oracle.eimp_q2(m["GGT"], kv[a], kv[b])) for a, b in zip(k1, k2)])

    n0 = f.sum() / (info.nk_full * V)  # nv = 0: electrons only
    E1, _ = oracle.ee_model_P(k1, k2, U, kv, e, f, vq_of(model(n0)), km, sig, ethr, info.nk_full, True)
    scale_ee = float(f"{np.abs(P1).max() / np.abs(E1).max():.6g}")
    I1, _ = oracle.eimp_model_P(k1, k2, U, kv, e, model(n0, 1e-6))
    n_i = 1e-6 * 0.5 * np.abs(P1).max() / np.abs(I1).max()
    scale_ei, freq, nsteps = 0.7, 2, 5
    (d / "param.in").write_text("alg_ode_method = rk4\nBzpert = 2\ntstep = 1\ntend = 5\nfreq_measure = 1\n"
                                f"eeMode = Pee_update\nfreq_update_ee_model = {freq}\nscale_ee = {scale_ee}\n"
                                f"scale_ei = {scale_ei}\nimpurity_density = {n_i / BOHR2CM ** 3:.17g}\n"
                                "impMode = model_ionized\nscrMode = medium\nscrFormula = debye\n"
                                f"epsilon_background = {eps}\n")
    r = subprocess.run([RUN, str(d)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-800:])
    assert r.returncode == 0 and "model e-e (Pee_update)" in r.stdout
    rho = oracle.dm_Bpert_1st(e, f, s3, np.array([0, 0, 2 * TESLA]), T, 0.0, True).reshape(nk, nb, nb)
    rho = (0.5 * (rho + rho.conj().transpose(0, 2, 1))).reshape(nk, nb * nb)
    prefac = 2 * np.pi / info.nk_full
    obs_prefac = 1.0 / (info.nk_full * V * BOHR2CM ** 3)
    spin = lambda y, t: oracle.spin(s3[:, 2], e, f, y, t, 0, nb, obs_prefac)
    P = {}

    def rebuild(it, t, y):
        if (it - 1) % freq:
            return
        occ = y.reshape(nk, nb, nb)[:, np.arange(nb), np.arange(nb)].real
        m = model(occ.sum() / (info.nk_full * V), n_i)  # update_screening: n_free of the state
        M1, M2 = oracle.eimp_model_P(k1, k2, U, kv, e, m)
        S = y.reshape(nk, nb, nb).copy()  # the Schroedinger picture of the state
        for i in range(nb):
            for j in range(i + 1, nb):
                S[:, i, j] *= np.exp(1j * (e[:, j] - e[:, i]) * t)
                S[:, j, i] = S[:, i, j].conj()
        S1 = -S.copy()
        S1[:, np.arange(nb), np.arange(nb)] += 1.0
        Q1, Q2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq_of(m), km, sig, ethr, info.nk_full, True,
                                   dm=S.reshape(nk, -1), dm1=S1.reshape(nk, -1))
        P["1"], P["2"] = P1 + scale_ei * M1 + scale_ee * Q1, P2 + scale_ei * M2 + scale_ee * Q2

    g = lambda t, y: oracle.evolve(k1, k2, P["1"], P["2"], e, y, t=t, prefac=prefac)
    ref = _rk4_run(g, rho, nsteps, 1 * FS, spin, rebuild)
    rows = np.loadtxt(d / "sz_elec_tot.out")
    print("S_z runner", rows[:, 1], "oracle", ref)
    assert np.abs(rows[:, 1] - ref).max() < 1e-10 * np.abs(ref).max()
    # the rebuilds matter: P_ee fixed at f differs well beyond that
    Q1, Q2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq_of(model(n0)), km, sig, ethr, info.nk_full, True)
    M1, M2 = oracle.eimp_model_P(k1, k2, U, kv, e, model(n0, n_i))
    P["1"], P["2"] = P1 + scale_ei * M1 + scale_ee * Q1, P2 + scale_ei * M2 + scale_ee * Q2
    fixed = _rk4_run(g, rho, nsteps, 1 * FS, spin)
    assert np.abs(fixed - ref).max() > 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("ode", ["rk4", "euler"])
def test_dmd_run_pee_update_inside_every_rhs(tmp_path, ode):
    """freq_update_ee_model < 0: compute() rebuilds P from every stage state
    (update_scatt_inside, DensityMatrix_Dynamics.h:262,281-289,307-310): P =
    P_eph + scale_ei P_imp + scale_ee P_ee(stage state) with the screening
    re-derived from the stage state, before each right-hand side of RK4 and of
    Euler (update_dm_euler)."""
    from denmat_dynm_development_amd import scatt
    from oracle import oracle
    print(f"case dmd_run Pee_update inside every rhs ({ode})", flush=True)
    d, info, U, e, kv, k1, k2, P1, P2, s3, R = _folder(tmp_path, 53)
    nk, nb = e.shape
    V = abs(np.linalg.det(R))
    T = info.temperature
    f = _fermi(T, info.mu, e)
    eps = 4.0
    km = tuple(info.kmesh)
    sig, ethr = info.degauss, info.degauss * info.ndegauss

    def model(n_free, n_i=1.0):
        return scatt.eimp_model("debye", n_free, T, V, scatt.reciprocal_metric(R), 5, n_ionized=n_i,
                                epsilon_background=eps, degauss=sig, ethr=ethr, kmesh=km,
                                nk_full=info.nk_full).as_dict()

    def vq_of(m):
        return np.array([oracle.eimp_vq(m, oracle.eimp_q2(m["GGT"], kv[a], kv[b])) for a, b in zip(k1, k2)])

    n0 = f.sum() / (info.nk_full * V)
    E1, _ = oracle.ee_model_P(k1, k2, U, kv, e, f, vq_of(model(n0)), km, sig, ethr, info.nk_full, True)
    scale_ee = float(f"{np.abs(P1).max() / np.abs(E1).max():.6g}")
    I1, _ = oracle.eimp_model_P(k1, k2, U, kv, e, model(n0, 1e-6))
    n_i = 1e-6 * 0.5 * np.abs(P1).max() / np.abs(I1).max()
    scale_ei, nsteps = 0.7, 4
    (d / "param.in").write_text(f"alg_ode_method = {ode}\nBzpert = 2\ntstep = 1\ntend = {nsteps}\n"
                                "freq_measure = 1\neeMode = Pee_update\nfreq_update_ee_model = -1\n"
                                f"scale_ee = {scale_ee}\nscale_ei = {scale_ei}\n"
                                f"impurity_density = {n_i / BOHR2CM ** 3:.17g}\n"
                                "impMode = model_ionized\nscrMode = medium\nscrFormula = debye\n"
                                f"epsilon_background = {eps}\n")
    r = subprocess.run([RUN, str(d)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-800:])
    assert r.returncode == 0 and "rebuilt inside every right-hand side" in r.stdout
    rho = oracle.dm_Bpert_1st(e, f, s3, np.array([0, 0, 2 * TESLA]), T, 0.0, True).reshape(nk, nb, nb)
    rho = (0.5 * (rho + rho.conj().transpose(0, 2, 1))).reshape(nk, nb * nb)
    prefac = 2 * np.pi / info.nk_full
    obs_prefac = 1.0 / (info.nk_full * V * BOHR2CM ** 3)
    spin = lambda y, t: oracle.spin(s3[:, 2], e, f, y, t, 0, nb, obs_prefac)

    def g(t, y):  # reset_scatt at the stage state, then evolve
        occ = y.reshape(nk, nb, nb)[:, np.arange(nb), np.arange(nb)].real
        m = model(occ.sum() / (info.nk_full * V), n_i)
        M1, M2 = oracle.eimp_model_P(k1, k2, U, kv, e, m)
        S = y.reshape(nk, nb, nb).copy()
        for i in range(nb):
            for j in range(i + 1, nb):
                S[:, i, j] *= np.exp(1j * (e[:, j] - e[:, i]) * t)
                S[:, j, i] = S[:, i, j].conj()
        S1 = -S.copy()
        S1[:, np.arange(nb), np.arange(nb)] += 1.0
        Q1, Q2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq_of(m), km, sig, ethr, info.nk_full, True,
                                   dm=S.reshape(nk, -1), dm1=S1.reshape(nk, -1))
        return oracle.evolve(k1, k2, P1 + scale_ei * M1 + scale_ee * Q1, P2 + scale_ei * M2 + scale_ee * Q2, e, y,
                             t=t, prefac=prefac)

    if ode == "rk4":
        ref = _rk4_run(g, rho, nsteps, 1 * FS, spin)
    else:  # update_dm_euler: rho += dt (ddmdt + ddmdt^+) / 2
        y, ref = rho.copy(), [spin(rho, 0.0)]
        for i in range(nsteps):
            k = g(i * FS, y).reshape(nk, nb, nb)
            y = y + 0.5 * FS * (k + k.conj().transpose(0, 2, 1)).reshape(nk, -1)
            ref.append(spin(y, (i + 1) * FS))
        ref = np.array(ref)
    rows = np.loadtxt(d / "sz_elec_tot.out")
    print("S_z runner", rows[:, 1], "oracle", ref)
    assert np.abs(rows[:, 1] - ref).max() < 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("freq", [2, -1])
def test_dmd_run_linearize_dPee_adds_the_dP_term(tmp_path, freq):
    """alg_linearize_dPee with eeMode = Pee_update: every rebuild takes the
    e-e P at f_eq and its change dP towards the state (calc_A1rho_dP), and
    every right-hand side adds dP's row term with its phase
    (evolve_eph.cpp:107-148); ddmdt_eq stays zero (DensityMatrix_Dynamics.h:
    60-62).  freq_update_ee_model = -1: the rebuild (P and dP) at every stage
    state."""
    from denmat_dynm_development_amd import scatt
    from oracle import oracle
    print(f"case dmd_run linearize_dPee freq={freq}", flush=True)
    d, info, U, e, kv, k1, k2, P1, P2, s3, R = _folder(tmp_path, 47)
    nk, nb = e.shape
    V = abs(np.linalg.det(R))
    T = info.temperature
    f = _fermi(T, info.mu, e)
    eps = 4.0
    km = tuple(info.kmesh)
    sig, ethr = info.degauss, info.degauss * info.ndegauss

    def model(n_free, n_i=1.0):
        return scatt.eimp_model("debye", n_free, T, V, scatt.reciprocal_metric(R), 5, n_ionized=n_i,
                                epsilon_background=eps, degauss=sig, ethr=ethr, kmesh=km,
                                nk_full=info.nk_full).as_dict()

    def vq_of(m):
        return np.array([oracle.eimp_vq(m, oracle.eimp_q2(m["GGT"], kv[a], kv[b])) for a, b in zip(k1, k2)])

    n0 = f.sum() / (info.nk_full * V)
    E1, _ = oracle.ee_model_P(k1, k2, U, kv, e, f, vq_of(model(n0)), km, sig, ethr, info.nk_full, True)
    scale_ee = float(f"{np.abs(P1).max() / np.abs(E1).max():.6g}")
    I1, _ = oracle.eimp_model_P(k1, k2, U, kv, e, model(n0, 1e-6))
    n_i = 1e-6 * 0.5 * np.abs(P1).max() / np.abs(I1).max()
    scale_ei, nsteps = 0.7, 5
    (d / "param.in").write_text("alg_ode_method = rk4\nBzpert = 2\ntstep = 1\ntend = 5\nfreq_measure = 1\n"
                                "alg_ddmdteq = 1\nalg_linearize_dPee = 1\n"
                                f"eeMode = Pee_update\nfreq_update_ee_model = {freq}\nscale_ee = {scale_ee}\n"
                                f"scale_ei = {scale_ei}\nimpurity_density = {n_i / BOHR2CM ** 3:.17g}\n"
                                "impMode = model_ionized\nscrMode = medium\nscrFormula = debye\n"
                                f"epsilon_background = {eps}\n")
    r = subprocess.run([RUN, str(d)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-800:])
    assert r.returncode == 0 and "alg_linearize_dPee = 1" in r.stdout and "ddmdt_eq not computed" in r.stdout
    rho = oracle.dm_Bpert_1st(e, f, s3, np.array([0, 0, 2 * TESLA]), T, 0.0, True).reshape(nk, nb, nb)
    rho = (0.5 * (rho + rho.conj().transpose(0, 2, 1))).reshape(nk, nb * nb)
    target = rho.reshape(nk, nb, nb)[:, np.arange(nb), np.arange(nb)].real.sum()
    feq = _fermi(T, _mu_conserving(target, e, T), e)
    prefac = 2 * np.pi / info.nk_full
    obs_prefac = 1.0 / (info.nk_full * V * BOHR2CM ** 3)
    spin = lambda y, t: oracle.spin(s3[:, 2], e, f, y, t, 0, nb, obs_prefac)
    P = {}

    def rebuild(it, t, y):
        if freq > 0 and (it - 1) % freq:
            return
        occ = y.reshape(nk, nb, nb)[:, np.arange(nb), np.arange(nb)].real
        m = model(occ.sum() / (info.nk_full * V), n_i)
        M1, M2 = oracle.eimp_model_P(k1, k2, U, kv, e, m)
        S = y.reshape(nk, nb, nb).copy()
        for i in range(nb):
            for j in range(i + 1, nb):
                S[:, i, j] *= np.exp(1j * (e[:, j] - e[:, i]) * t)
                S[:, j, i] = S[:, i, j].conj()
        S1 = -S.copy()
        S1[:, np.arange(nb), np.arange(nb)] += 1.0
        Q1, Q2, D1, D2 = oracle.ee_model_P(k1, k2, U, kv, e, feq, vq_of(m), km, sig, ethr, info.nk_full, True,
                                           dm=S.reshape(nk, -1), dm1=S1.reshape(nk, -1), with_dP=True)
        P["1"], P["2"] = P1 + scale_ei * M1 + scale_ee * Q1, P2 + scale_ei * M2 + scale_ee * Q2
        P["D1"], P["D2"] = D1, D2

    def g(t, y):
        if freq < 0:  # update_scatt_inside: the rebuild at the stage state
            rebuild(1, t, y)
        return (oracle.evolve(k1, k2, P["1"], P["2"], e, y, t=t, prefac=prefac)
                + oracle.ee_dP_term(k1, k2, P["D1"], P["D2"], e, feq, t=t, prefac=prefac))

    ref = _rk4_run(g, rho, nsteps, 1 * FS, spin, rebuild if freq > 0 else None)
    rows = np.loadtxt(d / "sz_elec_tot.out")
    print("S_z runner", rows[:, 1], "oracle", ref)
    assert np.abs(rows[:, 1] - ref).max() < 1e-10 * np.abs(ref).max()
    # the dP term matters: without it the run differs well beyond that
    def g0(t, y):
        if freq < 0:
            rebuild(1, t, y)
        return oracle.evolve(k1, k2, P["1"], P["2"], e, y, t=t, prefac=prefac)

    plain = _rk4_run(g0, rho, nsteps, 1 * FS, spin, rebuild if freq > 0 else None)
    assert np.abs(plain - ref).max() > 1e-6 * np.abs(ref).max()


def test_dmd_run_reference_linearized_eph_ei_ee_keys(tmp_path):
    from denmat_dynm_development_amd import scatt
    from oracle import oracle
    print("case dmd_run linearized e-ph + e-i + e-e (RPA, Pee_update)", flush=True)
    d, info, U, e, kv, k1, k2, P1, P2, s3, R = _folder(tmp_path, 43)
    nk, nb = e.shape
    V = abs(np.linalg.det(R))
    T = info.temperature
    f = _fermi(T, info.mu, e)
    eps = 4.0
    km = tuple(info.kmesh)
    sig, ethr = info.degauss, info.degauss * info.ndegauss
    rho = oracle.dm_Bpert_1st(e, f, s3, np.array([0, 0, 2 * TESLA]), T, 0.0, True).reshape(nk, nb, nb)
    rho = (0.5 * (rho + rho.conj().transpose(0, 2, 1))).reshape(nk, nb * nb)
    target = rho.reshape(nk, nb, nb)[:, np.arange(nb), np.arange(nb)].real.sum()
    feq = _fermi(T, _mu_conserving(target, e, T), e)

    def model(n_i):  # RPA: n_free enters only Debye / Bechstedt / heg
        return scatt.eimp_model("RPA", feq.sum() / (info.nk_full * V), T, V, scatt.reciprocal_metric(R), 5,
                                n_ionized=n_i, epsilon_background=eps, degauss=sig, ethr=ethr, kmesh=km,
                                nk_full=info.nk_full).as_dict()

    m1 = model(1.0)
    vq_pair, _, cells = oracle.eimp_rpa_vq(k1, k2, U, kv, e, f, m1, cells=True)  # q_scr^2 at f of mu
    code = lambda x: int(np.dot(np.rint(np.array(km) * (x - np.floor(x))).astype(int) % np.array(km),
                                [km[1] * km[2], km[2], 1]))
    vq_ee = np.array([cells[code(kv[int(a)] - kv[int(b)])] for a, b in zip(k1, k2)])
    Q1, Q2 = oracle.ee_model_P(k1, k2, U, kv, e, feq, vq_ee, km, sig, ethr, info.nk_full, True, vq_cell=cells)
    scale_ee = float(f"{np.abs(P1).max() / np.abs(Q1).max():.6g}")
    I1, _ = oracle.eimp_model_P(k1, k2, U, kv, e, model(1e-6), f=f)
    n_i = 1e-6 * 0.5 * np.abs(P1).max() / np.abs(I1).max()
    M1, M2 = oracle.eimp_model_P(k1, k2, U, kv, e, model(n_i), f=f)
    scale_ei = 0.9
    S1, S2 = P1 + scale_ei * M1 + scale_ee * Q1, P2 + scale_ei * M2 + scale_ee * Q2
    (d / "param.in").write_text("alg_ode_method = rk4\nBzpert = 2\ntstep = 1\ntend = 4\nfreq_measure = 1\n"
                                "alg_linearize = 1\neeMode = Pee_update\nfreq_update_ee_model = 10\n"
                                f"scale_ee = {scale_ee}\nscale_ei = {scale_ei}\n"
                                f"impurity_density = {n_i / BOHR2CM ** 3:.17g}\nimpMode = model_ionized\n"
                                f"scrMode = medium\nepsilon_background = {eps}\n")
    r = subprocess.run([RUN, str(d)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-800:])
    assert r.returncode == 0 and "scrFormula = RPA" in r.stdout and "linearized at" in r.stdout
    Lij, Lji = oracle.set_Lsc(k1, k2, S1, S2, feq)
    P3, P4, P5, P6 = oracle.ee_model_P3P6(k1, k2, U, kv, e, feq, km, sig, ethr, info.nk_full, m1, True, False,
                                          vq_cell=cells)
    Lij2, Lji2 = oracle.add_Lsc_ee(k1, k2, feq, Lij, Lji, P3456=tuple(scale_ee * x for x in (P3, P4, P5, P6)))
    prefac = 2 * np.pi / info.nk_full
    obs_prefac = 1.0 / (info.nk_full * V * BOHR2CM ** 3)
    spin = lambda y, t: oracle.spin(s3[:, 2], e, f, y, t, 0, nb, obs_prefac)
    g = lambda t, y: oracle.evolve_linear(k1, k2, Lij2, Lji2, e, y, feq, t=t, prefac=prefac)
    ref = _rk4_run(g, rho, 4, 1 * FS, spin)
    rows = np.loadtxt(d / "sz_elec_tot.out")
    print("S_z runner", rows[:, 1], "oracle", ref)
    assert np.abs(rows[:, 1] - ref).max() < 1e-10 * np.abs(ref).max()
    # the P3..P6 terms matter at that tolerance
    g0 = lambda t, y: oracle.evolve_linear(k1, k2, Lij, Lji, e, y, feq, t=t, prefac=prefac)
    assert np.abs(_rk4_run(g0, rho, 4, 1 * FS, spin) - ref).max() > 1e-6 * np.abs(ref).max()


def test_dmd_run_linearized_ee_with_exchange_doubles_the_pair_terms(tmp_path):
    """alg_linearize = 1, eeMode = Pee_fixed_at_eq, ee_antisymmetry = 1 (Debye):
    set_Lsc of P_eph + scale_ee P_ee (with the exchange term) at f_eq, plus
    the fac2 doubling of the e-e Lscij / Lscji that add_Lsc_contrib("ee")
    applies (add_Lscij_from_P1_P2 with fac2 = ee_antisymmetry,
    ElectronPhonon_linear.cpp:79,180-199; Lscii once); S_z against the
    oracle's RK4 of evolve_linear of that operator, and the doubling is
    visible at that tolerance."""
    from denmat_dynm_development_amd import scatt
    from oracle import oracle
    print("case dmd_run linearized e-e with exchange", flush=True)
    d, info, U, e, kv, k1, k2, P1, P2, s3, R = _folder(tmp_path, 47)
    nk, nb = e.shape
    V = abs(np.linalg.det(R))
    T = info.temperature
    f = _fermi(T, info.mu, e)
    eps = 4.0
    km = tuple(info.kmesh)
    sig, ethr = info.degauss, info.degauss * info.ndegauss
    rho = oracle.dm_Bpert_1st(e, f, s3, np.array([0, 0, 2 * TESLA]), T, 0.0, True).reshape(nk, nb, nb)
    rho = (0.5 * (rho + rho.conj().transpose(0, 2, 1))).reshape(nk, nb * nb)
    target = rho.reshape(nk, nb, nb)[:, np.arange(nb), np.arange(nb)].real.sum()
    feq = _fermi(T, _mu_conserving(target, e, T), e)
    m = scatt.eimp_model("debye", feq.sum() / (info.nk_full * V), T, V, scatt.reciprocal_metric(R), 5,
                         n_ionized=1.0, epsilon_background=eps, degauss=sig, ethr=ethr, kmesh=km,
                         nk_full=info.nk_full, exchange=True).as_dict()
    vq = np.array([oracle.eimp_vq(m, oracle.eimp_q2(m["GGT"], kv[a], kv[b])) for a, b in zip(k1, k2)])
    Q1, Q2 = oracle.ee_model_P(k1, k2, U, kv, e, feq, vq, km, sig, ethr, info.nk_full, True, antisym_model=m)
    scale_ee = float(f"{np.abs(P1).max() / np.abs(Q1).max():.6g}")
    (d / "param.in").write_text("alg_ode_method = rk4\nBzpert = 2\ntstep = 1\ntend = 4\nfreq_measure = 1\n"
                                "alg_linearize = 1\neeMode = Pee_fixed_at_eq\nee_antisymmetry = 1\n"
                                f"scale_ee = {scale_ee}\nscrMode = medium\nscrFormula = debye\n"
                                f"epsilon_background = {eps}\n")
    r = subprocess.run([RUN, str(d)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-800:])
    assert r.returncode == 0 and "ee_antisymmetry = 1" in r.stdout and "linearized at" in r.stdout
    Lij, Lji = oracle.set_Lsc(k1, k2, P1 + scale_ee * Q1, P2 + scale_ee * Q2, feq)
    Lij2, Lji2 = oracle.add_Lsc_ee(k1, k2, feq, Lij, Lji, extra=scale_ee, Pee=(Q1, Q2))
    prefac = 2 * np.pi / info.nk_full
    obs_prefac = 1.0 / (info.nk_full * V * BOHR2CM ** 3)
    spin = lambda y, t: oracle.spin(s3[:, 2], e, f, y, t, 0, nb, obs_prefac)
    g = lambda t, y: oracle.evolve_linear(k1, k2, Lij2, Lji2, e, y, feq, t=t, prefac=prefac)
    ref = _rk4_run(g, rho, 4, 1 * FS, spin)
    rows = np.loadtxt(d / "sz_elec_tot.out")
    print("S_z runner", rows[:, 1], "oracle", ref)
    assert np.abs(rows[:, 1] - ref).max() < 1e-10 * np.abs(ref).max()
    g0 = lambda t, y: oracle.evolve_linear(k1, k2, Lij, Lji, e, y, feq, t=t, prefac=prefac)
    assert np.abs(_rk4_run(g0, rho, 4, 1 * FS, spin) - ref).max() > 1e-6 * np.abs(ref).max()
