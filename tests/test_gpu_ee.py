"""The model e-e P at the equilibrium occupations generated on the device
(dmd_scatt_add_ee_model) against the oracle's elecelec_model::calc_P, for
nb = 1..4, Debye and static RPA screening, both scattering forms, set and add
modes, and the operator built on it against the oracle's evolve."""
import numpy as np
import pytest

from helpers import FS, random_P, random_rho, rel_err

pytestmark = pytest.mark.gpu
NB3_OFF = pytest.mark.skip(reason="nb = 3 e-e generator switched off after a device failure whose cause is open")


def _mods():
    from denmat_dynm_development_amd import scatt
    from oracle import oracle
    return scatt, oracle


def _case(nb, seed, km=(4, 3, 3), nkeep=26):
    rng = np.random.default_rng(seed)
    cells = np.array([(a, b, c) for a in range(km[0]) for b in range(km[1]) for c in range(km[2])])
    keep = np.sort(rng.choice(len(cells), nkeep, replace=False))
    kv = cells[keep] / np.array(km) + rng.integers(-1, 2, size=(nkeep, 3))
    e = np.sort(rng.uniform(0, 2e-3, size=(nkeep, nb)), axis=1)
    f = 1 / (np.exp((e - 1e-3) / 9.5e-4) + 1)
    nw = nb + 2
    U = rng.normal(size=(nkeep, nw, nb)) + 1j * rng.normal(size=(nkeep, nw, nb))
    k1, k2 = np.triu_indices(nkeep)
    sel = rng.uniform(size=len(k1)) < 0.5
    return rng, km, kv, e, f, U, k1[sel].astype(np.uint64), k2[sel].astype(np.uint64)


def _model(scatt, scr, nw, km, nk_full, kind):
    R = np.diag([7.0, 8.0, 9.0])
    return scatt.eimp_model(scr, 3e-6, 9.5e-4, abs(np.linalg.det(R)), scatt.reciprocal_metric(R), nw, n_ionized=1.0,
                            epsilon_background=3.0, degauss=4e-4, ethr=1.2e-3, scatt=kind, kmesh=km, nk_full=nk_full)


def _vq(oracle, k1, k2, U, kv, e, f, m):
    d = m.as_dict()
    if d["screening"] == 3:
        return oracle.eimp_rpa_vq(k1, k2, U, kv, e, f, d)[0]
    return np.array([oracle.eimp_vq(d, oracle.eimp_q2(d["GGT"], kv[a], kv[b])) for a, b in zip(k1, k2)])


@pytest.mark.parametrize("nb", [1, 2, pytest.param(3, marks=NB3_OFF), 4])
def test_ee_P_matches_the_oracle(nb):
    scatt, oracle = _mods()
    rng, km, kv, e, f, U, k1, k2 = _case(nb, 60 + nb)
    nk = len(kv)
    for scr in ("debye", "RPA"):
        for kind in ("lindblad", "conventional"):
            print(f"case ee model nb={nb} {scr} {kind}", flush=True)
            m = _model(scatt, scr, U.shape[1], km, 36.0, kind)
            vq = _vq(oracle, k1, k2, U, kv, e, f, m)
            R1, R2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, 4e-4, 1.2e-3, 36.0, kind == "lindblad")
            with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
                op.add_ee_model(k1, k2, U, kv, f, m, scale=0.6, add=False)
                A1, A2 = op.download_P()
            s = np.abs(R1).max()
            assert s > 0 and np.abs(A1 - 0.6 * R1).max() <= 1e-12 * s and np.abs(A2 - 0.6 * R2).max() <= 1e-12 * s


def test_ee_P_adds_to_P_and_drives_evolve():
    scatt, oracle = _mods()
    nb = 2
    rng, km, kv, e, f, U, k1, k2 = _case(nb, 99, nkeep=30)
    nk = len(kv)
    P1, P2 = random_P(len(k1), nb, rng)
    m = _model(scatt, "debye", U.shape[1], km, 36.0, "lindblad")
    vq = _vq(oracle, k1, k2, U, kv, e, f, m)
    R1, R2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, 4e-4, 1.2e-3, 36.0, True)
    a = np.abs(P1).max() / np.abs(R1).max()
    S1, S2 = P1 + a * R1, P2 + a * R2
    print("case ee model add + evolve", flush=True)
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(P1, P2)
        op.add_ee_model(k1, k2, U, kv, f, m, scale=a, add=True)
        for t_fs in (0.0, 23.0):
            rho = random_rho(nk, nb, rng)
            ref = oracle.evolve(k1, k2, S1, S2, e, rho, t=t_fs * FS, prefac=2 * np.pi / nk)
            assert rel_err(op.eval(t_fs * FS, rho), ref) < 1e-12


def test_ee_P_with_exchange():
    """ee_antisymmetry on the device against the oracle (Debye, nb = 2; nb = 3
    is switched off, and the other nb are covered in test_gpu_ee_dynamic.py)."""
    scatt, oracle = _mods()
    for nb in (2,):
        rng, km, kv, e, f, U, k1, k2 = _case(nb, 70 + nb)
        nk = len(kv)
        R = np.diag([7.0, 8.0, 9.0])
        m = scatt.eimp_model("debye", 3e-6, 9.5e-4, abs(np.linalg.det(R)), scatt.reciprocal_metric(R), U.shape[1],
                             n_ionized=1.0, epsilon_background=3.0, degauss=4e-4, ethr=1.2e-3, kmesh=km, nk_full=36.0,
                             exchange=True)
        vq = _vq(oracle, k1, k2, U, kv, e, f, m)
        R1, R2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, 4e-4, 1.2e-3, 36.0, True, antisym_model=m.as_dict())
        print(f"case ee model exchange nb={nb}", flush=True)
        with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
            op.add_ee_model(k1, k2, U, kv, f, m, add=False)
            A1, A2 = op.download_P()
        s_ = np.abs(R1).max()
        assert s_ > 0 and np.abs(A1 - R1).max() <= 1e-12 * s_ and np.abs(A2 - R2).max() <= 1e-12 * s_
