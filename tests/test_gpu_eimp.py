"""The model ionized-impurity P generated on the device
(dmd_scatt_add_eimp_model) against the oracle's elecimp_model::calc_P, for
nb = 1..4, every |q|-only screening form, both scattering forms, with and
without detailed balance; set and add modes; the reduced-precision layouts;
and the operator built on it against the oracle's evolve."""
import numpy as np
import pytest

from helpers import FS, random_P, random_pairs, random_rho, rel_err

pytestmark = pytest.mark.gpu


def _mods():
    from denmat_dynm_development_amd import _lib, scatt
    from oracle import oracle
    return _lib, scatt, oracle


def _case(nb, nw, seed, nk=40, npart=9):
    rng = np.random.default_rng(seed)
    k1, k2 = random_pairs(nk, npart, rng, diag=True)
    e = np.sort(rng.uniform(0, 2e-3, size=(nk, nb)), axis=1)
    U = rng.normal(size=(nk, nw, nb)) + 1j * rng.normal(size=(nk, nw, nb))
    kv = rng.uniform(-0.5, 0.5, size=(nk, 3))
    kv[:4] = [[0.5, 0, 0], [-0.5, 0.25, 0], [0.1, 0.1, 0.1], [0.1, 0.1, 0.1]]  # wrap boundary, q = 0
    return rng, nk, k1, k2, e, U, kv


def _model(scatt, screening, kind, db, nw):
    R = np.array([[0.0, 5.3, 5.3], [5.3, 0.0, 5.3], [5.3, 5.3, 0.0]])
    return scatt.eimp_model(screening, 3e-6, 9.5e-4, abs(np.linalg.det(R)), scatt.reciprocal_metric(R), nw, Z=1.3,
                            n_ionized=2e-6, epsilon_background=4.0, degauss=3e-4, ethr=1.5e-3, scatt=kind,
                            detail_balance=db)


@pytest.mark.parametrize("nb", [1, 2, 3, 4])
def test_model_P_matches_the_oracle(nb):
    _lib, scatt, oracle = _mods()
    nw = nb + 3
    rng, nk, k1, k2, e, U, kv = _case(nb, nw, 40 + nb)
    for screening in ("unscreened", "debye", "Bechstedt"):
        for kind in ("lindblad", "conventional"):
            for db in (False, True):
                print(f"case eimp model nb={nb} {screening} {kind} db={db}", flush=True)
                m = _model(scatt, screening, kind, db, nw)
                R1, R2 = oracle.eimp_model_P(k1, k2, U, kv, e, m.as_dict())
                assert np.abs(R1).max() > 0
                with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
                    op.add_eimp_model(k1, k2, U, kv, m, scale=0.8, add=False)
                    A1, A2 = op.download_P()
                    s = np.abs(R1).max()
                    assert np.abs(A1 - 0.8 * R1).max() <= 1e-13 * s and np.abs(A2 - 0.8 * R2).max() <= 1e-13 * s


def test_model_P_adds_to_P_and_drives_evolve():
    """add = 1 on an uploaded e-ph P gives P_eph + scale_ei P_imp, and the
    operator's ddm equals the oracle's evolve on that sum at t = 0 and with
    phases."""
    _lib, scatt, oracle = _mods()
    nb, nw = 2, 6
    rng, nk, k1, k2, e, U, kv = _case(nb, nw, 77, nk=64, npart=12)
    P1, P2 = random_P(len(k1), nb, rng)
    m = _model(scatt, "debye", "lindblad", False, nw)
    R1, R2 = oracle.eimp_model_P(k1, k2, U, kv, e, m.as_dict())
    a = np.abs(P1).max() / np.abs(R1).max()  # comparable magnitudes
    S1, S2 = P1 + 0.5 * a * R1, P2 + 0.5 * a * R2
    print("case eimp model add + evolve", flush=True)
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(P1, P2)
        op.add_eimp_model(k1, k2, U, kv, m, scale=0.5 * a, add=True)
        A1, A2 = op.download_P()
        assert np.abs(A1 - S1).max() <= 1e-13 * np.abs(S1).max() and np.abs(A2 - S2).max() <= 1e-13 * np.abs(S1).max()
        for t_fs in (0.0, 31.0):
            rho = random_rho(nk, nb, rng)
            ref = oracle.evolve(k1, k2, S1, S2, e, rho, t=t_fs * FS, prefac=2 * np.pi / nk)
            assert rel_err(op.eval(t_fs * FS, rho), ref) < 1e-12


@pytest.mark.parametrize("prec", ["c64", "bf16"])
def test_model_P_in_reduced_precision_layouts(prec):
    _lib, scatt, oracle = _mods()
    p = {"c64": _lib.DMD_PREC_C64, "bf16": _lib.DMD_PREC_BF16}[prec]
    nb, nw = 4, 6
    rng, nk, k1, k2, e, U, kv = _case(nb, nw, 91)
    m = _model(scatt, "Bechstedt", "lindblad", False, nw)
    R1, R2 = oracle.eimp_model_P(k1, k2, U, kv, e, m.as_dict())
    print(f"case eimp model {prec}", flush=True)
    with scatt.ScatteringOperator(nk, nb, k1, k2, e, precision=p) as op:
        op.add_eimp_model(k1, k2, U, kv, m, add=False)
        A1, A2 = op.download_P()
    tol = 1e-6 if prec == "c64" else 1e-2
    for A, R in ((A1, R1), (A2, R2)):
        per = np.maximum(np.abs(R).max(axis=1, keepdims=True), 1e-300)
        assert np.all(np.abs(A - R) <= tol * per)


def test_model_P_rejects_bad_input():
    _lib, scatt, oracle = _mods()
    nb, nw = 2, 3
    rng, nk, k1, k2, e, U, kv = _case(nb, nw, 5)
    m = _model(scatt, "debye", "lindblad", False, nw)
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        with pytest.raises(Exception, match="add = 0 sets it"):
            op.add_eimp_model(k1, k2, U, kv, m, add=True)  # no P yet
        bad = k2.copy()
        bad[3] = nk
        with pytest.raises(Exception, match="outside"):
            op.add_eimp_model(k1, bad, U, kv, m, add=False)
        with pytest.raises(ValueError):
            op.add_eimp_model(k1, k2, U[:, :2], kv, m, add=False)


def test_model_P_with_rpa_screening():
    """Static RPA screening summed on the device over a k mesh with missing
    points (mesh map, per-q sums, per-pair v(q)) against the oracle's
    qIndexMap / kIndexMap restatement, for nb = 2 and 4."""
    _lib, scatt, oracle = _mods()
    rng = np.random.default_rng(12)
    km = (5, 4, 3)
    cells = np.array([(a, b, c) for a in range(5) for b in range(4) for c in range(3)])
    for nb in (2, 4):
        nw = nb + 2
        keep = np.sort(rng.choice(len(cells), 45, replace=False))
        kv = cells[keep] / np.array(km) + rng.integers(-1, 2, size=(45, 3))
        nk = len(kv)
        k1, k2 = random_pairs(nk, 10, rng, diag=True)
        e = np.sort(rng.uniform(0, 2e-3, size=(nk, nb)), axis=1)
        f = 1 / (np.exp((e - 1e-3) / 9.5e-4) + 1)
        U = rng.normal(size=(nk, nw, nb)) + 1j * rng.normal(size=(nk, nw, nb))
        R = np.diag([7.0, 8.0, 9.0])
        m = scatt.eimp_model("RPA", 0.0, 9.5e-4, abs(np.linalg.det(R)), scatt.reciprocal_metric(R), nw, Z=1.1,
                             n_ionized=2e-6, epsilon_background=3.0, degauss=3e-4, ethr=1.5e-3, kmesh=km, nk_full=60.0)
        R1, R2 = oracle.eimp_model_P(k1, k2, U, kv, e, m.as_dict(), f=f)
        print(f"case eimp model RPA nb={nb}", flush=True)
        with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
            op.add_eimp_model(k1, k2, U, kv, m, add=False, f=f)
            A1, A2 = op.download_P()
        s = np.abs(R1).max()
        assert s > 0 and np.abs(A1 - R1).max() <= 1e-12 * s and np.abs(A2 - R2).max() <= 1e-12 * s
