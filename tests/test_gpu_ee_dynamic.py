"""The e-e model on the device beyond the fixed-at-equilibrium P, against the
oracle:
- P at a density matrix (the Pee_update rebuild, dmd_scatt_add_ee_model_state)
  for nb = 1..4, both scattering forms, Debye and the static RPA (per-cell v(q)
  table from dmd_scatt_ee_vq_cells, itself checked against the oracle's);
- the exchange term for every closed-form screening and the static RPA, both
  forms, nb = 1..4;
- the linearized operator with the e-e terms (set_Lsc + dmd_scatt_add_Lsc_ee:
  the fac2 doubling under exchange, P3/P4 with the reference's sign or the
  exact one, P5/P6): the converted store and evolve_linear."""
import numpy as np
import pytest

from helpers import FS, random_P, rel_err

pytestmark = pytest.mark.gpu
# the nb = 3 e-e generator is switched off (DMD_ERR_UNSUPPORTED): a device run of
# it failed and took the device down; its cause is not found yet
NB3_OFF = pytest.mark.skip(reason="nb = 3 e-e generator switched off after a device failure whose cause is open")
NBS = [1, 2, pytest.param(3, marks=NB3_OFF), 4]


def _mods():
    from denmat_dynm_development_amd import scatt
    from oracle import oracle
    return scatt, oracle


def _case(nb, seed, km=(4, 3, 3), nkeep=24):
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
    sel[k1 == k2] = True
    return rng, km, kv, e, f, U, k1[sel].astype(np.uint64), k2[sel].astype(np.uint64)


def _model(scatt, scr, nw, km, kind, exchange=False):
    R = np.diag([7.0, 8.0, 9.0])
    return scatt.eimp_model(scr, 3e-6, 9.5e-4, abs(np.linalg.det(R)), scatt.reciprocal_metric(R), nw, n_ionized=1.0,
                            epsilon_background=3.0, degauss=4e-4, ethr=1.2e-3, scatt=kind, kmesh=km, nk_full=36.0,
                            exchange=exchange)


def _state(rng, f, amp=0.04):
    nk, nb = f.shape
    A = rng.normal(size=(nk, nb, nb)) + 1j * rng.normal(size=(nk, nb, nb))
    r = amp * 0.5 * (A + A.conj().transpose(0, 2, 1))
    r[:, np.arange(nb), np.arange(nb)] += f
    return r.reshape(nk, nb * nb)


def _tables(scatt, oracle, op, k1, k2, U, kv, e, f, m):
    """(v per pair, v per cell or None) for the oracle, checking the device's cell table."""
    d = m.as_dict()
    if d["screening"] == 3:
        vq, _, cells = oracle.eimp_rpa_vq(k1, k2, U, kv, e, f, d, cells=True)
        dev = op.ee_vq_cells(U, kv, f, m)
        assert np.abs(dev - cells).max() <= 1e-12 * np.abs(cells).max()
        return vq, cells
    return np.array([oracle.eimp_vq(d, oracle.eimp_q2(d["GGT"], kv[a], kv[b])) for a, b in zip(k1, k2)]), None


@pytest.mark.parametrize("nb", NBS)
def test_ee_P_at_a_state_matches_the_oracle(nb):
    scatt, oracle = _mods()
    rng, km, kv, e, f, U, k1, k2 = _case(nb, 80 + nb)
    nk = len(kv)
    dm = _state(rng, f)
    dm1 = -dm.copy()
    dm1[:, ::nb + 1] += 1.0
    for scr in ("debye", "RPA"):
        for kind in ("lindblad", "conventional"):
            print(f"case ee state nb={nb} {scr} {kind}", flush=True)
            m = _model(scatt, scr, U.shape[1], km, kind)
            with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
                vq, cells = _tables(scatt, oracle, op, k1, k2, U, kv, e, f, m)
                R1, R2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, 4e-4, 1.2e-3, 36.0, kind == "lindblad",
                                           vq_cell=cells, dm=dm, dm1=dm1)
                op.add_ee_model_state(k1, k2, U, kv, f, m, dm=dm, vq_cell=cells, scale=0.7, add=False)
                A1, A2 = op.download_P()
                s = np.abs(R1).max()
                assert s > 0 and np.abs(A1 - 0.7 * R1).max() <= 1e-12 * s and np.abs(A2 - 0.7 * R2).max() <= 1e-12 * s
                # at dm = None the state form is the equilibrium P
                op.add_ee_model_state(k1, k2, U, kv, f, m, vq_cell=cells, add=False)
                E1, _ = op.download_P()
                Q1, _ = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, 4e-4, 1.2e-3, 36.0, kind == "lindblad",
                                          vq_cell=cells)
                assert np.abs(E1 - Q1).max() <= 1e-12 * np.abs(Q1).max()


@pytest.mark.parametrize("exchange", [False, True])
def test_ee_one_pass_per_form_matches_both_forms(exchange, monkeypatch):
    """The nb <= 2 generator forms P1 and P2 in one pass by default; one pass
    per form (DMD_EE_FORMS=1) gives the same P, at f and at a state."""
    scatt, oracle = _mods()
    print(f"case ee forms exchange={exchange}", flush=True)
    rng, km, kv, e, f, U, k1, k2 = _case(2, 140 + exchange)
    nk = len(kv)
    dm = _state(rng, f)
    m = _model(scatt, "debye", U.shape[1], km, "lindblad", exchange=exchange)
    out = {}
    for forms in ("2", "1"):
        monkeypatch.setenv("DMD_EE_FORMS", forms)
        with scatt.ScatteringOperator(nk, 2, k1, k2, e) as op:
            op.add_ee_model_state(k1, k2, U, kv, f, m, dm=dm, add=False)
            S = op.download_P()
            op.add_ee_model(k1, k2, U, kv, f, m, add=False)
            out[forms] = (S, op.download_P())
    for (a1, a2), (b1, b2) in zip(out["2"], out["1"]):
        s = np.abs(b1).max()
        assert s > 0 and np.abs(a1 - b1).max() <= 1e-13 * s and np.abs(a2 - b2).max() <= 1e-13 * s


@pytest.mark.parametrize("nb", NBS)
def test_ee_exchange_every_screening(nb):
    scatt, oracle = _mods()
    rng, km, kv, e, f, U, k1, k2 = _case(nb, 90 + nb)
    nk = len(kv)
    for scr in ("unscreened", "debye", "Bechstedt", "RPA"):
        for kind in ("lindblad", "conventional"):
            print(f"case ee exchange nb={nb} {scr} {kind}", flush=True)
            m = _model(scatt, scr, U.shape[1], km, kind, exchange=True)
            with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
                vq, cells = _tables(scatt, oracle, op, k1, k2, U, kv, e, f, m)
                if cells is not None:  # the direct term's v from the cell table too, as the device uses
                    code = lambda x: int(np.dot(np.rint(np.array(km) * (x - np.floor(x))).astype(int) % np.array(km),
                                                [km[1] * km[2], km[2], 1]))
                    vq = np.array([cells[code(kv[int(a)] - kv[int(b)])] for a, b in zip(k1, k2)])
                R1, R2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, 4e-4, 1.2e-3, 36.0, kind == "lindblad",
                                           antisym_model=m.as_dict(), vq_cell=cells)
                op.add_ee_model(k1, k2, U, kv, f, m, add=False)
                A1, A2 = op.download_P()
            s = np.abs(R1).max()
            assert s > 0 and np.abs(A1 - R1).max() <= 1e-12 * s and np.abs(A2 - R2).max() <= 1e-12 * s


@pytest.mark.parametrize("nb", NBS)
@pytest.mark.parametrize("exchange", [False, True])
def test_linearized_ee_terms_match_the_oracle(nb, exchange):
    scatt, oracle = _mods()
    rng, km, kv, e, f, U, k1, k2 = _case(nb, 100 + nb + 10 * exchange, nkeep=20)
    nk = len(kv)
    scr = "debye" if nb % 2 else "RPA"
    m = _model(scatt, scr, U.shape[1], km, "lindblad", exchange=exchange)
    P1, P2 = random_P(len(k1), nb, rng)
    P1, P2 = 1e-3 * P1, 1e-3 * P2
    rho = _state(rng, f, 0.02)
    t = 37 * FS
    prefac = 2 * np.pi / nk
    for coef34 in ((1.0, -1.0) if not exchange else (1.0,)):
        print(f"case linearized ee nb={nb} exchange={exchange} {scr} coef34={coef34}", flush=True)
        with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
            vq, cells = _tables(scatt, oracle, op, k1, k2, U, kv, e, f, m)
            if cells is not None:
                code = lambda x: int(np.dot(np.rint(np.array(km) * (x - np.floor(x))).astype(int) % np.array(km),
                                            [km[1] * km[2], km[2], 1]))
                vq = np.array([cells[code(kv[int(a)] - kv[int(b)])] for a, b in zip(k1, k2)])
            am = m.as_dict() if exchange else None
            R1, R2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, 4e-4, 1.2e-3, 36.0, True, antisym_model=am,
                                       vq_cell=cells)
            sc = 0.8
            Lij, Lji = oracle.set_Lsc(k1, k2, P1 + sc * R1, P2 + sc * R2, f)
            P3, P4, P5, P6 = oracle.ee_model_P3P6(k1, k2, U, kv, e, f, km, 4e-4, 1.2e-3, 36.0, m.as_dict(), True,
                                                  exchange, vq_cell=cells)
            Lij, Lji = oracle.add_Lsc_ee(k1, k2, f, Lij, Lji, extra=sc if exchange else 0.0, Pee=(R1, R2),
                                         P3456=(sc * coef34 * P3, sc * coef34 * P4, sc * P5, sc * P6))
            op.upload_P(P1, P2)
            op.add_ee_model_state(k1, k2, U, kv, f, m, vq_cell=cells, scale=sc, add=True)
            op.set_Lsc(f)
            op.add_Lsc_ee(k1, k2, U, kv, f, m, scale=sc, extra=1.0 if exchange else 0.0, coef34=coef34,
                          dP_terms=True, vq_cell=cells)
            A1, A2 = op.download_P()
            s = np.abs(Lij).max()
            assert np.abs(A1 - Lij).max() <= 1e-12 * s and np.abs(A2 - Lji).max() <= 1e-12 * s
            got = op.eval(t, rho)
            ref = oracle.evolve_linear(k1, k2, Lij, Lji, e, rho, f, t=t, prefac=prefac)
            assert rel_err(got, ref) < 1e-12


@pytest.mark.parametrize("nb", NBS)
def test_ee_dP_and_its_evolve_term_match_the_oracle(nb):
    """alg_linearize_dPee: the dP tensors (dmd_scatt_add_ee_dP) for Debye, the
    static RPA and exchange, both forms; the row term kept by
    dmd_scatt_set_ee_dP; evaluations in both pictures with it added;
    compute_ddmdt_eq without it; and the term cleared."""
    scatt, oracle = _mods()
    rng, km, kv, e, f, U, k1, k2 = _case(nb, 120 + nb, nkeep=20)
    nk = len(kv)
    dm = _state(rng, f)
    dm1 = -dm.copy()
    dm1[:, ::nb + 1] += 1.0
    dm1 = dm1 + _state(rng, np.zeros_like(f), 0.02)  # the two sides independent
    P1, P2 = random_P(len(k1), nb, rng)
    P1, P2 = 1e-3 * P1, 1e-3 * P2
    rho = _state(rng, f, 0.02)
    t = 23 * FS
    prefac = 2 * np.pi / nk
    for scr, kind, exchange in (("debye", "lindblad", False), ("RPA", "conventional", False),
                                ("debye", "lindblad", True)):
        print(f"case ee dP nb={nb} {scr} {kind} exchange={exchange}", flush=True)
        m = _model(scatt, scr, U.shape[1], km, kind, exchange=exchange)
        with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
            vq, cells = _tables(scatt, oracle, op, k1, k2, U, kv, e, f, m)
            _, _, D1, D2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, 4e-4, 1.2e-3, 36.0, kind == "lindblad",
                                             antisym_model=m.as_dict() if exchange else None, vq_cell=cells,
                                             dm=dm, dm1=dm1, with_dP=True)
            op.add_ee_dP(k1, k2, U, kv, f, m, dm, dm1=dm1, vq_cell=cells, scale=0.6, add=False)
            A1, A2 = op.download_P()
            s = np.abs(D1).max()
            assert s > 0 and np.abs(A1 - 0.6 * D1).max() <= 1e-12 * s and np.abs(A2 - 0.6 * D2).max() <= 1e-12 * s
            op.upload_P(P1, P2)
            op.set_ee_dP(k1, k2, U, kv, f, m, dm, dm1=dm1, vq_cell=cells)
            H = op.get_ee_dP()
            Href = oracle.ee_dP_term(k1, k2, D1, D2, e, f, t=0.0, prefac=prefac)
            assert np.abs(H - Href).max() <= 1e-12 * np.abs(Href).max()
            term = oracle.ee_dP_term(k1, k2, D1, D2, e, f, t=t, prefac=prefac)
            for schro in (False, True):  # the Schroedinger picture runs with expt = 0 (parameters.cpp:452-453)
                op.set_flags(expt=not schro, schrodinger=schro)
                got = op.eval(t, rho)
                ref = oracle.evolve(k1, k2, P1, P2, e, rho, t=t, prefac=prefac, expt=not schro,
                                    schrodinger=schro) + term
                assert rel_err(got, ref) < 1e-12
            op.set_flags(expt=True, schrodinger=False)
            # compute_ddmdt_eq is evolve with compute_eq: no dP term in it
            op.compute_ddmdt_eq(f)
            eq = oracle.evolve(k1, k2, P1, P2, e, _diag(f), dm1=_diag(1 - f), t=0.0, prefac=prefac)
            got = op.eval(t, rho)
            ref = oracle.evolve(k1, k2, P1, P2, e, rho, t=t, prefac=prefac, ddmdt_eq=eq) + term
            assert rel_err(got, ref) < 1e-12
            op.set_ddmdt_eq(None)
            op.set_ee_dP(k1, k2, U, kv, f, m, None)
            got = op.eval(t, rho)
            assert rel_err(got, oracle.evolve(k1, k2, P1, P2, e, rho, t=t, prefac=prefac)) < 1e-12


def _diag(v):
    d = np.zeros((v.shape[0], v.shape[1] ** 2), complex)
    d[:, ::v.shape[1] + 1] = v
    return d
