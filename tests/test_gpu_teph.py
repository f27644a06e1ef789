"""The C++ drop-in for the reference's Teph, electronphonon_mi355x<complex>,
run the way dm_dynamics drives it (src_v4.5.7/DensityMatrix_Dynamics.h:59-63,
70,266-319; ElectronPhonon.h:158): a compiled driver (tests/native/
teph_driver.cpp) builds it from an ldbd_data folder (6 file bands, density
matrix on [1, 5), e-ph window [2, 4), scale_eph 0.8), calls evolve_driver on
row-pointer tables without and with an explicit 1 - dm, then
compute_ddmdt_eq and evolve_driver again; each result against the oracle's
evolve on the window (evolve_eph.cpp:21-166, the eq form :3-19)."""
import os
import subprocess

import numpy as np
import pytest

from helpers import FS, random_P, random_pairs, random_rho, rel_err, write_ldbd_folder

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "native", "teph_driver")


def test_teph_driver_matches_the_oracle(tmp_path):
    from oracle import oracle
    print("case teph driver", flush=True)
    assert os.path.exists(DRIVER), "tests/native/teph_driver is built by __graft_entry__.build()"
    rng = np.random.default_rng(21)
    nk, nbf = 36, 6
    dmw, ephw = (1, 5), (2, 4)
    e = np.sort(rng.uniform(-0.02, 0.06, (nk, nbf)), axis=1)
    k1, k2 = random_pairs(nk, 6, rng, diag=True)
    P1, P2 = random_P(len(k1), 2, rng)
    folder = tmp_path / "ldbd_data"
    write_ldbd_folder(str(folder), 2, nbf, 1, dmw, ephw, e, k1, k2, P1, P2, nk_full=1000.0)
    nbd, b0 = dmw[1] - dmw[0], ephw[0] - dmw[0]
    e_eph = e[:, ephw[0]:ephw[1]]
    A = random_rho(nk, nbd, rng).reshape(nk, nbd, nbd)
    f = np.sort(rng.uniform(0.05, 0.6, (nk, nbd)), axis=1)[:, ::-1].copy()
    A.reshape(nk, -1).astype(np.complex128).tofile(tmp_path / "dm.bin")
    f.astype(np.float64).tofile(tmp_path / "f.bin")
    se, t = 0.8, 61 * FS
    r = subprocess.run([DRIVER, str(folder), repr(se), repr(t), str(tmp_path / "dm.bin"), str(tmp_path / "f.bin"),
                        str(tmp_path / "out")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    got = {k: np.fromfile(tmp_path / f"out_{k}.bin", np.complex128).reshape(nk, nbd * nbd)
           for k in ("plain", "dm1", "eq")}
    prefac = 2 * np.pi / 1000.0
    cut = A[:, b0:b0 + 2, b0:b0 + 2].reshape(nk, 4)
    fw = f[:, b0:b0 + 2]
    dm_eq = np.zeros((nk, 4), np.complex128)
    dm_eq[:, 0], dm_eq[:, 3] = fw[:, 0], fw[:, 1]
    eq = oracle.evolve(k1, k2, se * P1, se * P2, e_eph, dm_eq, t=0.0, prefac=prefac)

    def embed(inner):
        ref = np.zeros((nk, nbd, nbd), np.complex128)
        ref[:, b0:b0 + 2, b0:b0 + 2] = inner.reshape(nk, 2, 2)
        return ref.reshape(nk, nbd * nbd)

    plain = embed(oracle.evolve(k1, k2, se * P1, se * P2, e_eph, cut, t=t, prefac=prefac))
    with_eq = embed(oracle.evolve(k1, k2, se * P1, se * P2, e_eph, cut, t=t, prefac=prefac, ddmdt_eq=eq))
    assert rel_err(got["plain"], plain) < 1e-12
    assert rel_err(got["dm1"], plain) < 1e-12
    assert rel_err(got["eq"], with_eq) < 1e-12
    # the eq subtraction is visible: the two results differ by ddmdt_eq
    assert rel_err(got["eq"], got["dm1"]) > 1e-6
