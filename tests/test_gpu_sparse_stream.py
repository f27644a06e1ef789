"""P stored by its nonzero elements (dmd_scatt_compress_P, the streaming side
of the reference's sparse modes) through the C-ABI, nb = 2 tile kernel:
- evaluation on the compressed store equals the dense store's and the
  oracle's sparse evolve (per-pair element lists, evolve_eph.cpp:69-86) on the
  same P, at t = 0 and with phases, for densities from 100 % to 5 % and with
  every pair at a density of its own;
- RK4 on the compressed store follows the dense one;
- the stored bytes fall with the density (about half at half the density),
  and a P with more than 3/4 of its elements nonzero stays dense;
- operations that need the dense store are refused afterwards."""
import numpy as np
import pytest

from helpers import FS, random_P, random_energies, random_pairs, random_rho, rel_err, sparse_lists

pytestmark = pytest.mark.gpu


def _mods():
    from denmat_dynm_development_amd import _lib, scatt
    from oracle import oracle
    return _lib, scatt, oracle


@pytest.mark.parametrize("density", [1.0, 0.5, 0.25, 0.05, "mixed"])
def test_compressed_eval_matches_dense_and_oracle(density):
    _lib, scatt, oracle = _mods()
    print(f"case compressed P density={density}", flush=True)
    rng = np.random.default_rng(7 if density == "mixed" else int(100 * density))
    nk, nb = 700, 2
    k1, k2 = random_pairs(nk, 9, rng, diag=True)
    perm = rng.permutation(len(k1))
    k1, k2 = k1[perm], k2[perm]
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    if density == "mixed":  # every pair its own density, 0 .. 1
        d = rng.uniform(size=(len(k1), 1))
        P1 = np.where(rng.uniform(size=P1.shape) < d, P1, 0)
        P2 = np.where(rng.uniform(size=P2.shape) < d, P2, 0)
    else:
        P1 = np.where(rng.uniform(size=P1.shape) < density, P1, 0)
        P2 = np.where(rng.uniform(size=P2.shape) < density, P2, 0)
    prefac = 2 * np.pi / nk
    rhos = [random_rho(nk, nb, rng) for _ in range(2)]
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(P1, P2)
        dense_bytes = op.info()["p_bytes"]
        dense = [op.eval(t, r) for t, r in zip((0.0, 37 * FS), rhos)]
        stored = op.compress_P()
        sparse_bytes = op.info()["p_bytes"]
        comp = [op.eval(t, r) for t, r in zip((0.0, 37 * FS), rhos)]
        if density == 1.0:  # kept dense above 3/4 nonzero: P reads back unchanged
            D1, D2 = op.download_P()
            assert np.array_equal(D1, P1) and np.array_equal(D2, P2)
        else:
            with pytest.raises(_lib.DMDError) as ex:
                op.download_P()
            assert ex.value.code == _lib.DMD_ERR_STATE and "compressed" in str(ex.value)
            with pytest.raises(_lib.DMDError):
                op.upload_P(P1, P2)
        comp2 = op.eval(37 * FS, rhos[1])  # the refusals left the store intact
    s1, s2 = sparse_lists(P1), sparse_lists(P2)
    for (t, r), a, b in zip(zip((0.0, 37 * FS), rhos), dense, comp):
        ref = oracle.evolve_sparse(k1, k2, s1, s2, e, r, t=t, prefac=prefac)
        assert rel_err(b, a) < 1e-13, rel_err(b, a)
        assert rel_err(b, ref) < 1e-12, rel_err(b, ref)
    assert rel_err(comp2, comp[1]) < 1e-13
    nnz = np.count_nonzero(P1) + np.count_nonzero(P2)
    print(f"  nonzero fraction {nnz / P1.size / 2:.3f}: stored values {stored} for {nnz} nonzeros, "
          f"bytes {sparse_bytes} vs dense {dense_bytes} ({sparse_bytes / dense_bytes:.3f})", flush=True)
    assert stored == nnz
    if density == 0.5:
        assert sparse_bytes < 0.56 * dense_bytes
    if density == 1.0:
        assert sparse_bytes == dense_bytes


def test_compressed_rk4_follows_dense():
    _lib, scatt, oracle = _mods()
    print("case compressed P rk4", flush=True)
    rng = np.random.default_rng(3)
    nk, nb = 400, 2
    k1, k2 = random_pairs(nk, 6, rng, diag=True)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    P1 = np.where(rng.uniform(size=P1.shape) < 0.4, P1, 0)
    P2 = np.where(rng.uniform(size=P2.shape) < 0.4, P2, 0)
    rho0 = random_rho(nk, nb, rng, scale=0.02)
    s = random_rho(nk, nb, rng, scale=0.5, diag=0.0)
    f = np.full((nk, nb), 0.3)
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(P1, P2)
        op.set_observable(s, f, 0, nb, 1.0)
        rd, od = op.integrate_rk4(rho0, 0.0, 2 * FS, 12, report_every=3)
        op.compress_P()
        rc, oc = op.integrate_rk4(rho0, 0.0, 2 * FS, 12, report_every=3)
    assert rel_err(rc - rho0, rd - rho0) < 1e-12
    assert np.abs(oc - od).max() <= 1e-12 * np.abs(od).max()
