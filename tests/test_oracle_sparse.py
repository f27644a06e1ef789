"""The oracle's sparse-P evolution (the alg_sparseP / alg_Pin_is_sparse branch
of electronphonon::evolve, src_v4.5.7/evolve_eph.cpp:69-86) pinned to the dense
oracle: the same operator on the densified P for any element order, at t = 0
and with phases; a planted transposition fails it.  The numpy reader of the
sparse files checks its sizes."""
import numpy as np
import pytest

from helpers import FS, random_P, random_energies, random_pairs, random_rho, rel_err, sparse_lists, sparsify, \
    write_sP_files


def _case(nb, seed):
    rng = np.random.default_rng(seed)
    nk = 40 if nb <= 2 else 14
    k1, k2 = random_pairs(nk, 4, rng, diag=True)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    return rng, nk, k1, k2, e, sparsify(P1, rng, 0.6), sparsify(P2, rng, 0.4)


@pytest.mark.parametrize("nb", [1, 2, 3, 4])
def test_sparse_equals_dense_on_the_densified_P(nb):
    from oracle import oracle
    rng, nk, k1, k2, e, P1, P2 = _case(nb, 90 + nb)
    prefac = 2 * np.pi / nk
    for shuffle in (None, rng):
        s1, s2 = sparse_lists(P1, shuffle), sparse_lists(P2, shuffle)
        for t_fs, expt in ((0.0, True), (37.0, True), (37.0, False)):
            rho = random_rho(nk, nb, rng)
            got = oracle.evolve_sparse(k1, k2, s1, s2, e, rho, t=t_fs * FS, prefac=prefac, expt=expt)
            ref = oracle.evolve(k1, k2, P1, P2, e, rho, t=t_fs * FS, prefac=prefac, expt=expt)
            assert rel_err(got, ref) < 1e-14, (t_fs, expt, rel_err(got, ref))


def test_planted_transposition_fails():
    from oracle import oracle
    rng, nk, k1, k2, e, P1, P2 = _case(2, 97)
    s1, s2 = sparse_lists(P1), sparse_lists(P2)
    rho = random_rho(nk, 2, rng)
    ref = oracle.evolve(k1, k2, P1, P2, e, rho, t=20 * FS, prefac=1.0)
    wrong = (s1[0], s1[1], s1[3], s1[2])  # row and column lists exchanged
    assert rel_err(oracle.evolve_sparse(k1, k2, wrong, s2, e, rho, t=20 * FS, prefac=1.0), ref) > 1e-3
    with pytest.raises(ValueError):  # an index outside 0..nb^2-1
        bad = (s1[0], s1[1], s1[2] + 4, s1[3])
        oracle.evolve_sparse(k1, k2, bad, s2, e, rho)


def test_read_sP_checks_sizes(tmp_path):
    from denmat_dynm_development_amd import ldbd
    from helpers import write_ldbd_folder
    rng, nk, k1, k2, e, P1, P2 = _case(2, 98)
    d = tmp_path / "ldbd_data"
    write_ldbd_folder(d, 0, 2, 0, (0, 2), (0, 2), e, k1, k2)
    s1 = sparse_lists(P1, rng)
    write_sP_files(d, 1, s1)
    got = ldbd.read_sP(d, 1)
    for a, b in zip(got, s1):
        assert np.array_equal(a, b)
    write_sP_files(d, 1, (s1[0], s1[1][:-1], s1[2], s1[3]))
    with pytest.raises(ValueError):
        ldbd.read_sP(d, 1)
    write_sP_files(d, 1, (s1[0][:-1], s1[1], s1[2], s1[3]))
    with pytest.raises(ValueError):
        ldbd.read_sP(d, 1)
