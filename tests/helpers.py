"""Shared builders for test inputs (seeded)."""
import numpy as np

FS = 41.341373335182  # 1 fs in Hartree atomic time (CODATA); the reference's comm/Units.h:59-61 derives
# 41.34137417 from its base constants, which dmd_run uses for param.in times


def random_pairs(nk, npart, rng, diag=False):
    s = set()
    for k in range(nk):
        for _ in range(npart):
            kp = int(rng.integers(nk))
            if kp == k and not diag:
                continue
            s.add((min(k, kp), max(k, kp)))
    if diag:
        for k in range(0, nk, 3):
            s.add((k, k))
    pr = sorted(s)
    k1 = np.array([a for a, _ in pr], np.uint64)
    k2 = np.array([b for _, b in pr], np.uint64)
    return k1, k2


def random_energies(nk, nb, rng, emax=3.7e-3):
    return np.sort(rng.uniform(0, emax, (nk, nb)), axis=1)


def random_P(npairs, nb, rng, structured=True):
    """P1/P2 either in the Lindblad Gram form sum_l A (x) conj(A) (the form the
    initializer writes) or unstructured complex numbers."""
    n2, n4 = nb * nb, nb ** 4
    if not structured:
        P1 = rng.normal(size=(npairs, n4)) + 1j * rng.normal(size=(npairs, n4))
        P2 = rng.normal(size=(npairs, n4)) + 1j * rng.normal(size=(npairs, n4))
        return 1e-3 * P1, 1e-3 * P2
    P1 = np.zeros((npairs, nb, nb, nb, nb), np.complex128)
    P2 = np.zeros_like(P1)
    for _ in range(3):
        A = (rng.normal(size=(npairs, nb, nb)) + 1j * rng.normal(size=(npairs, nb, nb))) * 0.03
        B = (rng.normal(size=(npairs, nb, nb)) + 1j * rng.normal(size=(npairs, nb, nb))) * 0.03
        # P1[(i1 i2),(i3 i4)] = A[i1,i3] conj(A[i2,i4]); P2 = B[i3,i1] conj(B[i4,i2])
        P1 += np.einsum("pac,pbd->pabcd", A, A.conj())
        P2 += np.einsum("pca,pdb->pabcd", B, B.conj())
    return P1.reshape(npairs, n4), P2.reshape(npairs, n4)


def random_rho(nk, nb, rng, scale=0.05, diag=0.3):
    A = rng.normal(size=(nk, nb, nb)) + 1j * rng.normal(size=(nk, nb, nb))
    H = scale * (A + A.conj().transpose(0, 2, 1)) / 2
    H += np.eye(nb) * diag
    return H.reshape(nk, nb * nb)


def rel_err(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def gaas_spin_fixture():
    """The reference GaAs run's final state and S_z (tests/golden/make_gaas_spin.py).

    Returns (e_dm, f, s_z, dm, i0, i1, t, prefac, sz, sz_amp) with s_z and dm
    full nb_dm x nb_dm per k (zero outside what the electron S_z reads).  The
    run's sz_elec_tot.out holds totals per unit cell: prefac = 1/nk_full.
    """
    import os
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gaas_spin_restart.npz"))
    e = d["e_dm"]
    nk, nb, nv = e.shape[0], int(d["nb_dm"]), int(d["nv_dm"])
    f = 1.0 / (np.exp((e - float(d["mu"])) / float(d["T"])) + 1.0)  # electron.h:119-124
    s = np.zeros((nk, nb, nb), np.complex128)
    s[:, nv:, :] = d["sz_rows"]
    dm = np.zeros((nk, nb, nb), np.complex128)
    dm[:, :, nv:] = d["dm_cols"]
    return (e, f, s.reshape(nk, nb * nb), dm.reshape(nk, nb * nb), nv, nb, float(d["t"]),
            1.0 / float(d["nk_full"]), float(d["sz"]), float(d["sz_amp"]))


def quantise_P(P1, P2, precision):
    """Stored values of the reduced-precision forms: complex64 (round to
    nearest), or BF16 re/im (round to nearest even) of P / s with the per-pair
    power-of-two scale s = 2^e, max |re|, |im| of the pair in [2^(e-1), 2^e)."""
    P1 = np.asarray(P1, np.complex128)
    P2 = np.asarray(P2, np.complex128)
    if precision == 1:
        return P1.astype(np.complex64).astype(np.complex128), P2.astype(np.complex64).astype(np.complex128)
    both = np.concatenate([P1, P2], axis=1)
    mx = np.maximum(np.abs(both.real), np.abs(both.imag)).max(axis=1)
    _, ex = np.frexp(np.where(mx > 0, mx, 1.0))
    sc = np.where(mx > 0, np.ldexp(1.0, ex), 1.0).astype(np.float32).astype(np.float64)[:, None]

    def bf16(x):
        u = (x / sc).astype(np.float32).view(np.uint32).astype(np.uint64)
        u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
        return u.astype(np.uint32).view(np.float32).astype(np.float64) * sc

    q = bf16(both.real) + 1j * bf16(both.imag)
    n = P1.shape[1]
    return q[:, :n], q[:, n:]


def write_ldbd_folder(d, kind, nb, nv, dm, eph, e, k1, k2, P1=None, P2=None, nk_full=None, kmesh=(0, 0, 0),
                      modes=(0, 6, 0, 0), T=9.5004e-4, mu=(0.05, 0.05, 0.05), degauss=(1.8e-4, 1000),
                      ewin=(-0.1, 0.1, -0.05, 0.08, 0.0, 0.06), omega_max=1.9e-3, scatt="lindblad", extra_P=()):
    """An ldbd_data folder in the layouts of the initializer's saveSize / saveEk
    / P writers (lindbladInit.../electron_save.cpp:16-65): kind 0 conduction
    only, 1 valence only (hole files and fields), 2 both.  dm and eph are
    [bottom, top) band ranges of the ldbd_ek.bin bands; extra_P holds
    (impurity index, P1, P2) written as the _D<i> files."""
    import os
    os.makedirs(d, exist_ok=True)
    nk = e.shape[0]
    nk_full = float(nk if nk_full is None else nk_full)
    np_ = len(k1)
    first = ["There are scatterings only for conduction electrons", "There are scatterings only for valence electrons",
             "There are scatterings for both valence and conduction electrons"][kind]
    if kind == 1:
        line2 = "%d %d %d %d %d %d %d %d %d %d %d # nb nv bBot_dm bTop_dm bBot_eph_elec bTop_eph_elec bBot_eph_hole " \
                "bTop_eph_hole nb_wannier bBot_probe band_dft_skipped" % (nb, nv, dm[0], dm[1], 0, 0, eph[0], eph[1],
                                                                           nb + 4, 0, 0)
        line4 = "%d %d # nkpair_elec nkpair_hole" % (0, np_)
    else:
        line2 = "%d %d %d %d %d %d %d %d %d # nb nv bBot_dm bTop_dm bBot_eph bTop_eph nb_wannier bBot_probe " \
                "band_dft_skipped" % (nb, nv, dm[0], dm[1], eph[0], eph[1], nb + 4, 0, 0)
        line4 = "%d # nkpair" % np_
    lines = [first, line2, "%21.14le %d %d %d %d # nk_full nk kmesh" % (nk_full, nk, *kmesh), line4,
             "%d %d %d %d# modeStart modeStp modeSkipStart modeSkipStop" % tuple(modes),
             "%21.14le # T" % T, "%21.14le %21.14le %21.14le # muMin, muMax mu (given carrier density)" % tuple(mu),
             "%g %g # degauss" % tuple(degauss),
             "%14.7le %14.7le %14.7le %14.7le %14.7le %14.7le # EBot_probe, ETop_probe, EBot_dm, ETop_dm, EBot_eph, "
             "ETop_eph" % tuple(ewin)]
    text = "\n".join(lines) + "\n" + "%14.7le # omegaMax" % omega_max
    with open(os.path.join(d, "ldbd_size.dat"), "w") as f:
        f.write(text)
    np.ascontiguousarray(e, np.float64).tofile(os.path.join(d, "ldbd_ek.bin"))
    sfx = "_hole" if kind == 1 else ""
    np.ascontiguousarray(k1, np.uint64).tofile(os.path.join(d, f"ldbd_kpair_k1st{sfx}.bin"))
    np.ascontiguousarray(k2, np.uint64).tofile(os.path.join(d, f"ldbd_kpair_k2nd{sfx}.bin"))
    if P1 is not None:
        np.ascontiguousarray(P1, np.complex128).tofile(os.path.join(d, f"ldbd_P1_{scatt}{sfx}.bin"))
        np.ascontiguousarray(P2, np.complex128).tofile(os.path.join(d, f"ldbd_P2_{scatt}{sfx}.bin"))
    for i, Q1, Q2 in extra_P:
        np.ascontiguousarray(Q1, np.complex128).tofile(os.path.join(d, f"ldbd_P1_{scatt}_D{i}{sfx}.bin"))
        np.ascontiguousarray(Q2, np.complex128).tofile(os.path.join(d, f"ldbd_P2_{scatt}_D{i}{sfx}.bin"))


def sparse_lists(P, rng=None, dup_frac=0.0):
    """COO lists (ns, s, i, j) of the nonzero elements of P (npairs x nb^4),
    row index i1*nb+i2 and column index i3*nb+i4, in row-major order or, with
    rng, shuffled within each pair; dup_frac of the pairs also list one element
    twice, first with a wrong value (a reader keeps the last)."""
    npairs, n4 = P.shape
    nb2 = int(round(np.sqrt(n4)))
    ns, s, ii, jj = [], [], [], []
    for p in range(npairs):
        idx = np.flatnonzero(P[p] != 0)
        if rng is not None:
            idx = rng.permutation(idx)
        vals = list(P[p, idx])
        rows, cols = list(idx // nb2), list(idx % nb2)
        if rng is not None and len(idx) and rng.uniform() < dup_frac:
            q = int(rng.integers(len(idx)))
            vals.insert(0, vals[q] * 3 + 1)
            rows.insert(0, rows[q])
            cols.insert(0, cols[q])
        ns.append(len(vals))
        s += vals
        ii += rows
        jj += cols
    return (np.array(ns, np.int32), np.array(s, np.complex128).reshape(-1), np.array(ii, np.int32),
            np.array(jj, np.int32))


def sparsify(P, rng, zero_frac=0.5):
    """P with about zero_frac of its elements set to exactly zero."""
    return np.where(rng.uniform(size=P.shape) < zero_frac, 0, P)


def write_sP_files(d, which, lists, scatt="lindblad", hole=False):
    """sP<which>_<scatt>[_hole]_{ns,s,i,j}.bin as sparse_mat::write_to_files
    lays them out (comm/sparse_matrix.h:37-42): int counts, complex128 values,
    int row and int column indices, pair after pair."""
    import os
    ns, s, i, j = lists
    base = os.path.join(str(d), f"sP{which}_{scatt}{'_hole' if hole else ''}_")
    np.asarray(ns, np.int32).tofile(base + "ns.bin")
    np.asarray(s, np.complex128).tofile(base + "s.bin")
    np.asarray(i, np.int32).tofile(base + "i.bin")
    np.asarray(j, np.int32).tofile(base + "j.bin")
