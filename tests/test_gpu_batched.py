"""Batched trajectories (B states sharing one pass over P on the matrix cores,
nb = 4, BF16 or complex64 P) through the C-ABI, against the FP64 oracle on
the stored P and against the unbatched reduced-precision path, one
trajectory at a time.

The batched form carries drho = rho - diag(f) as two BF16 terms (a 16-bit
significand) with FP32 sums, and complex64 P as BF16 hi + lo; everything
else matches the unbatched forms.  For the BF16 store the visit records hold
each folded entry as one FP16 word and drho as FP16 hi + lo.  A folded entry
is a BF16 word or the sum or difference of two, exact in FP16's 11-bit
significand when the two exponents differ by at most 2 and rounded by at
most 2^-11 relative otherwise, so the drho-driven part carries well under
2^-11 of rounding (measured 4e-6 .. 9e-6 here, 7e-7 at the nb = 4, nk = 32768
workload) against 2^-9 for the BF16 store itself; its tolerance is 1e-4, the
complex64 form's 5e-6.  DMD_BATCH_HALF=0 keeps BF16 hi + lo records for the
BF16 store.
Tolerances are stated against the part of ddm that drho drives.
"""
import numpy as np
import pytest

from helpers import FS, random_P, random_energies, random_pairs, random_rho, rel_err

pytestmark = pytest.mark.gpu
BF16 = 2
TOL_DRHO = {1: 5e-6, 2: 1e-4}  # complex64 (BF16 hi + lo records) | BF16 store (FP16 records)


def _mods():
    from denmat_dynm_development_amd import scatt
    from oracle import oracle
    return scatt, oracle


def _hub_pairs(nk, npart, rng):
    """random pairs plus a hub row 0 linked to every k (a row of nk visits, several segments)"""
    k1, k2 = random_pairs(nk, npart, rng, diag=True)
    s = set(zip(k1.tolist(), k2.tolist())) | {(0, k) for k in range(nk)} | {(k, nk - 1) for k in range(nk)}
    pr = sorted(s)
    return np.array([a for a, _ in pr], np.uint64), np.array([b for _, b in pr], np.uint64)


def _states(B, nk, nb, f, rng, amp=1e-3):
    rho = np.zeros((B, nk, nb, nb), np.complex128)
    rho[:, :, np.arange(nb), np.arange(nb)] = f
    rho = rho.reshape(B, nk, nb * nb)
    for b in range(B):
        rho[b] += random_rho(nk, nb, rng, scale=amp * (1 + b % 5), diag=0.0)
    return rho


@pytest.mark.parametrize("prec,B", [(2, 3), (2, 40), (2, 64), (1, 40), (1, 64)])
def test_eval_batched_matches_oracle_and_unbatched(prec, B):
    scatt, oracle = _mods()
    print(f"case batched eval prec={prec} B={B}", flush=True)
    rng = np.random.default_rng(100 + B)
    nk, nb = 300, 4
    k1, k2 = _hub_pairs(nk, 3, rng)
    perm = rng.permutation(len(k1))
    k1, k2 = k1[perm], k2[perm]
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    f = np.sort(rng.uniform(0.05, 0.6, (nk, nb)), axis=1)[:, ::-1].copy()
    rho = _states(B, nk, nb, f, rng)
    eq = random_rho(nk, nb, rng, scale=1e-6, diag=1e-7)
    t = 77 * FS
    prefac = 2 * np.pi / nk
    with scatt.ScatteringOperator(nk, nb, k1, k2, e, precision=prec) as op:
        op.upload_P(P1, P2)
        op.set_split(f)
        op.set_ddmdt_eq(eq)
        Q1, Q2 = op.download_P()
        got = op.eval_batched(t, rho)
        one = np.stack([op.eval(t, rho[b]) for b in range(B)])
        rho_f = np.zeros((nk, nb * nb), np.complex128)
        rho_f[:, ::nb + 1] = f
        base = op.eval(t, rho_f)
    worst = 0.0
    for b in sorted({0, B // 2, B - 1}):
        ref = oracle.evolve(k1, k2, Q1, Q2, e, rho[b], t=t, prefac=prefac, ddmdt_eq=eq)
        ref_f = oracle.evolve(k1, k2, Q1, Q2, e, rho_f, t=t, prefac=prefac, ddmdt_eq=eq)
        worst = max(worst, rel_err(got[b] - base, ref - ref_f))
        assert rel_err(got[b], ref) < (1e-7 if prec == 1 else 1e-5)
    d_one = rel_err(got - base[None], one - base[None])
    print(f"prec={prec} B={B}: drho part vs oracle {worst:.2e}, vs unbatched {d_one:.2e}")
    # 16-bit drho operand (and, for complex64, 16-bit hi + lo P), FP32 sums;
    # BF16 store: FP16 entries of M
    assert worst < TOL_DRHO[prec] and d_one < TOL_DRHO[prec]
    if prec == BF16:
        assert worst > 1e-7  # the FP16 records are in use (hi + lo records reach ~1e-7)


def test_fp16_records_match_hi_lo_records(monkeypatch):
    """The BF16 store's FP16 records against its BF16 hi + lo records
    (DMD_BATCH_HALF=0) on the same handle set-up, with states whose drho spans
    nine decades across k and trajectories of very different size (each
    trajectory's FP16 operand has its own power-of-two scale)."""
    scatt, oracle = _mods()
    print("case batched fp16 records", flush=True)
    rng = np.random.default_rng(77)
    nk, nb, B = 220, 4, 4
    k1, k2 = _hub_pairs(nk, 3, rng)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    P1[: len(P1) // 3] *= 1e-7  # pairs far below the largest scale of the store
    f = np.sort(rng.uniform(0.05, 0.6, (nk, nb)), axis=1)[:, ::-1].copy()
    rho = _states(B, nk, nb, f, rng)
    decay = np.logspace(0, -9, nk)[:, None]
    for b in range(B):  # drho x decay(k) x 1e-3b; the FP64 assembly resolves all of them
        d = rho[b].copy()
        d[:, ::nb + 1] -= f
        d *= decay * 10.0 ** (-3 * b)
        d[:, ::nb + 1] += f
        rho[b] = d
    t = 29 * FS
    rho_f = np.zeros((nk, nb * nb), np.complex128)
    rho_f[:, ::nb + 1] = f
    out = {}
    for half in ("1", "0"):
        monkeypatch.setenv("DMD_BATCH_HALF", half)
        with scatt.ScatteringOperator(nk, nb, k1, k2, e, precision=BF16) as op:
            op.upload_P(P1, P2)
            op.set_split(f)
            out[half] = op.eval_batched(t, rho) - op.eval(t, rho_f)[None]
    for b in range(B):
        err = rel_err(out["1"][b], out["0"][b])
        print(f"trajectory {b} (drho x 1e-{3 * b}): FP16 records vs hi + lo {err:.2e}")
        assert np.linalg.norm(out["0"][b]) > 0 and err < TOL_DRHO[BF16]


@pytest.mark.parametrize("prec", [1, 2])
def test_record_free_pass_matches_the_record_pass(prec, monkeypatch):
    """The pass that forms each visit's folded M from the stored P (used when
    the visit records do not fit; DMD_BATCH_RECORDS=0 forces it) against the
    record pass on the same handle set-up: the same products, so the results
    differ only by the order of the FP32 row-sum atomics."""
    scatt, oracle = _mods()
    print(f"case batched record-free prec={prec}", flush=True)
    rng = np.random.default_rng(7 + prec)
    nk, nb, B = 250, 4, 40
    k1, k2 = _hub_pairs(nk, 3, rng)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    f = np.sort(rng.uniform(0.05, 0.6, (nk, nb)), axis=1)[:, ::-1].copy()
    rho = _states(B, nk, nb, f, rng)
    t = 53 * FS
    out = {}
    monkeypatch.setenv("DMD_BATCH_HALF", "0")  # the record pass with the same (hi + lo) products
    for rec in ("1", "0"):
        monkeypatch.setenv("DMD_BATCH_RECORDS", rec)
        with scatt.ScatteringOperator(nk, nb, k1, k2, e, precision=prec) as op:
            op.upload_P(P1, P2)
            op.set_split(f)
            rho_f = np.zeros((nk, nb * nb), np.complex128)
            rho_f[:, ::nb + 1] = f
            out[rec] = (op.eval_batched(t, rho), op.eval(t, rho_f))
    (a, base), (b, _) = out["1"], out["0"]
    err = rel_err(b - base[None], a - base[None])
    print(f"prec={prec}: record-free vs records, drho part {err:.2e}")
    assert err < 1e-6


def test_eval_batched_fp64_handle_is_per_trajectory():
    scatt, oracle = _mods()
    rng = np.random.default_rng(7)
    nk, nb, B = 50, 4, 3
    k1, k2 = random_pairs(nk, 4, rng, diag=True)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    rho = np.stack([random_rho(nk, nb, rng) for _ in range(B)])
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(P1, P2)
        got = op.eval_batched(11 * FS, rho)
        for b in range(B):
            assert rel_err(got[b], op.eval(11 * FS, rho[b])) < 1e-13  # FP64 atomics: rounding order only


@pytest.mark.parametrize("half", ["0", "1"])
def test_rk4_batched_follows_unbatched(half, monkeypatch):
    scatt, oracle = _mods()
    print(f"case batched rk4 half={half}", flush=True)
    monkeypatch.setenv("DMD_BATCH_HALF", half)
    tol = 1e-4 if half == "0" else TOL_DRHO[BF16]
    rng = np.random.default_rng(9)
    nk, nb, B = 160, 4, 5
    k1, k2 = _hub_pairs(nk, 4, rng)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    P1, P2 = 0.2 * P1, 0.2 * P2
    f = np.sort(rng.uniform(0.05, 0.6, (nk, nb)), axis=1)[:, ::-1].copy()
    rho0 = _states(B, nk, nb, f, rng)
    s = random_rho(nk, nb, rng, scale=0.5, diag=0.0)
    dt, n, rep = 2 * FS, 24, 6
    with scatt.ScatteringOperator(nk, nb, k1, k2, e, precision=BF16) as op:
        op.upload_P(P1, P2)
        op.set_split(f)
        op.set_observable(s, f, 0, nb, 1.0)
        rb, ob = op.integrate_rk4_batched(rho0, 0.0, dt, n, report_every=rep)
        monkeypatch.setenv("DMD_GRAPH", "0")
        rb2, ob2 = op.integrate_rk4_batched(rho0, 0.0, dt, n, report_every=rep)
        monkeypatch.delenv("DMD_GRAPH")
        singles = [op.integrate_rk4(rho0[b], 0.0, dt, n, report_every=rep) for b in range(B)]
    r1 = np.stack([x[0] for x in singles])
    o1 = np.stack([x[1] for x in singles])
    sig = np.abs(o1 - o1[:, :1]).max()
    d_state = rel_err(rb - rho0, r1 - rho0)
    d_obs = np.abs(ob - o1).max() / sig
    print(f"batched rk4: state change {d_state:.2e}, S(t) {d_obs:.2e} of the signal change {sig:.2e}")
    assert ob.shape == (B, n // rep + 1)
    assert d_state < tol and d_obs < tol
    # graph replay and eager stepping agree (FP32 atomics: same to rounding)
    assert rel_err(rb2 - rho0, rb - rho0) < 1e-5


def test_partner_block_segment_orders_agree(monkeypatch):
    """The segment order the autotuner may pick (row order, or visits grouped
    in blocks of partner rows so the partner operands share the L2) changes
    only the FP32 summation order: every forced order matches the oracle on
    the stored P, and the orders match each other."""
    scatt, oracle = _mods()
    print("case batched partner blocks", flush=True)
    rng = np.random.default_rng(140)
    nk, nb, B = 200, 4, 32
    k1, k2 = _hub_pairs(nk, 4, rng)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    f = np.sort(rng.uniform(0.05, 0.6, (nk, nb)), axis=1)[:, ::-1].copy()
    rho = _states(B, nk, nb, f, rng)
    t = 31 * FS
    prefac = 2 * np.pi / nk
    rho_f = np.zeros((nk, nb * nb), np.complex128)
    rho_f[:, ::nb + 1] = f
    results = []
    for pblock in (0, 7, 64):
        monkeypatch.setenv("DMD_BATCH_PBLOCK", str(pblock))
        with scatt.ScatteringOperator(nk, nb, k1, k2, e, precision=1) as op:
            op.upload_P(P1, P2)
            op.set_split(f)
            Q1, Q2 = op.download_P()
            got = op.eval_batched(t, rho)
            base = op.eval(t, rho_f)
        results.append(got - base[None])
    monkeypatch.delenv("DMD_BATCH_PBLOCK")
    ref_f = oracle.evolve(k1, k2, Q1, Q2, e, rho_f, t=t, prefac=prefac)
    for b in (0, B - 1):
        ref = oracle.evolve(k1, k2, Q1, Q2, e, rho[b], t=t, prefac=prefac)
        for r in results:
            assert rel_err(r[b], ref - ref_f) < 5e-6
    for r in results[1:]:
        assert rel_err(r, results[0]) < 2e-6


@pytest.mark.parametrize("prec", [1, 2])
def test_persistent_pass_matches_the_per_segment_grid(prec, monkeypatch):
    """The persistent pass (resident wavefronts taking segments from a queue
    counter, the default) against one wavefront per segment
    (DMD_BATCH_PERSIST=0): the same products per segment, so the results
    differ only by the order of the FP32 row-sum atomics; two evaluations in a
    row check that the queue counters start from zero every time."""
    scatt, oracle = _mods()
    print(f"case batched persistent prec={prec}", flush=True)
    rng = np.random.default_rng(31 + prec)
    nk, nb, B = 230, 4, 40
    k1, k2 = _hub_pairs(nk, 3, rng)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    f = np.sort(rng.uniform(0.05, 0.6, (nk, nb)), axis=1)[:, ::-1].copy()
    rho = _states(B, nk, nb, f, rng)
    t = 41 * FS
    rho_f = np.zeros((nk, nb * nb), np.complex128)
    rho_f[:, ::nb + 1] = f
    out = {}
    for persist in ("1", "0"):
        monkeypatch.setenv("DMD_BATCH_PERSIST", persist)
        with scatt.ScatteringOperator(nk, nb, k1, k2, e, precision=prec) as op:
            op.upload_P(P1, P2)
            op.set_split(f)
            base = op.eval(t, rho_f)
            first = op.eval_batched(t, rho) - base[None]
            second = op.eval_batched(t, rho) - base[None]
            out[persist] = (first, second)
    (a1, a2), (g1, _) = out["1"], out["0"]
    err, rep = rel_err(a1, g1), rel_err(a2, a1)
    print(f"prec={prec}: persistent vs grid {err:.2e}, repeat {rep:.2e}")
    assert np.linalg.norm(g1) > 0 and err < 1e-6 and rep < 1e-6


@pytest.mark.parametrize("prec,half", [(1, "1"), (2, "1"), (2, "0")])
def test_packed_pass_matches_the_32_wide_pass(prec, half, monkeypatch):
    """B <= 16: the packed pass (x_hi and x_lo of 16 trajectories in the 32
    MFMA columns, halves added at the flush) against the 32-wide pass
    (DMD_BATCH_PACK=0) on the same stored P, for FP16 and hi + lo records; B = 16
    and B = 5 on one handle (buffers sized for 16 serve 5), then B = 40 (the
    32-wide pass) after them."""
    scatt, oracle = _mods()
    print(f"case batched packed prec={prec} half={half}", flush=True)
    monkeypatch.setenv("DMD_BATCH_HALF", half)
    rng = np.random.default_rng(61 + prec)
    nk, nb = 230, 4
    k1, k2 = _hub_pairs(nk, 3, rng)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    f = np.sort(rng.uniform(0.05, 0.6, (nk, nb)), axis=1)[:, ::-1].copy()
    rho = _states(40, nk, nb, f, rng)
    t = 41 * FS
    rho_f = np.zeros((nk, nb * nb), np.complex128)
    rho_f[:, ::nb + 1] = f
    out = {}
    for pack in ("1", "0"):
        monkeypatch.setenv("DMD_BATCH_PACK", pack)
        with scatt.ScatteringOperator(nk, nb, k1, k2, e, precision=prec) as op:
            op.upload_P(P1, P2)
            op.set_split(f)
            base = op.eval(t, rho_f)
            out[pack] = [op.eval_batched(t, rho[:B]) - base[None] for B in (16, 5, 40)]
    for B, a, g in zip((16, 5, 40), out["1"], out["0"]):
        err = rel_err(a, g)
        print(f"prec={prec} half={half} B={B}: packed vs 32-wide {err:.2e}")
        assert np.linalg.norm(g) > 0 and err < 1e-6


@pytest.mark.parametrize("prec", [1, 2])
def test_batched_with_no_stored_pairs(prec):
    """An operator with no stored pairs: no visits, no segments, an empty
    queue for the persistent pass; every trajectory's ddm is zero, as the
    unbatched evaluation's is."""
    scatt, oracle = _mods()
    print(f"case batched no pairs prec={prec}", flush=True)
    rng = np.random.default_rng(3)
    nk, nb, B = 12, 4, 5
    e = random_energies(nk, nb, rng)
    f = np.sort(rng.uniform(0.05, 0.6, (nk, nb)), axis=1)[:, ::-1].copy()
    rho = _states(B, nk, nb, f, rng)
    empty = np.zeros(0, np.uint64)
    with scatt.ScatteringOperator(nk, nb, empty, empty, e, precision=prec) as op:
        op.upload_P(np.zeros((0, nb ** 4), np.complex128), np.zeros((0, nb ** 4), np.complex128))
        op.set_split(f)
        got = op.eval_batched(13 * FS, rho)
        one = op.eval(13 * FS, rho[0])
    assert got.shape == (B, nk, nb * nb) and np.all(got == 0) and np.all(one == 0)
