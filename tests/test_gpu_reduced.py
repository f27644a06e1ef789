"""The nb = 4 tile kernel and the reduced-precision forms (complex64 P, BF16
P with a per-pair scale) through the C-ABI.

- Storage: the P the library holds (download_P) is exactly the quantised
  input (quantise_P) and within the rounding bound of each format.
- Arithmetic: an evaluation equals the FP64 oracle run on the stored
  (quantised) P to FP32-product accuracy; with rho split around diag(f) only
  drho enters FP32 and the error falls by the size of drho.
- Dynamics: RK4 trajectories of the reduced forms follow the FP64 one; the
  spin signal differs by the quantisation of P.
"""
import numpy as np
import pytest

from helpers import FS, quantise_P, random_P, random_energies, random_pairs, random_rho, rel_err

pytestmark = pytest.mark.gpu
PRECS = [1, 2]  # DMD_PREC_C64, DMD_PREC_BF16


def _mods():
    from denmat_dynm_development_amd import scatt
    from oracle import oracle
    return scatt, oracle


def _near_eq_state(nk, nb, rng, amp=1e-3):
    """rho = diag(f) + drho with |drho| ~ amp (the regime of spin relaxation)."""
    f = np.sort(rng.uniform(0.05, 0.6, (nk, nb)), axis=1)[:, ::-1].copy()
    rho = np.zeros((nk, nb, nb), np.complex128)
    rho[:, np.arange(nb), np.arange(nb)] = f
    rho = rho.reshape(nk, nb * nb) + random_rho(nk, nb, rng, scale=amp, diag=0.0)
    return f, rho


@pytest.mark.parametrize("prec", PRECS)
def test_stored_P_is_the_quantised_input(prec):
    scatt, oracle = _mods()
    print(f"case quantise prec={prec}", flush=True)
    rng = np.random.default_rng(5)
    nk, nb = 70, 4
    k1, k2 = random_pairs(nk, 4, rng, diag=True)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    P1[3] *= 1e-30  # a pair far below the others: the per-pair scale keeps its digits
    Q1, Q2 = quantise_P(P1, P2, prec)
    with scatt.ScatteringOperator(nk, nb, k1, k2, e, precision=prec) as op:
        op.upload_P(P1, P2)
        D1, D2 = op.download_P()
    assert np.array_equal(D1, Q1) and np.array_equal(D2, Q2)
    # rounding bound of the format, element-wise relative to the pair's largest component
    u = 2.0 ** -24 if prec == 1 else 2.0 ** -8
    for A, B in ((P1, D1), (P2, D2)):
        mx = np.maximum(np.abs(np.concatenate([P1, P2], 1).real), np.abs(np.concatenate([P1, P2], 1).imag)).max(1)
        err = np.maximum(np.abs((A - B).real), np.abs((A - B).imag)).max(1)
        assert np.all(err <= u * mx * 1.0000001)


@pytest.mark.parametrize("prec", PRECS)
def test_reduced_eval_matches_oracle_on_stored_P(prec):
    scatt, oracle = _mods()
    print(f"case reduced eval prec={prec}", flush=True)
    rng = np.random.default_rng(11 + prec)
    nk, nb = 120, 4
    k1, k2 = random_pairs(nk, 7, rng, diag=True)
    perm = rng.permutation(len(k1))
    k1, k2 = k1[perm], k2[perm]
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    f, rho = _near_eq_state(nk, nb, rng)
    t = 300 * FS
    prefac = 2 * np.pi / nk
    with scatt.ScatteringOperator(nk, nb, k1, k2, e, precision=prec) as op:
        op.upload_P(P1, P2)
        Q1, Q2 = op.download_P()
        ref = oracle.evolve(k1, k2, Q1, Q2, e, rho, t=t, prefac=prefac)
        full = oracle.evolve(k1, k2, P1, P2, e, rho, t=t, prefac=prefac)
        no_split = op.eval(t, rho)
        op.set_split(f)
        split = op.eval(t, rho)
        # the drho part alone: ddm(rho) - ddm(diag f) isolates what FP32 touches
        rho_f = np.zeros_like(rho)
        rho_f[:, ::nb + 1] = f
        ref_d = ref - oracle.evolve(k1, k2, Q1, Q2, e, rho_f, t=t, prefac=prefac)
        got_d = split - op.eval(t, rho_f)
    e_no, e_split, e_d = rel_err(no_split, ref), rel_err(split, ref), rel_err(got_d, ref_d)
    print(f"prec={prec}: no split {e_no:.2e}, split {e_split:.2e}, drho part {e_d:.2e}, "
          f"vs unquantised P {rel_err(split, full):.2e}")
    assert e_no < 2e-6          # FP32 products of the whole rho
    assert e_split < 1e-8       # FP32 products of drho (~1e-3 of rho) only
    assert e_d < 2e-6           # the drho response itself to FP32 accuracy
    assert e_split < e_no / 20


@pytest.mark.parametrize("prec", PRECS)
def test_reduced_generate_matches_host_generator(prec):
    scatt, oracle = _mods()
    p = scatt.synth_params(0)
    p.nb, p.nk, p.partners = 4, 40, 5
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    P1, P2 = scatt.synth_P(p, e, k1, k2)
    with scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, precision=prec) as op:
        op.generate_P(p)
        D1, D2 = op.download_P()
    Q1, Q2 = quantise_P(P1, P2, prec)
    # device and host transcendental functions may differ in the last bits of
    # the FP64 value (elements that cancel to rounding noise differ entirely),
    # which can move a quantised value by one step
    step = 2.0 ** -23 if prec == 1 else 2.0 ** -7
    assert rel_err(D1, Q1) < step and rel_err(D2, Q2) < step


def test_nb4_kernel_switch_and_dm1():
    """FP64 nb = 4: tiles <-> rows relayout keeps P bit-exact; explicit 1-rho on
    the tile kernel."""
    scatt, oracle = _mods()
    rng = np.random.default_rng(23)
    nk, nb = 80, 4
    k1, k2 = random_pairs(nk, 5, rng, diag=True)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    rho = random_rho(nk, nb, rng)
    dm1 = oracle.one_minus(rho, nb) + 1e-3 * random_rho(nk, nb, rng, diag=0.0)
    ref = oracle.evolve(k1, k2, P1, P2, e, rho, dm1, t=50 * FS, prefac=0.3)
    with scatt.ScatteringOperator(nk, nb, k1, k2, e, nk_full=2 * np.pi / 0.3) as op:
        op.upload_P(P1, P2)
        for kernel in (3, 1, 3, 0):
            op.set_flags(kernel=kernel)
            D1, D2 = op.download_P()
            assert np.array_equal(D1, P1) and np.array_equal(D2, P2)
            assert rel_err(op.eval(50 * FS, rho, dm1), ref) < 1e-12


@pytest.mark.parametrize("prec", PRECS)
def test_reduced_rk4_spin_follows_fp64(prec):
    """RK4 with the observable: the reduced forms against FP64 on the same
    inputs, and against FP64 on their own stored P (arithmetic only)."""
    scatt, oracle = _mods()
    print(f"case reduced rk4 prec={prec}", flush=True)
    rng = np.random.default_rng(40 + prec)
    nk, nb = 96, 4
    k1, k2 = random_pairs(nk, 6, rng)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    P1, P2 = 0.2 * P1, 0.2 * P2
    f, rho0 = _near_eq_state(nk, nb, rng)
    s = random_rho(nk, nb, rng, scale=0.5, diag=0.0)
    dt, n = 2 * FS, 40

    def run(precision, Pa, Pb):
        with scatt.ScatteringOperator(nk, nb, k1, k2, e, precision=precision) as op:
            op.upload_P(Pa, Pb)
            op.set_split(f)
            op.set_observable(s, f, 0, nb, 1.0)
            rho, obs = op.integrate_rk4(rho0, 0.0, dt, n, report_every=4)
            return rho, obs, (op.download_P() if precision else None)

    r64, o64, _ = run(0, P1, P2)
    rq, oq, (Q1, Q2) = run(prec, P1, P2)
    rq64, oq64, _ = run(0, Q1, Q2)
    sig = np.abs(o64 - o64[0]).max()  # the change of the signal over the run
    d_arith = np.abs(oq - oq64).max() / sig
    d_total = np.abs(oq - o64).max() / sig
    e_state = rel_err(rq - rho0, rq64 - rho0)  # relative to the change of rho over the run
    print(f"prec={prec}: S(t) change {sig:.3e}; arithmetic {d_arith:.2e}, with quantisation {d_total:.2e}; "
          f"state change {e_state:.2e}")
    assert e_state < 1e-6
    assert d_arith < 1e-6
    assert d_total < (1e-5 if prec == 1 else 5e-3)
