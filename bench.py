"""Benchmark of the Lindblad scattering hot path on one MI355X.

Default mode (the contract line): a step is one fixed-step RK4 step (4
evaluations of dRho/dt|scatt over every stored k-pair, plus the stage
combinations) of the on-device integrator at the one-GPU FP64 workload:
nb = 2, nk = 65536, ~900 partners drawn per k (58.2 M stored pairs, 29.8 GB
of complex128 P resident in HBM), T = 300 K, dt = 1 fs, synthetic ldbd_data
generated from seed 2 on the device.  value = stored k-pair contractions per
second (pairs x 4 x steps / time), whole job.  Also reported: the
contraction kernel's HBM GB/s against the MI355X peak (roofline),
trajectory-steps/s, the CPU oracle timed on this host's cores on a bounded
sample of the same workload, and (mixed_batched) trajectory-steps/s of the
batched reduced-precision forms against the FP64 kernel at the nb = 4
power-law workload (nk = 32768, 10 .. 5000 partners per k), with the fitted
spin lifetime of the same runs compared.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batched]
  python bench.py --mode accuracy   # FP64 / complex64 / BF16 S_z(t) and tau_s, nb = 4 with e-i
  python bench.py --mode adjoint    # forward + adjoint RK4 with 7 parameters, nb = 2, nk = 16384
  python bench.py --mode linear     # the linearized operator (set_Lsc + evolve_linear), nb = 2 and nb = 4
  python bench.py --mode ode        # the adaptive rkf45 driver and compute_rhs against fixed-step RK4, nb = 2
  python bench.py --mode eimp       # model e-i P (Debye, static RPA) and e-e P generated on the device, nb = 2
  python bench.py --mode sparse     # compressed P against the dense store at several densities, nb = 2
  python bench.py --mode panels     # the nb = 4 panel kernel alone in FP64, complex64 and BF16
  python bench.py --mode batched [--B 3,16,64] [--batched-forms c64,bf16]  # the batched forms alone
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FS = 41.341373335182
HBM_PEAK_GBPS = 8000.0
# HBM bytes per launch of the dominant kernel, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE runs of this script (calibrated against a 1 GiB elementwise kernel:
# FETCH_SIZE x 2.0, WRITE_SIZE x 1.0); key (preset, kernel kind) -> (bytes, profile file)
MEASURED_TRAFFIC = {
    (2, "tiles"): (31.402e9, "profiles/k_tiles_soa_nb2_nk65536_58Mpairs.txt"),
}
PREC_NAME = {0: "fp64", 1: "c64", 2: "bf16"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def initial_state(p, e, s3, axis=2, amp=1e-3):
    from denmat_dynm_development_amd import scatt
    return scatt.synth_initial_state(p, e, s3, axis, amp)


def batch_states(p, e, s3, B):
    from denmat_dynm_development_amd import scatt
    return scatt.synth_batch_states(p, e, s3, B)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, e, k1, k2, rho, chunk=2_000_000):
    """The oracle (oracle/liboracle.so: the reference's evolve loop with
    per-element compute_Pt phases, OpenMP over pairs) timed on this host.

    - All threads: ONE full right-hand-side evaluation of the workload, every
      stored pair, streamed in chunks of `chunk` pairs (the operator is a sum
      over pairs, so the chunks' times add up to one evaluation's); P of each
      chunk is generated untimed.  The per-evaluation rate is extrapolated to
      the run's 4 x steps evaluations, and labelled so.
    - One core: 3 evaluations of the first 1 M stored pairs.
    Threads: the OpenMP default of this process (OMP_NUM_THREADS, which is
    this host's CPU share per GPU), next to nproc and the CPU model."""
    from denmat_dynm_development_amd import scatt
    from oracle import oracle
    nthreads = max(1, oracle.max_threads())
    prefac = 2 * np.pi / p.nk
    t_eval = 0.0
    n = len(k1)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        P1, P2 = scatt.synth_P(p, e, k1[a:b], k2[a:b], nthreads)
        t0 = time.perf_counter()
        oracle.evolve(k1[a:b], k2[a:b], P1, P2, e, rho, t=17 * FS, prefac=prefac, nthreads=nthreads)
        t_eval += time.perf_counter() - t0
        del P1, P2
    n1 = min(n, 1_000_000)
    P1, P2 = scatt.synth_P(p, e, k1[:n1], k2[:n1], nthreads)
    t0 = time.perf_counter()
    for r in range(3):
        oracle.evolve(k1[:n1], k2[:n1], P1, P2, e, rho, t=(17 + r) * FS, prefac=prefac, nthreads=1)
    t_one = time.perf_counter() - t0
    return {"value": n / t_eval, "unit": "pairs/s", "cores": nthreads, "kind": "port",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "one_core": {"value": 3 * n1 / t_one, "unit": "pairs/s", "cores": 1,
                         "sample": f"3 evaluations of the first {n1} stored pairs, {t_one:.1f} s"},
            "eval_s": t_eval,
            "sample": f"one full right-hand-side evaluation of the workload ({n} stored pairs, in chunks of "
                      f"{chunk}) on {nthreads} threads, {t_eval:.2f} s; oracle evolve with per-element phases",
            "extrapolated": "the rate of one full evaluation stands for the run's evaluations (not timed in full)"}


def alg_bytes(n_pairs, nk, nb, esize=16, extra=0):
    """algorithmic bytes of one evaluation: P once + 4-byte index per stored
    pair (+ extra per pair), rho in and ddmdt out"""
    return n_pairs * (2 * nb ** 4 * esize + 4 + extra) + 2 * nk * nb ** 2 * 16


def bench_plain(args, torch, scatt):
    p = scatt.synth_params(args.preset)
    t0 = time.perf_counter()
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    s3 = scatt.synth_spin(p)
    rho0, f = initial_state(p, e, s3)
    log(f"preset {args.preset}: nb={p.nb} nk={p.nk} pairs={len(k1)} (host setup {time.perf_counter() - t0:.1f}s)")
    op = scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, nk_full=p.nk)
    op.set_flags(expt=True, kernel=args.kernel)
    t1 = time.perf_counter()
    op.generate_P(p)
    log(f"P generated on device in {time.perf_counter() - t1:.1f}s")
    rho_d = torch.from_numpy(rho0).cuda()
    dt = 1.0 * FS
    n_pairs = len(k1)
    if args.warmup > 0:
        op.integrate_rk4(rho_d, 0.0, dt, args.warmup)  # also exercises graph capture
    torch.cuda.synchronize()
    ts = time.perf_counter()
    rho_out, _ = op.integrate_rk4(rho_d, 0.0, dt, args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - ts
    assert torch.isfinite(torch.view_as_real(rho_out)).all()
    ms_per_step = el / args.steps * 1e3
    value = n_pairs * 4 * args.steps / el
    log(f"timed {args.steps} RK4 steps: {el:.3f}s, {ms_per_step:.3f} ms/step")
    # dominant kernel: the contraction, timed with HIP events on the handle's stream
    op.eval(0.0, rho_d)
    kernel_ms, eval_ms = op.time_eval(17 * FS, 10)
    ab = alg_bytes(n_pairs, p.nk, p.nb)
    achieved = ab / (kernel_ms * 1e-3) / 1e9
    # the same evaluation with host buffers (PCIe in and out each call)
    th = time.perf_counter()
    for _ in range(5):
        op.eval(0.0, rho0)
    pcie_ms = (time.perf_counter() - th) / 5 * 1e3
    log(f"contraction kernel {kernel_ms:.3f} ms/launch ({eval_ms:.3f} ms with the rotation), "
        f"{achieved:.0f} GB/s algorithmic; host-buffer eval {pcie_ms:.3f} ms")
    if args.kernel == 1 or p.nb == 3:
        kind, kname = "rows", f"k_rows<{p.nb}>"
    elif args.kernel == 2 or p.nb == 1:
        kind, kname = "single_read", f"k_fwd_atomic<{p.nb}> (+k_assemble)"
    elif p.nb == 4:
        kind, kname = "tiles", "k_panels4<fp64> (+k_assemble_h4)"
    else:
        kind, kname = "tiles", "k_tiles_soa<false> (+k_assemble_h)"
    traffic = MEASURED_TRAFFIC.get((args.preset, kind))
    out = {
        "metric": "k-pair contractions/s",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain FP64 Lindblad e-ph, preset {args.preset}: nb={p.nb} nk={p.nk} "
                               f"stored_pairs={n_pairs} RK4 dt=1fs",
                   "nb": p.nb, "nk": p.nk, "stored_pairs": n_pairs, "precision": "fp64",
                   "kernel": kind},
        "hbm_GBps": achieved,
        "traj_steps_per_s": args.steps / el,
        "eval_ms_host_buffers": pcie_ms,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": traffic[0] if traffic else None,
                     "traffic_source": traffic[1] if traffic else None,
                     "kernel": kname,
                     "kernel_ms": kernel_ms, "alg_bytes_per_launch": ab},
    }
    if args.copy_calib:
        copy_calibration(torch)
    op.close()
    return out, (p, e, k1, k2, rho0)


def copy_calibration(torch):
    """one elementwise kernel of known traffic for scaling the HBM counters"""
    src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)  # 1 GiB
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    torch.mul(src, 1.0, out=dst)  # reads 1 GiB, writes 1 GiB
    torch.cuda.synchronize()
    log("copy calibration: 1 GiB read + 1 GiB written by one elementwise kernel")
    del src, dst


def run_traj(op, torch, rho, dt, steps, batched):
    """timed RK4 steps with the observable every step; (seconds, S series)"""
    torch.cuda.synchronize()
    t = time.perf_counter()
    if batched:
        _, obs = op.integrate_rk4_batched(rho, 0.0, dt, steps, report_every=1)
    else:
        _, obs = op.integrate_rk4(rho, 0.0, dt, steps, report_every=1)
    torch.cuda.synchronize()
    return time.perf_counter() - t, obs


def bench_mixed_batched(args, torch, scatt):
    """FP64 tile kernel (one trajectory) against the batched reduced-precision
    forms (B trajectories per pass over P) at the nb = 4 power-law workload."""
    from denmat_dynm_development_amd import taufit
    p = scatt.synth_params(args.batched_preset)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    s3 = scatt.synth_spin(p)
    Bs = [int(x) for x in args.B.split(",")]
    states, f = batch_states(p, e, s3, max(Bs))
    n_pairs = len(k1)
    dt, K = 1.0 * FS, args.batched_steps
    times = np.arange(K + 1) * dt
    res = {"workload": f"preset {args.batched_preset}: nb={p.nb} nk={p.nk} stored_pairs={n_pairs} "
                       f"(partners per k power law {p.partners_min}..{p.partners_max}), RK4 dt=1fs, {K} steps",
           "stored_pairs": n_pairs}
    log(f"batched workload: nb={p.nb} nk={p.nk} pairs={n_pairs}")
    # FP64, one trajectory
    op = scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, nk_full=p.nk)
    op.generate_P(p)
    op.set_observable(s3[:, 2], f, 0, p.nb, 1.0 / p.nk)
    run_traj(op, torch, states[0], dt, 1, False)
    el, obs = run_traj(op, torch, states[0], dt, K, False)
    tau64 = taufit.fit_tau(times, obs)
    op.eval(0.0, states[0])
    kms, _ = op.time_eval(0.0, 3)
    ab = alg_bytes(n_pairs, p.nk, p.nb)
    res["fp64"] = {"traj_steps_per_s": K / el, "ms_per_step": el / K * 1e3, "tau_s_ps": tau64,
                   "kernel": "k_panels4<fp64>", "kernel_ms": kms, "hbm_GBps": ab / kms / 1e6,
                   "roofline_frac": ab / kms / 1e6 / HBM_PEAK_GBPS}
    log(f"fp64: {el / K * 1e3:.1f} ms/step, kernel {kms:.2f} ms ({ab / kms / 1e6:.0f} GB/s), tau {tau64:.6g} ps")
    op.close()
    base = K / el
    # the batched pass stores every visit's matrices as BF16 hi + lo words for the
    # complex64 store (P rounded to a 16-bit significand) and as one FP16 word for
    # the BF16 store (DMD_BATCH_HALF=0: hi + lo words for it too)
    half = os.environ.get("DMD_BATCH_HALF", "1") != "0"
    forms = {"c64": 1, "bf16": 2}
    for prec in [forms[x] for x in args.batched_forms.split(",")]:
        name = {1: "c64_P_as_bf16_hi_lo", 2: "bf16"}[prec]
        try:
            op = scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, nk_full=p.nk, precision=prec)
            op.generate_P(p)
            op.set_split(f)
            op.set_observable(s3[:, 2], f, 0, p.nb, 1.0 / p.nk)
            per = {}
            for B in Bs:
                run_traj(op, torch, states[:B], dt, 1, True)  # builds visit records, tunes, captures
                el, obs = run_traj(op, torch, states[:B], dt, K, True)
                tau = taufit.fit_tau(times, obs[0])
                op.eval_batched(0.0, states[:B])
                kms = op.time_eval_batched(0.0, 3)
                # matrix-core work of one evaluation: every stored pair is two visits (k < k')
                # or one (k = k'); per visit and 32 trajectories 6 v_mfma_f32_32x32x16_bf16
                # (hi + lo records) or 4 v_mfma_f32_32x32x16_f16 (FP16 records); for B <= 16
                # the packed pass forms x_hi and x_lo of 16 trajectories in one MFMA: 4 or 2
                n_visits = 2 * n_pairs - int(np.count_nonzero(k1 == k2))
                n_mfma = 4 if (prec == 2 and half) else 6
                packed = B <= 16 and os.environ.get("DMD_BATCH_PACK", "1") != "0"
                n_mfma_run = n_mfma - 2 if packed else n_mfma
                mfma_flop = n_visits * (1 if packed else (B + 31) // 32) * n_mfma_run * 2 * 32 * 32 * 16
                # the pass is bound by its record stream: 2 KiB (FP16) or 4 KiB (hi + lo) per visit
                rec_bytes = n_visits * (2048 if n_mfma == 4 else 4096)
                per[str(B)] = {"traj_steps_per_s": B * K / el, "ms_per_step": el / K * 1e3,
                               "speedup_vs_fp64": B * K / el / base, "eval_ms": kms,
                               "mfma_tflops": mfma_flop / (kms * 1e-3) / 1e12,
                               "mfma_frac_of_dense_peak": mfma_flop / (kms * 1e-3) / 2.5e15,
                               "records": "fp16" if n_mfma == 4 else "bf16_hi_lo",
                               "record_GBps": rec_bytes / (kms * 1e-3) / 1e9,
                               "record_frac_of_hbm_peak": rec_bytes / (kms * 1e-3) / 1e9 / HBM_PEAK_GBPS,
                               "pair_traj_per_s": n_pairs * B / (kms * 1e-3),
                               "tau_s_ps": tau, "tau_rel_err_vs_fp64": abs(tau / tau64 - 1)}
                log(f"{name} B={B}: {el / K * 1e3:.1f} ms/step, {B * K / el:.1f} traj-steps/s "
                    f"({B * K / el / base:.1f}x fp64), eval {kms:.2f} ms, |dtau|/tau {abs(tau / tau64 - 1):.2e}")
            res[name] = per
            op.close()
        except Exception as ex:  # e.g. not enough HBM for both stores of a form
            res[name] = {"error": str(ex)}
            log(f"{name}: {ex}")
    return res


def mode_accuracy(args, torch, scatt):
    """S_z(t) and tau_s of FP64, complex64 and BF16 (unbatched and batched) on
    the nb = 4 workload with the e-i term."""
    from denmat_dynm_development_amd import taufit
    p = scatt.synth_params(3)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    s3 = scatt.synth_spin(p)
    states, f = batch_states(p, e, s3, 3)
    dt, K = 1.0 * FS, args.steps
    times = np.arange(K + 1) * dt
    out = {"mode": "accuracy", "workload": f"preset 3: nb=4 nk={p.nk} stored_pairs={len(k1)} with e-i, "
                                           f"RK4 dt=1fs {K} steps", "results": {}}
    ref = None
    for prec in (0, 1, 2):
        op = scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, nk_full=p.nk, precision=prec)
        op.generate_P(p)
        op.set_split(f)
        op.set_observable(s3[:, 2], f, 0, p.nb, 1.0 / p.nk)
        # ddm at t = 37 fs on the three initial states, against the FP64 kernel
        # (itself the oracle's to 1e-12 in the parity tests)
        t_ddm = 37 * FS
        ddm = [op.eval(t_ddm, states[b]) for b in range(3)]
        el, obs = run_traj(op, torch, states[0], dt, K, False)
        r = {"tau_s_ps": taufit.fit_tau(times, obs), "ms_per_step": el / K * 1e3}
        if prec == 0:
            ref = obs, r["tau_s_ps"], ddm
        else:
            dd = np.stack(ddm) - np.stack(ref[2])
            r["ddm_max_rel"] = float(np.abs(dd).max() / np.abs(np.stack(ref[2])).max())
            r["ddm_rms_rel"] = float(np.linalg.norm(dd) / np.linalg.norm(np.stack(ref[2])))
            r["tau_rel_err"] = abs(r["tau_s_ps"] / ref[1] - 1)
            r["S_max_abs_err_rel"] = float(np.abs(obs - ref[0]).max() / np.abs(ref[0]).max())
            db = op.eval_batched(t_ddm, states) - np.stack(ref[2])
            r["batched_B3_ddm_max_rel"] = float(np.abs(db).max() / np.abs(np.stack(ref[2])).max())
            r["batched_B3_ddm_rms_rel"] = float(np.linalg.norm(db) / np.linalg.norm(np.stack(ref[2])))
            elb, ob = run_traj(op, torch, states, dt, K, True)
            r["batched_B3_tau_rel_err"] = abs(taufit.fit_tau(times, ob[0]) / ref[1] - 1)
            r["batched_B3_S_max_abs_err_rel"] = float(np.abs(ob[0] - ref[0]).max() / np.abs(ref[0]).max())
        out["results"][PREC_NAME[prec]] = r
        log(f"{PREC_NAME[prec]}: {r}")
        op.close()
    return out


def mode_adjoint(args, torch, scatt):
    """Forward RK4 and the checkpointed adjoint with the six branch scales and
    the impurity density as parameters, nb = 2, nk = 16384; the gradient of
    tau_s checked against central differences (h = 1e-6) of the forward."""
    from denmat_dynm_development_amd.adjoint import SpinLifetimeProblem
    p = scatt.synth_params(5)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    s3 = scatt.synth_spin(p)
    rho0, f = initial_state(p, e, s3)
    masks = [1 << a for a in range(p.n_modes)] + [1 << 31]
    params = []
    for m in masks:
        q = scatt.synth_params(5)
        q.branch_mask = m
        params.append(q)
    K = args.steps
    # one segment (every stage state kept, 4 x 1 MiB per step at nk = 16384)
    # unless a checkpoint interval is given
    C = args.checkpoint_every if args.checkpoint_every > 0 else K
    pr = SpinLifetimeProblem.from_device_generator(
        params, p.nk, p.nb, k1, k2, e, rho0=rho0, s_obs=s3[:, 2], f=f, i0=0, i1=p.nb, obs_prefac=1.0 / p.nk,
        t0=0.0, dt=1.0 * FS, n_steps=K, report_every=1, checkpoint_every=C, nk_full=p.nk)
    theta = np.ones(len(masks))
    try:
        pr.tau(theta)  # warm (graph capture)
        t = time.perf_counter()
        tau = pr.tau(theta)
        t_fwd = time.perf_counter() - t
        t = time.perf_counter()
        tau2, grad = pr.tau_and_grad(theta)
        t_adj = time.perf_counter() - t - t_fwd
        fd = []
        for c in range(len(theta) if args.fd_all else 2):
            tp, tm = theta.copy(), theta.copy()
            tp[c] += 1e-6
            tm[c] -= 1e-6
            fd.append((pr.tau(tp) - pr.tau(tm)) / 2e-6)
        rel = [abs(grad[c] / fd[c] - 1) for c in range(len(fd))]
        gnorm = float(np.linalg.norm(fd))
        rel_norm = float(max(abs(grad[c] - fd[c]) for c in range(len(fd))) / gnorm)
    finally:
        pr.close()
    fwd_ms, adj_ms = t_fwd / K * 1e3, t_adj / K * 1e3
    # HBM streams of one adjoint step: the forward sweep's 4 evaluations and
    # the reverse sweep's 4 one-pass VJPs read P (32 complex128 per pair), the
    # parameter fold reads and writes Gc (same size) once per 4 steps; segments
    # that are recomputed add 4 evaluations
    p_bytes = len(k1) * 32 * 16
    streams = 4 + 4 + 0.5 + (4 if C < K else 0)
    out = {"mode": "adjoint", "workload": f"preset 5: nb=2 nk={p.nk} stored_pairs={len(k1)}, 6 branches + e-i, "
                                          f"RK4 dt=1fs {K} steps, checkpoints every {C}",
           "tau_s_ps": tau, "dtau_dtheta": list(map(float, grad)), "fd_central_h1e-6": fd,
           "adjoint_rel_err_vs_fd": rel, "adjoint_max_err_over_fd_norm": rel_norm,
           "forward_s": t_fwd, "adjoint_s": t_adj, "forward_ms_per_step": fwd_ms,
           "adjoint_ms_per_step": adj_ms, "adjoint_over_forward": adj_ms / fwd_ms,
           "roofline": {"bound": "hbm", "achieved": streams * p_bytes / (adj_ms * 1e-3) / 1e9, "peak": 8000.0,
                        "unit": "GB/s", "frac": streams * p_bytes / (adj_ms * 1e-3) / 8e12,
                        "traffic": None, "streams_per_step": streams, "p_bytes": p_bytes}}
    log(json.dumps(out))
    return out


def mode_sparse(args, torch, scatt):
    """The sparse modes' store (dmd_scatt_compress_P: per pair a nonzero mask,
    an offset and the nonzero values) on the plain nb = 2 workload: P kept at
    about 100, 50 and 25 % of its elements (alg_sparseP's threshold, chosen
    from the magnitudes of a sample of pairs), bytes per pair and contraction
    time of the compressed store against the dense one, and the compressed
    evaluation against the dense store holding the same thresholded P."""
    p = scatt.synth_params(args.preset)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    s3 = scatt.synth_spin(p)
    rho0, f = initial_state(p, e, s3)
    n_pairs = len(k1)
    rho_d = torch.from_numpy(rho0).cuda()
    op = scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, nk_full=p.nk)
    op.generate_P(p)
    S1, S2 = op.download_P(0, min(n_pairs, 20000))
    mags = np.abs(np.concatenate([S1.ravel(), S2.ravel()]))
    op.eval(0.0, rho_d)
    dense_ms, _ = op.time_eval(17 * FS, 10)
    dense_bytes = op.info()["p_bytes"]
    op.close()
    rows = []
    for keep in (1.0, 0.5, 0.25):
        thr = 0.0 if keep == 1.0 else float(np.quantile(mags, 1.0 - keep))
        ref = scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, nk_full=p.nk)
        ref.generate_P(p)
        if thr > 0:
            ref.threshold_P(thr)
        want = ref.eval(17 * FS, rho_d).cpu().numpy()
        ref.close()
        op = scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, nk_full=p.nk)
        op.generate_P(p)
        if thr > 0:
            op.threshold_P(thr)
        stored = op.compress_P()
        got = op.eval(17 * FS, rho_d).cpu().numpy()
        kernel_ms, _ = op.time_eval(17 * FS, 10)
        nbytes = op.info()["p_bytes"]
        op.close()
        err = float(np.abs(got - want).max() / max(np.abs(want).max(), 1e-300))
        r = {"kept_elements": stored / (n_pairs * 32), "threshold": thr, "compressed": nbytes != dense_bytes,
             "p_bytes": nbytes,
             "bytes_per_pair": nbytes / n_pairs, "kernel_ms": kernel_ms,
             "GBps": nbytes / (kernel_ms * 1e-3) / 1e9, "vs_dense_time": kernel_ms / dense_ms,
             "vs_dense_bytes": nbytes / dense_bytes, "max_rel_err_vs_dense_store": err}
        log(f"sparse keep~{keep}: {r}")
        rows.append(r)
    return {"mode": "sparse", "workload": f"preset {args.preset}: nb={p.nb} nk={p.nk} stored_pairs={n_pairs}",
            "dense": {"p_bytes": dense_bytes, "bytes_per_pair": dense_bytes / n_pairs, "kernel_ms": dense_ms,
                      "GBps": dense_bytes / (dense_ms * 1e-3) / 1e9},
            "compressed": rows}


def mode_panels(args, torch, scatt):
    """The nb = 4 panel kernel alone (one trajectory, dmd_scatt_time_eval: HIP
    events over 10 launches) in FP64, complex64 and BF16 on one preset, with
    the algorithmic bytes of each store (the counters of a profiled run of
    this mode back the per-precision bounds)."""
    p = scatt.synth_params(args.preset)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    s3 = scatt.synth_spin(p)
    rho0, f = initial_state(p, e, s3)
    n_pairs = len(k1)
    out = {"mode": "panels", "workload": f"preset {args.preset}: nb={p.nb} nk={p.nk} stored_pairs={n_pairs}",
           "results": {}}
    esize = {0: 16, 1: 8, 2: 4}
    for prec in (0, 1, 2):
        op = scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, nk_full=p.nk, precision=prec)
        op.generate_P(p)
        if prec:
            op.set_split(f)
        rho_d = torch.from_numpy(rho0).cuda()
        op.eval(17 * FS, rho_d)
        kms, ems = op.time_eval(17 * FS, 10)
        ab = alg_bytes(n_pairs, p.nk, p.nb, esize=esize[prec], extra=4 if prec == 2 else 0)
        r = {"kernel_ms": kms, "eval_ms": ems, "alg_bytes": ab, "GBps": ab / kms / 1e6,
             "roofline_frac": ab / kms / 1e6 / HBM_PEAK_GBPS}
        out["results"][PREC_NAME[prec]] = r
        log(f"{PREC_NAME[prec]}: {r}")
        op.close()
    return out


def mode_linear(args, torch, scatt):
    """The linearized operator (alg_linearize): set_Lsc at the Fermi occupations,
    then evolve_linear timed per evaluation and per RK4 step, FP64, at the
    default nb = 2 workload and the nb = 4 power-law workload."""
    out = {"mode": "linear", "results": {}}
    for preset in (2, 4):
        p = scatt.synth_params(preset)
        e = scatt.synth_energies(p)
        k1, k2 = scatt.synth_kpairs(p)
        s3 = scatt.synth_spin(p)
        rho0, f = initial_state(p, e, s3)
        op = scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, nk_full=p.nk)
        op.generate_P(p)
        t0 = time.perf_counter()
        op.set_Lsc(f)
        t_lsc = time.perf_counter() - t0
        rho_d = torch.from_numpy(rho0).cuda()
        op.eval(0.0, rho_d)
        kms, ems = op.time_eval(17 * FS, 5)
        ab = alg_bytes(len(k1), p.nk, p.nb)
        op.integrate_rk4(rho_d, 0.0, FS, 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.integrate_rk4(rho_d, 0.0, FS, args.steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        kname = {2: "k_tiles_soa<LIN>", 4: "k_panels4<LIN>"}.get(p.nb, "k_lin_rows")
        r = {"nb": p.nb, "nk": p.nk, "stored_pairs": len(k1), "set_Lsc_s": t_lsc, "kernel": kname,
             "kernel_ms": kms, "hbm_GBps": ab / kms / 1e6, "roofline_frac": ab / kms / 1e6 / HBM_PEAK_GBPS,
             "ms_per_step": el / args.steps * 1e3, "pairs_per_s": len(k1) * 4 * args.steps / el}
        out["results"][f"preset{preset}"] = r
        log(f"linear preset {preset}: {r}")
        op.close()
    return out


def mode_ode(args, torch, scatt):
    """The reference's default integrator on the device: rkf45 (evolve_gsl,
    GSL's standard control) at the default nb = 2 workload with the
    reference's defaults (tstep 1 fs, ode_hstart 1e-3 fs, ode_hmax = tstep,
    ode_epsabs 1e-8) over --steps tsteps after a first tstep that ramps the step
    up; steps, rejections and right-hand sides per tstep, ms per right-hand side
    against the fixed-step RK4's, and one compute_rhs (what report_tau calls)."""
    p = scatt.synth_params(2)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    s3 = scatt.synth_spin(p)
    rho0, f = initial_state(p, e, s3)
    op = scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, nk_full=p.nk)
    op.generate_P(p)
    rho = torch.from_numpy(rho0).cuda()
    t, h = 0.0, 1e-3 * FS
    t0 = time.perf_counter()
    y, t, h, st = op.rkf45_apply(rho, t, FS, h, 0.0, FS, 1e-8)
    torch.cuda.synchronize()
    first = {"s": time.perf_counter() - t0, **st}
    log(f"ode: first tstep {first}")
    n = {"steps": 0, "rejected": 0, "rhs_evals": 0}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y, t, h, st = op.rkf45_apply(y, t, t + FS, h, 0.0, FS, 1e-8)
        for k in n:
            n[k] += st[k]
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    op.integrate_rk4(rho.clone(), 0.0, FS, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op.integrate_rk4(rho.clone(), 0.0, FS, args.steps)
    torch.cuda.synchronize()
    el4 = time.perf_counter() - t0
    out = torch.empty_like(y)
    op.compute_rhs(t, y, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        op.compute_rhs(t, y, out)
    torch.cuda.synchronize()
    rhs_ms = (time.perf_counter() - t0) / 5 * 1e3
    r = {"mode": "ode", "workload": "nb=2 nk=65536 58.2M stored pairs FP64", "tsteps": args.steps,
         "first_tstep": first, "rkf45": {**n, "ms_per_tstep": el / args.steps * 1e3,
                                          "ms_per_rhs": el / max(n["rhs_evals"], 1) * 1e3},
         "rk4": {"ms_per_step": el4 / args.steps * 1e3, "ms_per_rhs": el4 / (4 * args.steps) * 1e3},
         "compute_rhs_ms": rhs_ms, "t_end_fs": t / FS}
    log(f"ode: {r}")
    op.close()
    return r


def mode_eimp(args, torch, scatt):
    """The model ionized-impurity P (add_scatt_contrib("eimp") with
    elecimp_model::calc_P) generated on the device for every stored pair of the
    default workload (nb = 2, nk = 65536 on a 64 x 32 x 32 k mesh, 58 M pairs,
    nb_wannier = 4): Debye screening (closed form per pair) and the static RPA
    screening (a k-mesh sum for every distinct q, then v(q) per pair).  Wall
    time of the whole call; the kernels' own times come from rocprofv3."""
    p = scatt.synth_params(args.preset)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    km = (64, 32, 32)
    assert p.nk == km[0] * km[1] * km[2]
    g = np.stack(np.meshgrid(np.arange(km[0]), np.arange(km[1]), np.arange(km[2]), indexing="ij"), -1).reshape(-1, 3)
    kv = g / np.array(km, dtype=np.float64)
    rng = np.random.default_rng(1)
    nw = 4
    U = (rng.normal(size=(p.nk, nw, p.nb)) + 1j * rng.normal(size=(p.nk, nw, p.nb))) / 2
    T = p.temperature
    f = 1 / (np.exp((e - (e.min() + 0.02 * (e.max() - e.min()))) / T) + 1)
    R = np.array([[0.0, 5.34145, 5.34145], [5.34145, 0.0, 5.34145], [5.34145, 5.34145, 0.0]])
    V = abs(np.linalg.det(R))
    n_free = f.sum() / (p.nk * V)
    out = {"mode": "eimp", "workload": f"preset {args.preset}: nb={p.nb} nk={p.nk} stored_pairs={len(k1)}, "
                                       f"k mesh {km}, nb_wannier={nw}", "results": {}}
    op = scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, nk_full=p.nk)
    for scr in ("debye", "RPA"):
        m = scatt.eimp_model(scr, n_free, T, V, scatt.reciprocal_metric(R), nw, n_ionized=n_free,
                             epsilon_background=12.9, degauss=p.degauss,
                             ethr=1e-3, kmesh=km, nk_full=float(p.nk))
        op.add_eimp_model(k1, k2, U, kv, m, add=False, f=f)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.add_eimp_model(k1, k2, U, kv, m, add=False, f=f)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        wb = len(k1) * 2 * p.nb ** 4 * 16
        r = {"wall_s": el, "pairs_per_s": len(k1) / el, "P_written_GB": wb / 1e9, "P_GBps_wall": wb / el / 1e9}
        out["results"][scr] = r
        log(f"eimp {scr}: {r}")
    op.close()
    # the model e-e P at equilibrium: a sum over k3 per pair (O(pairs x nk x nb^6)); 100,000 pairs of the
    # nb = 2, nk = 4096 workload on a 16 x 16 x 16 mesh
    p1 = scatt.synth_params(1)
    e1 = scatt.synth_energies(p1)
    a1, b1 = scatt.synth_kpairs(p1)
    a1, b1 = a1[:100000], b1[:100000]
    km1 = (16, 16, 16)
    g1 = np.stack(np.meshgrid(*[np.arange(n) for n in km1], indexing="ij"), -1).reshape(-1, 3) / 16.0
    U1 = (rng.normal(size=(p1.nk, nw, p1.nb)) + 1j * rng.normal(size=(p1.nk, nw, p1.nb))) / 2
    f1 = 1 / (np.exp((e1 - (e1.min() + 0.02 * (e1.max() - e1.min()))) / p1.temperature) + 1)
    m = scatt.eimp_model("debye", f1.sum() / (p1.nk * V), p1.temperature, V, scatt.reciprocal_metric(R), nw,
                         n_ionized=1.0, epsilon_background=12.9, degauss=p1.degauss, ethr=1e-3, kmesh=km1,
                         nk_full=float(p1.nk))
    op = scatt.ScatteringOperator(p1.nk, p1.nb, a1, b1, e1, nk_full=p1.nk)
    op.add_ee_model(a1, b1, U1, g1, f1, m, add=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op.add_ee_model(a1, b1, U1, g1, f1, m, add=False)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    terms = len(a1) * p1.nk * p1.nb ** 6
    out["results"]["ee_debye"] = {"workload": f"nb={p1.nb} nk={p1.nk} (16^3 mesh), {len(a1)} pairs", "wall_s": el,
                                  "pairs_per_s": len(a1) / el, "pair_k3_terms_per_s": terms / el}
    log(f"ee debye: {out['results']['ee_debye']}")
    op.close()
    # the Pee_update rebuild (P_ee at the current density matrix) on every stored pair of that workload,
    # against the RK4 step it interleaves with (eeMode = Pee_update rebuilds every freq_update_ee_model steps)
    a1, b1 = scatt.synth_kpairs(p1)
    s3 = scatt.synth_spin(p1)
    rho1, _ = initial_state(p1, e1, s3)
    dm = rho1.copy()
    dm[:, ::p1.nb + 1] = f1  # the window's state: f on the diagonal plus the spin imbalance
    with scatt.ScatteringOperator(p1.nk, p1.nb, a1, b1, e1, nk_full=p1.nk) as op:
        op.generate_P(p1)
        op.add_ee_model_state(a1, b1, U1, g1, f1, m, dm=dm, add=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.add_ee_model_state(a1, b1, U1, g1, f1, m, dm=dm, add=False)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        rho_d = torch.from_numpy(rho1).cuda()
        op.integrate_rk4(rho_d, 0.0, 1.0 * FS, 2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.integrate_rk4(rho_d, 0.0, 1.0 * FS, 20)
        torch.cuda.synchronize()
        step = (time.perf_counter() - t0) / 20
    out["results"]["ee_rebuild"] = {"workload": f"nb={p1.nb} nk={p1.nk} (16^3 mesh), all {len(a1)} stored pairs, "
                                                f"P_ee at a density matrix", "wall_s": el,
                                    "pairs_per_s": len(a1) / el, "rk4_step_ms": step * 1e3,
                                    "rebuild_in_rk4_steps": el / step}
    log(f"ee rebuild: {out['results']['ee_rebuild']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--preset", type=int, default=2)
    ap.add_argument("--kernel", type=int, default=0)
    ap.add_argument("--mode", default="default", choices=["default", "accuracy", "adjoint", "linear", "ode", "eimp", "sparse",
                                                        "panels", "batched"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-preset", type=int, default=4)
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--B", default="3,16,64")
    ap.add_argument("--batched-forms", default="c64,bf16", help="batched stores to time (c64, bf16)")
    ap.add_argument("--fd-all", action="store_true", help="adjoint mode: check every parameter")
    ap.add_argument("--checkpoint-every", type=int, default=0,
                    help="adjoint mode: steps per checkpoint segment (0: one segment, every stage state kept)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a plain 1 GiB device copy (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the scattering operator runs on a single MI355X here "
              "(sharding over several GPUs is a design, not launched by this benchmark)", file=sys.stderr)
        sys.exit(2)

    import torch
    from denmat_dynm_development_amd import scatt

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.mode == "accuracy":
        print(json.dumps(mode_accuracy(args, torch, scatt)), flush=True)
        return
    if args.mode == "adjoint":
        print(json.dumps(mode_adjoint(args, torch, scatt)), flush=True)
        return
    if args.mode == "sparse":
        print(json.dumps(mode_sparse(args, torch, scatt)), flush=True)
        return
    if args.mode == "panels":
        print(json.dumps(mode_panels(args, torch, scatt)), flush=True)
        return
    if args.mode == "linear":
        print(json.dumps(mode_linear(args, torch, scatt)), flush=True)
        return
    if args.mode == "ode":
        print(json.dumps(mode_ode(args, torch, scatt)), flush=True)
        return
    if args.mode == "eimp":
        print(json.dumps(mode_eimp(args, torch, scatt)), flush=True)
        return
    if args.mode == "batched":
        if args.copy_calib:
            copy_calibration(torch)
        print(json.dumps({"mode": "batched", **bench_mixed_batched(args, torch, scatt)}), flush=True)
        return
    out, (p, e, k1, k2, rho0) = bench_plain(args, torch, scatt)
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(p, e, k1, k2, rho0)
        log(f"cpu baseline {out['cpu_baseline']['value']:.3e} pairs/s on {out['cpu_baseline']['cores']} threads")
    if not args.no_batched:
        out["mixed_batched"] = bench_mixed_batched(args, torch, scatt)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
