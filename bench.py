"""Benchmark of the Lindblad scattering hot path on one MI355X.

A step is one fixed-step RK4 step (4 evaluations of dRho/dt|scatt over every
stored k-pair, plus the stage combinations) of the on-device integrator, at
the one-GPU workload: nb = 2, nk = 65536, ~900 partners drawn per k
(~59 M stored pairs, 30.7 GB of complex128 P resident in HBM), T = 300 K,
dt = 1 fs, FP64 throughout, synthetic ldbd_data generated from seed 2 on the
device.  value = stored k-pair contractions per second (pairs x 4 x steps /
time), whole job.  Also reported: HBM GB/s of the contraction kernel against
the MI355X peak (roofline), trajectory-steps/s, and the CPU oracle timed on
this host's cores on a bounded sample of the same workload.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--preset 2] [--no-cpu]
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
    (2, "tiles"): (31.697e9, "profiles/k_tiles_soa_nb2_nk65536_58Mpairs.txt"),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def initial_state(p, e, s3):
    """rho0 = f + (df/de) b S_z (first order, electron.cpp:62-98) with b set so
    that Tr(S_z drho) / sum Tr f = 1e-3; mu = CBM - 0.1 eV."""
    nk, nb = e.shape
    T = p.temperature
    mu = -0.1 / 27.21138505
    f = 1.0 / (np.exp((e - mu) / T) + 1.0)
    sz = s3[:, 2].reshape(nk, nb, nb)
    de = e[:, :, None] - e[:, None, :]
    fd = f[:, :, None] - f[:, None, :]
    deg = np.abs(de) < 1e-8
    favg = 0.5 * (f[:, :, None] + f[:, None, :])
    dfde = np.where(deg, favg * (favg - 1) / T, fd / np.where(deg, 1.0, de))
    d1 = dfde * sz  # drho for b = 1
    s_per_b = np.real(np.einsum("kij,kji->", sz, d1))
    b = 1e-3 * f.sum() / s_per_b
    rho = (np.einsum("ki,ij->kij", f, np.eye(nb)) + b * d1).reshape(nk, nb * nb)
    return rho.astype(np.complex128), f


def cpu_baseline(p, e, k1, k2, rho, budget_s=12.0):
    """Oracle (oracle/liboracle.so, OpenMP over pairs) on the first rows of the
    same workload; returns pairs/s, threads used and the sample description."""
    from denmat_dynm_development_amd import scatt
    from oracle import oracle
    nthreads = max(1, min(os.cpu_count() or 1, 16))
    n = min(len(k1), 2_000_000)
    # the first n stored pairs close over rows [0, k1[n-1]]
    P1, P2 = scatt.synth_P(p, e, k1[:n], k2[:n], nthreads)
    prefac = 2 * np.pi / p.nk
    t0 = time.perf_counter()
    evals = 0
    while True:
        oracle.evolve(k1[:n], k2[:n], P1, P2, e, rho, t=evals * FS, prefac=prefac, nthreads=nthreads)
        evals += 1
        el = time.perf_counter() - t0
        if el > budget_s or evals >= 50:
            break
    rate = n * evals / el
    return {"value": rate, "unit": "pairs/s", "cores": nthreads, "kind": "port",
            "sample": f"{evals} RHS evaluations over the first {n} stored pairs of the workload "
                      f"(oracle evolve, per-element phases, OpenMP over pairs), {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--preset", type=int, default=2)
    ap.add_argument("--kernel", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a plain 1 GiB device copy (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the scattering operator runs on a single MI355X here "
              "(multi-GPU sharding is described in DESIGN.md, not launched)", file=sys.stderr)
        sys.exit(2)

    import torch
    from denmat_dynm_development_amd import scatt

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
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
    log(f"P generated on device in {time.perf_counter() - t1:.1f}s; create+generate {time.perf_counter() - t0:.1f}s")
    rho_d = torch.from_numpy(rho0).cuda()
    dt = 1.0 * FS
    n_pairs = len(k1)
    # warmup (also exercises graph capture)
    if args.warmup > 0:
        op.integrate_rk4(rho_d, 0.0, dt, args.warmup)
    torch.cuda.synchronize()
    log("warmup done")
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
    alg_bytes = n_pairs * (2 * p.nb ** 4 * 16 + 4) + 2 * p.nk * p.nb ** 2 * 16
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    log(f"contraction kernel {kernel_ms:.3f} ms/launch ({eval_ms:.3f} ms with the rotation), "
        f"{achieved:.0f} GB/s algorithmic")
    if args.kernel == 1 or p.nb > 2:
        kind, kname = "rows", f"k_rows<{p.nb}>"
    elif args.kernel == 2 or p.nb == 1:
        kind, kname = "single_read", f"k_fwd_atomic<{p.nb}> (+k_assemble)"
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
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": traffic[0] if traffic else None,
                     "traffic_source": traffic[1] if traffic else None,
                     "kernel": kname,
                     "kernel_ms": kernel_ms, "alg_bytes_per_launch": alg_bytes},
    }
    if args.copy_calib:
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)  # 1 GiB
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 1.0, out=dst)  # one elementwise kernel: reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        log("copy calibration: 1 GiB read + 1 GiB written by one elementwise kernel")
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(p, e, k1, k2, rho0)
        log(f"cpu baseline {out['cpu_baseline']['value']:.3e} pairs/s on {out['cpu_baseline']['cores']} threads")
    op.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
