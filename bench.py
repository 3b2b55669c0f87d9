"""Benchmark of the FP64 IMEX step on one MI355X.

Workload: the production configuration (sparse-Chebyshev integration,
n_r_max = 97, m_max = 128 so n_phi = 384, Ekman 1e-5, Ra 3e7, Pr 1,
thermal + compositional with Sc 3, DIRK BPR353, dt 5e-6), seeded random
T and xi perturbations of amplitude 1e-4.  One "step" is one full IMEX
time step (4 BPR353 stages: 3 radial loops, 4 m-loop solves) with all
state resident in HBM.

Prints one JSON line: steps/s, the roofline figure of the dominant kernel
(the m-loop stage kernel k_mphase, timed with HIP events on the library
stream), the ensemble throughput factor of the uneven 64-run collocation
ensemble (batched launches versus the plain step run by run), the cost of
the adjoint (forward and backward sweeps over the plain forward) and a CPU
baseline of the oracle (the C restatement, on all host cores and on one)
timed on a bounded sample of the same workload, and the steps/s of the
finite-difference workload (n_r 257, n_phi 2046, SBDF3).
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from pizza_amd import workloads as W  # noqa: E402
from pizza_amd.runtime import Run  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
PROFILE = os.path.join(HERE, "profiles", "k_mphase_integ_bpr353_nr97_nphi384.md")
COPY_BYTES = 64 << 20  # calibration copy of the counter runs


def mphase_bytes(cfg, nstage_slots=1):
    """Algorithmic HBM bytes of one k_mphase launch (stage 1, explicit part on).

    Per azimuthal wavenumber m, with F = 16 n_r bytes (one complex radial
    column):
      reads : dVs x3, current explicit slot x3, psi/om/up/us (Ekman pumping),
              old x3 and impl(stage 1) x3 for the IMEX right-hand side,
              preconditioners x3 (8 n_r), bordered LU factors of T, xi
              (A4 13 x (n_r-2) + A2, A3 2 x (n_r-2) reals + pivots) and psi
              (A4 25 x (n_r-4) + A2, A3 4 x (n_r-4) complex + pivots), the
              per-m bands RHSI_T (5 n_r reals), RHSI_psi (17 n_r reals) and
              RHSIL_psi (17 n_r complex);
      writes: explicit slot x3, T, xi, hats x3, psi/us/up/om, impl x3.
    """
    N = cfg.n_r_max
    nm = cfg.m_max // cfg.minc + 1
    F = 16 * N
    nb2, nb4 = N - 2, N - 4
    reads = 16 * F + 3 * 8 * N
    reads += 2 * (13 * nb2 * 8 + 4 * nb2 * 8 + 4 * nb2 + 4 * 8 + 8)
    reads += 25 * nb4 * 16 + 8 * nb4 * 16 + 4 * nb4 + 16 * 16 + 16
    reads += 5 * N * 8 + 17 * N * 8 + 17 * N * 16
    writes = 15 * F
    return nm * (reads + writes)


def measured_traffic():
    """HBM bytes per k_mphase launch from the committed counter summary:
    FETCH_SIZE and WRITE_SIZE of the kernel, each scaled by the same counter
    read for a plain device copy of COPY_BYTES (rocprofv3 --pmc runs of
    --pmc-sample).  None when no summary is present."""
    if not os.path.exists(PROFILE):
        return None
    for line in open(PROFILE):
        line = line.strip()
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":", 1)[1]))
    return None


def pmc_sample():
    """Short sample for the counter runs: 5 k_mphase launches of the C2
    workload after two steps, then 3 plain copies of COPY_BYTES."""
    import torch
    cfg = W.c2_production()
    run = Run(cfg)
    W.seed_run(run, cfg)
    run.start()
    run.step(2)
    run.time()
    run.time_kernel(0, 5)
    x = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    run.close()
    print(json.dumps({"pmc_sample": "k_mphase x5 (stage 1, C2)", "copy_bytes": COPY_BYTES, "copies": 3}))


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, seconds=15.0, max_steps=200):
    """Oracle (the C restatement of the reference's step, OpenMP over the
    radii of the radial loop and the modes of the m-loop) steps/s on a
    bounded sample of the same workload, on all the host threads OpenMP is
    given (OMP_NUM_THREADS, else every core), and beside it on one thread.
    The sample stays inside the workload's 200 steps: a run that reaches
    them starts again from the seeded state (set-up and its first step,
    which factorises the matrices, untimed).  Past about 380 steps this
    configuration's Courant limit cuts dt and the run diverges, on the
    oracle as on the GPU."""
    from oracle.oracle import Oracle, threads

    def rate(nthreads, secs):
        threads(nthreads)
        n, busy = 0, 0.0
        while busy < secs:
            o = Oracle(cfg)
            W.seed_run(o, cfg)
            o.start()
            o.step(1)  # includes the first matrix factorisation
            t0 = time.perf_counter()
            for _ in range(max_steps - 1):
                o.step(1)
                n += 1
                if busy + time.perf_counter() - t0 >= secs:
                    break
            busy += time.perf_counter() - t0
        return n, busy

    all_threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    n, dt = rate(all_threads, seconds)
    used = threads()
    n1, dt1 = rate(1, seconds / 2)
    threads(all_threads)
    return {"value": n / dt, "unit": "steps/s", "cores": used, "kind": "port",
            "nproc": os.cpu_count(), "cpu": _cpu_model(),
            "single_thread_value": n1 / dt1,
            "sample": f"{n} steps of the same workload (n_r=97, n_phi=384, BPR353; steps 2-200 of fresh runs), "
                      f"{dt:.1f} s on {used} OpenMP threads; one thread: {n1} steps in {dt1:.1f} s"}


def ensemble_bench(nruns=64, steps=100, warmup=3):
    """Ensemble throughput x: the uneven 64-run collocation ensemble (CNAB2),
    advanced `steps` steps after `warmup` steps, once run by run through
    pz_step (the plain sequential GPU step) and once through one
    pz_ensemble_step; two identical sets of runs, so both time the same
    step window (including any Courant-driven matrix rebuilds)."""
    from pizza_amd.runtime import Ensemble
    cfgs = W.c3_ensemble(nruns)

    def members():
        runs = []
        for k, c in enumerate(cfgs):
            r = Run(c)
            W.seed_member(r, c, k)
            r.start()
            r.step(warmup)
            r.time()
            runs.append(r)
        return runs

    seq = members()
    t0 = time.perf_counter()
    for r in seq:
        r.step(steps)
        r.time()
    t_seq = time.perf_counter() - t0
    for r in seq:
        r.close()
    ens_runs = members()
    ens = Ensemble(ens_runs)
    ens.step(1)                         # autotuning and graph capture
    ens_runs[0].time()
    policy, trial_ms = ens.policy()
    t0 = time.perf_counter()
    ens.step(steps)
    for r in ens_runs:
        r.time()
    t_ens = time.perf_counter() - t0
    ens.close()
    total = nruns * steps
    return {"runs": nruns, "steps_per_run": steps, "x": t_seq / t_ens,
            "ensemble_run_steps_per_s": total / t_ens, "sequential_run_steps_per_s": total / t_seq,
            "tuned_policy": policy, "policy_trial_ms": [float(v) for v in trial_ms],
            "workload": "64 uneven collocation runs, n_r in {33,65,97,129}, n_phi in {96,192,384,510}, "
                        "Ra log-uniform [1e5,1e8], E in {1e-4,1e-5}, CNAB2, seeded T perturbation 1e-3"}


def adjoint_bench(steps=400, ckpt=20, reps=2):
    """Adjoint cost: the collocation CNAB2 case (n_r 65, n_phi 255, E 1e-4,
    Ra 1e6), gradient of the outer-wall Nusselt number averaged over
    `steps` steps with respect to the initial temperature, snapshots every
    `ckpt` steps.  cost = (forward sweep + backward sweep) / plain forward
    (pz_step over the same steps from the same state); best of `reps`."""
    import numpy as np
    cfg = W.c4_adjoint()
    run = Run(cfg)
    W.seed_member(run, cfg, 0)
    T0 = run.get_field("temp")
    z = np.zeros_like(T0)

    def reset():
        run.set_field("us", z)
        run.set_field("up", z)
        run.set_field("temp", T0)
        run.start()

    reset()
    run.step(5)
    run.time()
    t_fwd = t_adj = float("inf")
    for _ in range(reps):
        reset()
        t0 = time.perf_counter()
        run.step(steps)
        run.time()
        t_fwd = min(t_fwd, time.perf_counter() - t0)
    J = None
    for _ in range(reps):
        reset()
        t0 = time.perf_counter()
        J = run.adjoint_forward("nutop", steps, ckpt)
        g, _ = run.adjoint_backward()
        t_adj = min(t_adj, time.perf_counter() - t0)
    run.close()
    return {"workload": "colloc-CNAB2 n_r=65 n_phi=255 E=1e-4 Ra=1e6 Pr=1 dt=1e-5, d<NuTop>/dT0",
            "steps": steps, "checkpoint_every": ckpt, "forward_ms": 1e3 * t_fwd,
            "forward_plus_adjoint_ms": 1e3 * t_adj, "cost": t_adj / t_fwd, "objective": J,
            "grad_max_abs": float(np.abs(g).max())}


def fd_bench(steps=50, warmup=5):
    """The finite-difference workload: n_r 257 on the stretched grid,
    n_phi 2046, E 1e-6, Ra 1e9, Pr 0.1, SBDF3, dt 1e-7, seeded T
    perturbation 1e-3; steps/s of pz_step over `steps` steps."""
    cfg = W.c5_fd()
    run = Run(cfg)
    W.seed_run(run, cfg, amp_t=1e-3)
    run.start()
    run.step(warmup)
    run.time()
    t0 = time.perf_counter()
    run.step(steps)
    run.time()
    elapsed = time.perf_counter() - t0
    run.close()
    return {"workload": "FD-SBDF3 n_r=257 n_phi=2046 E=1e-6 Ra=1e9 Pr=0.1 dt=1e-7", "steps": steps,
            "steps_per_s": steps / elapsed, "ms_per_step": 1e3 * elapsed / steps}


def adjoint_fd_bench(steps=100, ckpt=10, reps=2):
    """Adjoint cost of the C5 gradient: finite differences at n_r 257,
    n_phi 2046, SBDF3, the mean kinetic energy over `steps` steps with
    respect to the seeded 16-mode outer-wall heat-flux pattern (and the
    initial temperature), snapshots every `ckpt` steps.  cost = (forward
    sweep + backward sweep) / plain forward; best of `reps`."""
    import numpy as np
    cfg = W.c5_gradient()
    run = Run(cfg)
    W.seed_run(run, cfg, amp_t=1e-3)
    T0 = run.get_field("temp")
    z = np.zeros_like(T0)

    def reset():
        run.set_field("us", z)
        run.set_field("up", z)
        run.set_field("temp", T0)
        run.start()

    reset()
    run.step(5)
    run.time()
    t_fwd = t_adj = float("inf")
    for _ in range(reps):
        reset()
        t0 = time.perf_counter()
        run.step(steps)
        run.time()
        t_fwd = min(t_fwd, time.perf_counter() - t0)
    J = gbc = None
    for _ in range(reps):
        reset()
        t0 = time.perf_counter()
        J = run.adjoint_forward("ekin", steps, ckpt)
        _, gbc = run.adjoint_backward()
        t_adj = min(t_adj, time.perf_counter() - t0)
    run.close()
    return {"workload": "FD-SBDF3 n_r=257 n_phi=2046 E=1e-6 Ra=1e9 Pr=0.1 dt=1e-7, d<E_kin>/d(outer heat-flux "
                        "pattern, m=1..16)",
            "steps": steps, "checkpoint_every": ckpt, "forward_ms": 1e3 * t_fwd,
            "forward_plus_adjoint_ms": 1e3 * t_adj, "cost": t_adj / t_fwd, "objective": J,
            "grad_pattern_max_abs": float(np.abs(gbc[1:17, 0]).max())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-only", action="store_true", help="measure only the ensemble factor")
    ap.add_argument("--ensemble-steps", type=int, default=100)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost")
    ap.add_argument("--adjoint-only", action="store_true", help="measure only the adjoint cost")
    ap.add_argument("--fd-only", action="store_true",
                    help="measure only the finite-difference workload (n_r 257, n_phi 2046, SBDF3)")
    ap.add_argument("--pmc-sample", action="store_true",
                    help="run only the short sample profiled by the counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs each simulation on a single "
              f"MI355X (requested --gpus {args.gpus}); ensembles shard whole runs per GPU.",
              file=sys.stderr)
        sys.exit(2)
    if args.pmc_sample:
        pmc_sample()
        return
    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_bench(steps=args.ensemble_steps)}))
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_bench(), "adjoint_fd": adjoint_fd_bench()}))
        return
    if args.fd_only:
        print(json.dumps({"fd": fd_bench()}))
        return

    cfg = W.c2_production()
    run = Run(cfg)
    W.seed_run(run, cfg)
    run.start()
    run.step(args.warmup)
    run.time()  # synchronises the library stream
    t0 = time.perf_counter()
    run.step(args.steps)
    t_end, dt = run.time()
    elapsed = time.perf_counter() - t0
    steps_per_s = args.steps / elapsed

    # dominant kernel: k_mphase, timed on a second handle so the timed run
    # keeps a valid state
    krun = Run(cfg)
    W.seed_run(krun, cfg)
    krun.start()
    krun.step(2)
    krun.time()
    reps = 200
    avg_ms = krun.time_kernel(0, reps)
    bytes_launch = mphase_bytes(cfg)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    krun.close()

    line = {
        "metric": "FP64 IMEX steps/s (n_r=97,n_phi=384)",
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded T/xi perturbation, amplitude 1e-4)",
        "config": {"workload": "integ-BPR353 n_r=97 n_phi=384 E=1e-5 Ra=3e7 Pr=1 Sc=3 Raxi=3e7 "
                               "Ekman pumping, dt=5e-6",
                   "n_r_max": cfg.n_r_max, "n_phi_max": 3 * cfg.m_max, "m_max": cfg.m_max,
                   "time_scheme": "BPR353", "sim_time_end": t_end, "dt": dt},
        "roofline": {"bound": "hbm", "kernel": "k_mphase (m-loop stage kernel)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "bytes_per_launch": bytes_launch, "avg_launch_ms": avg_ms},
    }
    if not args.no_ensemble:
        line["ensemble"] = ensemble_bench(steps=args.ensemble_steps)
    if not args.no_adjoint:
        line["adjoint"] = adjoint_bench()
        line["adjoint_fd"] = adjoint_fd_bench()
    line["fd"] = fd_bench()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(cfg, args.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
