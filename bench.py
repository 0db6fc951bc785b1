"""Benchmark: qubit*Trotter-steps/s of the 100k-qubit random 3-regular QUBO anneal on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--chi 8] [--precision single|double] [--no-accuracy]
                    [--no-cpu] [--serve-jobs 48]

Workload: random 3-regular graph on 100,000 qubits (configuration model,
seed 3), QUBO weights uniform integers in [-10, 10] on edges and diagonal,
mapped to Ising form; linear anneal T = 20 over 200 Trotter steps; bond
dimension chi = 8 (headline) and chi = 4 (reported beside it); BP with the
reference defaults bp_eps = 1e-6, at most 75 sweeps, pinv_eps = 1e-6.
A "step" is one Trotter step (simple update + local gates + gauge + BP to
convergence).  W untimed warmup steps (the chi ramp) precede K timed steps;
the schedule has max(200, W + K) steps.  Inputs are resident in HBM before
the timed region starts.

The FP32 path is timed (--precision double times the FP64 path instead).
Accuracy: the same anneal in FP64 on the GPU (wall clock reported beside the
FP32 line) gives the per-qubit <Z> error, rounded-spin disagreements and the
energy error.  cpu_baseline: the FP64 NumPy oracle on a bounded sample of the
same workload (3,000-qubit instance from the same generator, six Trotter
steps once the bond dimension reaches chi, one core).  serving: a seeded mix
of smaller jobs arriving over time through bqa_amd.serve (jobs/h, p50/p99
time-to-solution).  roofline: the chi_max BP-sweep kernel against the FP32
matrix peak with the reference contraction schedule's flops, the HBM
fraction beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "qubit·steps/s, 100k-qubit 3-reg anneal @⟨Z⟩ err≤1e-4; jobs/h + p50/p99 latency"
HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK_TFLOPS = 157.3
FP64_MFMA_PEAK_TFLOPS = 78.6
N_QUBITS = 100_000
TOTAL_TIME = 20.0
BASE_STEPS = 200
SEED = 3
BP_EPS = 1e-6
BP_ITERS = 75
PINV_EPS = 1e-6
SERVE_SEED = 5


def make_config(n, chi, steps, precision):
    from bqa_amd import instances
    nodes, edges, _ = instances.regular_qubo(n, 3, seed=SEED, qmax=10)
    return {"nodes": nodes, "edges": edges, "max_bond_dim": chi, "bp_eps": BP_EPS,
            "max_bp_iter_number": BP_ITERS, "pinv_eps": PINV_EPS, "damping": 0.0,
            "schedule": instances.anneal_schedule(TOTAL_TIME, steps), "backend": "mi355x",
            "precision": precision}


def run_gpu(ctx, precision, warmup, steps, timed):
    """Run the anneal; with `timed`, the BP-kernel HIP-event timing covers the timed steps that run at the
    final bond dimension ctx.max_bond_dim (the launches of the chi = max kernel)."""
    from bqa_amd.native import Job
    job = Job(ctx, precision)
    try:
        job.init_state()
        layers = [i for i in ctx.instructions if isinstance(i, dict)]
        stats = []

        def step(ins):
            return job.run_layer(ins["xtime"], ins["ztime"], ctx.pinv_eps, ctx.max_bp_iters_number,
                                 ctx.bp_eps, ctx.damping)

        for ins in layers[:warmup]:
            stats.append(step(ins))
        job.density()                         # synchronises the stream
        timing = timed and bool(stats) and stats[-1]["bond_dim"] >= ctx.max_bond_dim
        if timing:
            job.timing_enable(True)
        t0 = time.perf_counter()
        for ins in layers[warmup:warmup + steps]:
            stats.append(step(ins))
            if timed and not timing and stats[-1]["bond_dim"] >= ctx.max_bond_dim:
                job.timing_enable(True)
                timing = True
        t1 = time.perf_counter()
        kern = job.timing_read() if timed else (0, 0.0, 0.0)
        job.timing_enable(False)
        rho, bloch = job.density()
        energy, spins = job.energy()
        return {"seconds": t1 - t0, "stats": stats, "bloch": bloch, "energy": energy, "spins": spins,
                "kernel": kern, "device_bytes": job.device_bytes()}
    finally:
        job.close()


def load_traffic():
    """HBM bytes per BP-sweep launch measured with rocprofv3 counters (profiles/)."""
    path = os.path.join(ROOT, "profiles", "bp_sweep_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def cpu_baseline(chi):
    """FP64 NumPy oracle, one thread, 3,000-qubit sample of the same workload."""
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:              # pragma: no cover
        threadpool_limits = None
    from oracle import bqa_oracle as O
    n = 3000
    cfg = make_config(n, chi, BASE_STEPS, "double")
    ctx = O.make_context(cfg)
    st = O.init_state(ctx)
    layers = [i for i in ctx.instructions if isinstance(i, dict)]
    ctl = threadpool_limits(limits=1) if threadpool_limits else None
    try:
        ramp = 0
        while ramp < 80 and st.chi < chi:        # untimed: run until the bond dimension reaches chi
            ins = layers[ramp]
            O.run_layer(ctx, st, ins["xtime"], ins["ztime"])
            ramp += 1
        s_timed = 6
        t0 = time.perf_counter()
        for ins in layers[ramp:ramp + s_timed]:
            O.run_layer(ctx, st, ins["xtime"], ins["ztime"])
        dt = time.perf_counter() - t0
    finally:
        if ctl is not None:
            ctl.unregister() if hasattr(ctl, "unregister") else None
    sweeps = st.bp_iters[ramp:ramp + s_timed]
    return {"value": n * s_timed / dt, "unit": "qubit·steps/s", "cores": 1, "kind": "port",
            "sample": (f"FP64 NumPy oracle, {n}-qubit random 3-regular QUBO (same generator), chi={chi}, "
                       f"Trotter steps {ramp + 1}-{ramp + s_timed} of {BASE_STEPS}, the first at bond dimension {chi}, "
                       f"BP sweeps per step {sweeps}, {dt:.1f} s"),
            "seconds_per_step": dt / s_timed, "host_cpus": len(os.sched_getaffinity(0))}


def accuracy(fast, ref):
    z32, z64 = fast["bloch"][:, 2], ref["bloch"][:, 2]
    dz = np.abs(z32 - z64)
    decided = np.abs(z64) >= 1e-3
    spin_mismatch = int(np.sum((fast["spins"] != ref["spins"]) & decided))
    e_rel = abs(fast["energy"] - ref["energy"]) / max(1.0, abs(ref["energy"]))
    return {"max_abs_dz": float(dz.max()), "mean_abs_dz": float(dz.mean()),
            "p99_abs_dz": float(np.percentile(dz, 99)), "p999_abs_dz": float(np.percentile(dz, 99.9)),
            "qubits_dz_above_1e-4": int(np.sum(dz > 1e-4)), "qubits": int(dz.size),
            "max_abs_dbloch": float(np.abs(fast["bloch"] - ref["bloch"]).max()),
            "spin_mismatches_decided": spin_mismatch, "energy_fp32": fast["energy"],
            "energy_fp64": ref["energy"], "energy_rel_err": e_rel}


def summarize(run, n, steps):
    stats = run["stats"][-steps:] if steps else []
    sweeps = [s["sweeps"] for s in stats]
    launches, ms, bytes_ = run["kernel"]
    return {"qubit_steps_per_s": n * steps / run["seconds"], "ms_per_step": 1e3 * run["seconds"] / steps,
            "bp_sweeps_total": int(sum(sweeps)), "bp_sweeps_per_step_mean": float(np.mean(sweeps)),
            "bond_dim_final": stats[-1]["bond_dim"] if stats else None, "bp_kernel_launches": launches,
            "bp_kernel_ms": ms, "bp_kernel_bytes": bytes_, "device_bytes": run["device_bytes"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=195)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--chi", type=int, default=8)
    ap.add_argument("--qubits", type=int, default=N_QUBITS)
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-chi4", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=48, help="jobs in the serving sample (0 skips it)")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="BYTES",
                    help="also run the plain copy kernel on BYTES bytes (HBM counter calibration)")
    ap.add_argument("--precision", choices=("single", "double"), default="single",
                    help="timed path: FP32 storage with FP64 eigen work (default), or FP64 throughout, which is "
                         "its own accuracy reference")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(independent jobs would shard one per GPU with no communication)", file=sys.stderr)
        sys.exit(2)
    from bqa_amd.config import config_to_context
    from bqa_amd.native import device_info
    n_dev, arch_ok, arch = device_info()
    if n_dev < 1:
        raise RuntimeError("no HIP device")
    n = args.qubits
    total_steps = max(BASE_STEPS, args.warmup + args.steps)
    copy_ms = None
    if args.copy_calibrate:
        from bqa_amd.native import copy_calibrate
        copy_ms = copy_calibrate(args.copy_calibrate, 5)
    log = lambda m: print(m, file=sys.stderr, flush=True)

    tag = "fp32" if args.precision == "single" else "fp64"

    def one(chi, want_acc):
        ctx32 = config_to_context(make_config(n, chi, total_steps, args.precision))
        log(f"[bench] chi={chi} {tag}: {args.warmup} warmup + {args.steps} timed Trotter steps, n={n}")
        fast = run_gpu(ctx32, args.precision, args.warmup, args.steps, timed=True)
        s = summarize(fast, n, args.steps)
        log(f"[bench] chi={chi} {tag}: {s['qubit_steps_per_s']:.4g} qubit*steps/s, {s['ms_per_step']:.2f} ms/step, "
            f"BP sweeps/step {s['bp_sweeps_per_step_mean']:.1f}")
        out = {"summary": s, "run": fast, "ctx": ctx32}
        if want_acc and args.precision == "single":
            ctx64 = config_to_context(make_config(n, chi, total_steps, "double"))
            log(f"[bench] chi={chi} fp64 reference run (wall clock only)")
            ref = run_gpu(ctx64, "double", args.warmup, args.steps, timed=False)
            out["accuracy"] = accuracy(fast, ref)
            s64 = summarize(ref, n, args.steps)
            out["fp64"] = {"qubit_steps_per_s": round(s64["qubit_steps_per_s"], 2),
                           "ms_per_step": round(s64["ms_per_step"], 4),
                           "bp_sweeps_per_step": s64["bp_sweeps_per_step_mean"],
                           "note": "the FP64 path on the same K steps: the accuracy reference itself"}
            log(f"[bench] chi={chi} fp64: {s64['qubit_steps_per_s']:.4g} qubit*steps/s, "
                f"{s64['ms_per_step']:.2f} ms/step")
            log(f"[bench] chi={chi} accuracy {out['accuracy']}")
        return out

    head = one(args.chi, not args.no_accuracy)
    s = head["summary"]
    launches, ms, bytes_ = head["run"]["kernel"]
    ctx = head["ctx"]
    # Reference-schedule flops of one BP sweep (d outputs x [(d - 1) mode products + 1 Gram] x 2 chi^(d+1)
    # complex MACs per node, 8 real flops each), spread over the per-class launches of the sweep.
    classes = [(d, lay.node_ids.size) for d, lay in ctx.degree_to_layout.items() if d > 0 and lay.node_ids.size]
    flops_sweep = sum(8 * nd * d * d * 2 * args.chi ** (d + 1) for d, nd in classes)
    flops_launch = flops_sweep / max(1, len(classes))
    sec = ms / 1e3
    tflops = flops_launch * launches / sec / 1e12 if sec > 0 else 0.0
    gbs = (bytes_ / 1e9) / sec if sec > 0 else 0.0
    # the HBM counters were taken on the FP32 sweep kernel
    traffic = load_traffic() if args.precision == "single" else None
    peak = FP32_MFMA_PEAK_TFLOPS if args.precision == "single" else FP64_MFMA_PEAK_TFLOPS
    roofline = {"bound": "mfma", "achieved": round(tflops, 3), "peak": peak, "unit": "TFLOP/s",
                "frac": round(tflops / peak, 5),
                "traffic": (traffic or {}).get("hbm_bytes_per_launch") if traffic else None,
                "kernel": f"BP sweep at chi={args.chi} (one launch per degree class per sweep)",
                "launches": launches, "avg_launch_us": round(1e6 * sec / max(1, launches), 3),
                "algorithmic_flops_per_launch": flops_launch,
                "algorithmic_bytes_per_launch": bytes_ / max(1, launches),
                "flop_per_byte": round(flops_launch / max(1.0, bytes_ / max(1, launches)), 2),
                "hbm_achieved_GBps": round(gbs, 2), "hbm_peak_GBps": HBM_PEAK_GBS,
                "hbm_frac": round(gbs / HBM_PEAK_GBS, 5)}
    extra = {}
    if not args.no_chi4 and args.chi != 4:
        c4 = one(4, not args.no_accuracy)
        extra["chi4"] = dict(c4["summary"], accuracy=c4.get("accuracy"), fp64=c4.get("fp64"))
    serving = None
    if args.serve_jobs > 0:
        from bqa_amd.serve import mixed_requests, serve
        log(f"[bench] serving: {args.serve_jobs} mixed jobs, Poisson arrivals")
        reqs = mixed_requests(args.serve_jobs, seed=SERVE_SEED, n_min=100, n_max=5000, rate_per_s=8.0,
                              degrees=(3,), chis=(4, 8))
        sv = serve(reqs, max_concurrent=8)
        keep = lambda d: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items() if k != "outcomes"}
        serving = keep(sv)
        serving["workload"] = (f"{args.serve_jobs} jobs, Poisson arrivals at 8/s (seed {SERVE_SEED}); random 3-regular, "
                               "n log-uniform in [100, 5000], chi in {4, 8}, J ~ U[-1,1], h = 0, T = 10, "
                               "100-300 Trotter steps growing with log n; FP32; at most 8 resident jobs")
        serving["note"] = ("jobs_per_h and p50/p99 at the offered load above; capacity: the same jobs all queued "
                           "at t = 0")
        for r in reqs:
            r.arrival = 0.0
        cap = serve(reqs, max_concurrent=8)
        serving["capacity"] = {k: v for k, v in keep(cap).items() if k in ("jobs", "failed", "makespan_s", "jobs_per_h",
                                                                        "qubit_steps_per_s", "p50_s", "p99_s")}
        # the same queued jobs one at a time (the plain FP32 path, no overlap)
        seq = serve(reqs, max_concurrent=1)
        serving["one_at_a_time"] = {k: v for k, v in keep(seq).items() if k in ("jobs", "failed", "makespan_s",
                                                                               "jobs_per_h", "p50_s", "p99_s")}
        serving["capacity_over_one_at_a_time"] = round(cap["jobs_per_h"] / max(1e-12, seq["jobs_per_h"]), 3)
        log(f"[bench] serving: {serving['jobs_per_h']:.4g} jobs/h offered, p50 {serving['p50_s']:.3f} s, "
            f"p99 {serving['p99_s']:.3f} s; capacity {cap['jobs_per_h']:.4g} jobs/h; failed {serving['failed']}; "
            f"one at a time {seq['jobs_per_h']:.4g} jobs/h ({serving['capacity_over_one_at_a_time']:.2f}x)")
    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline (FP64 NumPy oracle)")
        cpu = cpu_baseline(args.chi)
        log(f"[bench] cpu baseline {cpu['value']:.4g} qubit*steps/s")
    E = head["ctx"].edges_number
    result = {
        "metric": METRIC,
        "value": round(s["qubit_steps_per_s"], 2),
        "unit": "qubit·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": tag,
        "data": "synthetic",
        "config": {"workload": f"{n}-qubit random 3-regular QUBO anneal (ints in [-10,10]), chi={args.chi}, "
                               f"T={TOTAL_TIME:g}, {total_steps} Trotter steps",
                   "qubits": n, "edges": E // 2, "chi": args.chi, "bp_eps": BP_EPS,
                   "max_bp_iter_number": BP_ITERS, "pinv_eps": PINV_EPS,
                   "precision": ("fp32 tensors/messages (the reference CuPy backend's complex64), fp64 "
                                 "SVD/lambdas/energy" if args.precision == "single" else "fp64 throughout"),
                   "device": arch},
        "message_updates_per_s": round(E * s["bp_sweeps_total"] / max(1e-12, s["ms_per_step"] * args.steps / 1e3), 1),
        "bp_sweeps_per_step": s["bp_sweeps_per_step_mean"],
        "roofline": roofline,
        "cpu_baseline": cpu,
        "accuracy": head.get("accuracy"),
        "fp64": head.get("fp64"),
        "serving": serving,
    }
    result.update(extra)
    if copy_ms is not None:
        result["copy_calibration"] = {"bytes": args.copy_calibrate, "ms_per_copy": copy_ms,
                                      "GB_per_s": 2 * args.copy_calibrate / 1e6 / copy_ms}
    print(json.dumps(result))


if __name__ == "__main__":
    main()
