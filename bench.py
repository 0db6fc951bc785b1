"""Benchmark: qubit*Trotter-steps/s of the 100k-qubit random 3-regular QUBO anneal on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--chi 8] [--precision double|single]
                    [--no-accuracy] [--no-cpu] [--no-chi4] [--no-grid] [--serve-jobs 48]

Workload: random 3-regular graph on 100,000 qubits (configuration model,
seed 3), QUBO weights uniform integers in [-10, 10] on edges and diagonal,
mapped to Ising form; linear anneal T = 20 over 200 Trotter steps; bond
dimension chi = 8 (headline) and chi = 4 (reported beside it); BP with the
reference defaults bp_eps = 1e-6, at most 75 sweeps, pinv_eps = 1e-6.

A "step" is one Trotter step (simple update + local gates + gauge + BP to
convergence).  The anneal runs untimed for at least W steps and until the
bond dimension has reached chi (the ramp 1, 2, 4, 8 takes about 27 steps at
100k qubits); then exactly K steps are timed, every one of them at bond
dimension chi (their chi is listed in the JSON); then the anneal runs to its
end, untimed, so that the accuracy figures cover all 200 steps.  Inputs are
resident in HBM before the timed region starts.

The timed path is FP64 (--precision double, the default): it is the path whose
<Z> agrees with the FP64 oracle within the metric's 1e-4.  The FP32 path
(complex64 tensors and messages, FP64 eigen work) runs the same anneal beside
it; its throughput and its <Z>/energy error against the FP64 path over all 200
steps are reported, since it does not meet 1e-4 on every qubit of this
anneal.  The FP64 path's own agreement with the FP64 NumPy oracle is measured
on a 100-qubit instance of the same generator over all 200 steps, the same
run that gives cpu_baseline.

roofline: the chi = 8 BP sweep kernel's algorithmic HBM bytes per launch over
its average HIP-event launch time (events on the job's stream), against 8 TB/s;
the matrix-peak fraction with the reference contraction schedule's flops is
reported beside it.  traffic: HBM bytes per launch from rocprofv3 counters,
read from the committed summary under profiles/ (labelled as such; counters
cannot be read inside this process).
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
MFMA_PEAK_TFLOPS = {"single": 157.3, "double": 78.6}
N_QUBITS = 100_000
TOTAL_TIME = 20.0
BASE_STEPS = 200
RAMP_ALLOWANCE = 40          # steps reserved for the bond-dimension ramp when K is large
SEED = 3
BP_EPS = 1e-6
BP_ITERS = 75
PINV_EPS = 1e-6
SERVE_SEED = 5
SERVE_SLOTS = 32             # resident jobs of the batched serving executor (four batches of eight)
SAMPLE_QUBITS = 100          # oracle sample: full 200-step anneal on the CPU (~30 s, one core)
GRID_SIDE = 100
GRID_STEPS = 10              # timed chi = 8 steps of the 100 x 100 grid line
HHEX_CELLS = 64              # heavy-hex lattice of 64 x 64 cells: 10,144 qubits, degrees 2 and 3
HHEX_CHI = 16
HHEX_STEPS = 300
HHEX_TIMED = 10
HHEX32_TIMED = 3             # chi = 32 runs the generic kernels (seconds per step)
TAG = {"single": "fp32", "double": "fp64"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def qubo_config(n, chi, steps, precision):
    from bqa_amd import instances
    nodes, edges, _ = instances.regular_qubo(n, 3, seed=SEED, qmax=10)
    return {"nodes": nodes, "edges": edges, "max_bond_dim": chi, "bp_eps": BP_EPS,
            "max_bp_iter_number": BP_ITERS, "pinv_eps": PINV_EPS, "damping": 0.0,
            "schedule": instances.anneal_schedule(TOTAL_TIME, steps), "backend": "mi355x",
            "precision": precision}


def grid_config(side, chi, steps, precision):
    from bqa_amd import instances
    nodes, edges = instances.grid_gaussian(side, side, seed=2, field_sigma=0.1)
    return {"nodes": nodes, "edges": edges, "max_bond_dim": chi, "bp_eps": BP_EPS,
            "max_bp_iter_number": BP_ITERS, "pinv_eps": PINV_EPS, "damping": 0.0,
            "schedule": instances.anneal_schedule(TOTAL_TIME, steps), "backend": "mi355x",
            "precision": precision}


def hhex_config(chi, steps, precision):
    from bqa_amd import instances
    nodes, edges = instances.heavy_hex_gaussian(HHEX_CELLS, HHEX_CELLS, seed=4, field_sigma=0.5)
    return {"nodes": nodes, "edges": edges, "max_bond_dim": chi, "bp_eps": BP_EPS,
            "max_bp_iter_number": BP_ITERS, "pinv_eps": PINV_EPS, "damping": 0.0,
            "schedule": instances.anneal_schedule(30.0, steps), "backend": "mi355x",
            "precision": precision}


def run_anneal(ctx, precision, warmup, steps, finish=True, timed=True):
    """Run one anneal on the GPU.

    Untimed until at least `warmup` steps have run and the bond dimension equals ctx.max_bond_dim; then
    `steps` timed steps (wall clock around them, BP-kernel HIP events inside them); then, with `finish`,
    the rest of the schedule untimed.  Every Trotter step returns after a stream synchronisation (the BP
    status readback), so the wall clock brackets completed GPU work."""
    from bqa_amd.native import Job
    layers = [i for i in ctx.instructions if isinstance(i, dict)]
    chi = ctx.max_bond_dim
    job = Job(ctx, precision)
    try:
        job.init_state()
        stats, step_s = [], []

        def step(k):
            ins = layers[k]
            t = time.perf_counter()
            out = job.run_layer(ins["xtime"], ins["ztime"], ctx.pinv_eps, ctx.max_bp_iters_number,
                                ctx.bp_eps, ctx.damping)
            step_s.append(time.perf_counter() - t)
            stats.append(out)
            if (k + 1) % 50 == 0:
                log(f"[bench]   step {k + 1}/{len(layers)}: chi {out['bond_dim']}, {out['sweeps']} sweeps, "
                    f"{1e3 * step_s[-1]:.1f} ms")

        t_all = time.perf_counter()
        k = 0
        while k < len(layers) and (k < warmup or not stats or stats[-1]["bond_dim"] < chi):
            step(k)
            k += 1
        start = k
        if start + steps > len(layers):
            raise RuntimeError(f"the bond dimension reached {chi} at step {start}: {len(layers) - start} steps "
                               f"are left, {steps} were asked for")
        if timed:
            job.timing_enable(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            step(k)
            k += 1
        t1 = time.perf_counter()
        kern = job.timing_read() if timed else (0, 0.0, 0.0)
        job.timing_enable(False)
        if finish:
            while k < len(layers):
                step(k)
                k += 1
        t_end = time.perf_counter()
        rho, bloch = job.density()
        energy, spins = job.energy()
        return {"seconds": t1 - t0, "start": start, "stats": stats, "step_s": step_s, "bloch": bloch,
                "energy": energy, "spins": spins, "kernel": kern, "device_bytes": job.device_bytes(),
                "steps_run": k, "anneal_seconds": t_end - t_all}
    finally:
        job.close()


def summarize(run, n, steps):
    win = run["stats"][run["start"]:run["start"] + steps]
    sweeps = [s["sweeps"] for s in win]
    launches, ms, bytes_ = run["kernel"]
    full = run["steps_run"]
    return {"qubit_steps_per_s": n * steps / run["seconds"], "ms_per_step": 1e3 * run["seconds"] / steps,
            "timed_steps": [run["start"], run["start"] + steps], "chi_timed": sorted({s["bond_dim"] for s in win}),
            "bp_sweeps_total": int(sum(sweeps)), "bp_sweeps_per_step_mean": float(np.mean(sweeps)),
            "bp_kernel_launches": launches, "bp_kernel_ms": ms, "bp_kernel_bytes": bytes_,
            "device_bytes": run["device_bytes"],
            "full_anneal": {"steps": full, "seconds": round(run["anneal_seconds"], 4),
                            "qubit_steps_per_s": round(n * full / run["anneal_seconds"], 2),
                            "bp_sweeps_total": int(sum(s["sweeps"] for s in run["stats"]))}}


def accuracy(fast, ref):
    """<Z> / Bloch / spin / energy agreement of `fast` with `ref` (the FP64 result) at the end of the anneal."""
    za, zb = fast["bloch"][:, 2], ref["bloch"][:, 2]
    dz = np.abs(za - zb)
    decided = np.abs(zb) >= 1e-3
    spin_mismatch = int(np.sum((fast["spins"] != ref["spins"]) & decided))
    e_rel = abs(fast["energy"] - ref["energy"]) / max(1e-300, abs(ref["energy"]))
    return {"max_abs_dz": float(dz.max()), "mean_abs_dz": float(dz.mean()),
            "p99_abs_dz": float(np.percentile(dz, 99)), "p999_abs_dz": float(np.percentile(dz, 99.9)),
            "qubits_dz_above_1e-4": int(np.sum(dz > 1e-4)), "qubits_dz_above_1e-5": int(np.sum(dz > 1e-5)),
            "qubits": int(dz.size), "max_abs_dbloch": float(np.abs(fast["bloch"] - ref["bloch"]).max()),
            "spin_mismatches_decided": spin_mismatch, "energy": fast["energy"], "energy_ref": ref["energy"],
            "energy_rel_err": e_rel}


def load_traffic(precision, chi):
    """HBM bytes per BP-sweep launch of the chi = 8 (or chi = 4) d = 3 kernel, measured with rocprofv3 counters
    (committed under profiles/); None for other shapes."""
    suffix = {8: "", 4: "_chi4"}.get(chi)
    if suffix is None:
        return None, None
    path = os.path.join(ROOT, "profiles", f"bp_sweep_traffic_{TAG[precision]}{suffix}.json")
    if not os.path.exists(path):
        return None, None
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("hbm_bytes_per_launch"), f"profiles/{os.path.basename(path)} (rocprofv3 --pmc, calibrated)"
    except (OSError, ValueError):
        return None, None


def roofline_of(summary, ctx, chi, precision, kernel):
    launches, ms = summary["bp_kernel_launches"], summary["bp_kernel_ms"]
    bytes_ = summary["bp_kernel_bytes"]
    sec = ms / 1e3
    classes = [(d, lay.node_ids.size) for d, lay in ctx.degree_to_layout.items() if d > 0 and lay.node_ids.size]
    # reference contraction schedule: d outputs x [(d - 1) mode products + 1 Gram] x 2 chi^(d+1) complex MACs
    # per node, 8 real flops each; spread over the per-class launches of a sweep
    flops_sweep = sum(8 * nd * d * d * 2 * chi ** (d + 1) for d, nd in classes)
    flops_launch = flops_sweep / max(1, len(classes))
    gbs = (bytes_ / 1e9) / sec if sec > 0 else 0.0
    tflops = flops_launch * launches / sec / 1e12 if sec > 0 else 0.0
    traffic, src = load_traffic(precision, chi)
    peak = MFMA_PEAK_TFLOPS[precision]
    return {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_source": src,
            "kernel": kernel, "launches": launches, "avg_launch_us": round(1e6 * sec / max(1, launches), 3),
            "algorithmic_bytes_per_launch": bytes_ / max(1, launches),
            "algorithmic_flops_per_launch": flops_launch,
            "mfma_achieved_TFLOPs": round(tflops, 3), "mfma_peak_TFLOPs": peak,
            "mfma_frac": round(tflops / peak, 5)}


def cpu_sample(chi, precisions):
    """FP64 NumPy oracle, one thread, on a SAMPLE_QUBITS-qubit instance of the same generator over the full
    200-step anneal (cpu_baseline: the steps at bond dimension chi), and the GPU paths on the same instance
    (their agreement with the oracle)."""
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:              # pragma: no cover
        threadpool_limits = None
    from bqa_amd.config import config_to_context
    from oracle import bqa_oracle as O
    n = SAMPLE_QUBITS
    cfg = qubo_config(n, chi, BASE_STEPS, "double")
    ctx = O.make_context(cfg)
    st = O.init_state(ctx)
    layers = [i for i in ctx.instructions if isinstance(i, dict)]
    ctl = threadpool_limits(limits=1) if threadpool_limits else None
    try:
        t_chi = 0.0
        steps_chi = 0
        for ins in layers:
            t = time.perf_counter()
            O.run_layer(ctx, st, ins["xtime"], ins["ztime"])
            if st.chi >= chi:
                t_chi += time.perf_counter() - t
                steps_chi += 1
    finally:
        if ctl is not None and hasattr(ctl, "unregister"):
            ctl.unregister()
    bloch = O.bloch_vectors(O.get_density_matrices(ctx, st))
    spins = np.where(bloch[:, 2] > 0, 1, -1).astype(np.int8)
    ref = {"bloch": bloch, "spins": spins, "energy": O.ising_energy(ctx, spins)}
    cpu = {"value": n * steps_chi / t_chi, "unit": "qubit·steps/s", "cores": 1, "kind": "port",
           "sample": (f"FP64 NumPy oracle (one thread), {n}-qubit random 3-regular QUBO from the same generator, "
                      f"chi={chi}, all {BASE_STEPS} Trotter steps run; timed: the {steps_chi} steps at bond "
                      f"dimension {chi} ({t_chi:.1f} s, {sum(st.bp_iters)} BP sweeps in the anneal)"),
           "seconds_per_step": t_chi / max(1, steps_chi), "host_cpus": len(os.sched_getaffinity(0))}
    parity = {}
    for prec in precisions:
        g = run_anneal(config_to_context(dict(cfg, precision=prec)), prec, 0, 1, finish=True, timed=False)
        a = accuracy(g, ref)
        a["bp_sweeps_equal_oracle"] = [s["sweeps"] for s in g["stats"]] == list(st.bp_iters)
        a["chi_equal_oracle"] = [s["bond_dim"] for s in g["stats"]] == list(st.chis)
        parity[TAG[prec]] = a
    return cpu, parity


def serving_sample(jobs, degrees, chi_cap):
    """The serving figures for one job mix: offered load, capacity (all queued at t = 0) with the batched
    executor and with one thread per job, and the same jobs one at a time."""
    from bqa_amd.serve import mixed_requests, serve
    log(f"[bench] serving: {jobs} mixed jobs, degrees {degrees}, Poisson arrivals")
    reqs = mixed_requests(jobs, seed=SERVE_SEED, n_min=100, n_max=5000, rate_per_s=8.0, degrees=degrees,
                          chis=(4, 8), chi_cap=chi_cap)
    sv = serve(reqs, max_concurrent=SERVE_SLOTS)
    keep = lambda d: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items() if k != "outcomes"}
    pick = ("jobs", "failed", "makespan_s", "jobs_per_h", "qubit_steps_per_s", "p50_s", "p99_s")
    out = keep(sv)
    deg = ", ".join(str(d) for d in degrees)
    cap_txt = "" if not chi_cap else " (" + ", ".join(f"chi {c} at d = {d}" for d, c in chi_cap.items()) + ")"
    out["workload"] = (f"{jobs} jobs, Poisson arrivals at 8/s (seed {SERVE_SEED}); random d-regular with d in "
                       f"{{{deg}}}, n log-uniform in [100, 5000], chi in {{4, 8}}{cap_txt}, J ~ U[-1,1], h = 0, "
                       "T = 10, 100-300 Trotter steps growing with log n; FP32; at most "
                       f"{SERVE_SLOTS} resident jobs, batched executor (four host threads, each stepping its "
                       "resident jobs together through its own bqa_batch: canonicalisers and BP sweeps in launches "
                       "shared per shape)")
    out["note"] = ("jobs_per_h and p50/p99 at the offered load above; capacity: the same jobs all queued at t = 0; "
                   "threaded: the same queue on one host thread and stream per job (8 resident); one_at_a_time: "
                   "each job alone on its own bqa_job (no overlap, no sharing)")
    out["mix"] = {f"d{d}_chi{c}": sum(1 for r in reqs if r.tag["degree"] == d and r.tag["chi"] == c)
                  for d in degrees for c in (4, 8)}
    for r in reqs:
        r.arrival = 0.0
    cap = serve(reqs, max_concurrent=SERVE_SLOTS)
    out["capacity"] = {k: v for k, v in keep(cap).items() if k in pick}
    thr = serve(reqs, max_concurrent=8, batched=False)
    out["threaded"] = {k: v for k, v in keep(thr).items() if k in pick}
    seq = serve(reqs, max_concurrent=1, batched=False)
    out["one_at_a_time"] = {k: v for k, v in keep(seq).items() if k in pick}
    # a ratio only when every job of both runs completed
    complete = cap["failed"] == 0 and seq["failed"] == 0 and seq["jobs_per_h"]
    out["capacity_over_one_at_a_time"] = round(cap["jobs_per_h"] / seq["jobs_per_h"], 3) if complete else None
    f = lambda v, spec: format(v, spec) if isinstance(v, (int, float)) else "n/a"
    log(f"[bench] serving d in {degrees}: {f(out['jobs_per_h'], '.4g')} jobs/h offered, p50 {f(out['p50_s'], '.3f')} s, "
        f"p99 {f(out['p99_s'], '.3f')} s; capacity {f(cap['jobs_per_h'], '.4g')} jobs/h (threaded "
        f"{f(thr['jobs_per_h'], '.4g')}); failed {out['failed']} / {cap['failed']} / {thr['failed']} / {seq['failed']}; "
        f"one at a time {f(seq['jobs_per_h'], '.4g')} jobs/h ({f(out['capacity_over_one_at_a_time'], '.2f')}x)")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=150)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--chi", type=int, default=8)
    ap.add_argument("--qubits", type=int, default=N_QUBITS)
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-chi4", action="store_true")
    ap.add_argument("--no-grid", action="store_true")
    ap.add_argument("--no-hhex", action="store_true", help="skip the heavy-hex chi=16 line")
    ap.add_argument("--no-hhex32", action="store_true", help="skip the heavy-hex chi=32 line")
    ap.add_argument("--hhex-precisions", default="single,double",
                    help="precisions of the heavy-hex line (comma-separated; the error needs both)")
    ap.add_argument("--grid-precisions", default=None,
                    help="precisions of the grid line (comma-separated; default: the timed one, then the other)")
    ap.add_argument("--serve-jobs", type=int, default=48, help="jobs in the serving sample (0 skips it)")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="BYTES",
                    help="also run the plain copy kernel on BYTES bytes (HBM counter calibration)")
    ap.add_argument("--precision", choices=("single", "double"), default="double",
                    help="timed path: FP64 throughout (default; meets the metric's <Z> tolerance) or FP32 "
                         "storage with FP64 eigen work")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(independent jobs would shard one per GPU with no communication)", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        raise SystemExit("--steps must be at least 1")
    from bqa_amd.config import config_to_context
    from bqa_amd.native import device_info
    n_dev, arch_ok, arch = device_info()
    if n_dev < 1:
        raise RuntimeError("no HIP device")
    n = args.qubits
    total_steps = max(BASE_STEPS, args.warmup + args.steps + RAMP_ALLOWANCE)
    copy_ms = None
    if args.copy_calibrate:
        from bqa_amd.native import copy_calibrate
        copy_ms = copy_calibrate(args.copy_calibrate, 5)
    other = "single" if args.precision == "double" else "double"

    def one(chi, prec):
        ctx = config_to_context(qubo_config(n, chi, total_steps, prec))
        log(f"[bench] chi={chi} {TAG[prec]}: anneal of {total_steps} steps, n={n}; {args.steps} timed at chi={chi}")
        run = run_anneal(ctx, prec, args.warmup, args.steps, finish=not args.no_accuracy)
        s = summarize(run, n, args.steps)
        log(f"[bench] chi={chi} {TAG[prec]}: {s['qubit_steps_per_s']:.4g} qubit*steps/s, {s['ms_per_step']:.2f} "
            f"ms/step over steps {s['timed_steps']}, BP sweeps/step {s['bp_sweeps_per_step_mean']:.1f}; full "
            f"anneal {s['full_anneal']['seconds']:.2f} s")
        return {"summary": s, "run": run, "ctx": ctx}

    head = one(args.chi, args.precision)
    s = head["summary"]
    if s["bp_kernel_launches"] == 0:
        raise RuntimeError("no BP sweep kernel was timed at the target bond dimension")
    if s["chi_timed"] != [args.chi]:
        raise RuntimeError(f"timed steps ran at bond dimensions {s['chi_timed']}, not {args.chi}")
    kname = {"double": "k_mfma8_bp64_3p<4> (FP64 MFMA, six contractions per node)", "single": "k_mfma8<3,kModeBP,1> (FP32 MFMA)"}
    roofline = roofline_of(s, head["ctx"], args.chi, args.precision,
                           f"BP sweep at chi={args.chi}, d=3: {kname[args.precision] if args.chi == 8 else 'per shape'}"
                           " (one launch per degree class per sweep)")
    accuracy_block, other_line = None, None
    if not args.no_accuracy:
        o = one(args.chi, other)
        fp64_run, fp32_run = (head["run"], o["run"]) if args.precision == "double" else (o["run"], head["run"])
        acc = accuracy(fp32_run, fp64_run)
        accuracy_block = {"fp32_vs_fp64_gpu": acc,
                          "over": f"all {total_steps} steps of the anneal (end-of-anneal <Z>, spins, energy)"}
        log(f"[bench] fp32 vs fp64 over {total_steps} steps: {acc}")
        os_ = o["summary"]
        other_line = {"qubit_steps_per_s": round(os_["qubit_steps_per_s"], 2), "ms_per_step": round(os_["ms_per_step"], 4),
                      "timed_steps": os_["timed_steps"], "bp_sweeps_per_step": os_["bp_sweeps_per_step_mean"],
                      "full_anneal": os_["full_anneal"],
                      "roofline": roofline_of(os_, o["ctx"], args.chi, other, "BP sweep, same shape")}
    extra = {}
    if not args.no_chi4 and args.chi != 4:
        c4 = {}
        runs = {}
        for prec in (args.precision, other):
            r = one(4, prec)
            runs[prec] = r["run"]
            c4[TAG[prec]] = dict(r["summary"], roofline=roofline_of(r["summary"], r["ctx"], 4, prec, "BP sweep chi=4"))
        if not args.no_accuracy:
            c4["fp32_vs_fp64_gpu"] = accuracy(runs["single"], runs["double"])
        extra["chi4"] = c4
    if not args.no_grid:
        g = {}
        grid_precs = ([p for p in args.grid_precisions.split(",") if p] if args.grid_precisions
                      else [args.precision, other])
        for prec in grid_precs:
            ctx = config_to_context(grid_config(GRID_SIDE, args.chi, BASE_STEPS, prec))
            log(f"[bench] grid {GRID_SIDE}x{GRID_SIDE} chi={args.chi} {TAG[prec]}: {GRID_STEPS} timed steps")
            r = run_anneal(ctx, prec, args.warmup, GRID_STEPS, finish=False)
            gs = summarize(r, GRID_SIDE * GRID_SIDE, GRID_STEPS)
            gs.pop("full_anneal")
            g[TAG[prec]] = gs
            log(f"[bench] grid {TAG[prec]}: {gs['qubit_steps_per_s']:.4g} qubit*steps/s, {gs['ms_per_step']:.2f} ms/step")
        g["workload"] = (f"{GRID_SIDE}x{GRID_SIDE} open grid (degrees 2/3/4), J ~ N(0,1), h ~ N(0,0.1), T=20, "
                         f"200-step schedule, chi={args.chi}; {GRID_STEPS} steps timed once the bond dimension is "
                         f"{args.chi}")
        extra["grid100"] = g
    if not args.no_hhex:
        hx, runs = {}, {}
        nq = None
        for prec in [p for p in args.hhex_precisions.split(",") if p]:
            ctx = config_to_context(hhex_config(HHEX_CHI, HHEX_STEPS, prec))
            nq = ctx.nodes_number
            log(f"[bench] heavy-hex {nq} qubits chi={HHEX_CHI} {TAG[prec]}: {HHEX_STEPS}-step anneal, "
                f"{HHEX_TIMED} steps timed at chi={HHEX_CHI}")
            r = run_anneal(ctx, prec, args.warmup, HHEX_TIMED, finish=not args.no_accuracy)
            hs = summarize(r, nq, HHEX_TIMED)
            kern = ("k_mfma16_bp3h<8> or k_mfma16_bp<3,8> (autotuned) / k_mfma16_bp<2,1> (FP32 MFMA "
                    "v_mfma_f32_16x16x4f32)" if prec == "single"
                    else "k_mfma16_bp64<3,8> / <2,1> (FP64 MFMA v_mfma_f64_16x16x4f64)")
            hs["roofline"] = roofline_of(hs, ctx, HHEX_CHI, prec, f"BP sweep at chi={HHEX_CHI}, d=2,3: {kern}")
            hx[TAG[prec]] = hs
            runs[prec] = r
            log(f"[bench] heavy-hex {TAG[prec]}: {hs['qubit_steps_per_s']:.4g} qubit*steps/s, "
                f"{hs['ms_per_step']:.2f} ms/step, BP MFMA frac {hs['roofline']['mfma_frac']:.3f}; full anneal "
                f"{hs['full_anneal']['seconds']:.1f} s")
        if not args.no_accuracy and "single" in runs and "double" in runs:
            hx["fp32_vs_fp64_gpu"] = accuracy(runs["single"], runs["double"])
            log(f"[bench] heavy-hex fp32 vs fp64: {hx['fp32_vs_fp64_gpu']}")
        hx["workload"] = (f"heavy-hex lattice of {HHEX_CELLS}x{HHEX_CELLS} cells ({nq} qubits, degrees 2/3), "
                          f"J ~ N(0,1), h ~ N(0,0.5), seed 4, T=30, {HHEX_STEPS} Trotter steps, chi={HHEX_CHI}; "
                          f"{HHEX_TIMED} steps timed once the bond dimension is {HHEX_CHI}; fp32 = complex64 "
                          "tensors/messages on FP32 MFMA with FP64 canonicalisers, fp64 = complex128 throughout")
        extra["heavy_hex_chi16"] = hx
        if not args.no_hhex32:
            # the same lattice at chi = 32, FP32 storage: no chi = 32 kernels, so the generic node
            # kernels and the generic canonicaliser run it; a short timed window
            ctx = config_to_context(hhex_config(32, HHEX_STEPS, "single"))
            log(f"[bench] heavy-hex {ctx.nodes_number} qubits chi=32 fp32 (generic kernels): {HHEX32_TIMED} steps "
                f"timed at chi=32")
            r = run_anneal(ctx, "single", args.warmup, HHEX32_TIMED, finish=False)
            h32 = summarize(r, ctx.nodes_number, HHEX32_TIMED)
            h32.pop("full_anneal")
            h32["workload"] = (f"the heavy-hex lattice above at chi=32, fp32 storage, {HHEX32_TIMED} steps timed once "
                               "the bond dimension is 32; generic node kernels and canonicaliser")
            extra["heavy_hex_chi32"] = h32
            log(f"[bench] heavy-hex chi=32 fp32: {h32['qubit_steps_per_s']:.4g} qubit*steps/s, "
                f"{h32['ms_per_step']:.1f} ms/step")
    serving = None
    if args.serve_jobs > 0:
        serving = serving_sample(args.serve_jobs, degrees=(3, 4, 5), chi_cap={5: 4})
        serving["small_jobs_d3"] = serving_sample(args.serve_jobs, degrees=(3,), chi_cap=None)
    cpu, oracle_parity = None, None
    if not args.no_cpu:
        log(f"[bench] cpu baseline and oracle parity sample ({SAMPLE_QUBITS} qubits, {BASE_STEPS} steps)")
        cpu, oracle_parity = cpu_sample(args.chi, (args.precision, other))
        log(f"[bench] cpu baseline {cpu['value']:.4g} qubit*steps/s; oracle parity {oracle_parity}")
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
        "dtype": "f64" if args.precision == "double" else "f32",
        "data": "synthetic",
        "config": {"workload": f"{n}-qubit random 3-regular QUBO anneal (ints in [-10,10]), chi={args.chi}, "
                               f"T={TOTAL_TIME:g}, {total_steps} Trotter steps",
                   "qubits": n, "edges": E // 2, "chi": args.chi, "bp_eps": BP_EPS,
                   "max_bp_iter_number": BP_ITERS, "pinv_eps": PINV_EPS,
                   "precision": ("fp64 throughout (complex128 tensors and messages)" if args.precision == "double"
                                 else "fp32 tensors/messages (complex64), fp64 SVD/lambdas/energy"),
                   "device": arch},
        "timed_steps": s["timed_steps"],
        "chi_timed": s["chi_timed"],
        "chi_per_timed_step": [st["bond_dim"] for st in head["run"]["stats"][s["timed_steps"][0]:s["timed_steps"][1]]],
        "full_anneal": s["full_anneal"],
        "message_updates_per_s": round(E * s["bp_sweeps_total"] / max(1e-12, s["ms_per_step"] * args.steps / 1e3), 1),
        "bp_sweeps_per_step": s["bp_sweeps_per_step_mean"],
        "roofline": roofline,
        "cpu_baseline": cpu,
        "oracle_parity_sample": oracle_parity,
        "accuracy": accuracy_block,
        TAG[other]: other_line,
        "serving": serving,
    }
    result.update(extra)
    from bqa_amd.native import tuning_report
    result["autotuner"] = [{"precision": r["precision"], "degree": r["degree"], "chi": r["chi"], "n_log2": r["n_log2"],
                            "chosen": r["variants"][r["chosen"]]["name"],
                            "us": {v["name"]: v["us"] for v in r["variants"]}} for r in tuning_report()]
    if copy_ms is not None:
        result["copy_calibration"] = {"bytes": args.copy_calibrate, "ms_per_copy": copy_ms,
                                      "GB_per_s": 2 * args.copy_calibrate / 1e6 / copy_ms}
    print(json.dumps(result))


if __name__ == "__main__":
    main()
