"""Serving loop: annealing jobs of mixed sizes arriving over time on one GPU.

Each request is a reference config (the keys ``run_qa`` takes, core.py:13-35).
Jobs are admitted in arrival order once they have arrived, a slot is free
(at most ``max_concurrent`` jobs resident) and their device footprint
(``bqa_job_estimate_bytes``) fits the remaining HBM budget.  Every admitted
job runs its instruction stream on its own host thread with its own
``bqa_job`` handle and HIP stream, so the sweeps of concurrent jobs overlap on
the GPU; a job's results are those it gives alone (the kernels are
deterministic and jobs share no device state).

``serve`` returns each job's results with its arrival, start and finish
times, and the aggregate throughput (jobs/h, qubit-steps/s) with p50/p99
time-to-solution (arrival to result).
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass

import numpy as np

from .config import config_to_context
from . import instances

DEFAULT_HBM_BUDGET = 200 * 2**30  # bytes of the 288 GB of HBM on one MI355X given to resident jobs


@dataclass
class Request:
    config: dict
    arrival: float = 0.0      # seconds after the loop starts
    tag: object = None


@dataclass
class Outcome:
    tag: object
    results: list
    arrival: float
    start: float
    finish: float
    qubits: int
    steps: int
    device_bytes: int
    error: str | None = None

    @property
    def latency(self) -> float:
        return self.finish - self.arrival


def run_job(ctx, precision: str) -> tuple[list, int, int]:
    """Run one context's instruction stream on its own job; returns (results, steps, device bytes)."""
    from .native import Job
    job = Job(ctx, precision)
    try:
        job.init_state()
        results, steps = [], 0
        for ins in ctx.instructions:
            if isinstance(ins, dict):
                job.run_layer(ins["xtime"], ins["ztime"], ctx.pinv_eps, ctx.max_bp_iters_number, ctx.bp_eps,
                              ctx.damping)
                steps += 1
            elif ins == "get_bloch_vectors":
                results.append(["bloch_vectors", job.density()[1].tolist()])
            else:
                raise NotImplementedError(f"instruction {ins!r} is not served")
        return results, steps, job.device_bytes()
    finally:
        job.close()


def _percentile(v, q):
    return float(np.percentile(np.asarray(v, dtype=float), q)) if len(v) else None


def serve(requests, max_concurrent: int = 8, hbm_budget: int = DEFAULT_HBM_BUDGET, precision: str = "single",
          runner=None, footprint=None) -> dict:
    """Run the requests through the admission loop; see the module docstring.

    ``runner(ctx, precision) -> (results, steps, device_bytes)`` and
    ``footprint(ctx, precision) -> bytes`` default to the GPU job and
    ``bqa_job_estimate_bytes``; they are parameters so the scheduling can be
    exercised without a device.
    """
    if max_concurrent < 1:
        raise ValueError("max_concurrent must be at least 1")
    if runner is None:
        runner = run_job
    if footprint is None:
        from .native import estimate_bytes
        footprint = lambda ctx, prec: estimate_bytes(ctx, ctx.max_bond_dim, prec)  # noqa: E731
    pending = sorted(enumerate(requests), key=lambda x: (x[1].arrival, x[0]))
    prepared = []
    for idx, req in pending:
        ctx = config_to_context(dict(req.config))
        need = int(footprint(ctx, precision))
        if need > hbm_budget:
            raise MemoryError(f"request {req.tag!r} needs {need} bytes, more than the HBM budget {hbm_budget}")
        prepared.append((idx, req, ctx, need))
    outcomes: dict[int, Outcome] = {}
    lock = threading.Condition()
    state = {"running": 0, "bytes": 0}
    t0 = time.perf_counter()

    def worker(idx, req, ctx, need, start):
        err, res, steps, dbytes = None, [], 0, 0
        try:
            res, steps, dbytes = runner(ctx, precision)
        except Exception as e:  # recorded per job; the loop keeps serving the others
            err = f"{type(e).__name__}: {e}"
        fin = time.perf_counter() - t0
        with lock:
            outcomes[idx] = Outcome(req.tag, res, req.arrival, start, fin, ctx.nodes_number, steps, dbytes, err)
            state["running"] -= 1
            state["bytes"] -= need
            lock.notify_all()

    threads = []
    queue = list(prepared)
    with lock:
        while queue or state["running"]:
            now = time.perf_counter() - t0
            while queue and queue[0][1].arrival <= now and state["running"] < max_concurrent and \
                    state["bytes"] + queue[0][3] <= hbm_budget:
                idx, req, ctx, need = queue.pop(0)
                state["running"] += 1
                state["bytes"] += need
                th = threading.Thread(target=worker, args=(idx, req, ctx, need, now), daemon=True)
                th.start()
                threads.append(th)
            timeout = None
            if queue and queue[0][1].arrival > now:
                timeout = queue[0][1].arrival - now
            if queue or state["running"]:
                lock.wait(timeout=timeout if timeout is not None else 1.0)
    for th in threads:
        th.join()
    outs = [outcomes[i] for i in sorted(outcomes)]
    ok = [o for o in outs if o.error is None]
    makespan = max((o.finish for o in outs), default=0.0)
    lat = [o.latency for o in ok]
    qsteps = sum(o.qubits * o.steps for o in ok)
    return {
        "outcomes": outs,
        "jobs": len(outs),
        "failed": len(outs) - len(ok),
        "makespan_s": makespan,
        "jobs_per_h": 3600.0 * len(ok) / makespan if makespan > 0 else 0.0,
        "qubit_steps_per_s": qsteps / makespan if makespan > 0 else 0.0,
        "p50_s": _percentile(lat, 50),
        "p99_s": _percentile(lat, 99),
        "max_concurrent": max_concurrent,
    }


def mixed_requests(count: int, seed: int = 5, n_min: int = 100, n_max: int = 5000, rate_per_s: float = 8.0,
                   degrees=(3, 4), chis=(4, 8), total_time: float = 10.0) -> list[Request]:
    """Seeded mixed workload: Poisson arrivals at ``rate_per_s``; n log-uniform in [n_min, n_max] (n d even);
    degree and bond dimension drawn from ``degrees`` and ``chis``; J ~ U[-1, 1], h = 0; Trotter steps
    100..300 growing with log n (``instances.serving_steps``)."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    for k in range(count):
        t += rng.exponential(1.0 / rate_per_s)
        d = int(rng.choice(degrees))
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        if (n * d) % 2:
            n += 1
        chi = int(rng.choice(chis))
        nodes, edges = instances.regular_uniform(n, d, seed=int(rng.integers(1 << 30)), field_sigma=0.0)
        sched = instances.anneal_schedule(total_time, instances.serving_steps(n))
        sched["actions"][0]["type"] = "real_time_evolution"
        cfg = {"nodes": nodes, "edges": edges, "max_bond_dim": chi, "schedule": sched, "backend": "mi355x",
               "default_field": 0.0, "measurement_threshold": 0.95, "seed": 42}
        out.append(Request(cfg, t, tag={"n": n, "degree": d, "chi": chi}))
    return out
