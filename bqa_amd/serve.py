"""Serving loop: annealing jobs of mixed sizes arriving over time on one GPU.

Each request is a reference config (the keys ``run_qa`` takes, core.py:13-35).
Jobs are admitted in arrival order once they have arrived, a slot is free
(at most ``max_concurrent`` jobs resident) and their device footprint
(``bqa_job_estimate_bytes``) fits the remaining HBM budget.  Two executors:

* batched (the default): a few host threads, each stepping its resident jobs
  together through its own ``bqa_batch`` -- one call per Trotter step of all of
  them, with their canonicalisers and BP sweeps in launches shared per shape; a
  job joins a batch at its next step once admitted and leaves it after its last
  instruction;
* threaded: every job runs its instruction stream on its own host thread with
  its own ``bqa_job`` handle and HIP stream, so the sweeps of concurrent jobs
  overlap on the GPU.

Either way a job's results are those it gives alone with the same BP launch
variants (the kernels are deterministic and jobs share no device state).  The
shared launches compute what variant 0 computes; with the autotuner on
(``BQA_AUTOTUNE``, the default), a job alone may take a variant that orders a
node's sums differently, and then agrees with its served run to rounding.

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


class NativeBackend:
    """Jobs and batches of ``libbqa_hip.so`` for the batched executor."""

    @staticmethod
    def make_job(ctx, precision):
        from .native import Job
        job = Job(ctx, precision)
        try:
            job.init_state()
        except Exception:
            job.close()
            raise
        return job

    @staticmethod
    def make_batch(max_jobs):
        from .native import Batch
        return Batch(max_jobs)


def serve(requests, max_concurrent: int = 8, hbm_budget: int = DEFAULT_HBM_BUDGET, precision: str = "single",
          runner=None, footprint=None, batched: bool = True, backend=None, batch_workers: int = 4) -> dict:
    """Run the requests through the admission loop; see the module docstring.

    ``footprint(ctx, precision) -> bytes`` defaults to ``bqa_job_estimate_bytes``.  The
    batched executor takes its jobs and batch from ``backend`` (``NativeBackend``); the
    threaded one (``batched=False``, or any ``runner``) runs
    ``runner(ctx, precision) -> (results, steps, device_bytes)``, by default ``run_job``.
    Both are parameters so the scheduling can be exercised without a device.
    """
    if max_concurrent < 1:
        raise ValueError("max_concurrent must be at least 1")
    if runner is not None:
        batched = False
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
    if batched:
        outs = _serve_batched(prepared, max_concurrent, hbm_budget, precision, backend or NativeBackend,
                              workers=batch_workers)
        return _summary(outs, max_concurrent)
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
    return _summary([outcomes[i] for i in sorted(outcomes)], max_concurrent)


def _serve_batched(prepared, max_concurrent, hbm_budget, precision, backend, workers: int = 4) -> list:
    """The batched executor: ``workers`` host threads, each stepping its own resident jobs together through
    its own ``Batch`` (one ``Batch.run_layer`` call per round; readouts between steps run on the job itself),
    so one batch's kernels run while another waits on its latency-bound phases.  Admission (arrival order,
    the slots, the shared HBM budget) is common to all of them."""
    workers = max(1, min(workers, max_concurrent))
    caps = [max_concurrent // workers + (1 if w < max_concurrent % workers else 0) for w in range(workers)]
    lock = threading.Lock()
    queue = list(prepared)
    outcomes: dict[int, Outcome] = {}
    used = [0]
    t0 = time.perf_counter()
    errors = []

    def record(idx, outcome, need):
        with lock:
            outcomes[idx] = outcome
            used[0] -= need

    def worker(cap):
        batch = backend.make_batch(cap)
        resident = []

        def finish(r, err):
            dbytes = 0
            try:
                dbytes = r["job"].device_bytes()
            except Exception:
                pass
            try:
                r["job"].close()
            except Exception as e:  # reported with the job
                err = err or f"{type(e).__name__}: {e}"
            req, ctx = r["req"], r["ctx"]
            record(r["idx"], Outcome(req.tag, r["results"] if err is None else [], req.arrival, r["start"],
                                     time.perf_counter() - t0, ctx.nodes_number, r["steps"], dbytes, err), r["need"])
            resident.remove(r)

        try:
            while True:
                now = time.perf_counter() - t0
                taken, wait = [], None
                with lock:
                    while queue and queue[0][1].arrival <= now and len(resident) + len(taken) < cap and \
                            used[0] + queue[0][3] <= hbm_budget:
                        item = queue.pop(0)
                        used[0] += item[3]
                        taken.append(item)
                    if not queue and not resident and not taken:
                        break
                    if queue and queue[0][1].arrival > now:
                        wait = queue[0][1].arrival - now
                for idx, req, ctx, need in taken:
                    try:
                        job = backend.make_job(ctx, precision)
                    except Exception as e:  # recorded per job; the loop keeps serving the others
                        record(idx, Outcome(req.tag, [], req.arrival, now, time.perf_counter() - t0,
                                            ctx.nodes_number, 0, 0, f"{type(e).__name__}: {e}"), need)
                        continue
                    resident.append({"idx": idx, "req": req, "ctx": ctx, "need": need, "start": now, "job": job,
                                     "pc": 0, "steps": 0, "results": []})
                if not resident:
                    time.sleep(min(0.01, wait) if wait is not None else 0.002)
                    continue
                stepping = []
                for r in list(resident):
                    ins = r["ctx"].instructions
                    try:
                        while r["pc"] < len(ins) and not isinstance(ins[r["pc"]], dict):
                            if ins[r["pc"]] != "get_bloch_vectors":
                                raise NotImplementedError(f"instruction {ins[r['pc']]!r} is not served")
                            r["results"].append(["bloch_vectors", r["job"].density()[1].tolist()])
                            r["pc"] += 1
                    except Exception as e:
                        finish(r, f"{type(e).__name__}: {e}")
                        continue
                    if r["pc"] >= len(ins):
                        finish(r, None)
                    else:
                        stepping.append(r)
                if not stepping:
                    continue
                lays = [r["ctx"].instructions[r["pc"]] for r in stepping]
                cx = [r["ctx"] for r in stepping]
                try:
                    outs = batch.run_layer([r["job"] for r in stepping], [l["xtime"] for l in lays],
                                           [l["ztime"] for l in lays], [c.pinv_eps for c in cx],
                                           [c.max_bp_iters_number for c in cx], [c.bp_eps for c in cx],
                                           [c.damping for c in cx])
                except Exception as e:  # the whole step failed: every job in it is reported
                    for r in stepping:
                        finish(r, f"{type(e).__name__}: {e}")
                    continue
                for r, o in zip(stepping, outs):
                    if o.get("error"):
                        finish(r, f"AssertionError: {o['error']}")
                    else:
                        r["pc"] += 1
                        r["steps"] += 1
        except Exception as e:  # an executor failure: its jobs are reported below
            errors.append(f"{type(e).__name__}: {e}")
        finally:
            for r in list(resident):
                finish(r, errors[-1] if errors else "RuntimeError: the serving loop stopped")
            batch.close()

    threads = [threading.Thread(target=worker, args=(c,), daemon=True) for c in caps]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    with lock:
        for idx, req, ctx, need in queue:  # only left if every executor failed
            outcomes[idx] = Outcome(req.tag, [], req.arrival, 0.0, time.perf_counter() - t0, ctx.nodes_number, 0, 0,
                                    errors[-1] if errors else "RuntimeError: not served")
    return [outcomes[i] for i in sorted(outcomes)]


def _summary(outs, max_concurrent) -> dict:
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
                   degrees=(3, 4), chis=(4, 8), total_time: float = 10.0, chi_cap=None) -> list[Request]:
    """Seeded mixed workload: Poisson arrivals at ``rate_per_s``; n log-uniform in [n_min, n_max] (n d even);
    degree and bond dimension drawn from ``degrees`` and ``chis`` (the bond dimension then capped at
    ``chi_cap[degree]`` where given); J ~ U[-1, 1], h = 0; Trotter steps 100..300 growing with log n
    (``instances.serving_steps``)."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    for k in range(count):
        t += rng.exponential(1.0 / rate_per_s)
        d = int(rng.choice(degrees))
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        if (n * d) % 2:
            n += 1
        chi = int(rng.choice(chis))
        if chi_cap and d in chi_cap:
            chi = min(chi, int(chi_cap[d]))
        nodes, edges = instances.regular_uniform(n, d, seed=int(rng.integers(1 << 30)), field_sigma=0.0)
        sched = instances.anneal_schedule(total_time, instances.serving_steps(n))
        sched["actions"][0]["type"] = "real_time_evolution"
        cfg = {"nodes": nodes, "edges": edges, "max_bond_dim": chi, "schedule": sched, "backend": "mi355x",
               "default_field": 0.0, "measurement_threshold": 0.95, "seed": 42}
        out.append(Request(cfg, t, tag={"n": n, "degree": d, "chi": chi}))
    return out
