"""Serving loop: admission (concurrency, HBM budget, arrival order) with a stand-in runner on the CPU,
and on the GPU the results of jobs served together equal those of each job run alone."""
import threading
import time

import numpy as np
import pytest

from bqa_amd import instances
from bqa_amd.serve import Request, mixed_requests, serve


def _cfg(n, seed, chi=2, steps=4):
    nodes, edges = instances.regular_uniform(n, 3, seed=seed, field_sigma=0.3)
    return {"nodes": nodes, "edges": edges, "max_bond_dim": chi, "backend": "mi355x",
            "schedule": instances.anneal_schedule(1.0, steps)}


class FakeRunner:
    """Sleeps per job and records the peak concurrency and device bytes in use."""

    def __init__(self, footprint_by_n):
        self.lock = threading.Lock()
        self.running = 0
        self.bytes = 0
        self.peak = 0
        self.peak_bytes = 0
        self.fp = footprint_by_n

    def footprint(self, ctx, prec):
        return self.fp(ctx.nodes_number)

    def __call__(self, ctx, prec):
        with self.lock:
            self.running += 1
            self.bytes += self.fp(ctx.nodes_number)
            self.peak = max(self.peak, self.running)
            self.peak_bytes = max(self.peak_bytes, self.bytes)
        time.sleep(0.05)
        with self.lock:
            self.running -= 1
            self.bytes -= self.fp(ctx.nodes_number)
        if ctx.nodes_number == 13:
            raise RuntimeError("planted failure")
        return [["bloch_vectors", [[0.0, 0.0, 1.0]] * ctx.nodes_number]], 4, 100


def test_admission_respects_slots_budget_and_arrivals():
    reqs = [Request(_cfg(10 + 2 * k, k), arrival=0.01 * k, tag=k) for k in range(12)]
    fr = FakeRunner(lambda n: 1000 * n)
    out = serve(reqs, max_concurrent=3, hbm_budget=60_000, runner=fr, footprint=fr.footprint)
    assert out["jobs"] == 12 and out["failed"] == 0
    assert fr.peak <= 3
    assert fr.peak_bytes <= 60_000
    for o in out["outcomes"]:
        assert o.start >= o.arrival - 1e-9 and o.finish >= o.start
    assert out["p50_s"] <= out["p99_s"]
    assert out["jobs_per_h"] > 0


def test_failed_job_is_reported_and_others_complete():
    reqs = [Request(_cfg(n, n), tag=n) for n in (10, 12, 14)]
    reqs.append(Request(_cfg(14, 99), tag="bad"))
    reqs[3].config["nodes"] = {i: 0.0 for i in range(13)}  # 13 nodes: the stand-in runner fails on it
    reqs[3].config["edges"] = {(i, (i + 1) % 13): 1.0 for i in range(13)}
    fr = FakeRunner(lambda n: 10)
    out = serve(reqs, max_concurrent=2, runner=fr, footprint=fr.footprint)
    bad = [o for o in out["outcomes"] if o.tag == "bad"]
    assert len(bad) == 1 and "planted failure" in bad[0].error
    assert out["failed"] == 1
    assert sum(o.error is None for o in out["outcomes"]) == 3


def test_oversized_request_is_rejected():
    fr = FakeRunner(lambda n: 10**12)
    with pytest.raises(MemoryError):
        serve([Request(_cfg(10, 1))], hbm_budget=10**9, runner=fr, footprint=fr.footprint)


def test_mixed_requests_are_seeded_and_valid():
    a = mixed_requests(5, seed=3, n_max=300)
    b = mixed_requests(5, seed=3, n_max=300)
    assert [r.tag for r in a] == [r.tag for r in b]
    assert all(x.arrival < y.arrival for x, y in zip(a, a[1:]))
    assert all(100 <= r.tag["n"] <= 301 and (r.tag["n"] * r.tag["degree"]) % 2 == 0 for r in a)


class FakeJob:
    def __init__(self, ctx, fail_bp=False):
        self.ctx, self.fail_bp, self.closed = ctx, fail_bp, False

    def density(self):
        return None, np.zeros((self.ctx.nodes_number, 3))

    def device_bytes(self):
        return 100

    def close(self):
        self.closed = True


class FakeBackend:
    """Stand-in jobs and batch: records how many jobs each batched step held and the bytes resident."""

    def __init__(self, footprint_by_n, fail_create=(), fail_bp=()):
        self.fp, self.fail_create, self.fail_bp = footprint_by_n, set(fail_create), set(fail_bp)
        self.jobs, self.widths, self.peak_bytes = [], [], 0
        outer = self

        class FakeBatch:
            def __init__(self, max_jobs):
                self.max_jobs, self.closed = max_jobs, False

            def run_layer(self, jobs, xtime, ztime, pinv_eps, max_iters, bp_eps, damping):
                assert len(jobs) <= self.max_jobs and len(set(map(id, jobs))) == len(jobs)
                assert all(not j.closed for j in jobs)
                outer.widths.append(len(jobs))
                live = sum(outer.fp(j.ctx.nodes_number) for j in outer.jobs if not j.closed)
                outer.peak_bytes = max(outer.peak_bytes, live)
                time.sleep(0.002)
                return [{"bond_dim": 1, "trunc_err": 0.0, "sweeps": 1, "converged": True, "dist": 0.0,
                         "error": "BP produced no finite message distance" if j.fail_bp else None} for j in jobs]

            def close(self):
                self.closed = True

        self.batch_cls = FakeBatch

    def make_job(self, ctx, precision):
        if ctx.nodes_number in self.fail_create:
            raise MemoryError("planted allocation failure")
        job = FakeJob(ctx, fail_bp=ctx.nodes_number in self.fail_bp)
        self.jobs.append(job)
        return job

    def make_batch(self, max_jobs):
        self.batch = self.batch_cls(max_jobs)
        return self.batch


def test_batched_executor_admission_and_failures():
    """The batched executor steps at most max_concurrent jobs per call within the HBM budget, in
    arrival order; a job whose creation or BP fails is reported and the others complete."""
    reqs = [Request(_cfg(10 + 2 * k, k, steps=3 + k % 3), arrival=0.005 * k, tag=k) for k in range(10)]
    fb = FakeBackend(lambda n: 1000 * n, fail_create={14}, fail_bp={20})
    out = serve(reqs, max_concurrent=3, hbm_budget=70_000, footprint=lambda ctx, p: 1000 * ctx.nodes_number,
                backend=fb)
    assert out["jobs"] == 10 and out["failed"] == 2
    errs = {o.tag: o.error for o in out["outcomes"] if o.error}
    assert "planted allocation failure" in errs[2] and "no finite" in errs[5]
    assert max(fb.widths) <= 3 and fb.peak_bytes <= 70_000
    assert all(j.closed for j in fb.jobs) and fb.batch.closed
    for r, o in zip(reqs, out["outcomes"]):
        assert o.start >= o.arrival - 1e-9 and o.finish >= o.start
        if o.error is None:
            steps = r.config["schedule"]["actions"][0]["steps_number"]
            assert o.steps == steps and len(o.results) == 1 and o.results[0][0] == "bloch_vectors"


def _mix_cfg(n, d, chi, seed, steps):
    nodes, edges = instances.regular_uniform(n, d, seed=seed, field_sigma=0.3)
    return {"nodes": nodes, "edges": edges, "max_bond_dim": chi, "backend": "mi355x",
            "schedule": instances.anneal_schedule(2.0, steps)}


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("batched", [True, False])
def test_served_jobs_equal_jobs_alone(batched, precision, monkeypatch):
    """Jobs of degrees 3, 4 and 5 at chi 4 and 8 served together (batched: shared BP launches;
    threaded: one stream each) give the Bloch vectors of each job run alone, bit for bit.

    The autotuner is off: with it on, a job run alone may take a BP variant that splits a
    node's Grams over waves differently from the shared launches (same sums, another order),
    and the two then agree to rounding only."""
    from bqa_amd.config import config_to_context
    monkeypatch.setenv("BQA_AUTOTUNE", "0")
    from bqa_amd.serve import run_job
    mix = ((24, 3, 4), (40, 3, 8), (60, 3, 4), (30, 4, 8), (36, 4, 4), (20, 5, 4), (26, 3, 8))
    reqs = [Request(_mix_cfg(n, d, c, 40 + k, 10 + 2 * k), arrival=0.0, tag=(n, d, c))
            for k, (n, d, c) in enumerate(mix)]
    print(f"[gpu] serving {len(reqs)} jobs together, batched={batched}, {precision}", flush=True)
    out = serve(reqs, max_concurrent=5, precision=precision, batched=batched)
    assert out["failed"] == 0, [o.error for o in out["outcomes"]]
    for r, o in zip(reqs, out["outcomes"]):
        alone, steps, _ = run_job(config_to_context(dict(r.config)), precision)
        assert o.steps == steps
        assert np.array_equal(np.array(alone[0][1]), np.array(o.results[0][1])), r.tag


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["single", "double"])
def test_autotuned_served_jobs_use_npw_shared_launch(precision, monkeypatch):
    """With the autotuner on, the batch runs degree-2/3 classes at chi <= 4 through
    k_small_npw_multi (four nodes per wave): each served job then equals, bit for bit, the job
    run alone with its BP forced to the four-nodes-per-wave variant (variant 4 at those shapes;
    the other shapes fall back to variant 0, which the other shared launches compute).  FP64
    jobs stay at chi <= 4, where every BP launch of the batch is a shared one."""
    from bqa_amd.config import config_to_context
    from bqa_amd.serve import run_job
    monkeypatch.setenv("BQA_AUTOTUNE", "1")
    monkeypatch.delenv("BQA_BP_VARIANT", raising=False)
    cap = 8 if precision == "single" else 4
    mix = ((24, 3, 4), (40, 3, cap), (62, 3, 4), (30, 2, 4), (36, 2, cap), (150, 3, 4))
    reqs = [Request(_mix_cfg(n, d, c, 70 + k, 10 + 2 * k), arrival=0.0, tag=(n, d, c))
            for k, (n, d, c) in enumerate(mix)]
    print(f"[gpu] serving {len(reqs)} autotuned jobs together, batched, {precision}", flush=True)
    out = serve(reqs, max_concurrent=6, precision=precision, batched=True, batch_workers=1)
    assert out["failed"] == 0, [o.error for o in out["outcomes"]]
    monkeypatch.setenv("BQA_BP_VARIANT", "4")
    for r, o in zip(reqs, out["outcomes"]):
        alone, steps, _ = run_job(config_to_context(dict(r.config)), precision)
        assert o.steps == steps
        assert np.array_equal(np.array(alone[0][1]), np.array(o.results[0][1])), r.tag
