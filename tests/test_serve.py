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


@pytest.mark.gpu
def test_served_jobs_equal_jobs_alone():
    from bqa_amd.config import config_to_context
    from bqa_amd.serve import run_job
    reqs = [Request(_cfg(n, 40 + n, chi=c, steps=8), arrival=0.0, tag=n) for n, c in ((24, 4), (40, 8), (60, 4), (30, 8))]
    print("[gpu] serving 4 jobs together", flush=True)
    out = serve(reqs, max_concurrent=4)
    assert out["failed"] == 0
    for r, o in zip(reqs, out["outcomes"]):
        alone, _, _ = run_job(config_to_context(dict(r.config)), "single")
        assert np.array_equal(np.array(alone[0][1]), np.array(o.results[0][1]))
