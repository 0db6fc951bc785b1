"""Graph-captured BP chunks (through the C ABI).

Once a job has run BP eight times in a row at one bond dimension, its sweeps are
replays of a HIP graph holding four sweeps of every degree class.  The replays launch the same
kernels with the same launch variants as the eager path, so an anneal gives the
same Bloch vectors and BP sweep counts bit for bit with graphs on and off; jobs
capturing on several host threads at once give what each gives alone.
"""
import os
import threading

import numpy as np
import pytest

from bqa_amd import instances
from bqa_amd.config import config_to_context
from bqa_amd.native import Job

pytestmark = pytest.mark.gpu

CASES = {
    "rr60_chi8": lambda: instances.regular_qubo(60, 3, seed=21, qmax=10)[:2],
    "heavyhex_chi8": lambda: instances.heavy_hex_gaussian(3, 4, seed=23, field_sigma=0.5),
    "rr40_chi4": lambda: instances.regular_uniform(40, 3, seed=24, field_sigma=0.4),
    "grid5x7_chi8": lambda: instances.grid_gaussian(5, 7, seed=22, field_sigma=0.1),
}
# The heavy-hex patch has degree-1 nodes, whose sweep runs the generic kernel: its BP stays eager.
USES_GRAPHS = {"rr60_chi8": True, "heavyhex_chi8": False, "rr40_chi4": True, "grid5x7_chi8": True}


def _ctx(name, steps=40):
    nodes, edges = CASES[name]()
    chi = 4 if name.endswith("chi4") else 8
    return config_to_context({"nodes": nodes, "edges": edges, "max_bond_dim": chi, "bp_eps": 1e-9,
                              "max_bp_iter_number": 60, "schedule": instances.anneal_schedule(4.0, steps),
                              "backend": "mi355x"})


def _make_job(ctx, precision, graphs):
    old = os.environ.get("BQA_GRAPHS")
    os.environ["BQA_GRAPHS"] = "1" if graphs else "0"
    try:
        return Job(ctx, precision)
    finally:
        if old is None:
            os.environ.pop("BQA_GRAPHS", None)
        else:
            os.environ["BQA_GRAPHS"] = old


def _anneal(job, ctx):
    job.init_state()
    sweeps, chis = [], []
    for ins in ctx.instructions:
        if isinstance(ins, dict):
            out = job.run_layer(ins["xtime"], ins["ztime"], ctx.pinv_eps, ctx.max_bp_iters_number, ctx.bp_eps,
                                ctx.damping)
            sweeps.append(out["sweeps"])
            chis.append(out["bond_dim"])
    _, bloch = job.density()
    return bloch, sweeps, chis


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_graph_chunks_match_eager(name, precision):
    ctx = _ctx(name)
    res = {}
    for graphs in (True, False):
        job = _make_job(ctx, precision, graphs)
        try:
            res[graphs] = _anneal(job, ctx) + (job.graph_stats(),)
        finally:
            job.close()
    (bg, sg, cg, stats_g), (be, se, ce, stats_e) = res[True], res[False]
    print(f"[gpu] graphs {name} {precision}: graph stats {stats_g}, eager {stats_e}, sweeps {sum(sg)}", flush=True)
    assert stats_e == (0, 0, 0)
    assert stats_g[2] == 0
    if USES_GRAPHS[name]:
        assert stats_g[0] > 0 and stats_g[1] > 0
    elif USES_GRAPHS[name] is False:
        assert stats_g == (0, 0, 0)
    assert sg == se and cg == ce
    assert np.array_equal(bg, be)


def test_concurrent_capture_matches_alone():
    """Four jobs annealing on four host threads at once, each capturing its own graphs while the
    others create jobs, allocate and copy, give the results each gives alone, with no failed
    capture."""
    work = [("rr60_chi8", "single"), ("rr60_chi8", "double"), ("rr40_chi4", "single"),
            ("rr40_chi4", "double")]
    ctxs = [_ctx(n, steps=60) for n, _ in work]
    alone = []
    for ctx, (_, prec) in zip(ctxs, work):
        job = _make_job(ctx, prec, True)
        try:
            alone.append(_anneal(job, ctx))
        finally:
            job.close()
    together = [None] * len(work)
    errors = []

    def run(k):
        try:
            job = _make_job(ctxs[k], work[k][1], True)
            try:
                together[k] = _anneal(job, ctxs[k]) + (job.graph_stats(),)
            finally:
                job.close()
        except Exception as e:  # reported below, with the job it came from
            errors.append((k, repr(e)))

    threads = [threading.Thread(target=run, args=(k,)) for k in range(len(work))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for k, (a, t) in enumerate(zip(alone, together)):
        print(f"[gpu] concurrent {work[k]}: graph stats {t[3]}", flush=True)
        assert t[3][1] > 0 and t[3][2] == 0
        assert t[1] == a[1] and t[2] == a[2]
        assert np.array_equal(t[0], a[0])
