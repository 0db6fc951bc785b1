"""BP launch variants and the autotuner that chooses between them (through the C ABI).

Every variant of a shape (grid policies of the grid-stride node kernels, and the
kernels that split a node over a different number of waves) must produce the same
sweep as variant 0: bit for bit for a grid policy (each node is computed by one
block with the same code), to rounding for a different wave split.  The states
are random tensors and random positive messages at the target bond dimension on
instances large enough that every grid policy gives a different grid.
"""
import os

import numpy as np
import pytest

from bqa_amd import instances, native
from bqa_amd.config import config_to_context
from bqa_amd.native import Job

pytestmark = pytest.mark.gpu

# name: (instance, chi, precision, variants, tolerance of the wave-split variants relative to max |m|)
CASES = {
    # chi <= 4, degree 2 and 3: grid policies, then k_small_npw with 2 and 4 nodes per wave (another
    # fibre split of the Grams)
    "rr12000_chi4_fp32": (lambda: instances.regular_uniform(12000, 3, seed=41, field_sigma=0.3), 4, "single", 5,
                          2e-6),
    "rr12000_chi4_fp64": (lambda: instances.regular_uniform(12000, 3, seed=41, field_sigma=0.3), 4, "double", 5,
                          1e-13),
    "rr9000_chi3_fp32": (lambda: instances.regular_uniform(9000, 3, seed=45, field_sigma=0.3), 3, "single", 5, 2e-6),
    "grid90x90_chi2_fp64": (lambda: instances.grid_gaussian(90, 90, seed=46, field_sigma=0.1), 2, "double", 5,
                            1e-13),
    "rr6000_chi8_fp32": (lambda: instances.regular_uniform(6000, 3, seed=42, field_sigma=0.3), 8, "single", 3, 0.0),
    "rr6000_chi8_fp64": (lambda: instances.regular_uniform(6000, 3, seed=42, field_sigma=0.3), 8, "double", 4, 1e-13),
    "heavyhex_chi8_fp64": (lambda: instances.heavy_hex_gaussian(24, 24, seed=43, field_sigma=0.5), 8, "double", 4,
                           1e-13),
    "heavyhex_chi16_fp32": (lambda: instances.heavy_hex_gaussian(6, 6, seed=44, field_sigma=0.5), 16, "single", 3,
                            2e-6),
    # degree 4 at chi = 8: k_bp_d4c8 with 256 or 512 threads per node (another fibre split of the Grams),
    # then for FP64 k_bp_d4c8m64 (FP64 MFMA) with four or eight waves per node, and for FP32
    # k_bp_d4c8m (FP32 MFMA, FP32 Gram chains added in FP64)
    "grid24x24_chi8_fp64": (lambda: instances.grid_gaussian(24, 24, seed=47, field_sigma=0.1), 8, "double", 4,
                            1e-13),
    "grid24x24_chi8_fp32": (lambda: instances.grid_gaussian(24, 24, seed=47, field_sigma=0.1), 8, "single", 3,
                            2e-6),
}


def _config(nodes, edges, chi):
    return {"nodes": nodes, "edges": edges, "max_bond_dim": chi, "schedule": instances.anneal_schedule(1.0, 2),
            "backend": "mi355x"}


def _random_state(ctx, chi, seed):
    rng = np.random.default_rng(seed)
    tensors = []
    for d, lay in ctx.degree_to_layout.items():
        n = lay.node_ids.size
        shape = (n, 2) + (chi,) * d
        tensors.append((rng.normal(size=shape) + 1j * rng.normal(size=shape)) / chi ** (d / 2))
    b = rng.normal(size=(ctx.edges_number, chi, chi)) + 1j * rng.normal(size=(ctx.edges_number, chi, chi))
    m = b @ np.conj(np.swapaxes(b, 1, 2)) + chi * np.eye(chi)
    m /= np.trace(m, axis1=1, axis2=2).real[:, None, None]
    return tensors, m


def _job(ctx, precision, state, env):
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        job = Job(ctx, precision)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    try:
        job.init_state()
        tensors, msgs = state
        for ci, t in enumerate(tensors):
            job.set_tensors(ci, t)
        job.set_messages(msgs)
    except Exception:
        job.close()
        raise
    return job


@pytest.mark.parametrize("name", sorted(CASES))
def test_bp_variants_match_variant0(name):
    make, chi, precision, nvar, tol = CASES[name]
    nodes, edges = make()
    ctx = config_to_context(_config(nodes, edges, chi))
    state = _random_state(ctx, chi, seed=7)
    print(f"[gpu] bp variants {name}: {ctx.nodes_number} nodes, {nvar} variants", flush=True)
    out = []
    for v in range(nvar):
        job = _job(ctx, precision, state, {"BQA_BP_VARIANT": str(v)})
        try:
            sweeps, _, dist = job.run_bp(1, 0.0, 0.0)
            out.append((job.get_messages(), dist))
        finally:
            job.close()
    ref, dref = out[0]
    scale = np.abs(ref).max()
    for v, (m, dist) in enumerate(out[1:], start=1):
        diff = np.abs(m - ref).max() / scale
        print(f"[gpu]   variant {v}: max relative message difference {diff:.2e}, distance {dist:.6e} / {dref:.6e}",
              flush=True)
        assert diff <= tol
        assert abs(dist - dref) <= max(tol, 1e-15) * 10 * abs(dref)


def test_autotuner_table_and_results():
    """With the autotuner on, the first BP run at a shape times its variants and records the
    fastest; the run's messages equal those of variant 0 to rounding."""
    nodes, edges = instances.regular_uniform(6000, 3, seed=42, field_sigma=0.3)
    ctx = config_to_context(_config(nodes, edges, 8))
    state = _random_state(ctx, 8, seed=8)
    native.tuning_reset()
    runs = []
    for env in ({"BQA_AUTOTUNE": "1"}, {"BQA_AUTOTUNE": "0"}):
        job = _job(ctx, "double", state, env)
        try:
            sweeps, _, dist = job.run_bp(4, 0.0, 0.0)
            runs.append((sweeps, dist, job.get_messages()))
        finally:
            job.close()
    table = native.tuning_report()
    print(f"[gpu] autotuner table: {table}", flush=True)
    rows = [r for r in table if r["precision"] == "double" and r["degree"] == 3 and r["chi"] == 8]
    assert len(rows) == 1
    row = rows[0]
    assert row["n_log2"] == int(np.ceil(np.log2(6000)))
    assert len(row["variants"]) == 4 and 0 <= row["chosen"] < 4
    assert all(v["us"] > 0.0 for v in row["variants"])
    assert row["variants"][row["chosen"]]["us"] == min(v["us"] for v in row["variants"])
    (s0, d0, m0), (s1, d1, m1) = runs
    assert s0 == s1 == 4
    assert np.abs(m0 - m1).max() <= 1e-12 * np.abs(m1).max()
    native.tuning_reset()
    assert native.tuning_report() == []
