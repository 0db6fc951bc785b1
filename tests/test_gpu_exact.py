"""GPU exact state-vector validator (src/bqa/exact_sim.py:14-88) against the NumPy one and closed forms.

* n <= 18: amplitudes and Bloch vectors equal oracle/exact_sv.py (1e-12).
* 26 qubits, beyond the NumPy simulator: on a tree, BP with chi = 8 is exact for
  three Trotter steps (bond dimension at most 2^3), so the FP64 BP path and the
  exact state vector must agree (1e-9) -- each validates the other.
* 30 qubits, zero couplings: every qubit evolves alone, so each Bloch vector has a
  2 x 2 closed form; this reaches every tile pass of the X layer.
* 30 qubits, zero fields: the global spin flip commutes with the schedule, so every
  <Z> vanishes.
"""
import math

import numpy as np
import pytest

from bqa_amd import instances
from bqa_amd.config import config_to_context
from bqa_amd.exact import run_qa_exact, state_for
from oracle import exact_sv

pytestmark = pytest.mark.gpu


def random_tree(n, seed, max_degree=3):
    """Edges of a random tree on n nodes with degrees at most max_degree (parent chosen among earlier nodes)."""
    rng = np.random.default_rng(seed)
    deg = np.zeros(n, dtype=int)
    edges = []
    for v in range(1, n):
        open_ = np.flatnonzero(deg[:v] < max_degree)
        u = int(rng.choice(open_))
        deg[u] += 1
        deg[v] += 1
        edges.append((u, v))
    return edges


def tree_config(n, seed, steps):
    rng = np.random.default_rng(seed + 1)
    edges = {e: float(rng.normal()) for e in random_tree(n, seed)}
    nodes = {v: float(rng.normal(0.0, 0.5)) for v in range(n)}
    return {"nodes": nodes, "edges": edges, "max_bond_dim": 8, "bp_eps": 1e-13, "max_bp_iter_number": 200,
            "pinv_eps": 1e-12, "schedule": instances.anneal_schedule(0.3 * steps, steps)}


@pytest.mark.parametrize("case", ["rr12", "grid3x6"])
def test_matches_numpy_state_vector(case):
    if case == "rr12":
        nodes, edges = instances.regular_uniform(12, 3, seed=51, field_sigma=0.5)
    else:
        nodes, edges = instances.grid_gaussian(3, 6, seed=52, field_sigma=0.3)
    cfg = {"nodes": nodes, "edges": edges, "schedule": instances.anneal_schedule(3.0, 15)}
    print(f"[gpu] exact sv vs numpy {case}", flush=True)
    want = np.array(exact_sv.run_qa_exact(cfg)[0][1])
    got = np.array(run_qa_exact(dict(cfg, backend="mi355x"))[0][1])
    assert np.abs(got - want).max() < 1e-12
    # amplitudes after two steps
    ctx = config_to_context(dict(cfg, backend="mi355x"))
    sv = state_for(ctx)
    try:
        sv.init_state()
        for ins in [i for i in ctx.instructions if isinstance(i, dict)][:2]:
            sv.run_layer(ins["xtime"], ins["ztime"])
        psi = sv.state()
    finally:
        sv.close()
    octx = exact_sv.make_context(cfg)
    n = octx.nodes_number
    e = exact_sv._z_energy(octx)
    ref = np.where(np.array([bin(x).count("1") % 2 for x in range(1 << n)]) == 1, -1.0, 1.0) * 2.0 ** (-n / 2)
    ref = ref.astype(np.complex128)
    for ins in [i for i in octx.instructions if isinstance(i, dict)][:2]:
        ref = exact_sv._apply_x_all(ref * np.exp(-1j * ins["ztime"] * e), n, ins["xtime"])
    assert np.abs(psi - ref).max() < 1e-13


def test_bp_on_a_tree_equals_exact_state_at_26_qubits():
    from bqa_amd import run_qa
    cfg = tree_config(26, 61, 3)
    print("[gpu] exact sv vs BP on a 26-qubit tree, 3 steps", flush=True)
    exact = np.array(run_qa_exact(dict(cfg, backend="mi355x"))[0][1])
    bp = np.array(run_qa(dict(cfg, backend="mi355x", precision="double"))[0][1])
    err = np.abs(bp - exact).max()
    print(f"[gpu]   max |bp - exact| = {err:.2e}", flush=True)
    assert err < 1e-9


def single_qubit_bloch(h, schedule_steps):
    """Bloch vector of one uncoupled qubit from |->: per step exp(-i ztime h Z) then exp(-i xtime X)."""
    psi = np.array([1.0, -1.0], dtype=np.complex128) / math.sqrt(2.0)
    for xt, zt in schedule_steps:
        psi = np.array([np.exp(-1j * zt * h), np.exp(1j * zt * h)]) * psi
        c, s = math.cos(xt), math.sin(xt)
        psi = np.array([[c, -1j * s], [-1j * s, c]]) @ psi
    rho = np.outer(psi, psi.conj())
    return np.array([(rho[0, 1] + rho[1, 0]).real, (rho[1, 0] - rho[0, 1]).imag, (rho[0, 0] - rho[1, 1]).real])


def test_uncoupled_30_qubits_match_the_closed_form():
    n = 30
    rng = np.random.default_rng(71)
    h = rng.normal(0.0, 0.7, size=n)
    nodes = {v: float(h[v]) for v in range(n)}
    edges = {(v, v + 1): 0.0 for v in range(n - 1)}
    cfg = {"nodes": nodes, "edges": edges, "schedule": instances.anneal_schedule(2.0, 6), "backend": "mi355x"}
    ctx = config_to_context(cfg)
    steps = [(i["xtime"], i["ztime"]) for i in ctx.instructions if isinstance(i, dict)]
    print("[gpu] exact sv, 30 uncoupled qubits", flush=True)
    got = np.array(run_qa_exact(cfg)[0][1])
    want = np.array([single_qubit_bloch(h[v], steps) for v in range(n)])
    assert np.abs(got - want).max() < 1e-10


def test_z2_symmetric_30_qubits_have_zero_z():
    nodes, edges = instances.regular_uniform(30, 3, seed=72, field_sigma=0.0)
    cfg = {"nodes": nodes, "edges": edges, "schedule": instances.anneal_schedule(2.0, 4), "backend": "mi355x"}
    print("[gpu] exact sv, 30-qubit Z2-symmetric instance", flush=True)
    got = np.array(run_qa_exact(cfg)[0][1])
    assert np.abs(got[:, 2]).max() < 1e-12
    assert np.all(np.linalg.norm(got, axis=1) <= 1.0 + 1e-12)
    assert np.abs(got[:, :2]).max() > 1e-3          # the state moved
