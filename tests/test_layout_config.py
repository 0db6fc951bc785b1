"""Degree-grouped layout compiler and config DSL (subsystem (a)).

The integer KAT is the reference's tests/test_config_to_context.py:35-165,
restated for both the product compiler (bqa_amd.config) and the oracle's.
"""
import math

import numpy as np
import pytest

from bqa_amd import instances
from bqa_amd.config import ConfigSyntaxError, config_to_context
from oracle import bqa_oracle as O

KAT_CONFIG = {
    "edges": {(0, 2): 1.0, (1, 0): -1, (3, 0): 0.1, (2, 4): 1.1, (1, 4): 0, (3, 1): 1},
    "nodes": {2: 1, 6: -1.1},
    "default_field": -0.5,
    "schedule": {
        "starting_mixing": 0.8,
        "total_time": 5,
        "actions": [
            {"weight": 0.4, "final_mixing": 0.3, "steps_number": 8},
            "measure",
            {"type": "imag_time_evolution", "weight": 0.6, "final_mixing": 0.11, "steps_number": 10},
            "get_bloch_vectors",
        ],
    },
    "damping": 0.3,
}

# tests/test_config_to_context.py:68-119
KAT = {
    0: dict(ids=[5, 6], inp=[], outp=[], lp=[], h=[-0.5, -1.1], j=[]),
    2: dict(ids=[2, 3, 4], inp=[[9, 8, 3], [0, 11, 4]], outp=[[3, 2, 9], [6, 5, 10]],
            lp=[[3, 2, 3], [0, 5, 4]], h=[1., -0.5, -0.5], j=[[1.1, 0.1, 1.1], [1., 1., 0.]]),
    3: dict(ids=[0, 1], inp=[[6, 7], [1, 10], [2, 5]], outp=[[0, 1], [7, 4], [8, 11]],
            lp=[[0, 1], [1, 4], [2, 5]], h=[-0.5, -0.5], j=[[1., -1.], [-1., 0.], [0.1, 1.]]),
}


def _check_kat(ids, inp, outp, lp, h, j, want):
    assert list(np.asarray(ids)) == want["ids"]
    assert np.asarray(inp).reshape(-1).tolist() == np.asarray(want["inp"], dtype=int).reshape(-1).tolist()
    assert np.asarray(outp).reshape(-1).tolist() == np.asarray(want["outp"], dtype=int).reshape(-1).tolist()
    assert np.asarray(lp).reshape(-1).tolist() == np.asarray(want["lp"], dtype=int).reshape(-1).tolist()
    assert np.allclose(np.asarray(h), want["h"])
    assert np.allclose(np.asarray(j).reshape(-1), np.asarray(want["j"]).reshape(-1))


def test_product_layout_kat():
    ctx = config_to_context(KAT_CONFIG)
    assert ctx.edges_number == 12 and ctx.nodes_number == 7
    assert ctx.max_bp_iters_number == 75 and math.isclose(ctx.bp_eps, 1e-6)
    assert math.isclose(ctx.damping, 0.3) and ctx.max_bond_dim == 4
    assert list(ctx.degree_to_layout) == [3, 2, 0]          # order of first node id
    for d, want in KAT.items():
        lay = ctx.degree_to_layout[d]
        _check_kat(lay.node_ids, lay.input_msgs_position, lay.output_msgs_position, lay.lmbds_position,
                   lay.node_ampls, lay.edge_ampls, want)
    for node_id, (degree, pos) in ctx.path_to_tensors.items():
        assert ctx.degree_to_layout[degree].node_ids[pos] == node_id


def test_oracle_layout_kat():
    ctx = O.make_context(KAT_CONFIG)
    for d, want in KAT.items():
        lay = ctx.classes[d]
        _check_kat(lay.node_ids, lay.input_pos, lay.output_pos, lay.lmbd_pos, lay.node_ampls, lay.edge_ampls, want)


def test_schedule_kat():
    """Time sums and linear mixing (test_config_to_context.py:120-162)."""
    ctx = config_to_context(KAT_CONFIG)
    steps = [i for i in ctx.instructions if isinstance(i, dict)]
    assert ctx.instructions[8] == "measure" and ctx.instructions[-1] == "get_bloch_vectors"
    assert math.isclose(sum(s["xtime"] + s["ztime"] for s in steps), 5.0)
    assert math.isclose(sum(s["xtime"] + s["ztime"] for s in steps[:8]), 2.0)
    mix = [s["xtime"] / (s["xtime"] + s["ztime"]) for s in steps]
    assert math.isclose(mix[0], 0.8) and math.isclose(mix[8], 0.3)
    d1 = np.diff(mix[:8] + [mix[8]])
    d2 = np.diff(mix[8:] + [0.11])
    assert np.allclose(d1, d1[0]) and np.allclose(d2, d2[0])
    o = O.make_context(KAT_CONFIG)
    assert o.instructions == ctx.instructions


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_product_layout_equals_oracle_on_random_graphs(seed):
    rng = np.random.default_rng(seed)
    n = 40
    edges = {}
    while len(edges) < 70:
        a, b = (int(x) for x in rng.integers(0, n, 2))
        if a != b and (a, b) not in edges and (b, a) not in edges:
            edges[(a, b)] = float(rng.normal())
    nodes = {int(v): float(rng.normal()) for v in rng.choice(n, 10, replace=False)}
    cfg = {"edges": edges, "nodes": nodes, "default_field": 0.25}
    p = config_to_context(cfg)
    o = O.make_context(cfg)
    assert list(p.degree_to_layout) == list(o.classes)
    for d in o.classes:
        a, b = p.degree_to_layout[d], o.classes[d]
        assert np.array_equal(a.node_ids, b.node_ids)
        assert np.array_equal(a.input_msgs_position, b.input_pos)
        assert np.array_equal(a.output_msgs_position, b.output_pos)
        assert np.array_equal(a.lmbds_position, b.lmbd_pos)
        assert np.array_equal(a.node_ampls, b.node_ampls)
        assert np.array_equal(a.edge_ampls, b.edge_ampls)
    assert p.graph() == o.graph


@pytest.mark.parametrize("bad", [
    {"edges": {(0, 0): 1.0}},                       # self loop
    {"edges": {(0, 1): 1.0, (1, 0): 2.0}},          # duplicated edge
    {"edges": {(0, 1): "x"}},                       # non-number coupling
    {"edges": {(0, 1): 1.0}, "max_bond_dim": 0},
    {"edges": {(0, 1): 1.0}, "pinv_eps": 2.0},
    {"edges": {(0, 1): 1.0}, "damping": -0.1},
    {"edges": {(0, 1): 1.0}, "measurement_threshold": 0.3},
    {"edges": {(0, 1): 1.0}, "backend": "cupy"},
    {"edges": {(0, 1): 1.0}, "nodes": {(1,): 0.5}},
    {"edges": {(0, 1): 1.0}, "schedule": {"actions": [{"weight": 0.5}]}},    # weights must sum to 1
    {"edges": {(0, 1): 1.0}, "schedule": {"actions": [{"weight": 1.0, "initial_mixing": 0.3}]}},
    {"edges": {(0, 1): 1.0}, "precision": "half"},
    [("edges", 1)],
    {"nodes": {0: 1.0}},                            # edges are required
])
def test_invalid_configs_raise_config_syntax_error(bad):
    with pytest.raises(ConfigSyntaxError):
        config_to_context(bad)
    assert issubclass(ConfigSyntaxError, ValueError)


def test_defaults_and_ignored_keys():
    ctx = config_to_context({"edges": [((0, 1), 1.0), ((1, 2), -1.0)], "runtime_limit": 1, "description": "x"})
    assert ctx.backend == "mi355x" and ctx.max_bond_dim == 4 and ctx.max_bp_iters_number == 75
    assert ctx.pinv_eps == 1e-6 and ctx.damping == 0.0 and ctx.seed == 42
    assert len([i for i in ctx.instructions if isinstance(i, dict)]) == 100
    assert ctx.instructions[-1] == "get_bloch_vectors"


def test_instances_are_simple_regular_graphs():
    e = instances.random_regular_edges(1000, 3, seed=1)
    assert e.shape == (1500, 2) and np.all(e[:, 0] < e[:, 1])
    assert np.unique(e[:, 0] * 1000 + e[:, 1]).size == 1500
    assert np.all(np.bincount(e.ravel(), minlength=1000) == 3)
    assert np.array_equal(e, instances.random_regular_edges(1000, 3, seed=1))


def test_qubo_to_ising_energy_identity():
    rng = np.random.default_rng(0)
    n = 12
    e = instances.random_regular_edges(n, 3, seed=2)
    qe = rng.integers(-10, 11, size=len(e))
    qd = rng.integers(-10, 11, size=n)
    j, h, const = instances.qubo_to_ising(n, e, qe, qd)
    for _ in range(20):
        x = rng.integers(0, 2, size=n)
        s = 1 - 2 * x
        f = float(qd @ x + sum(q * x[a] * x[b] for (a, b), q in zip(e, qe)))
        ising = float(h @ s + sum(jj * s[a] * s[b] for (a, b), jj in zip(e, j))) + const
        assert math.isclose(f, ising, abs_tol=1e-9)


def test_heavy_hex_degrees():
    n, e = instances.heavy_hex_edges(40, 100)
    deg = np.bincount(e.ravel(), minlength=n)
    assert 9500 < n < 10500
    assert set(np.unique(deg)) <= {1, 2, 3}
    assert 0.55 < np.mean(deg == 2) < 0.65
