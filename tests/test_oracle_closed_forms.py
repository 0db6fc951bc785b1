"""Closed-form pins of the FP64 oracle (no reference import needed).

* BP on a tree converges to the exact branch environments for any chi
  (reference test_core_subroutines.py:191-204 checks chi = 1 only);
* BP densities on a tree equal those of the fully contracted state (:207-220);
* Vidal lambdas equal the Schmidt values across every edge (:141-158, 223-256), and the
  Vidal-gauge state (tensors with lambda on each edge) is the same state with identity
  branch environments (test_vg_and_sg_small_tree_circuit, :223-256);
* gate identities (test_gpt_generated_gates_application.py:26-101) and the
  ZZ split identity g0^2 I + g1^2 ZZ = exp(-i theta ZZ);
* end to end against the exact state vector (test_small_circuit_final_density.py:9-20);
* the exact simulator itself against dense matrix exponentials.
"""
import math

import numpy as np
import pytest
import scipy.linalg

from oracle import bqa_oracle as O
from oracle.exact_sv import run_qa_exact
from treeutil import (SMALL_TREE_CONFIG, branch_message, densities, full_state, per_node,
                      random_tree_tensors, schmidt_values)


def _tree_state(chi, seed):
    ctx = O.make_context(SMALL_TREE_CONFIG)
    st = O.init_state(ctx)
    rng = np.random.default_rng(seed)
    st.tensors = random_tree_tensors(ctx, chi, rng)
    st.chi = chi
    st.lmbds = np.ones((ctx.lmbds_number, chi), dtype=O.DTYPE)
    st.msgs = O.msgs_from_lmbds(ctx, st.lmbds)
    return ctx, st


@pytest.mark.parametrize("chi", [1, 2, 3])
def test_bp_tree_messages_equal_branch_contraction(chi):
    ctx, st = _tree_state(chi, 42 + chi)
    O.run_bp(ctx, st)
    tens = per_node(ctx, st.tensors)
    for (u, v), pos in ctx.edge_to_msg_pos.items():
        want = branch_message(ctx.graph, tens, u, v)
        assert np.abs(st.msgs[pos] - want).max() < 1e-10


@pytest.mark.parametrize("chi", [1, 3])
def test_bp_tree_densities_equal_full_contraction(chi):
    ctx, st = _tree_state(chi, 7 + chi)
    O.run_bp(ctx, st)
    rho = O.get_density_matrices(ctx, st)
    psi = full_state(ctx.graph, per_node(ctx, st.tensors))
    assert np.abs(rho - densities(psi)).max() < 1e-10


def test_tree_lambdas_equal_schmidt_values():
    ctx, st = _tree_state(2, 43)
    O.run_bp(ctx, st)
    lam, canon = O.canonicalizers(st.msgs, ctx.pinv_eps)
    psi = full_state(ctx.graph, per_node(ctx, st.tensors))
    for (u, v), pos in list(ctx.edge_to_msg_pos.items())[: ctx.lmbds_number]:
        want = schmidt_values(ctx.graph, psi, u, v)
        r = min(want.size, lam.shape[1])
        assert np.abs(lam[pos, :r].real - want[:r]).max() < 1e-9
        assert np.abs(want[r:]).max(initial=0.0) < 1e-9


def test_symmetric_gauge_preserves_the_tree_state():
    """Canonicalise (lambda and C absorbed), then the symmetric gauge: same state up to norm."""
    ctx, st = _tree_state(2, 44)
    O.run_bp(ctx, st)
    psi0 = full_state(ctx.graph, per_node(ctx, st.tensors))
    lam, canon = O.canonicalizers(st.msgs, ctx.pinv_eps)
    for d in st.tensors:
        lay = ctx.classes[d]
        for k in range(d):
            st.tensors[d] = O.right_product(st.tensors[d], canon[lay.input_pos[k]], k + 2)
    st.lmbds = lam
    O.set_symmetric_gauge(ctx, st)
    psi1 = full_state(ctx.graph, per_node(ctx, st.tensors))
    ov = abs(np.vdot(psi0, psi1)) / (np.linalg.norm(psi0) * np.linalg.norm(psi1))
    assert abs(ov - 1.0) < 1e-10


def _one_qubit_state(amp):
    ctx = O.make_context({"edges": {(0, 1): 0.0}})
    st = O.init_state(ctx)
    st.tensors = {1: np.array(amp, dtype=O.DTYPE).reshape(1, 2, 1).repeat(2, axis=0)}
    return ctx, st


def test_gate_identities():
    ctx, st = _one_qubit_state([0.6, 0.8])
    t0 = st.tensors[1].copy()
    O.apply_x_layer(ctx, st, math.pi)                       # Rx(pi) = -i X ... = -I up to X: cos pi = -1
    assert np.allclose(st.tensors[1], -t0)
    ctx, st = _one_qubit_state([math.sqrt(0.5), -math.sqrt(0.5)])
    t0 = st.tensors[1].copy()
    O.apply_x_layer(ctx, st, 0.37)                           # |-> is an X eigenstate
    r0 = np.einsum("np,nq->npq", t0[:, :, 0], t0[:, :, 0].conj())
    r1 = np.einsum("np,nq->npq", st.tensors[1][:, :, 0], st.tensors[1][:, :, 0].conj())
    assert np.allclose(r0, r1, atol=1e-12)
    ctx = O.make_context({"edges": {(0, 1): 0.0}, "nodes": {0: 1.0, 1: 1.0}})
    st = O.init_state(ctx)
    st.tensors = {1: np.array([0.6, 0.8], dtype=O.DTYPE).reshape(1, 2, 1).repeat(2, axis=0)}
    t0 = st.tensors[1].copy()
    O.apply_z_layer(ctx, st, math.pi)                        # Rz(pi) = -I
    assert np.allclose(st.tensors[1], -t0)


@pytest.mark.parametrize("theta", [0.0, 0.3, -0.7, 1.2, 2.0])
def test_zz_split_reconstructs_the_gate(theta):
    g0, g1 = O.zz_split(np.array([theta]))
    z = np.diag([1.0, -1.0])
    gate = g0[0] ** 2 * np.eye(4) + g1[0] ** 2 * np.kron(z, z)
    assert np.allclose(gate, scipy.linalg.expm(-1j * theta * np.kron(z, z)), atol=1e-12)


def test_conditional_z_doubles_bond_and_zero_coupling_is_identity():
    t = np.array([[[0.5], [0.8]]], dtype=O.DTYPE)            # (1, 2, 1)
    ext = O.extend_axis(t, 2, np.array([0.0]))
    assert ext.shape == (1, 2, 2)
    assert np.allclose(ext[:, :, 0], t[:, :, 0]) and np.allclose(ext[:, :, 1], 0.0)


SIX_NODE_CONFIG = {   # tests/test_small_circuit_final_density.py:9-15 of the reference
    "nodes": {0: 1., 1: -1., 2: 0.5, 3: -0.5, 4: 1.1, 5: 0.4},
    "edges": {(0, 1): 1., (2, 1): -1., (1, 3): 1., (4, 3): -1., (5, 3): 1.},
    "pinv_eps": 1e-9,
    "bp_eps": 1e-9,
    "max_bond_dim": 8,
}


def test_oracle_matches_exact_state_vector_on_tree():
    a = np.array(O.run_qa(SIX_NODE_CONFIG)[0][1])
    b = np.array(run_qa_exact(SIX_NODE_CONFIG)[0][1])
    assert np.abs(a - b).max() < 1e-5


def test_exact_sv_against_dense_evolution():
    cfg = {"nodes": {0: 0.3, 1: -0.8, 2: 0.5}, "edges": {(0, 1): 0.7, (1, 2): -1.1, (0, 2): 0.4},
           "schedule": {"total_time": 1.5, "starting_mixing": 1.0,
                        "actions": [{"weight": 1.0, "steps_number": 6, "final_mixing": 0.0},
                                    "get_bloch_vectors"]}}
    got = np.array(run_qa_exact(cfg)[0][1])
    ctx = O.make_context(cfg)
    n = 3
    x, z, i2 = np.array([[0, 1], [1, 0]]), np.diag([1.0, -1.0]), np.eye(2)

    def op(single, q):      # qubit q is bit q: kron order is (q = n-1, ..., 0)
        mats = [single if k == q else i2 for k in reversed(range(n))]
        out = mats[0]
        for m in mats[1:]:
            out = np.kron(out, m)
        return out

    hz = sum(ctx.node_to_ampl[v] * op(z, v) for v in range(n))
    for (a, b), j in list(ctx.edge_to_ampl.items())[:3]:
        hz = hz + j * op(z, a) @ op(z, b)
    hx = sum(op(x, v) for v in range(n))
    minus = np.array([1.0, -1.0]) / math.sqrt(2)
    psi = minus
    for _ in range(n - 1):
        psi = np.kron(psi, minus)
    for ins in ctx.instructions:
        if isinstance(ins, dict):
            psi = scipy.linalg.expm(-1j * ins["xtime"] * hx) @ (scipy.linalg.expm(-1j * ins["ztime"] * hz) @ psi)
    want = np.empty((n, 3))
    for v in range(n):
        want[v] = [np.vdot(psi, op(x, v) @ psi).real, np.vdot(psi, op(np.array([[0, -1j], [1j, 0]]), v) @ psi).real,
                   np.vdot(psi, op(z, v) @ psi).real]
    assert np.abs(got - want).max() < 1e-12


def test_z2_symmetric_instance_has_zero_magnetisation():
    """h = 0: the global spin flip commutes with every gate and |->^n is flip-even."""
    from bqa_amd import instances
    nodes, edges = instances.regular_pm1(12, 3, seed=3)
    cfg = {"nodes": nodes, "edges": edges, "max_bond_dim": 4, "bp_eps": 1e-9, "max_bp_iter_number": 100,
           "schedule": instances.anneal_schedule(2.0, 20)}
    ex = np.array(run_qa_exact(cfg)[0][1])
    orc = np.array(O.run_qa(cfg)[0][1])
    assert np.abs(ex[:, 2]).max() < 1e-12
    assert np.abs(orc[:, 2]).max() < 1e-8


def vidal_branch(graph, tensors, lmbds_of, u, v):
    """m_{u->v} of the Vidal-gauge state: u's branch with lambda on every edge inside it
    (test_core_subroutines.py:37-70 with is_vidal_gauge)."""
    from treeutil import side_nodes
    scaled = {}
    nodes = side_nodes(graph, u, v)
    for w in nodes:
        t = tensors[w]
        for k, x in enumerate(graph[w]):
            if x in nodes and w < x:     # each inner edge scaled once
                lam = lmbds_of(w, x)
                shape = [1] * t.ndim
                shape[k + 1] = lam.size
                t = t * lam.reshape(shape)
        scaled[w] = t
    full = [scaled.get(w, tensors[w]) for w in range(len(graph))]
    return branch_message(graph, full, u, v)


@pytest.mark.parametrize("chi", [2, 3])
def test_vidal_gauge_on_a_tree(chi):
    """test_vg_and_sg_small_tree_circuit restated: after BP and the Vidal gauge, lambdas are the Schmidt
    values, every branch environment (with lambdas inside) is the identity / rank on the kept block, the
    Vidal-gauge state with lambda on each edge is the original state, and the symmetric gauge then BP
    gives back the same densities."""
    ctx, st = _tree_state(chi, 45 + chi)
    O.run_bp(ctx, st)
    psi = full_state(ctx.graph, per_node(ctx, st.tensors))
    rho0 = O.get_density_matrices(ctx, st)
    O.set_to_vidal_gauge(ctx, st)
    tens = per_node(ctx, st.tensors)
    lam_of = lambda a, b: st.lmbds[ctx.edge_to_msg_pos[(a, b)] % ctx.lmbds_number].real
    for (u, v), pos in ctx.edge_to_msg_pos.items():
        want = schmidt_values(ctx.graph, psi, u, v)
        lam = lam_of(u, v)
        r = int(np.sum(want > 1e-9))
        assert np.abs(lam[:r] - want[:r]).max() < 1e-9 and np.abs(lam[r:]).max(initial=0.0) < 1e-9
        env = vidal_branch(ctx.graph, tens, lam_of, u, v)
        assert np.abs(env[:r, :r] - np.eye(r) / r).max() < 1e-9
        assert np.abs(env[r:]).max(initial=0.0) < 1e-9 and np.abs(env[:, r:]).max(initial=0.0) < 1e-9
    # Vidal-gauge state: one lambda per edge
    scaled = list(tens)
    done = set()
    for v in range(ctx.nodes_number):
        for k, w in enumerate(ctx.graph[v]):
            if (w, v) in done:
                continue
            done.add((v, w))
            shape = [1] * scaled[v].ndim
            shape[k + 1] = chi
            scaled[v] = scaled[v] * lam_of(v, w).reshape(shape)
    psi_vg = full_state(ctx.graph, scaled)
    ov = abs(np.vdot(psi, psi_vg)) / (np.linalg.norm(psi) * np.linalg.norm(psi_vg))
    assert abs(ov - 1.0) < 1e-10
    O.set_symmetric_gauge(ctx, st)
    O.run_bp(ctx, st)
    assert np.abs(O.get_density_matrices(ctx, st) - rho0).max() < 1e-10


def test_truncate_vidal_gauge_keeps_the_global_rank():
    """The truncation after the Vidal gauge keeps min(max_bond_dim, rank) columns on every bond."""
    ctx, st = _tree_state(3, 61)
    O.run_bp(ctx, st)
    O.set_to_vidal_gauge(ctx, st)
    lam = st.lmbds.real.copy()
    ctx.max_bond_dim = 2
    err = O.truncate_vidal_gauge(ctx, st)
    assert st.chi == 2 and st.lmbds.shape == (ctx.lmbds_number, 2)
    for d, t in st.tensors.items():
        assert t.shape[2:] == (2,) * d
    assert abs(err - np.sqrt(np.sum(lam.max(0)[2:] ** 2))) < 1e-12
