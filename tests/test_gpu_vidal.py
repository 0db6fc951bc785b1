"""GPU Vidal gauge (src/bqa/state.py:159-216) on a tree, against the oracle and closed forms.

The reference's own check (tests/test_core_subroutines.py:223-256) restated for
the GPU path: after BP and the Vidal gauge the lambdas are the Schmidt values,
the state rebuilt with one lambda per edge is the original state, and the
symmetric gauge followed by BP gives back the same densities.  The truncation
drops exactly the lambda columns that are zero on every edge.
"""
import numpy as np
import pytest

from bqa_amd.config import config_to_context
from bqa_amd.native import Job
from oracle import bqa_oracle as O
from treeutil import SMALL_TREE_CONFIG, densities, full_state, per_node, random_tree_tensors, schmidt_values

pytestmark = pytest.mark.gpu

TOL = {"double": 1e-10, "single": 2e-5}


def tree_state(chi, seed, pad_to=None):
    ctx_o = O.make_context(SMALL_TREE_CONFIG)
    st = O.init_state(ctx_o)
    st.tensors = random_tree_tensors(ctx_o, chi, np.random.default_rng(seed))
    st.chi = chi
    if pad_to:
        for d, t in st.tensors.items():
            out = np.zeros((t.shape[0], 2) + (pad_to,) * d, dtype=t.dtype)
            out[(slice(None), slice(None)) + (slice(0, chi),) * d] = t
            st.tensors[d] = out
        st.chi = chi = pad_to
    st.lmbds = np.ones((ctx_o.lmbds_number, chi), dtype=O.DTYPE)
    st.msgs = O.msgs_from_lmbds(ctx_o, st.lmbds)
    return ctx_o, st


def upload(job, st):
    for ci, d in enumerate(st.tensors):
        job.set_tensors(ci, st.tensors[d])
    job.set_lambdas(st.lmbds)
    job.set_messages(st.msgs)


def vidal_state(ctx_o, job):
    """Full state of the job's Vidal-gauge tensors with lambda once on every edge."""
    tens = per_node(ctx_o, {d: job.get_tensors(ci) for ci, d in enumerate(ctx_o.classes)})
    lam = job.get_lambdas().real
    done = set()
    for v in range(ctx_o.nodes_number):
        for k, w in enumerate(ctx_o.graph[v]):
            if (w, v) in done:
                continue
            done.add((v, w))
            l = lam[ctx_o.edge_to_msg_pos[(v, w)] % ctx_o.lmbds_number]
            shape = [1] * tens[v].ndim
            shape[k + 1] = l.size
            tens[v] = tens[v] * l.reshape(shape)
    return full_state(ctx_o.graph, tens)


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("chi", [2, 3])
def test_vidal_gauge_on_a_tree(chi, precision):
    ctx_o, st = tree_state(chi, 80 + chi)
    psi = full_state(ctx_o.graph, per_node(ctx_o, st.tensors))
    job = Job(config_to_context(dict(SMALL_TREE_CONFIG, max_bond_dim=4)), precision)
    tol = TOL[precision]
    print(f"[gpu] vidal gauge chi={chi} {precision}", flush=True)
    try:
        job.init_state()
        upload(job, st)
        job.run_bp(100, 1e-12 if precision == "double" else 1e-6, 0.0)
        rho0 = job.density()[0]
        job.set_vidal_gauge(ctx_o.pinv_eps)
        assert job.bond_dim == chi
        lam = job.get_lambdas().real
        for (u, v), pos in list(ctx_o.edge_to_msg_pos.items())[: ctx_o.lmbds_number]:
            want = schmidt_values(ctx_o.graph, psi, u, v)
            r = min(want.size, chi)
            assert np.abs(lam[pos, :r] - want[:r]).max() < tol
        psi_vg = vidal_state(ctx_o, job)
        ov = abs(np.vdot(psi, psi_vg)) / (np.linalg.norm(psi) * np.linalg.norm(psi_vg))
        assert abs(ov - 1.0) < tol
        job.set_symmetric_gauge()
        job.run_bp(100, 1e-12 if precision == "double" else 1e-6, 0.0)
        assert np.abs(job.density()[0] - rho0).max() < tol
        assert np.abs(rho0 - densities(psi)).max() < tol
    finally:
        job.close()


def test_truncate_vidal_gauge_drops_zero_columns():
    """A chi = 2 tree state zero-padded to 3: the Vidal lambdas' third column vanishes on every edge, so
    the truncation keeps 2 columns with zero error, and the state and its densities are unchanged."""
    ctx_o, st = tree_state(2, 90, pad_to=3)
    psi = full_state(ctx_o.graph, per_node(ctx_o, st.tensors))
    job = Job(config_to_context(dict(SMALL_TREE_CONFIG, max_bond_dim=4)), "double")
    print("[gpu] truncate vidal gauge chi 3 -> 2", flush=True)
    try:
        job.init_state()
        upload(job, st)
        job.run_bp(100, 1e-12, 0.0)
        rho0 = job.density()[0]
        job.set_vidal_gauge(ctx_o.pinv_eps)
        o2 = O.State(tensors={d: t.copy() for d, t in st.tensors.items()}, msgs=st.msgs.copy(),
                     lmbds=st.lmbds.copy(), chi=3)
        O.run_bp(ctx_o, o2)
        O.set_to_vidal_gauge(ctx_o, o2)
        want_err = O.truncate_vidal_gauge(ctx_o, o2)
        chi, err = job.truncate_vidal_gauge(ctx_o.pinv_eps)
        assert chi == o2.chi == 2
        assert abs(err - want_err) < 1e-10 and err < 1e-9
        assert job.get_lambdas().shape == (ctx_o.lmbds_number, 2)
        psi_vg = vidal_state(ctx_o, job)
        ov = abs(np.vdot(psi, psi_vg)) / (np.linalg.norm(psi) * np.linalg.norm(psi_vg))
        assert abs(ov - 1.0) < 1e-10
        job.set_symmetric_gauge()
        job.run_bp(100, 1e-12, 0.0)
        assert np.abs(job.density()[0] - rho0).max() < 1e-10
    finally:
        job.close()
