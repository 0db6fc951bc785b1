"""Write golden vectors by running the reference bqa (NumPy backend, FP64).

Run in the build container only (the reference tree is not on the GPU box):

    python tests/golden/make_golden.py [/path/to/reference]

The reference package ``bqa`` calls importlib.metadata.version("bqa") in its
``__init__`` (src/bqa/__init__.py:3) and is not pip-installable offline, so the
package object is registered in ``sys.modules`` with its source directory as
``__path__``; its submodules then import unchanged.

Each fixture is an ``.npz`` (no pickles) holding the instance (forward edge
list, couplings, fields), the run parameters as JSON text, and gauge-invariant
outputs only: Bloch vectors, the bond dimension after every Trotter step, the
number of BP sweeps in every step, the final lambdas (sorted per edge) and,
for schedules that end in ``"measure"``, the sampled outcomes (+1/-1 per qubit).
"""
from __future__ import annotations

import json
import logging
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from bqa_amd import instances  # noqa: E402


def load_reference(ref_root: str):
    os.environ.setdefault("BQA_PRECISION", "double")
    src = os.path.join(ref_root, "src", "bqa")
    pkg = types.ModuleType("bqa")
    pkg.__path__ = [src]
    sys.modules["bqa"] = pkg
    import bqa.state as state  # noqa: E402
    from bqa.config.core import config_to_context  # noqa: E402
    from bqa.utils import convert_density_matrix_to_bloch_vector  # noqa: E402
    return state, config_to_context, convert_density_matrix_to_bloch_vector


class _SweepCounter(logging.Handler):
    def __init__(self):
        super().__init__(logging.DEBUG)
        self.count = 0

    def emit(self, record):
        if record.getMessage().startswith("Iteration "):
            self.count += 1


def run_reference(ref, config: dict):
    state_mod, config_to_context, to_bloch = ref
    counter = _SweepCounter()
    lg = logging.getLogger("bqa.state")
    lg.setLevel(logging.DEBUG)
    lg.addHandler(counter)
    try:
        ctx = config_to_context(config)
        st = state_mod._initialize_state(ctx)
        chis, sweeps, bloch, outcomes = [], [], None, []
        for ins in ctx.instructions:
            if isinstance(ins, dict):
                counter.count = 0
                state_mod.run_layer(ctx, ins["xtime"], ins["ztime"], st)
                chis.append(st.lmbds.batch_shape[0])
                sweeps.append(counter.count)
            elif ins == "get_bloch_vectors":
                rho = state_mod.get_density_matrices(ctx, st)
                bloch = np.array([to_bloch(r) for r in rho])
            elif ins == "measure":
                # sequential sampling with the state's seeded generator (state.py:250-312)
                outcomes.append(state_mod.measure(ctx, st))
        lmbds = np.sort(st.lmbds.numpy.real, axis=1)[:, ::-1]
        return bloch, np.array(chis), np.array(sweeps), lmbds, np.array(outcomes, dtype=np.int8)
    finally:
        lg.removeHandler(counter)


def save(name: str, nodes: dict, edges: dict, params: dict, outputs):
    bloch, chis, sweeps, lmbds, outcomes = outputs
    e = np.array(list(edges.keys()), dtype=np.int64)
    j = np.array(list(edges.values()), dtype=np.float64)
    n = 1 + max(max(nodes.keys(), default=-1), int(e.max()))
    h = np.array([nodes.get(v, 0.0) for v in range(n)], dtype=np.float64)
    extra = {"outcomes": outcomes} if outcomes.size else {}
    np.savez_compressed(os.path.join(HERE, name + ".npz"), edges=e, couplings=j, fields=h,
                        params=np.array(json.dumps(params)), bloch=bloch, chis=chis,
                        sweeps=sweeps, lmbds=lmbds, **extra)
    print(f"{name}: n={n} E={len(e)} steps={len(chis)} chi_final={chis[-1]} "
          f"sweeps={int(sweeps.sum())}", flush=True)


def cases():
    """(name, nodes, edges, params) of every fixture."""
    out = []
    nodes, edges = instances.regular_pm1(100, 3, seed=1)
    p1 = {"max_bond_dim": 4, "bp_eps": 1e-8, "max_bp_iter_number": 200, "pinv_eps": 1e-6,
          "damping": 0.0, "schedule": instances.anneal_schedule(10.0, 100)}
    out.append(("rr100_pm1_chi4", nodes, edges, p1))
    rng = np.random.default_rng(11)
    nodes_h = {v: float(rng.normal(0.0, 0.1)) for v in range(100)}
    out.append(("rr100_fields_chi4", nodes_h, edges, p1))
    p_nc = dict(p1, max_bp_iter_number=3, schedule=instances.anneal_schedule(2.0, 12))
    out.append(("rr100_nonconverged", nodes_h, edges, p_nc))
    p_damp = dict(p1, damping=0.4, schedule=instances.anneal_schedule(2.0, 20))
    out.append(("rr100_damped", nodes_h, edges, p_damp))
    g_nodes, g_edges = instances.grid_gaussian(6, 6, seed=5, field_sigma=0.1)
    p_g = {"max_bond_dim": 8, "bp_eps": 1e-10, "max_bp_iter_number": 200, "pinv_eps": 1e-6,
           "damping": 0.0, "schedule": instances.anneal_schedule(2.0, 20)}
    out.append(("grid6x6_chi8", g_nodes, g_edges, p_g))
    s_nodes, s_edges = instances.regular_uniform(16, 3, seed=9, field_sigma=0.3)
    p_s = {"max_bond_dim": 4, "bp_eps": 1e-10, "max_bp_iter_number": 200, "pinv_eps": 1e-6,
           "damping": 0.0, "schedule": instances.anneal_schedule(3.0, 30)}
    out.append(("rr16_fields_chi4", s_nodes, s_edges, p_s))
    hh_nodes, hh_edges = instances.heavy_hex_gaussian(3, 5, seed=6, field_sigma=0.5)
    p_hh = {"max_bond_dim": 6, "bp_eps": 1e-10, "max_bp_iter_number": 200, "pinv_eps": 1e-6,
            "damping": 0.0, "schedule": instances.anneal_schedule(1.5, 15)}
    out.append(("heavyhex_chi6", hh_nodes, hh_edges, p_hh))
    # bond dimensions above 8 (the generic kernels): heavy-hex, degrees 2 and 3, with pinv_eps small enough
    # that the bond dimension ramps to 16 (and to 30 of at most 32) within the schedule
    h16_nodes, h16_edges = instances.heavy_hex_gaussian(3, 5, seed=12, field_sigma=0.5)
    p_16 = {"max_bond_dim": 16, "bp_eps": 1e-10, "max_bp_iter_number": 200, "pinv_eps": 1e-12,
            "damping": 0.0, "schedule": instances.anneal_schedule(3.0, 16)}
    out.append(("heavyhex_chi16", h16_nodes, h16_edges, p_16))
    p_32 = dict(p_16, max_bond_dim=32)
    out.append(("heavyhex_chi32", h16_nodes, h16_edges, p_32))
    # sequential measurement sampling after an anneal (seeded generator, state.py:70,250-312)
    m_nodes, m_edges = instances.regular_uniform(24, 3, seed=17, field_sigma=0.4)
    sched = instances.anneal_schedule(4.0, 20)
    sched["actions"].append("measure")
    p_m = {"max_bond_dim": 4, "bp_eps": 1e-10, "max_bp_iter_number": 200, "pinv_eps": 1e-6,
           "damping": 0.0, "schedule": sched, "seed": 1234, "measurement_threshold": 0.95}
    out.append(("rr24_measure_chi4", m_nodes, m_edges, p_m))
    g_sched = instances.anneal_schedule(2.0, 12)
    g_sched["actions"].append("measure")
    gm_nodes, gm_edges = instances.grid_gaussian(4, 5, seed=19, field_sigma=0.3)
    p_gm = {"max_bond_dim": 8, "bp_eps": 1e-10, "max_bp_iter_number": 200, "pinv_eps": 1e-6,
            "damping": 0.0, "schedule": g_sched, "seed": 7, "measurement_threshold": 0.9}
    out.append(("grid4x5_measure_chi8", gm_nodes, gm_edges, p_gm))
    return out


def main():
    """python tests/golden/make_golden.py [reference root] [fixture name ...] (default: every fixture)."""
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    only = set(sys.argv[2:])
    ref = load_reference(ref_root)
    for name, nodes, edges, params in cases():
        if only and name not in only:
            continue
        cfg = dict(params, nodes=nodes, edges=edges, backend="numpy")
        save(name, nodes, edges, params, run_reference(ref, cfg))


if __name__ == "__main__":
    main()
