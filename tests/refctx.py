"""Reference-shaped Context stand-ins for the boundary tests (no reference import).

The reference's ``Context`` (src/bqa/config/config_canonicalization.py:144-168)
holds per-degree ``Layout``s (:25-33) whose fields are backend tensors: node ids,
one tensor per bond of input/output message positions, lambda positions and
couplings, and complex node fields (backends.py:581-584 casts float lists to
complex).  These stand-ins carry the same attribute names and tensor shapes,
built either from the reference's own layout KAT table or from the FP64
oracle's context of a committed fixture.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


class FakeTensor:
    """Only what the reference's tensors expose to a layout reader: ``.numpy``, ``.batch_size``."""

    def __init__(self, arr):
        self._a = np.asarray(arr)

    @property
    def numpy(self):
        return self._a

    @property
    def batch_size(self):
        return self._a.shape[0]


@dataclass
class FakeLayout:
    node_ids: FakeTensor
    input_msgs_position: list
    output_msgs_position: list
    lmbds_position: list
    node_ampls: FakeTensor
    edge_ampls: list


@dataclass
class FakeContext:
    backend: object
    nodes_number: int
    edges_number: int
    degree_to_layout: dict
    edge_to_ampl: dict
    instructions: list
    bp_eps: float
    pinv_eps: float
    measurement_threshold: float
    max_bond_dim: int
    max_bp_iters_number: int
    seed: int
    damping: float

    @property
    def lmbds_number(self):
        return self.edges_number // 2


def layout(ids, inp, outp, lp, h, j) -> FakeLayout:
    """A reference Layout from plain lists ([bond][node] for the per-bond fields)."""
    ints = lambda rows: [FakeTensor(np.array(r, dtype=np.intp)) for r in rows]
    cplx = lambda rows: [FakeTensor(np.array(r, dtype=np.complex128)) for r in rows]
    return FakeLayout(FakeTensor(np.array(ids, dtype=np.intp)), ints(inp), ints(outp), ints(lp),
                      FakeTensor(np.array(h, dtype=np.complex128)), cplx(j))


def from_oracle(cfg: dict, backend=None) -> FakeContext:
    """The reference-shaped context of a config, with the oracle's layouts (pinned to the reference's KAT)."""
    from oracle import bqa_oracle as O
    ctx = O.make_context(cfg)
    lays = {d: layout(c.node_ids.tolist(), c.input_pos.tolist(), c.output_pos.tolist(), c.lmbd_pos.tolist(),
                      c.node_ampls.tolist(), c.edge_ampls.tolist()) for d, c in ctx.classes.items()}
    return FakeContext(backend=backend, nodes_number=ctx.nodes_number, edges_number=ctx.edges_number,
                       degree_to_layout=lays, edge_to_ampl=dict(ctx.edge_to_ampl), instructions=ctx.instructions,
                       bp_eps=ctx.bp_eps, pinv_eps=ctx.pinv_eps, measurement_threshold=ctx.measurement_threshold,
                       max_bond_dim=ctx.max_bond_dim, max_bp_iters_number=ctx.max_bp_iters, seed=ctx.seed,
                       damping=ctx.damping)
