"""Brute-force tree tensor-network contractions used as closed-form pins.

On a tree, BP converges to the exact environments, so its messages, densities
and lambdas have closed forms: contract every branch explicitly
(restating tests/test_core_subroutines.py:37-158 of the reference for any chi).
"""
from __future__ import annotations

import string

import numpy as np

# the 9-node tree of the reference's core-subroutine tests (test_core_subroutines.py:176-189)
SMALL_TREE_CONFIG = {
    "nodes": {1: 0.3, 3: -0.7, 5: 1., 6: -1., 7: 0.25},
    "edges": {(2, 0): 1., (1, 2): -1., (2, 4): 0.5, (4, 3): -0.5, (4, 5): 0.75,
              (4, 6): -0.75, (6, 7): 0.3, (8, 6): 0.6},
    "bp_eps": 1e-12,
    "pinv_eps": 1e-9,
    "default_field": 0.6,
    "max_bp_iter_number": 100,
}


def _letters():
    return iter(string.ascii_letters)


def _bond_letters(graph):
    it = _letters()
    bonds = {}
    for v, nbrs in enumerate(graph):
        for w in nbrs:
            key = (min(v, w), max(v, w))
            if key not in bonds:
                bonds[key] = next(it)
    return bonds, it


def full_state(graph, tensors) -> np.ndarray:
    """psi[p_0, ..., p_{n-1}] of a tree; tensors[v] has shape (2, chi...) in graph[v] bond order."""
    bonds, it = _bond_letters(graph)
    phys = [next(it) for _ in graph]
    subs = [phys[v] + "".join(bonds[(min(v, w), max(v, w))] for w in graph[v]) for v in range(len(graph))]
    return np.einsum(",".join(subs) + "->" + "".join(phys), *tensors, optimize=True)


def side_nodes(graph, u, v) -> list:
    """Nodes reachable from u without crossing the edge u-v."""
    seen, stack = {u}, [u]
    while stack:
        a = stack.pop()
        for b in graph[a]:
            if (a, b) in ((u, v),) or b in seen:
                continue
            seen.add(b)
            stack.append(b)
    return sorted(seen)


def branch_message(graph, tensors, u, v) -> np.ndarray:
    """m_{u->v}[c, e] = sum conj(B[.., c]) B[.., e] / trace, B = u's branch with bond u-v open."""
    nodes = side_nodes(graph, u, v)
    bonds, it = _bond_letters(graph)
    phys = {w: next(it) for w in nodes}
    subs = [phys[w] + "".join(bonds[(min(w, x), max(w, x))] for x in graph[w]) for w in nodes]
    open_b = bonds[(min(u, v), max(u, v))]
    out = "".join(phys[w] for w in nodes) + open_b
    b = np.einsum(",".join(subs) + "->" + out, *[tensors[w] for w in nodes], optimize=True)
    b = b.reshape(-1, b.shape[-1])
    m = b.conj().T @ b
    return m / np.trace(m)


def densities(psi: np.ndarray) -> np.ndarray:
    n = psi.ndim
    out = np.empty((n, 2, 2), dtype=np.complex128)
    for v in range(n):
        a = np.moveaxis(psi, v, 0).reshape(2, -1)
        r = a @ a.conj().T
        out[v] = r / np.trace(r)
    return out


def schmidt_values(graph, psi, u, v) -> np.ndarray:
    side = side_nodes(graph, u, v)
    rest = [w for w in range(psi.ndim) if w not in side]
    m = np.transpose(psi, side + rest).reshape(2 ** len(side), -1)
    s = np.linalg.svd(m, compute_uv=False)
    return s / np.linalg.norm(s)


def random_tree_tensors(ctx, chi: int, rng) -> dict:
    """Random complex node tensors per degree class, unit norm per node."""
    out = {}
    for d, lay in ctx_classes(ctx):
        shape = (len(lay_node_ids(lay)), 2) + (chi,) * d
        t = rng.normal(size=shape) + 1j * rng.normal(size=shape)
        t /= np.linalg.norm(t.reshape(shape[0], -1), axis=1).reshape((-1,) + (1,) * (t.ndim - 1))
        out[d] = t
    return out


def ctx_classes(ctx):
    """(degree, layout) pairs of an oracle or product context."""
    if hasattr(ctx, "classes"):
        return list(ctx.classes.items())
    return list(ctx.degree_to_layout.items())


def lay_node_ids(lay):
    return lay.node_ids


def per_node(ctx, tensors: dict) -> list:
    """Per-node tensor list (indexed by node id) from per-class arrays."""
    n = ctx.nodes_number
    out = [None] * n
    for d, lay in ctx_classes(ctx):
        for i, v in enumerate(lay_node_ids(lay)):
            out[int(v)] = tensors[d][i]
    return out
