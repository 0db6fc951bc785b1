"""Seeded synthetic Ising instances for the annealing workloads.

Every generator returns ``(nodes, edges)`` in the reference's config format:
``nodes`` maps qubit id -> local field h_i and ``edges`` maps (i, j) -> J_ij,
with edges listed once and in a fixed order (the order defines message
positions, src/bqa/config/config_syntax.py:118-154).

The reference builds its instances with networkx (src/bqa/benchmarking.py:39-72);
these generators are this package's own, seeded with ``numpy.random.default_rng``
so the same seed gives the same instance on every machine.
"""
from __future__ import annotations

import math

import numpy as np


def random_regular_edges(n: int, degree: int, seed: int, max_attempts: int = 100000) -> np.ndarray:
    """Uniform random simple ``degree``-regular graph on ``n`` nodes.

    Configuration model: pair the n*degree stubs by a random permutation and
    reject the whole pairing if it has a self-loop or a repeated edge.  Returns
    an (E, 2) int64 array of (i < j) pairs sorted lexicographically.
    """
    if (n * degree) % 2 or degree >= n:
        raise ValueError(f"no simple {degree}-regular graph on {n} nodes")
    rng = np.random.default_rng(seed)
    stubs = np.repeat(np.arange(n, dtype=np.int64), degree)
    for _ in range(max_attempts):
        pairs = rng.permutation(stubs).reshape(-1, 2)
        a = np.minimum(pairs[:, 0], pairs[:, 1])
        b = np.maximum(pairs[:, 0], pairs[:, 1])
        if np.any(a == b):
            continue
        key = a * n + b
        if np.unique(key).size != key.size:
            continue
        order = np.argsort(key)
        return np.stack([a[order], b[order]], axis=1)
    raise RuntimeError("configuration model did not produce a simple graph")


def _to_config(edge_arr: np.ndarray, couplings, fields) -> tuple[dict, dict]:
    edges = {(int(i), int(j)): float(c) for (i, j), c in zip(edge_arr, couplings)}
    nodes = {int(v): float(h) for v, h in enumerate(fields)}
    return nodes, edges


def regular_pm1(n: int, degree: int = 3, seed: int = 1) -> tuple[dict, dict]:
    """Random regular graph, J uniform in {-1, +1}, zero fields (Z2 symmetric)."""
    e = random_regular_edges(n, degree, seed)
    rng = np.random.default_rng(seed + 7919)
    j = rng.choice(np.array([-1.0, 1.0]), size=len(e))
    return _to_config(e, j, np.zeros(n))


def regular_uniform(n: int, degree: int, seed: int, field_sigma: float = 0.0) -> tuple[dict, dict]:
    """Random regular graph, J ~ U[-1, 1], h ~ N(0, field_sigma)."""
    e = random_regular_edges(n, degree, seed)
    rng = np.random.default_rng(seed + 7919)
    j = rng.uniform(-1.0, 1.0, size=len(e))
    h = rng.normal(0.0, field_sigma, size=n) if field_sigma > 0 else np.zeros(n)
    return _to_config(e, j, h)


def qubo_to_ising(n: int, edge_arr: np.ndarray, q_edge: np.ndarray, q_diag: np.ndarray):
    """Map min_x sum_i Q_ii x_i + sum_(ij) Q_ij x_i x_j over x in {0,1} to Ising form.

    With x = (1 - s) / 2: J_ij = Q_ij / 4, h_i = -Q_ii / 2 - sum_j Q_ij / 4,
    and an additive constant sum_i Q_ii / 2 + sum_(ij) Q_ij / 4, so that
    E(s) = sum_i h_i s_i + sum_(ij) J_ij s_i s_j + const.
    """
    j = q_edge.astype(np.float64) / 4.0
    h = -q_diag.astype(np.float64) / 2.0
    np.subtract.at(h, edge_arr[:, 0], q_edge / 4.0)
    np.subtract.at(h, edge_arr[:, 1], q_edge / 4.0)
    const = float(q_diag.sum() / 2.0 + q_edge.sum() / 4.0)
    return j, h, const


def regular_qubo(n: int, degree: int = 3, seed: int = 3, qmax: int = 10) -> tuple[dict, dict, float]:
    """Random regular QUBO with integer weights uniform in [-qmax, qmax], as Ising."""
    e = random_regular_edges(n, degree, seed)
    rng = np.random.default_rng(seed + 104729)
    q_edge = rng.integers(-qmax, qmax + 1, size=len(e))
    q_diag = rng.integers(-qmax, qmax + 1, size=n)
    j, h, const = qubo_to_ising(n, e, q_edge, q_diag)
    nodes, edges = _to_config(e, j, h)
    return nodes, edges, const


def grid_edges(rows: int, cols: int) -> np.ndarray:
    """Open-boundary grid, id = cols*row + col; per node in id order: right, then down."""
    out = []
    for r in range(rows):
        for c in range(cols):
            v = r * cols + c
            if c + 1 < cols:
                out.append((v, v + 1))
            if r + 1 < rows:
                out.append((v, v + cols))
    return np.array(out, dtype=np.int64)


def grid_gaussian(rows: int, cols: int, seed: int = 2, field_sigma: float = 0.1) -> tuple[dict, dict]:
    """Open grid, J ~ N(0, 1), h ~ N(0, field_sigma) (sigma is a standard deviation)."""
    e = grid_edges(rows, cols)
    rng = np.random.default_rng(seed)
    j = rng.normal(0.0, 1.0, size=len(e))
    h = rng.normal(0.0, field_sigma, size=rows * cols)
    return _to_config(e, j, h)


def heavy_hex_edges(rows: int, cols: int) -> tuple[int, np.ndarray]:
    """Heavy-hex lattice: a brick-wall honeycomb with one extra qubit on every edge.

    Honeycomb vertices (r, c), r < rows, c < cols, have id r*cols + c; horizontal
    edges join (r, c)-(r, c+1); vertical edges join (r, c)-(r+1, c) when r + c is
    even.  Each honeycomb edge u-w becomes u-m, m-w with a new midpoint qubit m,
    numbered after the vertices in edge order.  Interior vertices have degree 3,
    midpoints degree 2; every hexagonal cell holds 12 qubits.
    """
    base = []
    for r in range(rows):
        for c in range(cols):
            v = r * cols + c
            if c + 1 < cols:
                base.append((v, v + 1))
            if r + 1 < rows and (r + c) % 2 == 0:
                base.append((v, v + cols))
    n = rows * cols
    out = []
    for (u, w) in base:
        m = n
        n += 1
        out.append((u, m))
        out.append((m, w))
    return n, np.array(out, dtype=np.int64)


def heavy_hex_gaussian(rows: int, cols: int, seed: int = 4, field_sigma: float = 0.5) -> tuple[dict, dict]:
    """Heavy-hex lattice, J ~ N(0, 1), h ~ N(0, field_sigma)."""
    n, e = heavy_hex_edges(rows, cols)
    rng = np.random.default_rng(seed)
    j = rng.normal(0.0, 1.0, size=len(e))
    h = rng.normal(0.0, field_sigma, size=n)
    return _to_config(e, j, h)


def anneal_schedule(total_time: float, steps: int, readout: str = "get_bloch_vectors") -> dict:
    """Linear anneal from pure X mixing to the Ising Hamiltonian (reference keys)."""
    return {
        "total_time": float(total_time),
        "starting_mixing": 1.0,
        "actions": [
            {"weight": 1.0, "steps_number": int(steps), "final_mixing": 0.0},
            readout,
        ],
    }


def serving_steps(n: int) -> int:
    """Trotter steps for a serving job: 100..300, proportional to log(n) over [100, 50000]."""
    return int(round(100 + 200 * math.log(n / 100.0) / math.log(500.0)))
