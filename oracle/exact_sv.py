"""Exact state-vector simulator for <= 20 qubits (FP64 NumPy).

TEST INFRASTRUCTURE ONLY (see oracle/bqa_oracle.py).  Restates the semantics of
the reference's qem-based validator, src/bqa/exact_sim.py:14-88:

* start in |->^n (FLIPPED_HADAMARD |0>, exact_sim.py:14-17,45-49);
* each Trotter layer applies exp(-i ztime h_i Z_i) on every qubit,
  exp(-i ztime J_ij Z_i Z_j) on every undirected edge, then exp(-i xtime X_i)
  on every qubit (exact_sim.py:51-63);
* Bloch vectors from single-qubit reduced density matrices (exact_sim.py:69-70,
  utils.py:23-27).

The Z and ZZ gates are diagonal, so a layer's Z part is one phase vector.
Qubit v is bit v of the basis index (bit 0 = least significant).
"""
from __future__ import annotations

import math

import numpy as np

from .bqa_oracle import make_context


def _z_energy(ctx) -> np.ndarray:
    n = ctx.nodes_number
    idx = np.arange(1 << n, dtype=np.int64)
    spins = 1 - 2 * ((idx[:, None] >> np.arange(n)[None, :]) & 1)   # Z eigenvalue +1 for bit 0
    e = spins @ np.array([ctx.node_to_ampl[v] for v in range(n)], dtype=np.float64)
    n_l = ctx.lmbds_number
    for i, ((a, b), j) in enumerate(ctx.edge_to_ampl.items()):
        if i >= n_l:
            break
        e = e + j * spins[:, a] * spins[:, b]
    return e


def _apply_x_all(psi: np.ndarray, n: int, xtime: float) -> np.ndarray:
    c, s = math.cos(xtime), math.sin(xtime)
    for q in range(n):
        v = psi.reshape(-1, 2, 1 << q)
        a0 = v[:, 0, :].copy()
        a1 = v[:, 1, :].copy()
        v[:, 0, :] = c * a0 - 1j * s * a1
        v[:, 1, :] = c * a1 - 1j * s * a0
    return psi


def bloch_from_sv(psi: np.ndarray, n: int) -> np.ndarray:
    out = np.empty((n, 3))
    for q in range(n):
        v = psi.reshape(-1, 2, 1 << q)
        r00 = np.vdot(v[:, 0, :], v[:, 0, :]).real
        r11 = np.vdot(v[:, 1, :], v[:, 1, :]).real
        r10 = np.vdot(v[:, 0, :], v[:, 1, :])      # <0|rho|1>^* ... rho[1,0] = sum a1 conj(a0)
        tr = r00 + r11
        rho01 = np.conj(r10) / tr
        rho10 = r10 / tr
        out[q] = [(rho01 + rho10).real, (rho10 - rho01).imag, (r00 - r11) / tr]
    return out


def run_qa_exact(config: dict) -> list:
    """Bloch vectors after the config's schedule (``measure`` unsupported)."""
    ctx = make_context(config)
    n = ctx.nodes_number
    if n > 22:
        raise ValueError(f"{n} qubits is too many for the exact simulator")
    e = _z_energy(ctx)
    psi = np.full(1 << n, 1.0, dtype=np.complex128)
    # |->^n: amplitude prod_q (-1)^{bit_q} / sqrt(2)
    idx = np.arange(1 << n, dtype=np.int64)
    parity = np.zeros(1 << n, dtype=np.int64)
    for q in range(n):
        parity ^= (idx >> q) & 1
    psi = np.where(parity == 1, -1.0, 1.0).astype(np.complex128) * (2.0 ** (-n / 2))
    results = []
    for ins in ctx.instructions:
        if isinstance(ins, dict):
            psi = psi * np.exp(-1j * ins["ztime"] * e)
            psi = _apply_x_all(psi, n, ins["xtime"])
        elif ins == "get_bloch_vectors":
            results.append(["bloch_vectors", bloch_from_sv(psi, n).tolist()])
        else:
            raise ValueError(f"unsupported instruction {ins}")
    return results
