"""Exact state-vector validation on the GPU (src/bqa/exact_sim.py:14-88).

``run_qa_exact(config)`` runs a config's instruction stream on the full
2^n-amplitude state (n <= 32, FP64) with the reference validator's
semantics: start in |->^n; per Trotter step the Z gates exp(-i ztime h_q Z_q),
the ZZ gates exp(-i ztime J_ab Z_a Z_b) on every undirected edge, then
exp(-i xtime X_q) on every qubit; ``get_bloch_vectors`` reads every qubit's
reduced density matrix.  The reference's ``measure`` draws unseeded samples
from qem (exact_sim.py:65-67), so it has no reproducible result to match and
is not provided.
"""
from __future__ import annotations

from .config import config_to_context

MAX_QUBITS = 32


def state_for(context):
    """A GPU state vector for a context: its fields and its forward edges."""
    from .native import StateVector
    if context.nodes_number > MAX_QUBITS:
        raise ValueError(f"{context.nodes_number} qubits is too many for the exact simulator (at most {MAX_QUBITS})")
    e = context.edges
    return StateVector(context.nodes_number, context.node_fields, e.lhs, e.rhs, e.couplings)


def run_qa_exact(config) -> list:
    """``[["bloch_vectors", [[x, y, z], ...]], ...]`` of the exact Trotterised evolution."""
    context = config_to_context(config)
    sv = state_for(context)
    results = []
    try:
        sv.init_state()
        for instruction in context.instructions:
            if isinstance(instruction, dict):
                sv.run_layer(instruction["xtime"], instruction["ztime"])
            elif instruction == "get_bloch_vectors":
                results.append(["bloch_vectors", sv.density()[1].tolist()])
            elif instruction == "measure":
                raise NotImplementedError("exact sampling is not reproducible in the reference (unseeded qem "
                                          "draws, exact_sim.py:65-67) and is not provided")
            else:
                raise ValueError(f"Unknown instruction {instruction}")
    finally:
        sv.close()
    return results
