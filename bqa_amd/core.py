"""``run_qa``: the instruction interpreter (src/bqa/core.py:13-35)."""
from __future__ import annotations

import logging

from .config import config_to_context
from .state import _initialize_state, get_bloch_vectors, measure, run_layer

log = logging.getLogger(__name__)


def run_qa(config) -> list:
    """Run an annealing config; returns ``[["bloch_vectors", [[x, y, z], ...]], ...]``
    with one entry per ``get_bloch_vectors``/``measure`` instruction, like the reference."""
    context = config_to_context(config)
    n_instr = len(context.instructions)
    state = _initialize_state(context)
    results = []
    try:
        for number, instruction in enumerate(context.instructions):
            log.info(f"Instruction number {number} / {n_instr} started")
            if isinstance(instruction, dict):
                run_layer(context, instruction["xtime"], instruction["ztime"], state)
            elif instruction == "measure":
                results.append(["measurement_outcomes", measure(context, state)])
            elif instruction == "get_bloch_vectors":
                results.append(["bloch_vectors", get_bloch_vectors(context, state).tolist()])
            else:
                raise ValueError(f"Unknown instruction {instruction}")
    finally:
        state.backend.close()
    return results
