"""``run_qa``: the instruction interpreter (src/bqa/core.py:13-35)."""
from __future__ import annotations

from .config import config_to_context
from .state import run_instructions


def run_qa(config) -> list:
    """Run an annealing config; returns ``[["bloch_vectors", [[x, y, z], ...]], ...]`` and
    ``["measurement_outcomes", [+1/-1, ...]]`` entries, one per ``get_bloch_vectors`` / ``measure``
    instruction, like the reference."""
    return run_instructions(config_to_context(config))
