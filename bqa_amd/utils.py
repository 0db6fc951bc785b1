"""Readout helpers (src/bqa/utils.py)."""
from __future__ import annotations

import numpy as np


def convert_density_matrix_to_bloch_vector(density_matrix) -> list:
    """src/bqa/utils.py:23-27: x = Re(rho01 + rho10), y = Im(rho10 - rho01), z = Re(rho00 - rho11)."""
    x = float((density_matrix[0, 1] + density_matrix[1, 0]).real)
    y = float((density_matrix[1, 0] - density_matrix[0, 1]).imag)
    z = float((density_matrix[0, 0] - density_matrix[1, 1]).real)
    return [x, y, z]


def round_spins(bloch: np.ndarray) -> np.ndarray:
    """s_i = +1 iff <Z_i> > 0 (benchmarks_against_mqlib/utils.py:16-17)."""
    return np.where(np.asarray(bloch)[:, 2] > 0, 1, -1).astype(np.int8)
