"""Annealing state machine on the MI355X backend (src/bqa/state.py restated).

The reference keeps node tensors, messages and lambdas as backend arrays and
runs each phase as a sequence of batched primitive calls per degree class.
Here the state lives on the GPU inside a ``bqa_job``; each function below
names the reference function it mirrors and runs that phase as one group of
kernels.  ``run_layer`` runs the fused Trotter step; the unfused phase
functions exist for phase-by-phase parity checks.
"""
from __future__ import annotations

import logging
from dataclasses import dataclass, field

import numpy as np
from numpy.random import Generator, default_rng

from .backend import BACKEND_STR_TO_BACKEND

log = logging.getLogger(__name__)


@dataclass
class State:
    """src/bqa/state.py:24-33; tensors, messages and lambdas stay on the device."""
    backend: object
    rng: Generator
    chis: list = field(default_factory=list)
    bp_sweeps: list = field(default_factory=list)
    trunc_errs: list = field(default_factory=list)

    @property
    def bond_dim(self) -> int:
        return self.backend.bond_dim


def _initialize_state(context) -> State:
    """|-> product state, bond dimension 1 (src/bqa/state.py:67-74)."""
    cls = BACKEND_STR_TO_BACKEND[context.backend]
    backend = cls(context, context.precision)
    backend.init_state()
    return State(backend=backend, rng=default_rng(context.seed))


def _log_bp(context, sweeps: int, converged: bool, dist: float) -> None:
    if converged:
        log.debug(f"BP algorithm completed after {sweeps - 1} iterations")
    else:
        log.warning(f"BP algorithm exceeds iterations limit set to {context.max_bp_iters_number}, "
                    f"obtained bp_eps {dist}")


def _run_bp(context, state: State) -> None:
    """src/bqa/state.py:97-124."""
    sweeps, converged, dist = state.backend.run_bp(context.max_bp_iters_number, context.bp_eps, context.damping)
    state.bp_sweeps.append(sweeps)
    _log_bp(context, sweeps, converged, dist)


def _simple_update(context, time: float, state: State) -> None:
    """src/bqa/state.py:230-247."""
    chi, err = state.backend.simple_update(time, context.pinv_eps)
    state.chis.append(chi)
    state.trunc_errs.append(err)
    log.info(f"Truncation performed, per edge error upper bound: {err}")


def _apply_z_layer(context, ztime: float, state: State) -> None:
    """src/bqa/state.py:142-150."""
    state.backend.apply_local_gates(ztime, 0.0)


def _apply_x_layer(xtime: float, state: State) -> None:
    """src/bqa/state.py:153-156."""
    state.backend.apply_local_gates(0.0, xtime)


def _set_to_symmetric_gauge(context, state: State) -> None:
    """src/bqa/state.py:219-227."""
    state.backend.set_symmetric_gauge()


def run_layer(context, xtime: float, ztime: float, state: State, fused: bool = True) -> None:
    """One Trotter step (src/bqa/state.py:315-321)."""
    if fused:
        out = state.backend.run_layer(xtime, ztime, context.pinv_eps, context.max_bp_iters_number,
                                      context.bp_eps, context.damping)
        state.chis.append(out["bond_dim"])
        state.trunc_errs.append(out["trunc_err"])
        state.bp_sweeps.append(out["sweeps"])
        log.info(f"Truncation performed, per edge error upper bound: {out['trunc_err']}")
        _log_bp(context, out["sweeps"], out["converged"], out["dist"])
    else:
        _simple_update(context, ztime, state)
        _apply_z_layer(context, ztime, state)
        _apply_x_layer(xtime, state)
        _set_to_symmetric_gauge(context, state)
        _run_bp(context, state)
    log.info(f"Layer with ztime {ztime} and xtime {xtime} has been applied")


def get_density_matrices(context, state: State) -> np.ndarray:
    """(n, 2, 2) complex128 reduced density matrices (src/bqa/state.py:77-94)."""
    rho, _ = state.backend.density()
    log.info("Density matrices have been computed")
    return rho


def get_bloch_vectors(context, state: State) -> np.ndarray:
    """(n, 3) Bloch vectors computed on the device (utils.py:23-27)."""
    _, bloch = state.backend.density()
    return bloch


def measure(context, state: State) -> list:
    """Sequential projective sampling (src/bqa/state.py:250-312) is not provided by this backend."""
    raise NotImplementedError("the `measure` instruction is not supported by the mi355x backend; "
                              "use `get_bloch_vectors` and round the z components")
