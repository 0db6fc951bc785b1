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
    """|-> product state, bond dimension 1 (src/bqa/state.py:67-74).

    ``context.backend`` is a registered name (this package's Context) or the backend class itself (the
    reference's Context, config_canonicalization.py:211)."""
    b = context.backend
    cls = b if isinstance(b, type) else BACKEND_STR_TO_BACKEND[b]
    backend = cls(context)
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


def _set_to_vidal_gauge(context, state: State) -> None:
    """src/bqa/state.py:204-216 (unused by the reference's run_qa; kept for the gauge checks)."""
    state.backend.set_vidal_gauge(context.pinv_eps)
    log.debug("State has been set to Vidal gauge")


def _truncate_vidal_gauge(context, state: State) -> None:
    """src/bqa/state.py:160-168."""
    chi, err = state.backend.truncate_vidal_gauge(context.pinv_eps)
    log.info(f"Truncation performed, per edge error upper bound: {err}")


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
    """Sequential projective sampling (src/bqa/state.py:250-312).

    Until every qubit is measured: read the densities on the GPU; among the unmeasured qubits take the
    one whose ground probability p = rho[0, 0] has the largest |2p - 1| (the lowest id on ties, as
    ``max`` over the reference's id-ordered dict); draw u = rng.uniform(0, 1) from the state's seeded
    generator (state.py:70) and take outcome 0 iff p > u; project it (``bqa_project``) and run BP; then
    from one fresh set of densities project every unmeasured qubit with p > threshold on 0 and every one
    with p < 1 - threshold on 1; run BP again.  Returns +1 (outcome 0) / -1 (outcome 1) per qubit id.
    The draws, the choices and the thresholds run on the host in FP64, exactly as the reference's.
    """
    n = context.nodes_number
    thr = context.measurement_threshold
    rng = state.rng
    outcomes = np.zeros(n, dtype=np.int64)          # +1 / -1 once measured
    measured = np.zeros(n, dtype=bool)

    def ground_probs() -> np.ndarray:
        rho, _ = state.backend.density()
        return rho[:, 0, 0].real.astype(np.float64)

    def register(ids, outs) -> None:
        ids = np.asarray(ids, dtype=np.int64)
        outs = np.asarray(outs, dtype=np.int64)
        if ids.size:
            outcomes[ids] = 1 - 2 * outs
            measured[ids] = True
            state.backend.project(ids, outs)

    def bp() -> None:
        sweeps, converged, dist = state.backend.run_bp(context.max_bp_iters_number, context.bp_eps, context.damping)
        _log_bp(context, sweeps, converged, dist)

    log.debug("Measurement outcomes sampling started")
    while not measured.all():
        p = ground_probs()
        open_ids = np.flatnonzero(~measured)
        score = np.abs(2.0 * p[open_ids] - 1.0)
        v = int(open_ids[int(np.argmax(score))])      # first maximum in id order
        u = rng.uniform(0.0, 1.0)
        register([v], [0 if p[v] > u else 1])
        bp()
        p = ground_probs()
        open_ids = np.flatnonzero(~measured)
        up = open_ids[p[open_ids] > thr]
        down = open_ids[p[open_ids] < 1.0 - thr]
        register(up, np.zeros(up.size, dtype=np.int64))
        register(down, np.ones(down.size, dtype=np.int64))
        bp()
    log.info("Measurement outcomes sampling completed")
    return [int(x) for x in outcomes]


def run_instructions(context) -> list:
    """The instruction interpreter of src/bqa/core.py:13-35 over a built context of either kind."""
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
