"""Backend registry and the MI355X backend (the drop-in boundary).

The reference selects a backend with the config key ``"backend"``, validated
against ``BACKEND_STR_TO_BACKEND`` (src/bqa/backends.py:26,766,1023;
src/bqa/config/config_syntax.py:63-70).  This package registers ``"mi355x"``
when ``libbqa_hip.so`` loads, and otherwise warns and leaves it unregistered,
the pattern the reference uses for CuPy (backends.py:1025-1027).  There is no
CPU fallback: a config that names a backend that is not registered is
rejected with ``ConfigSyntaxError``.

Unlike the reference's per-primitive ``Tensor`` backends, this backend owns
whole phases of a Trotter step (``run_bp``, ``simple_update``, local gates,
gauge, ``density``); the state machine in ``bqa_amd.state`` calls them.  Each
phase names the reference composite it replaces.
"""
from __future__ import annotations

import logging

log = logging.getLogger(__name__)

BACKEND_STR_TO_BACKEND: dict = {}


class MI355XBackend:
    """Phase-level backend on one MI355X GPU via ``libbqa_hip.so``."""

    name = "mi355x"

    def __init__(self, ctx, precision: str = "single", max_bond_dim=None):
        from . import native
        self.job = native.Job(ctx, precision, max_bond_dim)

    def init_state(self):                       # state.py:67-74
        self.job.init_state()

    def run_bp(self, max_iters, bp_eps, damping):   # state.py:97-124
        return self.job.run_bp(max_iters, bp_eps, damping)

    def simple_update(self, ztime, pinv_eps):   # state.py:230-247
        return self.job.simple_update(ztime, pinv_eps)

    def apply_local_gates(self, ztime, xtime):  # state.py:142-156
        self.job.apply_local_gates(ztime, xtime)

    def set_symmetric_gauge(self):              # state.py:219-227
        self.job.set_symmetric_gauge()

    def run_layer(self, xtime, ztime, pinv_eps, max_iters, bp_eps, damping):   # state.py:315-321
        return self.job.run_layer(xtime, ztime, pinv_eps, max_iters, bp_eps, damping)

    def density(self):                          # state.py:77-94
        return self.job.density()

    def energy(self):
        return self.job.energy()

    @property
    def bond_dim(self):
        return self.job.bond_dim

    def close(self):
        self.job.close()


try:
    from . import native as _native  # noqa: F401
    BACKEND_STR_TO_BACKEND["mi355x"] = MI355XBackend
except (ImportError, OSError) as _e:
    log.warning("libbqa_hip.so is not available (%s): the mi355x backend is not registered", _e)
