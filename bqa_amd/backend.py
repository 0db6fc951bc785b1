"""Backend registry, the MI355X backend, and its adapter for the reference's own Context.

The reference selects a backend with the config key ``"backend"``, validated
against ``BACKEND_STR_TO_BACKEND`` (src/bqa/backends.py:26,766,1023;
src/bqa/config/config_syntax.py:63-70).  This package registers ``"mi355x"``
when ``libbqa_hip.so`` loads, and otherwise warns and leaves it unregistered,
the pattern the reference uses for CuPy (backends.py:1025-1027).  There is no
CPU fallback: a config that names a backend that is not registered is
rejected with ``ConfigSyntaxError``.

Unlike the reference's per-primitive ``Tensor`` backends, ``MI355XBackend``
owns whole phases of a Trotter step (``run_bp``, ``simple_update``, local
gates, gauge, ``density``, projective measurement) on one ``bqa_job``.  It is
built from either context kind:

* this package's ``bqa_amd.config.Context`` (NumPy layouts, [bond][node]);
* the reference's ``Context`` (src/bqa/config/config_canonicalization.py:144-168),
  whose ``Layout`` fields are backend tensors exposing ``.numpy`` and
  ``.batch_size``, with per-bond lists of position and coupling tensors.

For the second, the class also provides the tensor constructors the
reference's config canonicalisation calls on the registered backend class
(``make_from_list``, ``make_from_iter``, ``make_from_numpy``,
``make_constant``; backends.py:34-57,255-257), returning host arrays: the
layout stays on the host until ``bqa_job_create`` uploads it in one piece.
Registering ``MI355XBackend`` under ``"mi355x"`` in the reference's registry
and routing ``run_qa`` to ``MI355XBackend.run_qa(context)`` when
``context.backend.owns_phases`` is the whole binding.
"""
from __future__ import annotations

import logging
import os

import numpy as np

log = logging.getLogger(__name__)

BACKEND_STR_TO_BACKEND: dict = {}


class HostArray:
    """A host array with the part of the reference's ``Tensor`` interface that config canonicalisation and
    the layout reader use: ``.numpy``, ``.raw_tensor``, ``.raw_shape``, ``.batch_size``, ``.batch_shape``
    (src/bqa/backends.py:61-74,277-287)."""

    def __init__(self, arr: np.ndarray):
        self._a = arr

    @property
    def numpy(self) -> np.ndarray:
        return self._a

    @property
    def raw_tensor(self) -> np.ndarray:
        return self._a

    @property
    def raw_shape(self) -> tuple:
        return self._a.shape

    @property
    def batch_size(self) -> int:
        return self._a.shape[0]

    @property
    def batch_shape(self) -> tuple:
        return self._a.shape[1:]


def _host(x) -> np.ndarray:
    """NumPy view of a layout field: a NumPy array, or a reference tensor (``.numpy``)."""
    return x if isinstance(x, np.ndarray) else np.asarray(x.numpy)


def _per_bond(field, d: int, n: int, dtype) -> np.ndarray:
    """(d, n) array of a per-bond layout field: a (d, n) array, or the reference's list of d tensors of (n,)."""
    if isinstance(field, np.ndarray):
        out = field
    elif d == 0:
        out = np.zeros((0, n))
    else:
        out = np.stack([_host(t) for t in field])
    if np.iscomplexobj(out):
        out = out.real
    return np.ascontiguousarray(out, dtype=dtype).reshape(d, n)


def is_reference_context(context) -> bool:
    """The reference's Context keeps its forward edges in ``edge_to_ampl``; this package's in ``edges``."""
    return not hasattr(context, "edges") and hasattr(context, "edge_to_ampl")


def flatten_reference_context(context) -> dict:
    """Arrays for ``bqa_job_create`` from the reference's Context, in ``degree_to_layout`` order.

    Each class contributes its node ids, fields (the real part of the complex ``node_ampls``,
    backends.py:581-584), and [bond][node] input/output message positions, lambda positions and
    couplings (config_canonicalization.py:25-33,62-131).  The forward edges are the first E entries of
    ``edge_to_ampl`` (forward edges first, then the reversed ones, config_syntax.py:137-150).  The result
    has the keys and dtypes of ``bqa_amd.native.flatten_layouts``.
    """
    lays = list(context.degree_to_layout.items())
    sizes = [int(_host(l.node_ids).shape[0]) for _, l in lays]
    cat = lambda parts, dtype: np.ascontiguousarray(np.concatenate(parts) if parts else np.zeros(0), dtype=dtype)
    out = {
        "degree": np.array([d for d, _ in lays], dtype=np.int32),
        "size": np.array(sizes, dtype=np.int32),
        "ids": cat([_host(l.node_ids) for _, l in lays], np.int32),
        "inp": cat([_per_bond(l.input_msgs_position, d, n, np.int64).ravel() for (d, l), n in zip(lays, sizes)],
                   np.int32),
        "outp": cat([_per_bond(l.output_msgs_position, d, n, np.int64).ravel() for (d, l), n in zip(lays, sizes)],
                    np.int32),
        "lp": cat([_per_bond(l.lmbds_position, d, n, np.int64).ravel() for (d, l), n in zip(lays, sizes)], np.int32),
        "h": cat([np.real(_host(l.node_ampls)) for _, l in lays], np.float64),
        "j": cat([_per_bond(l.edge_ampls, d, n, np.float64).ravel() for (d, l), n in zip(lays, sizes)], np.float64),
    }
    fwd = list(context.edge_to_ampl.items())[: context.lmbds_number]
    out["lhs"] = np.array([a for (a, _), _ in fwd], dtype=np.int32)
    out["rhs"] = np.array([b for (_, b), _ in fwd], dtype=np.int32)
    out["jv"] = np.array([float(np.real(v)) for _, v in fwd], dtype=np.float64)
    return out


def context_precision(context) -> str:
    """The job precision: the context's ``precision`` (this package's key), else ``BQA_PRECISION``
    (single|double, default double), read like src/bqa/utils.py:10-20."""
    p = getattr(context, "precision", None)
    if p:
        return p
    p = (os.environ.get("BQA_PRECISION") or "double").lower()
    if p not in ("single", "double"):
        raise ValueError(f"Unknown value of BQA_PRECISION environment variable {p}")
    return p


class MI355XBackend:
    """Phase-level backend on one MI355X GPU via ``libbqa_hip.so``."""

    name = "mi355x"
    owns_phases = True          # the reference's state machine hands whole phases to this backend

    # -- the tensor constructors config canonicalisation calls on the registered class
    @classmethod
    def make_from_list(cls, lst: list) -> HostArray:
        """backends.py:581-584: integer lists stay integers, anything else becomes complex."""
        dtype = np.intp if lst and isinstance(lst[0], (int, np.integer)) else np.complex128
        return HostArray(np.array(lst, dtype=dtype))

    @classmethod
    def make_from_iter(cls, it) -> HostArray:
        """backends.py:255-257."""
        return cls.make_from_list(list(it))

    @classmethod
    def make_from_numpy(cls, arr: np.ndarray) -> HostArray:
        return HostArray(np.asarray(arr))

    @classmethod
    def make_constant(cls, const) -> HostArray:
        return HostArray(np.array(const))

    # -- one job
    def __init__(self, context, precision: str | None = None, max_bond_dim: int | None = None):
        from . import native
        self.precision = precision or context_precision(context)
        if is_reference_context(context):
            arrays = flatten_reference_context(context)
        else:
            arrays = native.flatten_layouts(context)
            arrays.update(lhs=native._i32(context.edges.lhs), rhs=native._i32(context.edges.rhs),
                          jv=native._f64(context.edges.couplings))
        chi = int(max_bond_dim if max_bond_dim is not None else context.max_bond_dim)
        self.job = native.Job.from_arrays(arrays, context.nodes_number, context.edges_number, chi, self.precision)

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

    def set_vidal_gauge(self, pinv_eps):        # state.py:204-216
        self.job.set_vidal_gauge(pinv_eps)

    def truncate_vidal_gauge(self, pinv_eps):   # state.py:160-168
        return self.job.truncate_vidal_gauge(pinv_eps)

    def run_layer(self, xtime, ztime, pinv_eps, max_iters, bp_eps, damping):   # state.py:315-321
        return self.job.run_layer(xtime, ztime, pinv_eps, max_iters, bp_eps, damping)

    def density(self):                          # state.py:77-94
        return self.job.density()

    def project(self, qubit_ids, outcomes):     # backends.py:309-310,729-734
        self.job.project(qubit_ids, outcomes)

    def energy(self):
        return self.job.energy()

    @property
    def bond_dim(self):
        return self.job.bond_dim

    def close(self):
        self.job.close()

    @classmethod
    def run_qa(cls, context) -> list:
        """The instruction interpreter of src/bqa/core.py:13-35 over an already built context of either
        kind: Trotter steps, ``"measure"`` and ``"get_bloch_vectors"`` on one job."""
        from .state import run_instructions
        return run_instructions(context)


try:
    from . import native as _native  # noqa: F401
    BACKEND_STR_TO_BACKEND["mi355x"] = MI355XBackend
except (ImportError, OSError) as _e:
    log.warning("libbqa_hip.so is not available (%s): the mi355x backend is not registered", _e)
