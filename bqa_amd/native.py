"""ctypes binding of libbqa_hip.so (declarations in include/bqa_hip.h).

The library is built in-tree (``bqa_amd/libbqa_hip.so``) by ``bqa_amd.build``.
If it is missing or fails to load, importing this module raises ImportError:
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import json
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int8, c_int32, c_int64, c_uint64, c_void_p

import numpy as np

LIB_NAME = "libbqa_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

BQA_OK = 0
BQA_ERR_INVALID = 1
BQA_ERR_MEMORY = 2
BQA_ERR_HIP = 3
BQA_ERR_NUMERIC = 4
PREC = {"single": 0, "double": 1}

_i32p = POINTER(c_int32)
_f64p = POINTER(c_double)

# name -> (restype, argtypes); every function returns an int status
SIGNATURES = {
    "bqa_last_error": (c_char_p, []),
    "bqa_device_info": (c_int, [POINTER(c_int), POINTER(c_int), c_char_p, c_int]),
    "bqa_job_create": (c_int, [POINTER(c_void_p), c_int32, c_int32, c_int32, _i32p, _i32p, _i32p, _i32p, _i32p,
                               _i32p, _f64p, _f64p, _i32p, _i32p, _f64p, c_int32, c_int32]),
    "bqa_job_destroy": (c_int, [c_void_p]),
    "bqa_job_device_bytes": (c_int, [c_void_p, POINTER(c_uint64)]),
    "bqa_job_estimate_bytes": (c_int, [c_int32, c_int32, _i32p, _i32p, c_int32, c_int32, POINTER(c_uint64)]),
    "bqa_init_state": (c_int, [c_void_p]),
    "bqa_run_bp": (c_int, [c_void_p, c_int32, c_double, c_double, _i32p, _i32p, _f64p]),
    "bqa_simple_update": (c_int, [c_void_p, c_double, c_double, _i32p, _f64p]),
    "bqa_apply_local_gates": (c_int, [c_void_p, c_double, c_double]),
    "bqa_set_vidal_gauge": (c_int, [c_void_p, c_double]),
    "bqa_truncate_vidal_gauge": (c_int, [c_void_p, c_double, _i32p, _f64p]),
    "bqa_set_symmetric_gauge": (c_int, [c_void_p]),
    "bqa_run_layer": (c_int, [c_void_p, c_double, c_double, c_double, c_int32, c_double, c_double, _i32p, _f64p,
                              _i32p, _i32p, _f64p]),
    "bqa_density": (c_int, [c_void_p, _f64p, _f64p]),
    "bqa_energy": (c_int, [c_void_p, _f64p, POINTER(c_int8)]),
    "bqa_project": (c_int, [c_void_p, c_int32, _i32p, _i32p]),
    "bqa_energy_noalloc": (c_int, [c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p]),
    "bqa_bond_dim": (c_int, [c_void_p, _i32p]),
    "bqa_get_tensors": (c_int, [c_void_p, c_int32, _f64p]),
    "bqa_set_tensors": (c_int, [c_void_p, c_int32, c_int32, _f64p]),
    "bqa_get_messages": (c_int, [c_void_p, _f64p]),
    "bqa_set_messages": (c_int, [c_void_p, c_int32, _f64p]),
    "bqa_get_lambdas": (c_int, [c_void_p, _f64p]),
    "bqa_set_lambdas": (c_int, [c_void_p, c_int32, _f64p]),
    "bqa_timing_enable": (c_int, [c_void_p, c_int32]),
    "bqa_timing_read": (c_int, [c_void_p, POINTER(c_int64), _f64p, _f64p]),
    "bqa_copy_calibrate": (c_int, [c_uint64, c_int32, _f64p]),
    "bqa_tuning_report": (c_int, [ctypes.c_char_p, c_int64, POINTER(c_int64)]),
    "bqa_tuning_reset": (c_int, []),
    "bqa_job_graph_stats": (c_int, [c_void_p, POINTER(c_int64), POINTER(c_int64), POINTER(c_int64)]),
    "bqa_batch_create": (c_int, [POINTER(c_void_p), c_int32]),
    "bqa_batch_destroy": (c_int, [c_void_p]),
    "bqa_batch_run_layer": (c_int, [c_void_p, c_int32, POINTER(c_void_p), _f64p, _f64p, _f64p, _i32p, _f64p, _f64p,
                                    _i32p, _f64p, _i32p, _i32p, _f64p, _i32p]),
    "bqa_sv_create": (c_int, [POINTER(c_void_p), c_int32, _f64p, c_int32, _i32p, _i32p, _f64p]),
    "bqa_sv_destroy": (c_int, [c_void_p]),
    "bqa_sv_init": (c_int, [c_void_p]),
    "bqa_sv_run_layer": (c_int, [c_void_p, c_double, c_double]),
    "bqa_sv_density": (c_int, [c_void_p, _f64p, _f64p]),
    "bqa_sv_get_state": (c_int, [c_void_p, _f64p]),
}


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        try:
            fn = getattr(lib, name)
        except AttributeError as e:     # a library built from older sources
            raise ImportError(f"{LIB_PATH} does not export {name}: rebuild it") from e
        fn.restype = res
        fn.argtypes = args
    return lib


LIB = _load()


def check(rc: int) -> None:
    """Map a status code to the reference's exception types (src/bqa/state.py:122 for NUMERIC)."""
    if rc == BQA_OK:
        return
    msg = LIB.bqa_last_error().decode(errors="replace")
    if rc == BQA_ERR_INVALID:
        raise ValueError(msg)
    if rc == BQA_ERR_MEMORY:
        raise MemoryError(msg)
    if rc == BQA_ERR_NUMERIC:
        raise AssertionError(msg)
    raise RuntimeError(msg)


def device_info() -> tuple[int, bool, str]:
    n, ok = c_int(0), c_int(0)
    name = ctypes.create_string_buffer(64)
    check(LIB.bqa_device_info(ctypes.byref(n), ctypes.byref(ok), name, 64))
    return n.value, bool(ok.value), name.value.decode()


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a: np.ndarray, typ):
    return a.ctypes.data_as(POINTER(typ))


def flatten_layouts(ctx) -> dict:
    """Concatenate a Context's degree-class layouts in class order ([bond][node] per class)."""
    lays = list(ctx.degree_to_layout.values())
    out = {
        "degree": _i32([l.degree for l in lays]),
        "size": _i32([l.size for l in lays]),
        "ids": _i32(np.concatenate([l.node_ids for l in lays])),
        "inp": _i32(np.concatenate([l.input_msgs_position.ravel() for l in lays])),
        "outp": _i32(np.concatenate([l.output_msgs_position.ravel() for l in lays])),
        "lp": _i32(np.concatenate([l.lmbds_position.ravel() for l in lays])),
        "h": _f64(np.concatenate([l.node_ampls for l in lays])),
        "j": _f64(np.concatenate([l.edge_ampls.ravel() for l in lays])),
    }
    return out


def estimate_bytes(ctx, max_bond_dim: int, precision: str) -> int:
    lays = list(ctx.degree_to_layout.values())
    deg = _i32([l.degree for l in lays])
    size = _i32([l.size for l in lays])
    out = c_uint64(0)
    check(LIB.bqa_job_estimate_bytes(ctx.edges_number, len(lays), _p(deg, c_int32), _p(size, c_int32),
                                     max_bond_dim, PREC[precision], ctypes.byref(out)))
    return int(out.value)


class Job:
    """One annealing instance resident on the GPU (a ``bqa_job`` handle)."""

    def __init__(self, ctx, precision: str = "single", max_bond_dim: int | None = None):
        f = flatten_layouts(ctx)
        f.update(lhs=_i32(ctx.edges.lhs), rhs=_i32(ctx.edges.rhs), jv=_f64(ctx.edges.couplings))
        chi_max = int(max_bond_dim if max_bond_dim is not None else ctx.max_bond_dim)
        self._create(f, ctx.nodes_number, ctx.edges_number, chi_max, precision)

    @classmethod
    def from_arrays(cls, arrays: dict, nodes_number: int, edges_number: int, max_bond_dim: int,
                    precision: str = "single") -> "Job":
        """A job from flattened layout arrays (the keys of ``flatten_layouts`` plus the forward edges
        ``lhs``, ``rhs``, ``jv``)."""
        self = cls.__new__(cls)
        self._create(arrays, nodes_number, edges_number, int(max_bond_dim), precision)
        return self

    def _create(self, f: dict, nodes_number: int, edges_number: int, chi_max: int, precision: str) -> None:
        self.handle = None
        if precision not in PREC:
            raise ValueError(f"unknown precision {precision}")
        self.precision = precision
        self.nodes_number = int(nodes_number)
        self.edges_number = int(edges_number)
        f = {k: (_i32(v) if k in ("degree", "size", "ids", "inp", "outp", "lp", "lhs", "rhs") else _f64(v))
             for k, v in f.items()}
        self.classes = list(zip(f["degree"].tolist(), f["size"].tolist()))
        self._keep = f
        h = c_void_p()
        check(LIB.bqa_job_create(
            ctypes.byref(h), self.nodes_number, self.edges_number, len(self.classes),
            _p(f["degree"], c_int32), _p(f["size"], c_int32), _p(f["ids"], c_int32), _p(f["inp"], c_int32),
            _p(f["outp"], c_int32), _p(f["lp"], c_int32), _p(f["h"], c_double), _p(f["j"], c_double),
            _p(f["lhs"], c_int32), _p(f["rhs"], c_int32), _p(f["jv"], c_double), chi_max, PREC[precision]))
        self.handle = h

    # lifecycle
    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            check(LIB.bqa_job_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_bytes(self) -> int:
        out = c_uint64(0)
        check(LIB.bqa_job_device_bytes(self.handle, ctypes.byref(out)))
        return int(out.value)

    # phases
    def init_state(self) -> None:
        check(LIB.bqa_init_state(self.handle))

    def run_bp(self, max_iters: int, bp_eps: float, damping: float) -> tuple[int, bool, float]:
        s, c, d = c_int32(0), c_int32(0), c_double(0.0)
        check(LIB.bqa_run_bp(self.handle, int(max_iters), float(bp_eps), float(damping),
                             ctypes.byref(s), ctypes.byref(c), ctypes.byref(d)))
        return s.value, bool(c.value), d.value

    def simple_update(self, ztime: float, pinv_eps: float) -> tuple[int, float]:
        chi, err = c_int32(0), c_double(0.0)
        check(LIB.bqa_simple_update(self.handle, float(ztime), float(pinv_eps), ctypes.byref(chi),
                                    ctypes.byref(err)))
        return chi.value, err.value

    def apply_local_gates(self, ztime: float, xtime: float) -> None:
        check(LIB.bqa_apply_local_gates(self.handle, float(ztime), float(xtime)))

    def set_symmetric_gauge(self) -> None:
        check(LIB.bqa_set_symmetric_gauge(self.handle))

    def set_vidal_gauge(self, pinv_eps: float) -> None:
        check(LIB.bqa_set_vidal_gauge(self.handle, float(pinv_eps)))

    def truncate_vidal_gauge(self, pinv_eps: float) -> tuple[int, float]:
        chi, err = c_int32(0), c_double(0.0)
        check(LIB.bqa_truncate_vidal_gauge(self.handle, float(pinv_eps), ctypes.byref(chi), ctypes.byref(err)))
        return chi.value, err.value

    def run_layer(self, xtime: float, ztime: float, pinv_eps: float, max_iters: int, bp_eps: float,
                  damping: float) -> dict:
        chi, err = c_int32(0), c_double(0.0)
        s, c, d = c_int32(0), c_int32(0), c_double(0.0)
        check(LIB.bqa_run_layer(self.handle, float(xtime), float(ztime), float(pinv_eps), int(max_iters),
                                float(bp_eps), float(damping), ctypes.byref(chi), ctypes.byref(err),
                                ctypes.byref(s), ctypes.byref(c), ctypes.byref(d)))
        return {"bond_dim": chi.value, "trunc_err": err.value, "sweeps": s.value,
                "converged": bool(c.value), "dist": d.value}

    def density(self) -> tuple[np.ndarray, np.ndarray]:
        rho = np.empty((self.nodes_number, 2, 2), dtype=np.complex128)
        bloch = np.empty((self.nodes_number, 3), dtype=np.float64)
        check(LIB.bqa_density(self.handle, rho.ctypes.data_as(_f64p), bloch.ctypes.data_as(_f64p)))
        return rho, bloch

    def energy(self) -> tuple[float, np.ndarray]:
        e = c_double(0.0)
        spins = np.empty(self.nodes_number, dtype=np.int8)
        check(LIB.bqa_energy(self.handle, ctypes.byref(e), spins.ctypes.data_as(POINTER(c_int8))))
        return e.value, spins

    def project(self, qubit_ids, outcomes) -> None:
        """Project qubits on outcomes 0 (|0>) / 1 (|1>) (``bqa_project``)."""
        ids, out = _i32(qubit_ids), _i32(outcomes)
        if ids.shape != out.shape:
            raise ValueError("qubit_ids and outcomes differ in length")
        check(LIB.bqa_project(self.handle, int(ids.size), _p(ids, c_int32), _p(out, c_int32)))

    @property
    def bond_dim(self) -> int:
        out = c_int32(0)
        check(LIB.bqa_bond_dim(self.handle, ctypes.byref(out)))
        return out.value

    # state transfer (complex128 host arrays)
    def get_tensors(self, class_index: int) -> np.ndarray:
        d, n = self.classes[class_index]
        chi = self.bond_dim
        out = np.empty((n, 2) + (chi,) * d, dtype=np.complex128)
        if out.size:
            check(LIB.bqa_get_tensors(self.handle, class_index, out.ctypes.data_as(_f64p)))
        return out

    def set_tensors(self, class_index: int, tensors: np.ndarray) -> None:
        d, n = self.classes[class_index]
        t = np.ascontiguousarray(tensors, dtype=np.complex128)
        chi = t.shape[2] if d > 0 else self.bond_dim
        if t.shape != (n, 2) + (chi,) * d:
            raise ValueError(f"tensor shape {t.shape} does not match class ({d}, {n})")
        check(LIB.bqa_set_tensors(self.handle, class_index, chi, t.ctypes.data_as(_f64p)))

    def get_messages(self) -> np.ndarray:
        chi = self.bond_dim
        out = np.empty((self.edges_number, chi, chi), dtype=np.complex128)
        check(LIB.bqa_get_messages(self.handle, out.ctypes.data_as(_f64p)))
        return out

    def set_messages(self, msgs: np.ndarray) -> None:
        m = np.ascontiguousarray(msgs, dtype=np.complex128)
        check(LIB.bqa_set_messages(self.handle, m.shape[1], m.ctypes.data_as(_f64p)))

    def get_lambdas(self) -> np.ndarray:
        chi = self.bond_dim
        out = np.empty((self.edges_number // 2, chi), dtype=np.complex128)
        check(LIB.bqa_get_lambdas(self.handle, out.ctypes.data_as(_f64p)))
        return out

    def set_lambdas(self, lmbds: np.ndarray) -> None:
        l = np.ascontiguousarray(lmbds, dtype=np.complex128)
        check(LIB.bqa_set_lambdas(self.handle, l.shape[1], l.ctypes.data_as(_f64p)))

    # measurement helpers
    def timing_enable(self, on: bool) -> None:
        check(LIB.bqa_timing_enable(self.handle, 1 if on else 0))

    def graph_stats(self) -> tuple[int, int, int]:
        """(graphs captured, graph replays, failed captures) of this job's BP chunks."""
        c, n, f = c_int64(0), c_int64(0), c_int64(0)
        check(LIB.bqa_job_graph_stats(self.handle, ctypes.byref(c), ctypes.byref(n), ctypes.byref(f)))
        return c.value, n.value, f.value

    def timing_read(self) -> tuple[int, float, float]:
        n, ms, b = c_int64(0), c_double(0.0), c_double(0.0)
        check(LIB.bqa_timing_read(self.handle, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)))
        return n.value, ms.value, b.value


class Batch:
    """A serving batch (``bqa_batch``): one Trotter step of several jobs per call, their phases on one
    stream and their BP sweeps in shared launches per (precision, degree, chi) shape.  Each job's
    results are those of ``Job.run_layer`` on it alone."""

    def __init__(self, max_jobs: int):
        h = c_void_p()
        check(LIB.bqa_batch_create(ctypes.byref(h), int(max_jobs)))
        self.handle = h
        self.max_jobs = int(max_jobs)

    def run_layer(self, jobs: list, xtime, ztime, pinv_eps, max_iters, bp_eps, damping) -> list:
        """One Trotter step of every job; per-job arguments are sequences (or scalars for all).  Returns one
        dict per job as ``Job.run_layer`` does, plus ``error`` (None, or the message of the job's BP failure:
        no finite distance)."""
        k = len(jobs)
        if k == 0:
            return []
        vec = lambda v, dt: np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=dt), (k,)))  # noqa: E731
        xt, zt, pe, be, dm = (vec(v, np.float64) for v in (xtime, ztime, pinv_eps, bp_eps, damping))
        it = vec(max_iters, np.int32)
        handles = (c_void_p * k)(*[j.handle for j in jobs])
        dim, sw, cv, st = (np.zeros(k, np.int32) for _ in range(4))
        err, dist = np.zeros(k), np.zeros(k)
        f, i = (lambda a: a.ctypes.data_as(_f64p)), (lambda a: a.ctypes.data_as(_i32p))
        check(LIB.bqa_batch_run_layer(self.handle, k, handles, f(xt), f(zt), f(pe), i(it), f(be), f(dm), i(dim),
                                      f(err), i(sw), i(cv), f(dist), i(st)))
        return [{"bond_dim": int(dim[q]), "trunc_err": float(err[q]), "sweeps": int(sw[q]),
                 "converged": bool(cv[q]), "dist": float(dist[q]),
                 "error": None if st[q] == BQA_OK else "BP produced no finite message distance"} for q in range(k)]

    def close(self) -> None:
        if getattr(self, "handle", None):
            check(LIB.bqa_batch_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def energy_noalloc(bloch_dev: int, nodes_number: int, fields_dev: int, edges_number: int, lhs_dev: int,
                   rhs_dev: int, j_dev: int, energy_dev: int, stream: int = 0) -> None:
    """``bqa_energy_noalloc`` on device pointers (integers, e.g. torch ``data_ptr()``)."""
    check(LIB.bqa_energy_noalloc(c_void_p(bloch_dev), int(nodes_number), c_void_p(fields_dev), int(edges_number),
                                 c_void_p(lhs_dev), c_void_p(rhs_dev), c_void_p(j_dev), c_void_p(energy_dev),
                                 c_void_p(stream)))


def tuning_report() -> list:
    """The BP launch autotuner's table (``bqa_tuning_report``): one dict per timed shape with the
    variants' names and microseconds per launch, and the index chosen."""
    need = c_int64(0)
    check(LIB.bqa_tuning_report(None, 0, ctypes.byref(need)))
    buf = ctypes.create_string_buffer(max(1, need.value))
    check(LIB.bqa_tuning_report(buf, len(buf), ctypes.byref(need)))
    return json.loads(buf.value.decode())


def tuning_reset() -> None:
    """Forget every autotuner choice of this process (``bqa_tuning_reset``)."""
    check(LIB.bqa_tuning_reset())


def copy_calibrate(nbytes: int, repeats: int = 20) -> float:
    ms = c_double(0.0)
    check(LIB.bqa_copy_calibrate(int(nbytes), int(repeats), ctypes.byref(ms)))
    return ms.value


class StateVector:
    """The exact state-vector validator on the GPU (a ``bqa_sv`` handle, include/bqa_hip.h)."""

    def __init__(self, qubits: int, fields, lhs, rhs, couplings):
        self.handle = None
        self.qubits = int(qubits)
        h = _f64(fields)
        a, b, j = _i32(lhs), _i32(rhs), _f64(couplings)
        if h.shape != (self.qubits,) or not (a.shape == b.shape == j.shape):
            raise ValueError("fields must have one entry per qubit and the edge arrays one per edge")
        hd = c_void_p()
        check(LIB.bqa_sv_create(ctypes.byref(hd), self.qubits, _p(h, c_double), int(a.size), _p(a, c_int32),
                                _p(b, c_int32), _p(j, c_double)))
        self.handle = hd

    def close(self) -> None:
        if self.handle is not None and self.handle.value:
            check(LIB.bqa_sv_destroy(self.handle))
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_state(self) -> None:
        check(LIB.bqa_sv_init(self.handle))

    def run_layer(self, xtime: float, ztime: float) -> None:
        check(LIB.bqa_sv_run_layer(self.handle, float(xtime), float(ztime)))

    def density(self) -> tuple[np.ndarray, np.ndarray]:
        rho = np.empty((self.qubits, 2, 2), dtype=np.complex128)
        bloch = np.empty((self.qubits, 3), dtype=np.float64)
        check(LIB.bqa_sv_density(self.handle, rho.ctypes.data_as(_f64p), bloch.ctypes.data_as(_f64p)))
        return rho, bloch

    def state(self) -> np.ndarray:
        out = np.empty(1 << self.qubits, dtype=np.complex128)
        check(LIB.bqa_sv_get_state(self.handle, out.ctypes.data_as(_f64p)))
        return out
