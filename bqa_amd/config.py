"""Config DSL, schedule desugaring and the degree-grouped layout compiler.

Same keys, defaults, validation rules and error type as the reference
(src/bqa/config/config_syntax.py:17-191, schedule_syntax.py:1-169,
config/utils.py:1-86): a bad config raises ``ConfigSyntaxError`` (a
``ValueError``), a missing optional key logs a warning and takes its default.
The backend key accepts the names in ``bqa_amd.backend.BACKEND_STR_TO_BACKEND``.

The layout compiler (subsystem (a) of the hot path) is vectorised NumPy: it
produces, per degree class, int32 position arrays laid out [bond][node] so a
device kernel reads bond k of node i at ``k * n_d + i``.  Message positions:
forward edges 0..E-1 in input order, then the reversed edges E..2E-1; lambda
position = message position mod E (config_canonicalization.py:189-205).
"""
from __future__ import annotations

import logging
import math
from dataclasses import dataclass, field

import numpy as np

log = logging.getLogger(__name__)


class ConfigSyntaxError(ValueError):
    """src/bqa/config/utils.py:3."""


# keys (config_syntax.py:17-39)
NODES_KEY = "nodes"
EDGES_KEY = "edges"
DEFAULT_FIELD_KEY = "default_field"
SCHEDULE_KEY = "schedule"
MAX_BOND_DIM_KEY = "max_bond_dim"
MAX_BP_ITER_NUMBER_KEY = "max_bp_iter_number"
SEED_KEY = "seed"
BP_EPS_KEY = "bp_eps"
PINV_EPS_KEY = "pinv_eps"
MEASUREMENT_THRESHOLD_KEY = "measurement_threshold"
DAMPING_KEY = "damping"
BACKEND_KEY = "backend"
# backend-specific keys of this package (ignored by the reference, which reads
# known keys only, config_syntax.py:157-171)
PRECISION_KEY = "precision"

# defaults (config_syntax.py:45-61)
DEFAULT_MAX_BOND_DIM = 4
DEFAULT_MAX_BP_ITERS_NUMBER = 75
DEFAULT_BP_EPS = 1e-6
DEFAULT_PINV_EPS = 1e-6
DEFAULT_BACKEND = "mi355x"
DEFAULT_DEFAULT_FIELD = 0.0
DEFAULT_MEASUREMENT_THRESHOLD = 0.95
DEFAULT_SEED = 42
DEFAULT_DAMPING = 0.0

# schedule (schedule_syntax.py:7-66)
FINAL_MIXING_KEY = "final_mixing"
INITIAL_MIXING_KEY = "initial_mixing"
STARTING_MIXING_KEY = "starting_mixing"
TYPE_KEY = "type"
TOTAL_TIME_KEY = "total_time"
ACTIONS_KEY = "actions"
WEIGHT_KEY = "weight"
STEPS_NUMBER_KEY = "steps_number"
GET_BLOCH_VECTORS = "get_bloch_vectors"
MEASURE = "measure"
REAL_TIME_EV_TYPE = "real_time_evolution"
IMAG_TIME_EV_TYPE = "imag_time_evolution"
SIMPLE_ACTION_TYPES = {MEASURE, GET_BLOCH_VECTORS}
COMPLEX_ACTION_TYPES = {REAL_TIME_EV_TYPE, IMAG_TIME_EV_TYPE}
ACTION_TYPES = SIMPLE_ACTION_TYPES | COMPLEX_ACTION_TYPES
DEFAULT_TOTAL_TIME = 10.0
DEFAULT_STEPS_NUMBER = 100
DEFAULT_STARTING_MIXING = 1.0
DEFAULT_SCHEDULE = {
    TOTAL_TIME_KEY: DEFAULT_TOTAL_TIME,
    STARTING_MIXING_KEY: DEFAULT_STARTING_MIXING,
    ACTIONS_KEY: [
        {TYPE_KEY: REAL_TIME_EV_TYPE, WEIGHT_KEY: 1.0, STEPS_NUMBER_KEY: DEFAULT_STEPS_NUMBER,
         FINAL_MIXING_KEY: 0.0},
        GET_BLOCH_VECTORS,
    ],
}


# --- atomic checks (config/utils.py:22-86) ----------------------------------

def _get_or_default_and_warn(dct, key, default, where):
    val = dct.get(key)
    if val is None:
        log.warning(f"`{key}` field is missing in {where}, set to default {default}")
        val = default
    return val


def _get_or_raise(dct, key, where):
    val = dct.get(key)
    if val is None:
        raise ConfigSyntaxError(f"`{key}` field is missing in {where}")
    return val


def _check(ok: bool, value, msg: str):
    if not ok:
        raise ConfigSyntaxError(f"Invalid value {value}, {msg}")
    return value


def _num(x) -> bool:
    # the reference accepts bool where it accepts int (isinstance(True, int))
    return isinstance(x, (float, int))


def analyse_number(v):
    return float(_check(_num(v), v, "must be a `float` or `int` number"))


def analyse_non_neg_number(v):
    return float(_check(_num(v) and v >= 0.0, v, "must be a non-negative `float` or `int` number"))


def analyse_0_to_1_number(v):
    return float(_check(_num(v) and 0.0 <= v <= 1.0, v, "must be a `float` or `int` number from [0, 1]"))


def analyse_half_to_1_number(v):
    return float(_check(_num(v) and 0.5 <= v <= 1.0, v, "must be a `float` or `int` number from [0.5, 1]"))


def analyse_non_neg_int(v):
    return _check(isinstance(v, int) and v >= 0, v, "must be a non-negative `int`")


def analyse_positive_int(v):
    return _check(isinstance(v, int) and v > 0, v, "must be a positive `int`")


# --- nodes / edges (config_syntax.py:63-154) --------------------------------

def analyse_backend(backend):
    from .backend import BACKEND_STR_TO_BACKEND
    if isinstance(backend, str):
        if backend not in BACKEND_STR_TO_BACKEND:
            raise ConfigSyntaxError(
                f"Unknown backend \"{backend}\", available backends {list(BACKEND_STR_TO_BACKEND)}")
        return backend
    raise ConfigSyntaxError(f"Invalid backend \"{backend}\"")


def analyse_nodes(nodes) -> dict:
    if not isinstance(nodes, (tuple, list, dict)):
        raise ConfigSyntaxError(f"Invalid nodes {nodes}")
    out = {}
    try:
        for node in (nodes.items() if isinstance(nodes, dict) else nodes):
            if not (isinstance(node, (tuple, list)) and len(node) == 2):
                raise ConfigSyntaxError(f"Invalid node {node}")
            try:
                node_id = analyse_non_neg_int(node[0])
                fld = analyse_number(node[1])
            except ConfigSyntaxError as e:
                raise ConfigSyntaxError(f"Invalid node {node}") from e
            if node_id in out:
                raise ConfigSyntaxError(f"Invalid node {node}") from ConfigSyntaxError(
                    f"Duplicated node ID {node_id}")
            out[node_id] = fld
    except ConfigSyntaxError as e:
        raise ConfigSyntaxError(f"Invalid nodes {nodes}") from e
    return out


def _analyse_edge_id(edge_id):
    if not (isinstance(edge_id, (list, tuple)) and len(edge_id) == 2):
        raise ConfigSyntaxError(f"Invalid edge ID {edge_id}")
    try:
        lhs = analyse_non_neg_int(edge_id[0])
        rhs = analyse_non_neg_int(edge_id[1])
        if lhs == rhs:
            raise ConfigSyntaxError(f"LHS and RHS of the edge ID must not be equal, got edge ID {edge_id}")
        return lhs, rhs
    except ConfigSyntaxError as e:
        raise ConfigSyntaxError(f"Invalid edge ID {edge_id}") from e


@dataclass
class Edges:
    """Forward edges in input order; message order is forward then reversed."""
    lhs: np.ndarray        # (E,) int64
    rhs: np.ndarray        # (E,) int64
    couplings: np.ndarray  # (E,) float64

    @property
    def count(self) -> int:
        return int(self.lhs.size)


def analyse_edges(edges) -> Edges:
    if not isinstance(edges, (dict, tuple, list)):
        raise ConfigSyntaxError(f"Invalid edges {edges}")
    seen = set()
    lhs_l, rhs_l, cpl = [], [], []
    try:
        for edge in (edges.items() if isinstance(edges, dict) else edges):
            if not (isinstance(edge, (tuple, list)) and len(edge) == 2):
                raise ConfigSyntaxError(f"Invalid edge {edge}")
            try:
                lhs, rhs = _analyse_edge_id(edge[0])
                coupling = analyse_number(edge[1])
                key = (lhs, rhs) if lhs < rhs else (rhs, lhs)
                if key in seen:
                    raise ConfigSyntaxError(f"Duplicated edge ID {(lhs, rhs)}, {(rhs, lhs)}")
            except ConfigSyntaxError as e:
                raise ConfigSyntaxError(f"Invalid edge {edge}") from e
            seen.add(key)
            lhs_l.append(lhs)
            rhs_l.append(rhs)
            cpl.append(coupling)
    except ConfigSyntaxError as e:
        raise ConfigSyntaxError(f"Invalid edges {edges}") from e
    return Edges(np.array(lhs_l, dtype=np.int64), np.array(rhs_l, dtype=np.int64),
                 np.array(cpl, dtype=np.float64))


# --- schedule (schedule_syntax.py:70-169) -----------------------------------

def _desugar_actions(actions, starting_mixing):
    current = starting_mixing
    for action in actions:
        if isinstance(action, dict):
            a = dict(action)
            if INITIAL_MIXING_KEY in a:
                raise ConfigSyntaxError(
                    f"`{INITIAL_MIXING_KEY}` must not be present in the action {a}, it is infered automatically")
            a[INITIAL_MIXING_KEY] = current
            if FINAL_MIXING_KEY not in a:
                a[FINAL_MIXING_KEY] = current
            else:
                current = a[FINAL_MIXING_KEY]
            yield a
        elif isinstance(action, str):
            yield {TYPE_KEY: action, WEIGHT_KEY: 0, INITIAL_MIXING_KEY: current,
                   FINAL_MIXING_KEY: current, STEPS_NUMBER_KEY: 1}
        else:
            raise TypeError(f"Invalid type {type(action)} of the action {action}")


def _analyse_actions(actions, mixing):
    if not isinstance(actions, (list, tuple)):
        raise ConfigSyntaxError(
            f"Actions must be either a list or a tuple, got {actions} of type {type(actions)}")
    try:
        out = []
        for a in _desugar_actions(actions, mixing):
            action_type = _get_or_default_and_warn(a, TYPE_KEY, REAL_TIME_EV_TYPE, "action")
            if action_type not in ACTION_TYPES:
                raise ConfigSyntaxError(f"Invalid action type {action_type} in action {a}")
            out.append({
                TYPE_KEY: action_type,
                WEIGHT_KEY: analyse_0_to_1_number(_get_or_raise(a, WEIGHT_KEY, "action")),
                STEPS_NUMBER_KEY: analyse_positive_int(
                    _get_or_default_and_warn(a, STEPS_NUMBER_KEY, DEFAULT_STEPS_NUMBER, "action")),
                INITIAL_MIXING_KEY: analyse_0_to_1_number(a[INITIAL_MIXING_KEY]),
                FINAL_MIXING_KEY: analyse_0_to_1_number(a[FINAL_MIXING_KEY]),
            })
        total = sum(a[WEIGHT_KEY] for a in out)
        if not math.isclose(total, 1.0):
            raise ConfigSyntaxError(
                f"`{WEIGHT_KEY}` fields must sum into 1. in actions {out}, but now it sums into {total}")
        return out
    except ConfigSyntaxError as e:
        raise ConfigSyntaxError(f"Invalid actions {actions}") from e


def analyse_schedule(schedule) -> dict:
    if not isinstance(schedule, dict):
        raise ConfigSyntaxError(f"Schedule must be a dict, got {schedule} of type {type(schedule)}")
    try:
        total_time = analyse_non_neg_number(
            _get_or_default_and_warn(schedule, TOTAL_TIME_KEY, DEFAULT_TOTAL_TIME, "schedule"))
        mixing = analyse_0_to_1_number(
            _get_or_default_and_warn(schedule, STARTING_MIXING_KEY, DEFAULT_STARTING_MIXING, "schedule"))
        actions = _analyse_actions(
            _get_or_default_and_warn(schedule, ACTIONS_KEY, DEFAULT_SCHEDULE[ACTIONS_KEY], "schedule"),
            mixing)
        return {TOTAL_TIME_KEY: total_time, ACTIONS_KEY: actions}
    except ConfigSyntaxError as e:
        raise ConfigSyntaxError(f"Invalid schedule {schedule}") from e


def canonicalize_schedule(schedule: dict) -> list:
    """Instruction stream (schedule_canonicalization.py:6-32): per step n of a
    complex action, p = p0 + n (p1 - p0) / N, dt = T * weight / N,
    xtime = p dt, ztime = (1 - p) dt; simple actions pass through as strings."""
    total_time = schedule[TOTAL_TIME_KEY]
    out = []
    for a in schedule[ACTIONS_KEY]:
        t = a[TYPE_KEY]
        if t in COMPLEX_ACTION_TYPES:
            steps = a[STEPS_NUMBER_KEY]
            dt = total_time * a[WEIGHT_KEY] / steps
            p0, p1 = a[INITIAL_MIXING_KEY], a[FINAL_MIXING_KEY]
            delta = (p1 - p0) / steps
            for n in range(steps):
                p = p0 + n * delta
                out.append({"type": t, "xtime": p * dt, "ztime": (1.0 - p) * dt})
        else:
            out.append(t)
    return out


# --- whole config (config_syntax.py:157-191) --------------------------------

def analyse_config(config) -> dict:
    if not isinstance(config, dict):
        raise ConfigSyntaxError(f"Config must be a dict, but got type {type(config)}")
    try:
        g = _get_or_default_and_warn
        nodes = g(config, NODES_KEY, {}, "config")
        edges = _get_or_raise(config, EDGES_KEY, "config")
        schedule = g(config, SCHEDULE_KEY, DEFAULT_SCHEDULE, "config")
        out = {
            MAX_BOND_DIM_KEY: analyse_positive_int(g(config, MAX_BOND_DIM_KEY, DEFAULT_MAX_BOND_DIM, "config")),
            MAX_BP_ITER_NUMBER_KEY: analyse_non_neg_int(
                g(config, MAX_BP_ITER_NUMBER_KEY, DEFAULT_MAX_BP_ITERS_NUMBER, "config")),
            BP_EPS_KEY: analyse_non_neg_number(g(config, BP_EPS_KEY, DEFAULT_BP_EPS, "config")),
            PINV_EPS_KEY: analyse_0_to_1_number(g(config, PINV_EPS_KEY, DEFAULT_PINV_EPS, "config")),
            BACKEND_KEY: None,
            DEFAULT_FIELD_KEY: analyse_number(g(config, DEFAULT_FIELD_KEY, DEFAULT_DEFAULT_FIELD, "config")),
            MEASUREMENT_THRESHOLD_KEY: analyse_half_to_1_number(
                g(config, MEASUREMENT_THRESHOLD_KEY, DEFAULT_MEASUREMENT_THRESHOLD, "config")),
            SEED_KEY: analyse_non_neg_int(g(config, SEED_KEY, DEFAULT_SEED, "config")),
            DAMPING_KEY: analyse_0_to_1_number(g(config, DAMPING_KEY, DEFAULT_DAMPING, "config")),
        }
        out[NODES_KEY] = analyse_nodes(nodes)
        out[EDGES_KEY] = analyse_edges(edges)
        out[SCHEDULE_KEY] = analyse_schedule(schedule)
        out[BACKEND_KEY] = analyse_backend(g(config, BACKEND_KEY, DEFAULT_BACKEND, "config"))
        precision = config.get(PRECISION_KEY)
        if precision is not None and precision not in ("single", "double"):
            raise ConfigSyntaxError(f"Invalid precision {precision}, must be `single` or `double`")
        out[PRECISION_KEY] = precision
        return out
    except ConfigSyntaxError as e:
        raise ConfigSyntaxError("Invalid config") from e


# --- layout compiler (config_canonicalization.py:62-249) --------------------

@dataclass
class Layout:
    """One degree class; every [bond][node] array is int32/float64, C-contiguous."""
    degree: int
    node_ids: np.ndarray             # (n_d,)
    input_msgs_position: np.ndarray  # (d, n_d): position of message w_k -> v
    output_msgs_position: np.ndarray  # (d, n_d): position of message v -> w_k
    lmbds_position: np.ndarray       # (d, n_d)
    node_ampls: np.ndarray           # (n_d,) h_v
    edge_ampls: np.ndarray           # (d, n_d) J_{v, w_k}

    @property
    def size(self) -> int:
        return int(self.node_ids.size)


@dataclass
class Context:
    """src/bqa/config/config_canonicalization.py:144-168 (host-side, NumPy)."""
    backend: str
    nodes_number: int
    edges_number: int                 # directed messages, 2E
    degree_to_layout: dict            # insertion order = degree of first node id
    instructions: list
    bp_eps: float
    pinv_eps: float
    measurement_threshold: float
    max_bond_dim: int
    max_bp_iters_number: int
    seed: int
    damping: float
    precision: str
    edges: Edges
    node_fields: np.ndarray           # (nodes_number,) float64
    path_to_tensors: dict = field(default_factory=dict)

    @property
    def lmbds_number(self) -> int:
        return self.edges_number // 2

    def graph(self) -> list:
        """Neighbour lists in bond order (config_canonicalization.py:182-186)."""
        g = [[] for _ in range(self.nodes_number)]
        for lhs, rhs in zip(np.concatenate([self.edges.lhs, self.edges.rhs]),
                            np.concatenate([self.edges.rhs, self.edges.lhs])):
            g[int(lhs)].append(int(rhs))
        return g


def compile_layouts(nodes_number: int, edges: Edges, fields: np.ndarray) -> dict:
    """Degree classes with their message/lambda positions, vectorised.

    Bond k of node v is the k-th directed edge leaving v in message-position
    order, so ``graph[v]`` lists forward neighbours first (config_canonicalization.py:182-186).
    """
    e = edges.count
    src = np.concatenate([edges.lhs, edges.rhs])
    dst = np.concatenate([edges.rhs, edges.lhs])
    cpl = np.concatenate([edges.couplings, edges.couplings])
    pos = np.arange(2 * e, dtype=np.int64)
    order = np.argsort(src, kind="stable")         # by source node, then by position
    deg = np.bincount(src, minlength=nodes_number)
    start = np.concatenate([[0], np.cumsum(deg)])
    out_pos_all = pos[order]
    in_pos_all = np.where(out_pos_all < e, out_pos_all + e, out_pos_all - e)
    cpl_all = cpl[order]
    first_seen = {}
    for v, d in enumerate(deg.tolist()):
        if d not in first_seen:
            first_seen[d] = v
    layouts = {}
    for d in sorted(first_seen, key=first_seen.get):
        ids = np.flatnonzero(deg == d)
        if d == 0:
            idx = np.zeros((0, ids.size), dtype=np.int64)
        else:
            idx = start[ids][None, :] + np.arange(d)[:, None]     # (d, n_d) into the sorted lists
        layouts[d] = Layout(
            degree=d,
            node_ids=ids.astype(np.int32),
            input_msgs_position=np.ascontiguousarray(in_pos_all[idx], dtype=np.int32),
            output_msgs_position=np.ascontiguousarray(out_pos_all[idx], dtype=np.int32),
            lmbds_position=np.ascontiguousarray(out_pos_all[idx] % max(e, 1), dtype=np.int32),
            node_ampls=fields[ids].astype(np.float64),
            edge_ampls=np.ascontiguousarray(cpl_all[idx], dtype=np.float64),
        )
    return layouts


def config_to_context(config) -> Context:
    """src/bqa/config/core.py:8-11."""
    cfg = analyse_config(config)
    edges = cfg[EDGES_KEY]
    if edges.count == 0:
        raise ConfigSyntaxError("Invalid config") from ConfigSyntaxError("`edges` must not be empty")
    nodes = cfg[NODES_KEY]
    nodes_number = 1 + max(max(nodes.keys(), default=-1), int(max(edges.lhs.max(), edges.rhs.max())))
    if nodes_number >= 2 ** 31 or 2 * edges.count >= 2 ** 31:
        raise ConfigSyntaxError("Invalid config") from ConfigSyntaxError("graph too large for int32 layout")
    fields = np.full(nodes_number, cfg[DEFAULT_FIELD_KEY], dtype=np.float64)
    if nodes:
        fields[np.fromiter(nodes.keys(), dtype=np.int64)] = np.fromiter(nodes.values(), dtype=np.float64)
    layouts = compile_layouts(nodes_number, edges, fields)
    precision = cfg[PRECISION_KEY] or _env_precision()
    ctx = Context(
        backend=cfg[BACKEND_KEY], nodes_number=nodes_number, edges_number=2 * edges.count,
        degree_to_layout=layouts, instructions=canonicalize_schedule(cfg[SCHEDULE_KEY]),
        bp_eps=cfg[BP_EPS_KEY], pinv_eps=cfg[PINV_EPS_KEY],
        measurement_threshold=cfg[MEASUREMENT_THRESHOLD_KEY], max_bond_dim=cfg[MAX_BOND_DIM_KEY],
        max_bp_iters_number=cfg[MAX_BP_ITER_NUMBER_KEY], seed=cfg[SEED_KEY],
        damping=cfg[DAMPING_KEY], precision=precision, edges=edges, node_fields=fields,
    )
    ctx.path_to_tensors = {int(v): (d, i) for d, lay in layouts.items()
                           for i, v in enumerate(lay.node_ids.tolist())}
    log.info("Context is built")
    return ctx


def _env_precision() -> str:
    """``BQA_PRECISION`` (single|double, default double), src/bqa/utils.py:10-20."""
    import os
    p = (os.environ.get("BQA_PRECISION") or "double").lower()
    if p not in ("single", "double"):
        raise ValueError(f"Unknown value of BQA_PRECISION environment variable {p}")
    return p
