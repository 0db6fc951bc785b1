"""FP64 NumPy restatement of LuchnikovI/bqa's BP tensor-network annealing path.

TEST INFRASTRUCTURE ONLY. This module is the correctness oracle. Only the
test-suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or as the timed CPU
baseline.  The product (``bqa_amd``) never imports it and has no CPU
fallback.

Every function names the reference code it restates (paths are relative to the
reference repository).  The arithmetic is complex128 throughout, like the
reference NumPy backend at ``BQA_PRECISION=double`` (src/bqa/utils.py:10-30).

Parity pins (see tests/test_oracle_*.py):
  * layout: integer KAT of tests/test_config_to_context.py:35-165;
  * BP fixed point and densities on trees for chi > 1: brute-force branch
    contraction (tests/test_core_subroutines.py:37-122 restated);
  * canonicalisers / lambdas on trees: Schmidt values from the SVD of the
    fully contracted state (tests/test_core_subroutines.py:141-158);
  * gates: algebraic identities (tests/test_gpt_generated_gates_application.py);
  * end to end: the exact state-vector simulator in ``oracle/exact_sv.py``
    (tests/test_small_circuit_final_density.py:9-20, 1e-5) and golden vectors
    written by the reference itself (tests/golden/).
"""
from __future__ import annotations

import logging
import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

log = logging.getLogger("bqa_oracle")

DTYPE = np.complex128

# src/bqa/backends.py:20-22 -- e^{-i pi/4} and e^{+i pi/4}
_EM = math.sqrt(2.0) / 2.0 - 1j * math.sqrt(2.0) / 2.0
_EP = math.sqrt(2.0) / 2.0 + 1j * math.sqrt(2.0) / 2.0

# src/bqa/config/config_syntax.py:45-61 and schedule_syntax.py:32-56
DEFAULTS = {
    "nodes": {},
    "max_bond_dim": 4,
    "max_bp_iter_number": 75,
    "bp_eps": 1e-6,
    "pinv_eps": 1e-6,
    "default_field": 0.0,
    "measurement_threshold": 0.95,
    "seed": 42,
    "damping": 0.0,
}
DEFAULT_SCHEDULE = {
    "total_time": 10.0,
    "starting_mixing": 1.0,
    "actions": [
        {"type": "real_time_evolution", "weight": 1.0, "steps_number": 100, "final_mixing": 0.0},
        "get_bloch_vectors",
    ],
}


# ---------------------------------------------------------------------------
# config -> context
# ---------------------------------------------------------------------------

@dataclass
class ClassLayout:
    """One degree class (src/bqa/config/config_canonicalization.py:25-33)."""
    degree: int
    node_ids: np.ndarray          # (n_d,) int64, ascending
    input_pos: np.ndarray         # (d, n_d)  position of message w_k -> v
    output_pos: np.ndarray        # (d, n_d)  position of message v -> w_k
    lmbd_pos: np.ndarray          # (d, n_d)  lambda position of edge (v, w_k)
    node_ampls: np.ndarray        # (n_d,)  h_v
    edge_ampls: np.ndarray        # (d, n_d) J_{v,w_k}


@dataclass
class Context:
    """src/bqa/config/config_canonicalization.py:144-168."""
    graph: list
    nodes_number: int
    edges_number: int             # number of directed messages (2E)
    classes: dict                 # degree -> ClassLayout, insertion order = first node id
    instructions: list
    bp_eps: float
    pinv_eps: float
    max_bond_dim: int
    max_bp_iters: int
    damping: float
    seed: int
    measurement_threshold: float
    edge_to_msg_pos: dict
    edge_to_ampl: dict
    node_to_ampl: dict

    @property
    def lmbds_number(self) -> int:
        return self.edges_number // 2


def _items(x):
    return x.items() if isinstance(x, dict) else x


def _directed_edges(edges) -> dict:
    """Forward edges in input order, then every edge reversed in the same order.

    src/bqa/config/config_syntax.py:118-154 (``forward_edges | backward_edges``).
    """
    fwd, bwd = {}, {}
    for (lhs, rhs), coupling in _items(edges):
        lhs, rhs = int(lhs), int(rhs)
        if lhs == rhs or (lhs, rhs) in fwd or (rhs, lhs) in fwd:
            raise ValueError(f"invalid or duplicated edge {(lhs, rhs)}")
        fwd[(lhs, rhs)] = float(coupling)
        bwd[(rhs, lhs)] = float(coupling)
    out = dict(fwd)
    out.update(bwd)
    return out


def canonical_schedule(schedule) -> list:
    """Flat instruction list.

    src/bqa/config/schedule_syntax.py:91-117 (mixing desugaring) and
    src/bqa/config/schedule_canonicalization.py:6-32 (left-Riemann linear
    interpolation; dt = total_time * weight / steps).
    """
    total_time = float(schedule.get("total_time", DEFAULT_SCHEDULE["total_time"]))
    mixing = float(schedule.get("starting_mixing", DEFAULT_SCHEDULE["starting_mixing"]))
    out = []
    for action in schedule.get("actions", DEFAULT_SCHEDULE["actions"]):
        if isinstance(action, str):
            out.append(action)
            continue
        start = mixing
        end = float(action.get("final_mixing", mixing))
        mixing = end
        steps = int(action.get("steps_number", 100))
        dt = total_time * float(action["weight"]) / steps
        delta = (end - start) / steps
        for n in range(steps):
            p = start + n * delta
            out.append({"type": action.get("type", "real_time_evolution"),
                        "xtime": p * dt, "ztime": (1.0 - p) * dt})
    return out


def make_context(config: dict) -> Context:
    """src/bqa/config/config_canonicalization.py:62-131,171-249."""
    cfg = dict(DEFAULTS)
    cfg.update({k: v for k, v in config.items() if v is not None})
    edges = _directed_edges(cfg["edges"])
    nodes = {int(k): float(v) for k, v in _items(cfg["nodes"])}
    nodes_number = 1 + max(max(nodes.keys(), default=-1),
                           max(max(l, r) for l, r in edges.keys()))
    graph = [[] for _ in range(nodes_number)]
    for lhs, rhs in edges:
        graph[lhs].append(rhs)
    msg_pos = {e: i for i, e in enumerate(edges)}
    n_lmbd = len(edges) // 2
    node_to_ampl = {v: nodes.get(v, float(cfg["default_field"])) for v in range(nodes_number)}
    acc = {}
    for v, nbrs in enumerate(graph):
        d = len(nbrs)
        a = acc.setdefault(d, {"ids": [], "in": [], "out": [], "lp": [], "h": [], "j": []})
        a["ids"].append(v)
        a["h"].append(node_to_ampl[v])
        a["in"].append([msg_pos[(w, v)] for w in nbrs])
        a["out"].append([msg_pos[(v, w)] for w in nbrs])
        a["lp"].append([msg_pos[(v, w)] % n_lmbd for w in nbrs])
        a["j"].append([edges[(v, w)] for w in nbrs])
    classes = {}
    for d, a in acc.items():
        def mat(key, dtype):
            return np.array(a[key], dtype=dtype).reshape(len(a["ids"]), d).T.copy()
        classes[d] = ClassLayout(
            degree=d,
            node_ids=np.array(a["ids"], dtype=np.int64),
            input_pos=mat("in", np.int64),
            output_pos=mat("out", np.int64),
            lmbd_pos=mat("lp", np.int64),
            node_ampls=np.array(a["h"], dtype=np.float64),
            edge_ampls=mat("j", np.float64),
        )
    return Context(
        graph=graph, nodes_number=nodes_number, edges_number=len(edges),
        classes=classes,
        instructions=canonical_schedule(cfg.get("schedule", DEFAULT_SCHEDULE)),
        bp_eps=float(cfg["bp_eps"]), pinv_eps=float(cfg["pinv_eps"]),
        max_bond_dim=int(cfg["max_bond_dim"]), max_bp_iters=int(cfg["max_bp_iter_number"]),
        damping=float(cfg["damping"]), seed=int(cfg["seed"]),
        measurement_threshold=float(cfg["measurement_threshold"]),
        edge_to_msg_pos=msg_pos, edge_to_ampl=edges, node_to_ampl=node_to_ampl,
    )


# ---------------------------------------------------------------------------
# batched tensor algebra (src/bqa/backends.py:329-404)
# ---------------------------------------------------------------------------

def mode_product(t: np.ndarray, m: np.ndarray, axis: int) -> np.ndarray:
    """Y[..., a, ...] = sum_b m[a, b] * t[..., b, ...] along ``axis`` (batch axis 0).

    The contraction of src/bqa/backends.py:386 (``tensordot(msg, [[1], [1]])``)
    with the new index put back in place of the old one.
    """
    tm = np.moveaxis(t, axis, -1)
    shp = tm.shape
    flat = tm.reshape(shp[0], -1, shp[-1])
    out = np.matmul(flat, np.swapaxes(m, 1, 2))
    out = out.reshape(shp[:-1] + (m.shape[1],))
    return np.moveaxis(out, -1, axis)


def right_product(t: np.ndarray, w: np.ndarray, axis: int) -> np.ndarray:
    """Y[..., n, ...] = sum_c t[..., c, ...] * w[c, n] (src/bqa/backends.py:430)."""
    tm = np.moveaxis(t, axis, -1)
    shp = tm.shape
    flat = tm.reshape(shp[0], -1, shp[-1])
    out = np.matmul(flat, w).reshape(shp[:-1] + (w.shape[2],))
    return np.moveaxis(out, -1, axis)


def gram(bra: np.ndarray, ket: np.ndarray, free_axis: int) -> np.ndarray:
    """out[c, e] = sum over all axes but ``free_axis`` of conj(bra)[..c..] ket[..e..].

    src/bqa/backends.py:400-402.
    """
    b = np.moveaxis(bra, free_axis, -1)
    k = np.moveaxis(ket, free_axis, -1)
    b = b.reshape(b.shape[0], -1, b.shape[-1])
    k = k.reshape(k.shape[0], -1, k.shape[-1])
    return np.matmul(np.swapaxes(b.conj(), 1, 2), k)


def trace_normalize(m: np.ndarray) -> np.ndarray:
    """src/bqa/backends.py:316-318."""
    tr = np.trace(m, axis1=-2, axis2=-1)
    return m * (1.0 / tr)[:, None, None]


def zz_split(theta: np.ndarray, conj: bool = False):
    """Ket-side (or bra-side) coefficients of exp(-i theta ZZ) split over a bond.

    g0 = sqrt(cos theta), g1 = e^{-i pi/4} sqrt(sin theta), principal complex
    square roots of complex theta (src/bqa/backends.py:519-526).
    """
    theta = np.asarray(theta, dtype=DTYPE)
    g0 = np.sqrt(np.cos(theta))
    g1 = np.sqrt(np.sin(theta))
    if conj:
        return g0.conj(), _EP * g1.conj()
    return g0, _EM * g1


def extend_axis(t: np.ndarray, axis: int, theta: np.ndarray, conj: bool = False) -> np.ndarray:
    """Double ``axis``: first half g0 * t, second half g1 * Z_phys t.

    src/bqa/backends.py:519-526 (``_apply_conditional_z_gate_to_single_axis``).
    """
    g0, g1 = zz_split(theta, conj)
    shape = (-1,) + (1,) * (t.ndim - 1)
    up = t * g0.reshape(shape)
    zt = t.copy()
    zt[:, 1] *= -1.0
    down = zt * g1.reshape(shape)
    return np.concatenate([up, down], axis=axis)


def pass_msgs(t: np.ndarray, msgs: list, thetas: Optional[list] = None) -> list:
    """All d outgoing messages of a degree class.

    src/bqa/backends.py:381-408: output k contracts conj(T) with T times every
    incoming message but the k-th; with ``thetas`` the output bond is extended
    by the ZZ split on both ket and bra (``_compute_msg``).
    """
    d = len(msgs)
    out = []
    for k in range(d):
        tm = t
        for j in range(d):
            if j != k:
                tm = mode_product(tm, msgs[j], j + 2)
        bra = t
        if thetas is not None:
            tm = extend_axis(tm, k + 2, thetas[k])
            # the reference extends conj(T) with conj(g) (backends.py:399);
            # ``gram`` conjugates its bra, so the bra is T extended with g.
            bra = extend_axis(t, k + 2, thetas[k])
        # gram over every axis but the bond (batch axis 0 excluded)
        out.append(trace_normalize(gram(bra, tm, k + 2)))
    return out


def density_matrices(t: np.ndarray, msgs: list) -> np.ndarray:
    """rho[p, q] = sum_a (T x_all m)[p, a] conj(T[q, a]), trace one.

    src/bqa/backends.py:440-448.  Degree-0 nodes (no messages) are handled as
    rho = T T^dagger; the reference asserts there (backends.py:330).
    """
    tm = t
    for j, m in enumerate(msgs):
        tm = mode_product(tm, m, j + 2)
    n = t.shape[0]
    a = tm.reshape(n, 2, -1)
    b = t.reshape(n, 2, -1)
    rho = np.matmul(a, np.swapaxes(b.conj(), 1, 2))
    return trace_normalize(rho)


def batched_svd(m: np.ndarray, pinv_eps: float):
    """src/bqa/backends.py:709-717: masked thin SVD."""
    u, s, vh = np.linalg.svd(m, full_matrices=False)
    mask = s > pinv_eps
    s = (s * mask).astype(DTYPE)
    return u * mask[..., None, :], s, vh * mask[..., None]


def pinv_vals(x: np.ndarray) -> np.ndarray:
    """src/bqa/backends.py:719-727: 1/x where x > machine eps else 0."""
    out = np.zeros_like(x)
    sel = x.real > np.finfo(DTYPE).eps
    out[sel] = 1.0 / x[sel]
    return out


def decompose_iden(msgs: np.ndarray, eps: float):
    """src/bqa/backends.py:483-490: ul = U S^{-1/2}, lu = S^{1/2} V^H."""
    u, s, vh = batched_svd(msgs, eps)
    ss = np.sqrt(s)
    lu = ss[:, :, None] * vh
    ul = u * pinv_vals(ss)[:, None, :]
    return ul, lu


def canonicalizers(msgs: np.ndarray, pinv_eps: float):
    """Vidal-gauge lambdas and bond canonicalisers (src/bqa/state.py:171-200).

    Returns (lambdas (E, n) L2-normalised, canonicalisers (2E, n, n)) with the
    backward canonicalisers first (state.py:182-183).
    """
    n_lmbd = msgs.shape[0] // 2
    ul, lu = decompose_iden(msgs, pinv_eps)
    fwd_lu, bwd_lu = lu[:n_lmbd], lu[n_lmbd:]
    fwd_ul, bwd_ul = ul[:n_lmbd], ul[n_lmbd:]
    ker = np.matmul(fwd_lu, np.swapaxes(bwd_lu, 1, 2))     # lu_f lu_b^T, no conj
    us, lam, vhs = batched_svd(ker, pinv_eps)
    vs = np.swapaxes(vhs, 1, 2)                              # transpose, no conj
    c_f = np.matmul(fwd_ul, us)
    c_b = np.matmul(bwd_ul, vs)
    norms = np.linalg.norm(lam, axis=1)
    lam = lam * (1.0 / norms)[:, None]
    return lam, np.concatenate([c_b, c_f], axis=0)


def truncate_lmbds(lmbds: np.ndarray, dim: int, eps: float):
    """src/bqa/backends.py:297-303: one global bond dimension for every edge."""
    max_l = np.abs(lmbds.max(0))
    rank = lmbds.shape[1] - int(np.sum(max_l[::-1] < eps))
    dim = min(rank, dim)
    err = float(np.sqrt(np.sum(max_l[dim:] ** 2)))
    return lmbds[:, :dim], dim, err


def get_dist(new: np.ndarray, old: np.ndarray) -> float:
    """src/bqa/backends.py:492-495."""
    return float(np.abs(new - old).max() / np.abs(new + old).max())


# ---------------------------------------------------------------------------
# state machine (src/bqa/state.py)
# ---------------------------------------------------------------------------

@dataclass
class State:
    """src/bqa/state.py:24-33; ``chi`` is tracked explicitly."""
    tensors: dict
    msgs: np.ndarray
    lmbds: np.ndarray
    chi: int = 1
    bp_iters: list = field(default_factory=list)
    chis: list = field(default_factory=list)
    trunc_errs: list = field(default_factory=list)
    rng: Optional[np.random.Generator] = None


def msgs_from_lmbds(ctx: Context, lmbds: np.ndarray) -> np.ndarray:
    """msgs[pos] = diag(lambda[pos mod E]) / sum(lambda) (src/bqa/state.py:56-57)."""
    idx = np.arange(ctx.edges_number) % ctx.lmbds_number
    lam = lmbds[idx]
    m = lam[:, :, None] * np.eye(lam.shape[1], dtype=DTYPE)[None]
    return trace_normalize(m)


def init_state(ctx: Context) -> State:
    """|-> on every qubit, all bonds of dimension 1 (src/bqa/state.py:21,36-74)."""
    minus = np.array([math.sqrt(0.5), -math.sqrt(0.5)], dtype=DTYPE)
    tensors = {}
    for d, lay in ctx.classes.items():
        shape = (lay.node_ids.size, 2) + (1,) * d
        tensors[d] = np.broadcast_to(minus.reshape((1, 2) + (1,) * d), shape).copy()
    lmbds = np.ones((ctx.lmbds_number, 1), dtype=DTYPE)
    return State(tensors=tensors, msgs=msgs_from_lmbds(ctx, lmbds), lmbds=lmbds,
                 chi=1, rng=np.random.default_rng(ctx.seed))


def run_bp(ctx: Context, st: State) -> int:
    """Jacobi BP with the reference's exit rules (src/bqa/state.py:97-124).

    * converged (dist < bp_eps): return WITHOUT applying the last update;
    * otherwise msgs <- damping * msgs + (1 - damping) * new;
    * iteration limit: msgs <- the last undamped iterate (``best_msgs`` aliases
      the reused ``new_msgs`` buffer, state.py:104,114-116).
    Returns the number of sweeps performed.
    """
    chi = st.chi
    best = math.inf
    have_best = False
    new = np.empty((ctx.edges_number, chi, chi), dtype=DTYPE)
    for it in range(ctx.max_bp_iters):
        for d, t in st.tensors.items():
            lay = ctx.classes[d]
            ins = [st.msgs[lay.input_pos[k]] for k in range(d)]
            outs = pass_msgs(t, ins)
            for k in range(d):
                new[lay.output_pos[k]] = outs[k]
        dist = get_dist(new, st.msgs)
        if dist < best:
            best = dist
            have_best = True
        log.debug("BP iteration %d dist %g", it, dist)
        if dist < ctx.bp_eps:
            st.bp_iters.append(it + 1)
            return it + 1
        st.msgs *= ctx.damping
        st.msgs += (1.0 - ctx.damping) * new
    if not have_best:
        raise AssertionError("BP produced no finite distance")
    st.msgs = new
    st.bp_iters.append(ctx.max_bp_iters)
    log.warning("BP exceeded %d iterations, best dist %g", ctx.max_bp_iters, best)
    return ctx.max_bp_iters


def extended_msgs(ctx: Context, st: State, ztime: float) -> np.ndarray:
    """src/bqa/state.py:127-139."""
    chi = st.chi
    new = np.empty((ctx.edges_number, 2 * chi, 2 * chi), dtype=DTYPE)
    for d, t in st.tensors.items():
        lay = ctx.classes[d]
        thetas = [lay.edge_ampls[k].astype(DTYPE) * ztime for k in range(d)]
        ins = [st.msgs[lay.input_pos[k]] for k in range(d)]
        outs = pass_msgs(t, ins, thetas)
        for k in range(d):
            new[lay.output_pos[k]] = outs[k]
    return new


def absorb(t: np.ndarray, canon: list, thetas: list) -> np.ndarray:
    """src/bqa/backends.py:416-432: per bond, extend by the ZZ split, contract C."""
    for k, (c, th) in enumerate(zip(canon, thetas)):
        t = right_product(extend_axis(t, k + 2, th), c, k + 2)
    return t


def simple_update(ctx: Context, st: State, ztime: float) -> None:
    """ZZ layer with truncation (src/bqa/state.py:230-247)."""
    ext = extended_msgs(ctx, st, ztime)
    lmbds, canon = canonicalizers(ext, ctx.pinv_eps)
    lmbds, chi, err = truncate_lmbds(lmbds, ctx.max_bond_dim, ctx.pinv_eps)
    st.lmbds = lmbds
    canon = canon[:, :, :chi]
    for d in st.tensors:
        lay = ctx.classes[d]
        cs = [canon[lay.input_pos[k]] for k in range(d)]
        thetas = [lay.edge_ampls[k].astype(DTYPE) * ztime for k in range(d)]
        st.tensors[d] = absorb(st.tensors[d], cs, thetas)
    st.chi = chi
    st.chis.append(chi)
    st.trunc_errs.append(err)


def set_to_vidal_gauge(ctx: Context, st: State) -> None:
    """Vidal gauge from the current messages (src/bqa/state.py:204-216): lambdas and canonicalisers of the
    chi x chi messages (state.py:171-200); every bond k of every tensor is contracted with the canonicaliser
    at its input message position (backends.py:410-414).  Messages are left as they are."""
    lam, canon = canonicalizers(st.msgs, ctx.pinv_eps)
    st.lmbds = lam
    for d in st.tensors:
        lay = ctx.classes[d]
        for k in range(d):
            st.tensors[d] = right_product(st.tensors[d], canon[lay.input_pos[k]], k + 2)


def truncate_vidal_gauge(ctx: Context, st: State) -> float:
    """src/bqa/state.py:160-168: the global truncation rule on the lambdas (backends.py:297-303) and every
    bond of every tensor cut to the new bond dimension (batch_truncate_all_but, backends.py:289-295).
    Returns the truncation error."""
    lam, chi, err = truncate_lmbds(st.lmbds, ctx.max_bond_dim, ctx.pinv_eps)
    for d, t in st.tensors.items():
        st.tensors[d] = t[(slice(None), slice(None)) + (slice(0, chi),) * d].copy()
    st.lmbds = lam
    st.chi = chi
    return err


def apply_z_layer(ctx: Context, st: State, ztime: float) -> None:
    """T <- (cos phi - i sin phi Z) T, phi = h ztime (src/bqa/state.py:142-150)."""
    for d, t in st.tensors.items():
        phi = ctx.classes[d].node_ampls.astype(DTYPE) * ztime
        shape = (-1,) + (1,) * (t.ndim - 1)
        zt = t.copy()
        zt[:, 1] *= -1.0
        st.tensors[d] = t * np.cos(phi).reshape(shape) - 1j * zt * np.sin(phi).reshape(shape)


def apply_x_layer(ctx: Context, st: State, xtime: float) -> None:
    """T <- (cos x - i sin x X) T (src/bqa/state.py:153-156, backends.py:506-507)."""
    for d, t in st.tensors.items():
        st.tensors[d] = math.cos(xtime) * t - 1j * math.sin(xtime) * t[:, ::-1]


def set_symmetric_gauge(ctx: Context, st: State) -> None:
    """T <- T prod_k sqrt(lambda_k), L2-normalised; msgs <- diag(lambda)/sum
    (src/bqa/state.py:219-227, backends.py:450-462)."""
    st.msgs = msgs_from_lmbds(ctx, st.lmbds)
    for d, t in st.tensors.items():
        lay = ctx.classes[d]
        for k in range(d):
            sq = np.sqrt(st.lmbds[lay.lmbd_pos[k]])
            shape = [sq.shape[0]] + [1] * (t.ndim - 1)
            shape[k + 2] = sq.shape[1]
            t = t * sq.reshape(shape)
        norms = np.linalg.norm(t.reshape(t.shape[0], -1), axis=1)
        st.tensors[d] = t * (1.0 / norms).reshape((-1,) + (1,) * (t.ndim - 1))


def run_layer(ctx: Context, st: State, xtime: float, ztime: float) -> None:
    """One Trotter step (src/bqa/state.py:315-321)."""
    simple_update(ctx, st, ztime)
    apply_z_layer(ctx, st, ztime)
    apply_x_layer(ctx, st, xtime)
    set_symmetric_gauge(ctx, st)
    run_bp(ctx, st)


def get_density_matrices(ctx: Context, st: State) -> np.ndarray:
    """(n, 2, 2) per-qubit reduced density matrices (src/bqa/state.py:77-94)."""
    out = np.empty((ctx.nodes_number, 2, 2), dtype=DTYPE)
    for d, t in st.tensors.items():
        lay = ctx.classes[d]
        ins = [st.msgs[lay.input_pos[k]] for k in range(d)]
        out[lay.node_ids] = density_matrices(t, ins)
    return out


def bloch_vectors(rho: np.ndarray) -> np.ndarray:
    """src/bqa/utils.py:23-27 vectorised: x, y, z per qubit."""
    x = (rho[:, 0, 1] + rho[:, 1, 0]).real
    y = (rho[:, 1, 0] - rho[:, 0, 1]).imag
    z = (rho[:, 0, 0] - rho[:, 1, 1]).real
    return np.stack([x, y, z], axis=1)


def project(ctx: Context, st: State, node: int, outcome: int) -> None:
    """Project qubit ``node`` on outcome 0 (|0>) or 1 (|1>): zero the other physical component of
    its tensor, then divide the whole degree-class batch by its Frobenius norm
    (src/bqa/backends.py:309-310,729-734)."""
    for d, lay in ctx.classes.items():
        hit = np.nonzero(lay.node_ids == node)[0]
        if hit.size:
            t = st.tensors[d]
            t[hit[0], 1 - outcome] = 0.0
            t /= np.linalg.norm(t)
            return
    raise KeyError(node)


def measure(ctx: Context, st: State) -> list:
    """Sequential projective sampling (src/bqa/state.py:250-312).

    Until every qubit is measured: among the unmeasured qubits take the one whose ground
    probability p = rho[0, 0] has the largest |2p - 1| (the first such id on ties, as
    Python's ``max`` over the id-ordered dict); draw u = rng.uniform(0, 1) from the state's
    generator and take outcome 0 iff p > u; project and run BP; then, from one fresh set of
    densities, project every unmeasured qubit with p > threshold on 0 and every one with
    p < 1 - threshold on 1; run BP again.  Returns +1 (outcome 0) / -1 (outcome 1) per id.
    """
    outcomes = {}
    thr = ctx.measurement_threshold
    n_layers = len(st.bp_iters)       # the sweep counts of this sampling are not Trotter steps

    def ground_probs():
        rho = get_density_matrices(ctx, st)
        return {v: float(rho[v, 0, 0].real) for v in range(ctx.nodes_number) if v not in outcomes}

    def register(v, outcome):
        outcomes[v] = 1 - 2 * outcome
        project(ctx, st, v, outcome)

    while len(outcomes) < ctx.nodes_number:
        probs = ground_probs()
        v, p = max(probs.items(), key=lambda kv: abs(2 * kv[1] - 1))
        u = st.rng.uniform(0.0, 1.0)
        register(v, 0 if p > u else 1)
        run_bp(ctx, st)
        probs = ground_probs()
        for w, q in list(probs.items()):
            if q > thr:
                register(w, 0)
        for w, q in probs.items():
            if q < 1.0 - thr:
                register(w, 1)
        run_bp(ctx, st)
    del st.bp_iters[n_layers:]
    return [outcomes[v] for v in range(len(outcomes))]


def run_qa(config: dict, state_out: Optional[list] = None) -> list:
    """src/bqa/core.py:13-35."""
    ctx = make_context(config)
    st = init_state(ctx)
    results = []
    for ins in ctx.instructions:
        if isinstance(ins, dict):
            run_layer(ctx, st, ins["xtime"], ins["ztime"])
        elif ins == "get_bloch_vectors":
            results.append(["bloch_vectors", bloch_vectors(get_density_matrices(ctx, st)).tolist()])
        elif ins == "measure":
            results.append(["measurement_outcomes", measure(ctx, st)])
        else:
            raise ValueError(f"Unknown instruction {ins}")
    if state_out is not None:
        state_out.append((ctx, st))
    return results


def ising_energy(ctx: Context, spins: np.ndarray) -> float:
    """E(s) = sum_i h_i s_i + sum_(i,j) J_ij s_i s_j over undirected edges, FP64.

    The sign convention is the one of the annealed Hamiltonian
    (src/bqa/exact_sim.py:36-63); the reference never sums it itself.
    """
    e = sum(h * spins[v] for v, h in ctx.node_to_ampl.items())
    n_l = ctx.lmbds_number
    for i, ((a, b), j) in enumerate(ctx.edge_to_ampl.items()):
        if i >= n_l:
            break
        e += j * spins[a] * spins[b]
    return float(e)
