"""The drop-in boundary: MI355XBackend built from the reference's own Context shape.

The adapter (bqa_amd.backend.flatten_reference_context) reads a reference
Context -- Layout fields as tensors with ``.numpy``, one tensor per bond --
and must hand bqa_job_create exactly the arrays the package's own layout
compiler produces (bqa_amd.native.flatten_layouts).  Stand-ins are built from
the reference's layout KAT (tests/test_config_to_context.py:68-119) and from
the oracle's context of committed fixtures; the reference is not imported.
"""
import numpy as np
import pytest

from conftest import golden_names, load_golden
from refctx import FakeContext, from_oracle, layout
from test_layout_config import KAT, KAT_CONFIG

native = pytest.importorskip("bqa_amd.native", reason="libbqa_hip.so is not built")

from bqa_amd.backend import MI355XBackend, flatten_reference_context, is_reference_context  # noqa: E402
from bqa_amd.config import config_to_context  # noqa: E402


def product_arrays(cfg):
    ctx = config_to_context(dict(cfg, backend="mi355x"))
    f = native.flatten_layouts(ctx)
    f.update(lhs=native._i32(ctx.edges.lhs), rhs=native._i32(ctx.edges.rhs), jv=native._f64(ctx.edges.couplings))
    return ctx, f


def assert_same_arrays(a, b):
    assert set(a) == set(b)
    for k in a:
        assert a[k].dtype == b[k].dtype, k
        assert a[k].shape == b[k].shape, k
        assert np.array_equal(a[k], b[k]), k


def kat_reference_context():
    """The reference's expected Layouts for the KAT config (tests/test_config_to_context.py:68-119)."""
    lays = {d: layout(w["ids"], w["inp"], w["outp"], w["lp"], w["h"], w["j"]) for d, w in KAT.items()}
    order = [3, 2, 0]                              # first node id of each degree class
    edges = {(int(a), int(b)): float(c) for (a, b), c in KAT_CONFIG["edges"].items()}
    directed = dict(edges)
    directed.update({(b, a): c for (a, b), c in edges.items()})
    return FakeContext(backend=MI355XBackend, nodes_number=7, edges_number=12,
                       degree_to_layout={d: lays[d] for d in order}, edge_to_ampl=directed, instructions=[],
                       bp_eps=1e-6, pinv_eps=1e-6, measurement_threshold=0.95, max_bond_dim=4,
                       max_bp_iters_number=75, seed=42, damping=0.3)


def test_kat_reference_context_flattens_like_the_product_compiler():
    ref = kat_reference_context()
    assert is_reference_context(ref)
    _, want = product_arrays(KAT_CONFIG)
    assert_same_arrays(flatten_reference_context(ref), want)


@pytest.mark.parametrize("name", golden_names())
def test_fixture_reference_context_flattens_like_the_product_compiler(name):
    cfg, _ = load_golden(name)
    ctx, want = product_arrays(cfg)
    ref = from_oracle(cfg, MI355XBackend)
    assert not is_reference_context(ctx) and is_reference_context(ref)
    assert_same_arrays(flatten_reference_context(ref), want)


def test_constructors_follow_the_reference_dtypes():
    """backends.py:581-584: integer lists stay integers (intp); other lists become complex."""
    a = MI355XBackend.make_from_list([3, 1, 2])
    assert a.numpy.dtype == np.intp and a.batch_size == 3
    b = MI355XBackend.make_from_list([0.5, -1.0])
    assert b.numpy.dtype == np.complex128 and b.batch_shape == ()
    c = MI355XBackend.make_from_iter(p % 3 for p in range(6))
    assert c.numpy.tolist() == [0, 1, 2, 0, 1, 2]
    assert MI355XBackend.make_constant(2.0).numpy.shape == ()
