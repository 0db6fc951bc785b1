"""GPU: the reference-Context adapter end to end, and sequential measurement sampling.

* A committed fixture's reference-shaped Context (tests/refctx.py) runs through
  MI355XBackend.run_qa -- the entry a reference run_qa dispatches to -- and
  reproduces the reference's Bloch vectors (FP64, 1e-8).
* ``measure`` (src/bqa/state.py:250-312): outcomes equal the reference's,
  sampled with the same seeded generator, on the two sampling fixtures.
"""
import numpy as np
import pytest

from bqa_amd import run_qa
from bqa_amd.backend import MI355XBackend
from conftest import load_golden
from refctx import from_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["rr100_fields_chi4", "grid6x6_chi8", "heavyhex_chi6"])
def test_reference_context_end_to_end(name, monkeypatch):
    cfg, want = load_golden(name)
    monkeypatch.setenv("BQA_PRECISION", "double")
    ctx = from_oracle(cfg, MI355XBackend)
    print(f"[gpu] adapter end to end {name}", flush=True)
    res = MI355XBackend.run_qa(ctx)
    bloch = np.array([r[1] for r in res if r[0] == "bloch_vectors"][0])
    assert np.abs(bloch - want["bloch"]).max() < 1e-8


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("name", ["rr24_measure_chi4", "grid4x5_measure_chi8"])
def test_measurement_outcomes_match_reference(name, precision):
    cfg, want = load_golden(name)
    print(f"[gpu] measure {name} {precision}", flush=True)
    res = run_qa(dict(cfg, backend="mi355x", precision=precision))
    kinds = [r[0] for r in res]
    assert kinds == ["bloch_vectors", "measurement_outcomes"]
    bloch = np.array(res[0][1])
    assert np.abs(bloch - want["bloch"]).max() < (1e-8 if precision == "double" else 1e-5)
    got = np.array(res[1][1])
    print(f"[gpu]   outcomes differ at {np.flatnonzero(got != want['outcomes'][0]).tolist()}", flush=True)
    assert np.array_equal(got, want["outcomes"][0])
