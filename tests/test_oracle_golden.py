"""The FP64 oracle against vectors written by the reference itself (tests/golden/make_golden.py).

These pin the loopy-graph trajectory, the global truncation rule, the BP exit
rules (convergence, damping, iteration limit) and mixed degree classes, which
no closed form covers.
"""
import numpy as np
import pytest

from conftest import golden_names, load_golden
from oracle import bqa_oracle as O


@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_reference_fixture(name):
    cfg, want = load_golden(name)
    box = []
    res = O.run_qa(cfg, state_out=box)
    _, st = box[0]
    bloch = np.array(res[0][1])
    assert np.abs(bloch - want["bloch"]).max() < 1e-10
    assert np.array_equal(np.array(st.chis), want["chis"])
    assert np.array_equal(np.array(st.bp_iters), want["sweeps"])
    lm = np.sort(st.lmbds.real, axis=1)[:, ::-1]
    assert lm.shape == want["lmbds"].shape
    assert np.abs(lm - want["lmbds"]).max() < 1e-10
    if "outcomes" in want:
        got = [r[1] for r in res if r[0] == "measurement_outcomes"]
        assert np.array_equal(np.array(got), want["outcomes"])


def test_fixtures_exercise_the_exit_rules():
    """The fixture set covers convergence, the iteration limit and damping."""
    names = golden_names()
    assert "rr100_nonconverged" in names and "rr100_damped" in names
    _, nc = load_golden("rr100_nonconverged")
    assert (nc["sweeps"] == 3).any()          # hit the limit of 3 sweeps
    cfg, _ = load_golden("rr100_damped")
    assert cfg["damping"] > 0


def test_measurement_fixtures_sample_both_outcomes():
    """The sampling fixtures draw both outcomes, so a flipped comparison would fail them."""
    for name in ("rr24_measure_chi4", "grid4x5_measure_chi8"):
        _, want = load_golden(name)
        out = want["outcomes"]
        assert out.shape[0] == 1 and set(np.unique(out)) == {-1, 1}
