"""The JSON command line (src/bqa/cli.py restated): options, errors and the result format."""
import io
import json

import numpy as np
import pytest

from bqa_amd import cli
from conftest import load_golden


def run(argv, stdin_text=""):
    out, err = io.StringIO(), io.StringIO()
    rc = cli.main(argv, stdin=io.StringIO(stdin_text), stdout=out, stderr=err)
    return rc, out.getvalue(), err.getvalue()


def json_config(cfg):
    """A fixture config in the reference's JSON form: [[lhs, rhs], J] edges, [id, h] nodes."""
    d = {k: v for k, v in cfg.items() if k not in ("nodes", "edges")}
    d["edges"] = [[[int(a), int(b)], float(j)] for (a, b), j in cfg["edges"].items()]
    d["nodes"] = [[int(v), float(h)] for v, h in cfg["nodes"].items()]
    return d


def test_help_exits_zero():
    rc, out, _ = run(["--help"])
    assert rc == 0 and "--input" in out and "--log-level" in out


def test_unknown_argument_exits_one():
    rc, out, _ = run(["--frobnicate"])
    assert rc == 1 and "Invalid command line argument --frobnicate" in out


@pytest.mark.parametrize("argv,msg", [(["-i"], "Input file is not provided after -i key"),
                                      (["-o"], "Output file is not provided after -o key"),
                                      (["-l", "LOUD"], "cannot be interpeted as a logging level"),
                                      (["-i", "config.txt"], "must have .json suffix")])
def test_option_errors(argv, msg):
    rc, _, err = run(argv)
    assert rc == 1 and msg in err


def test_bad_json_reports_the_cause_chain():
    rc, _, err = run([], stdin_text="{not json")
    assert rc == 1
    assert err.startswith("Error while parsing json data\ncaused by: ")


def test_unknown_backend_is_a_config_error():
    cfg, _ = load_golden("rr16_fields_chi4")
    rc, _, err = run(["-b", "numpy"], stdin_text=json.dumps(json_config(cfg)))
    assert rc == 1 and "numpy" in err


def test_json_config_round_trips_to_the_same_context():
    from bqa_amd.config import config_to_context
    cfg, _ = load_golden("heavyhex_chi6")
    a = config_to_context(dict(cfg, backend="mi355x"))
    b = config_to_context(dict(json.loads(json.dumps(json_config(cfg))), backend="mi355x"))
    for d in a.degree_to_layout:
        la, lb = a.degree_to_layout[d], b.degree_to_layout[d]
        assert np.array_equal(la.input_msgs_position, lb.input_msgs_position)
        assert np.array_equal(la.edge_ampls, lb.edge_ampls)


@pytest.mark.gpu
def test_cli_end_to_end(tmp_path, monkeypatch):
    """A fixture config through the CLI on the GPU: the result file holds the reference's Bloch vectors
    and sampled outcomes."""
    monkeypatch.setenv("BQA_PRECISION", "double")
    monkeypatch.chdir(tmp_path)
    cfg, want = load_golden("rr24_measure_chi4")
    (tmp_path / "cfg.json").write_text(json.dumps(json_config(cfg)))
    print("[gpu] cli end to end rr24_measure_chi4", flush=True)
    rc, _, err = run(["-i", "cfg.json", "-o", "res.json", "-b", "mi355x", "-l", "WARNING"])
    assert rc == 0, err
    res = json.loads((tmp_path / "res.json").read_text())
    assert [r[0] for r in res] == ["bloch_vectors", "measurement_outcomes"]
    assert np.abs(np.array(res[0][1]) - want["bloch"]).max() < 1e-8
    assert res[1][1] == want["outcomes"][0].tolist()
