import glob
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def golden_names():
    return sorted(os.path.splitext(os.path.basename(f))[0] for f in glob.glob(os.path.join(GOLDEN_DIR, "*.npz")))


def load_golden(name):
    """A reference-written fixture: (config dict without backend, arrays)."""
    z = np.load(os.path.join(GOLDEN_DIR, name + ".npz"))
    params = json.loads(str(z["params"]))
    edges = {(int(a), int(b)): float(c) for (a, b), c in zip(z["edges"], z["couplings"])}
    nodes = {v: float(h) for v, h in enumerate(z["fields"])}
    cfg = dict(params, nodes=nodes, edges=edges)
    keys = ("bloch", "chis", "sweeps", "lmbds", "outcomes")
    return cfg, {k: z[k] for k in keys if k in z.files}


@pytest.fixture(scope="session")
def golden():
    return load_golden
