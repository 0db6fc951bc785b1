"""The C-ABI library loads and exports every symbol include/bqa_hip.h declares (no compute calls)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "bqa_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bqa_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_phase_entry_points():
    names = declared_functions()
    for must in ("bqa_job_create", "bqa_run_bp", "bqa_simple_update", "bqa_apply_local_gates",
                 "bqa_set_symmetric_gauge", "bqa_run_layer", "bqa_density", "bqa_energy"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from bqa_amd import native
    lib = ctypes.CDLL(native.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(native.SIGNATURES) == set(declared_functions())


def test_backend_is_registered_when_the_library_loads():
    from bqa_amd.backend import BACKEND_STR_TO_BACKEND
    assert "mi355x" in BACKEND_STR_TO_BACKEND


def test_last_error_is_a_string():
    from bqa_amd import native
    assert isinstance(native.LIB.bqa_last_error(), bytes)


def test_estimate_bytes_scales_with_chi():
    from bqa_amd import instances, native
    from bqa_amd.config import config_to_context
    nodes, edges = instances.regular_pm1(100, 3, seed=1)
    ctx = config_to_context({"nodes": nodes, "edges": edges})
    b4 = native.estimate_bytes(ctx, 4, "single")
    b8 = native.estimate_bytes(ctx, 8, "single")
    assert b8 > 4 * b4 > 0
    assert b4 < native.estimate_bytes(ctx, 4, "double") < 2 * b4 + 1   # FP64 canonicalisers in both


def test_job_creation_without_gpu_fails_loudly():
    from bqa_amd import instances, native
    from bqa_amd.config import config_to_context
    n, _, _ = native.device_info()
    if n > 0:
        pytest.skip("a GPU is visible")
    nodes, edges = instances.regular_pm1(10, 3, seed=1)
    ctx = config_to_context({"nodes": nodes, "edges": edges})
    with pytest.raises(RuntimeError):
        native.Job(ctx)
