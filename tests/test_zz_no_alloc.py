"""bqa_energy_noalloc (the allocation-free entry point) against bqa_energy, through the C ABI.

Two calls in a row on the same GPU's scratch array, the second with a smaller
instance; each result must equal bqa_energy's bit for bit (same kernels, same
grid, same summation order).  The caller's device arrays come from the HIP
runtime the library links (libamdhip64.so.7) through ctypes: plumbing only.
"""
import numpy as np
import pytest

from bqa_amd import instances, native
from bqa_amd.config import config_to_context

pytestmark = pytest.mark.gpu


def annealed(n, seed):
    """A short anneal on a job: (Bloch vectors, energy from bqa_energy, context)."""
    nodes, edges = instances.regular_uniform(n, 3, seed=seed, field_sigma=0.5)
    ctx = config_to_context({"nodes": nodes, "edges": edges, "max_bond_dim": 4, "backend": "mi355x",
                             "schedule": instances.anneal_schedule(3.0, 12), "precision": "double"})
    job = native.Job(ctx, "double")
    try:
        job.init_state()
        for ins in [i for i in ctx.instructions if isinstance(i, dict)]:
            job.run_layer(ins["xtime"], ins["ztime"], ctx.pinv_eps, ctx.max_bp_iters_number, ctx.bp_eps, 0.0)
        _, bloch = job.density()
        e, _ = job.energy()
    finally:
        job.close()
    return bloch, e, ctx


class DeviceArrays:
    """Device buffers from the HIP runtime the library itself uses (test plumbing, through ctypes)."""

    H2D, D2H = 1, 2

    def __init__(self):
        import ctypes
        self.ct = ctypes
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.ptrs = []

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        p = self.ct.c_void_p()
        assert self.hip.hipMalloc(self.ct.byref(p), self.ct.c_size_t(max(a.nbytes, 8))) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(self.ct.c_void_p), self.ct.c_size_t(a.nbytes), self.H2D) == 0
        return p.value

    def download(self, ptr, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data_as(self.ct.c_void_p), self.ct.c_void_p(ptr),
                                  self.ct.c_size_t(out.nbytes), self.D2H) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def noalloc_energy(bloch, ctx):
    mem = DeviceArrays()
    try:
        b = mem.upload(np.asarray(bloch, dtype=np.float64))
        h = mem.upload(np.asarray(ctx.node_fields, dtype=np.float64))
        lhs = mem.upload(np.asarray(ctx.edges.lhs, dtype=np.int32))
        rhs = mem.upload(np.asarray(ctx.edges.rhs, dtype=np.int32))
        jv = mem.upload(np.asarray(ctx.edges.couplings, dtype=np.float64))
        out = mem.upload(np.array([np.nan]))
        native.energy_noalloc(b, ctx.nodes_number, h, ctx.edges.count, lhs, rhs, jv, out, 0)
        assert mem.hip.hipDeviceSynchronize() == 0
        return float(mem.download(out, np.zeros(1))[0])
    finally:
        mem.free()


def test_noalloc_energy_equals_bqa_energy_twice():
    b1, e1, c1 = annealed(60, 41)
    b2, e2, c2 = annealed(24, 42)
    print(f"[gpu] no-alloc energy: n={c1.nodes_number} then n={c2.nodes_number}", flush=True)
    got1 = noalloc_energy(b1, c1)
    got2 = noalloc_energy(b2, c2)          # smaller input on the same scratch
    print(f"[gpu]   {got1} vs {e1}; {got2} vs {e2}", flush=True)
    assert got1 == e1
    assert got2 == e2


def test_noalloc_energy_rejects_bad_arguments():
    with pytest.raises(ValueError):
        native.energy_noalloc(0, 10, 0, 0, 0, 0, 0, 0)           # null arrays
    with pytest.raises(ValueError):
        native.energy_noalloc(0, -1, 0, 0, 0, 0, 0, 8)
    with pytest.raises(ValueError):
        native.energy_noalloc(8, (1 << 30) + 1, 8, 0, 0, 0, 0, 8)   # beyond the stated maximum
