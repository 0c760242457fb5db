"""mpmgpu_solve_static (the entry point that allocates nothing) against the context path
(mpmgpu_create + upload + step + download) on the same inputs: two calls in a row on one
device, the second on a smaller input, both through the C ABI."""
import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
from helpers import compare

pytestmark = pytest.mark.gpu


def _context_run(prob, steps):
    g, mats, run, _, _ = prob.describe()
    d, i = prob.particles()
    sim = nm.Simulation(g, mats, run)
    sim.upload(d, i)
    sim.step(steps)
    out = sim.particles()
    sim.close()
    return out


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime the library itself links (a second
    runtime in the process, such as torch's, would not see the library's device)."""
    H2D, D2H = 1, 2

    def __init__(self):
        import ctypes
        nm.library()
        self.ct = ctypes
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipDeviceSynchronize.argtypes = []

    def upload(self, a):
        p = self.ct.c_void_p()
        assert self.lib.hipMalloc(self.ct.byref(p), a.nbytes) == 0
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D) == 0
        return p.value

    def download(self, ptr, like):
        out = np.empty_like(like)
        assert self.lib.hipMemcpy(out.ctypes.data, ptr, out.nbytes, self.D2H) == 0
        return out

    def free(self, ptr):
        assert self.lib.hipFree(ptr) == 0


def _static_run(hip, prob, steps):
    g, mats, run, n, _ = prob.describe()
    d, i = prob.particles()
    d = np.ascontiguousarray(d)
    i = np.ascontiguousarray(i)
    dd, di = hip.upload(d), hip.upload(i)
    try:
        nm.solve_static(g, mats, run, n, dd, di, steps)      # null stream; returns after the wait
        return hip.download(dd, d), hip.download(di, i)
    finally:
        hip.free(dd)
        hip.free(di)


@pytest.mark.parametrize("first,second", [
    (("j2", 16, 20), ("j2", 12, 7)),
    (("lcpdi", 16, 12), ("lcpdi", 10, 5)),
])
def test_static_solve_matches_context_path(first, second):
    def problem(kind, cells):
        lo, hi = cells // 4, cells - cells // 4
        if kind == "j2":
            return nm.Problem.from_xml(configs.j2_block_3d(cells=cells, lo=lo, hi=hi))
        return nm.Problem.from_xml(configs.neohookean_block_3d(cells=cells, lo=lo, hi=hi, vx=5000., shape="lCPDI"))
    hip = _Hip()
    for kind, cells, steps in (first, second):
        prob = problem(kind, cells)
        n = prob.describe()[3]
        print(f"static solve: {kind} block, {cells}^3 cells, {n} particles, {steps} steps", flush=True)
        ds, is_ = _static_run(hip, prob, steps)
        dc, ic = _context_run(prob, steps)
        errs = compare(ds, dc, 3)
        print(errs, "bitwise" if np.array_equal(ds, dc) else "not bitwise", flush=True)
        for k, v in errs.items():
            assert v <= 1e-9, (kind, k, v)
        assert np.array_equal(is_, ic)
        assert not np.array_equal(ds[nm.F["pos_x"]], prob.particles()[0][nm.F["pos_x"]])   # it moved
