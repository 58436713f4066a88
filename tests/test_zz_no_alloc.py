"""tpx_poisson_solve_device, the allocation-free form of the coefficient
Poisson solve (its residual lives in the library's static device scratch),
against tpx_poisson_solve through the C ABI: two calls in a row, the second
on a smaller grid, each equal to the allocating entry point's result (the
same kernels on the same data: the same iterations, and within the
floating-point parity bar 1e-10); a grid beyond the scratch is refused."""
import ctypes

import numpy as np
import pytest

from test_oracle_poisson_coef import manufactured

pytestmark = pytest.mark.gpu


class DeviceArrays:
    """Device copies of host arrays made with the HIP runtime directly (the
    test allocates; the entry point under test must not)."""

    def __init__(self, arrays):
        self.hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")  # the runtime libtpx is linked to
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs, self.sizes = [], []
        for a in arrays:
            a = np.ascontiguousarray(a, dtype=np.float64)
            p = ctypes.c_void_p()
            assert self.hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
            assert self.hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0  # host to device
            self.ptrs.append(p)
            self.sizes.append(a.shape)

    def get(self, k):
        out = np.empty(self.sizes[k])
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptrs[k], out.nbytes, 2) == 0  # to host
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _device_solve(s, C, P0, RHS):
    L = s.L
    L.tpx_poisson_solve_device.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 5 + [ctypes.POINTER(ctypes.c_int)]
    dev = DeviceArrays([C[0], C[1], C[2], P0, RHS])
    it = ctypes.c_int(0)
    try:
        rc = L.tpx_poisson_solve_device(s.h, *dev.ptrs, ctypes.byref(it))
        msg = L.tpx_last_error(s.h).decode()
        out = dev.get(3) if rc == 0 else None
    finally:
        dev.free()
    return rc, msg, out, it.value


@pytest.mark.parametrize("order", ["large_then_small"])
def test_device_solve_equals_the_allocating_solve(gpu_available, order):
    from tpx import Solver
    for shape in ((32, 64, 64), (16, 16, 32)):     # the second call is on a smaller grid
        p, _, C, RHS = manufactured(shape, seed=sum(shape), spread=0.5)
        p.epsilon, p.Iteration1, p.Iteration2 = 1e-10, 50, 50
        P0 = 0.1 * np.random.default_rng(len(shape)).standard_normal(shape)
        s = Solver(p, np_capacity=1)
        ref, its_ref = s.poisson_solve(C, P0, RHS)
        rc, msg, got, its = _device_solve(s, C, P0, RHS)
        s.close()
        assert rc == 0, msg
        e = float(np.max(np.abs(got - ref))) / float(np.max(np.abs(ref)))
        print(shape, "iterations", its, its_ref, f"rel diff {e:.2e}")
        assert its == its_ref and e <= 1e-10


def test_device_solve_refuses_a_grid_beyond_its_scratch(gpu_available):
    from tpx import Solver
    TPX_E_ARG = -1  # include/tpx.h
    p, _, C, RHS = manufactured((64, 128, 512), seed=1, spread=0.1)
    s = Solver(p, np_capacity=1)
    rc, msg, _, _ = _device_solve(s, C, np.zeros_like(RHS), RHS)
    s.close()
    assert rc == TPX_E_ARG and "static scratch" in msg, (rc, msg)
