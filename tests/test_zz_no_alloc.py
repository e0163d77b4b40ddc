"""fock_mo_grid_device_noalloc: the orbital-grid evaluation with no device allocation (its records
in the device's fixed 4 MiB array) against fock_mo_grid_device, which allocates them, on two
calls in a row through the C ABI, the second with a smaller input (fewer orbitals, smaller grid)
so that it runs over what the first call left in the array. Inputs whose records exceed the array
are rejected with FOCK_INVALID_ARG.
"""
import ctypes

import numpy as np
import pytest

from fock_amd import abi, molecule as m

pytestmark = pytest.mark.gpu


class DeviceArray:
    """Test-side device memory through the HIP runtime libfock.so has loaded (the entry point under
    test allocates nothing; the test does): float32, filled with all-ones bits (NaN) at creation."""

    def __init__(self, count):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.count = count
        self.p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.p), count * 4) == 0
        assert self.hip.hipMemset(self.p, 0xFF, count * 4) == 0
        assert self.hip.hipDeviceSynchronize() == 0

    @property
    def ptr(self):
        return self.p.value

    def host(self):
        out = np.empty(self.count, dtype=np.float32)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, self.count * 4, 2) == 0  # device to host
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = ctypes.c_void_p()


def test_noalloc_grid_matches_allocating_entry_point(ctx):
    from fock_amd.scf import try_scf
    print("no-alloc orbital grid: C8H18/6-31G", flush=True)
    atoms = m.octane_631g()
    r = try_scf(atoms, 300, 1e-8, 0, policy=abi.POLICY_ACCURATE, screen=abi.SCREEN_SCHWARZ, ctx=ctx,
                diis_start=2, diis_window=10, level_shift=0.5, level_shift_until=0.05)
    assert r is not None and r.converged
    fb, _ = m.flatten(atoms)
    C = r.coefficient_matrix
    cases = [(list(range(33)), (-12.0, -8.0, -6.0), (0.25, 0.25, 0.25), (96, 64, 48)),
             ([32, 33], (-3.0, -2.0, -1.0), (0.3, 0.2, 0.1), (20, 17, 9))]
    for orbs, origin, step, dims in cases:
        nv = len(orbs) * dims[0] * dims[1] * dims[2]
        a, b = DeviceArray(nv), DeviceArray(nv)
        try:
            ctx.mo_grid_device(fb, C, orbs, origin, step, dims, a.ptr)
            ctx.mo_grid_device_noalloc(fb, C, orbs, origin, step, dims, b.ptr)
            ha, hb = a.host(), b.host()
        finally:
            a.free()
            b.free()
        assert not np.isnan(hb).any()
        assert np.array_equal(ha, hb), float(np.abs(ha - hb).max())
        assert float(np.abs(hb).max()) > 0.05


def test_noalloc_grid_rejects_records_beyond_the_array(ctx):
    fb, _ = m.flatten(m.eicosane_631g())
    C = np.ones((fb.n, fb.n))  # every coefficient kept: records of 264 orbitals exceed 4 MiB
    out = DeviceArray(fb.n * 8)
    try:
        with pytest.raises(abi.FockError) as e:
            ctx.mo_grid_device_noalloc(fb, C, list(range(fb.n)), (0, 0, 0), (1, 1, 1), (2, 2, 2), out.ptr)
        assert e.value.status == abi.FOCK_INVALID_ARG
        # a smaller request on the same array still works afterwards
        ctx.mo_grid_device_noalloc(fb, C, [0], (0, 0, 0), (1, 1, 1), (2, 2, 2), out.ptr)
        h = out.host()
    finally:
        out.free()
    assert np.isfinite(h[:8]).all() and np.isnan(h[8:]).all()
