"""ou_coarse_prolong_noalloc: the learned preconditioner's prolongation on
device arrays, with its scratch in a fixed device array of the library, so
that it allocates nothing.  Checked through the C ABI against the allocating
entry point ou_coarse_prolong on the same inputs (the same kernels: equal
bit-for-bit), on two calls in a row, the second on a smaller problem and
grid, and the size limit."""
import ctypes

import numpy as np
import pytest

import ouranos_amd as ou
from ouranos_amd import problems as P

pytestmark = pytest.mark.gpu

_vp = ctypes.c_void_p


def _noalloc():
    L = ou.lib()
    f = L.ou_coarse_prolong_noalloc
    f.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp]
    return f


def _prolong_both(op, grid, seed):
    rng = np.random.default_rng(seed)
    corners = rng.standard_normal(grid ** 3 * 8).astype(np.float32)
    z0 = P.seeded_vector(op.n, seed)
    # the allocating entry point, host arrays
    z_ref = z0.copy()
    rc = ou.lib().ou_coarse_prolong(op._plan, grid,
                                    corners.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                    z_ref.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, ou.lib().ou_last_error()
    # the allocation-free one, device arrays
    d_c = ou.DeviceBuffer(corners.nbytes)
    d_z = ou.DeviceBuffer(z0.nbytes)
    d_c.upload(corners)
    d_z.upload(z0)
    rc = _noalloc()(op._plan, grid, d_c.ptr, d_z.ptr, None)
    assert rc == 0, ou.lib().ou_last_error()
    assert ou.lib().ou_device_synchronize() == 0
    z = np.empty_like(z0)
    d_z.download(z)
    d_c.free()
    d_z.free()
    return z, z_ref, z0


def test_noalloc_prolongation_equals_allocating_entry_point():
    big = P.thick_3d(n=32, block=4).operator(ou.FP64)
    small = P.thick_3d(n=8, block=2).operator(ou.FP64)
    for op, grid, seed in ((big, 16, 1), (small, 4, 2)):
        print(f"[gpu] no-alloc prolongation: mesh {round(op.n_dofs ** (1 / 3) / 2)}^3 Q1, "
              f"grid {grid}", flush=True)
        z, z_ref, z0 = _prolong_both(op, grid, seed)
        assert not np.array_equal(z, z0)  # the correction was added
        assert np.array_equal(z, z_ref)
    # beyond the fixed scratch: refused before any work
    rc = _noalloc()(small._plan, 128, None, None, None)
    assert rc == -2
    assert b"maximum" in ou.lib().ou_last_error()
    big.close()
    small.close()
