"""cvf_vdw_energy_noalloc (fixed device scratch, no allocation) against
cvf_vdw_energy (per-call device buffers), through the C ABI.

The two run the same kernel on the same staged values, so the energies
must be identical bit for bit: two calls in a row, the second smaller than
the first (fewer shells, fewer radii, the other precision), so a stale
scratch from the first call would show.  Sizes beyond the fixed maximum
are rejected with an error.
"""
import ctypes as C

import numpy as np
import pytest

from cvfbulkwater_amd import FP32_COMPENSATED, FP64, CVFError, load_library, vdw_energy
from oracle import py as O

pytestmark = pytest.mark.gpu


def test_noalloc_equals_allocating_entry_point_twice():
    d1, f1 = O.shell_table(96, 6.0)
    r1 = np.linspace(0.90, 1.40, 4096)  # the maximum number of radii
    a1 = vdw_energy(96, d1, f1, 6.0, r1, FP64)
    b1 = vdw_energy(96, d1, f1, 6.0, r1, FP64, noalloc=True)
    assert np.array_equal(a1, b1)
    ref = np.array([O.vdw_energy(x, 6.0, d1, f1) for x in r1[::64]])
    assert np.array_equal(b1[::64], ref)
    # second call: the L=8 table (28 shells), 300 radii, FP32-compensated
    d2, f2 = O.shell_table(8, 6.0)
    assert d2.size < d1.size
    r2 = np.linspace(0.95, 1.25, 300)
    a2 = vdw_energy(8, d2, f2, 6.0, r2, FP32_COMPENSATED)
    b2 = vdw_energy(8, d2, f2, 6.0, r2, FP32_COMPENSATED, noalloc=True)
    assert np.array_equal(a2, b2)
    assert not np.array_equal(b2, vdw_energy(96, d1, f1, 6.0, r2, FP32_COMPENSATED, noalloc=True))


def test_noalloc_rejects_sizes_beyond_its_scratch():
    lib = load_library()
    d, f = O.shell_table(32, 6.0)
    with pytest.raises(CVFError, match="exceed its fixed scratch"):
        vdw_energy(32, d, f, 6.0, np.linspace(0.9, 1.3, 4097), FP64, noalloc=True)
    big_d = np.resize(d, 513).astype(np.float32)
    big_f = np.resize(f, 513).astype(np.int32)
    r = np.array([1.0])
    e = np.zeros(1)
    rc = lib.cvf_vdw_energy_noalloc(0, 32, big_d.ctypes.data_as(C.POINTER(C.c_float)),
                                    big_f.ctypes.data_as(C.POINTER(C.c_int32)), 513, 6.0,
                                    r.ctypes.data_as(C.POINTER(C.c_double)), 1, FP64,
                                    e.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc < 0 and b"exceed" in lib.cvf_last_error()
    # the entry point still works after a rejected call
    assert np.array_equal(vdw_energy(32, d, f, 6.0, r, FP64, noalloc=True), vdw_energy(32, d, f, 6.0, r, FP64))
