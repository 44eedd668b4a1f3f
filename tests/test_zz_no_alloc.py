"""otm_conv_filter_noalloc (the filter for callers that cannot allocate)
against otm_conv_filter, called through the C ABI on device arrays: the
same kernels on the same inputs, so the outputs are bitwise equal. Two calls
in a row share the device scratch; the second has a smaller cell, so stale
weight sums of the first would show in it."""
import ctypes
import sys

import numpy as np
import pytest

import opentm_amd as otm
from opentm_amd import _capi

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def noalloc(x, kind, radius, periodic, adjoint):
    nz, ny, nx = x.shape
    reso = (ctypes.c_int * 3)(nx, ny, nz)
    din = otm.DeviceBuffer.from_numpy(x)
    dout = otm.DeviceBuffer(x.size)
    try:
        _capi.check(_capi.load().otm_conv_filter_noalloc(reso, int(kind), float(radius), int(periodic), int(adjoint),
                                                         din.ptr, dout.ptr, None),
                    "otm_conv_filter_noalloc")
        # otm_memcpy is a blocking copy on the default stream, behind the launches
        return dout.numpy().reshape(x.shape)
    finally:
        din.free()
        dout.free()


@pytest.mark.parametrize("kind,radius,periodic", [(0, 1.5, 0), (1, 2.0, 1), (1, 3.0, 0)])
def test_noalloc_filter_matches_allocating_entry_point(kind, radius, periodic):
    rng = np.random.default_rng(31)
    for shape in ((24, 20, 40), (8, 12, 16)):  # second call smaller
        x = rng.uniform(0.0, 1.0, shape).astype(np.float32)
        for adjoint in (0, 1):
            say("noalloc filter", kind, radius, periodic, "adjoint", adjoint, "shape", shape)
            ref = otm.conv_filter(x, kind, radius, bool(periodic), bool(adjoint))
            got = noalloc(x, kind, radius, periodic, adjoint)
            assert np.array_equal(got, ref)


def test_noalloc_filter_rejects_what_it_cannot_hold():
    lib = _capi.load()
    big = (ctypes.c_int * 3)(256, 256, 256)
    buf = otm.DeviceBuffer(64)
    try:
        st = lib.otm_conv_filter_noalloc(big, 0, 1.5, 0, 0, buf.ptr, buf.ptr, None)
        assert st == _capi.OTM_E_UNSUPPORTED
        assert b"scratch" in lib.otm_last_error()
        small = (ctypes.c_int * 3)(4, 4, 4)
        assert lib.otm_conv_filter_noalloc(small, 0, 1.5, 0, 0, buf.ptr, buf.ptr, None) == _capi.OTM_E_ARG
        assert b"overlap" in lib.otm_last_error()
    finally:
        buf.free()
