"""vof_diagnostics_noalloc (the entry point that allocates nothing) against
vof_diagnostics on the GPU.

The fields are the solver's own state after a few steps, copied into device
buffers the caller owns (allocated through the HIP runtime's C API); the
record is written to a caller-owned device buffer.  Two calls in a row, the
second on a smaller grid, so the shared static scratch is reused with fewer
partials than the first call left there; both must equal the solver's
vof_diagnostics bit for bit.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("F", "U", "V", "W", "P", "RHS")


class Hip:
    """The caller's own device memory, through the HIP runtime's C API (the
    entry point under test allocates nothing; its caller does)."""

    def __init__(self):
        try:  # the runtime libvof_hip.so is linked against, already loaded
            self.L = ctypes.CDLL("libamdhip64.so.7")
        except OSError:
            self.L = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def put(self, arr):
        a = np.ascontiguousarray(arr)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, nbytes):
        out = (ctypes.c_uint8 * nbytes)()
        assert self.L.hipMemcpy(ctypes.addressof(out), p, nbytes, 2) == 0  # device to host
        return bytes(out)

    def close(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _noalloc_record(vofmod, hip, params, fields):
    lib = vofmod.load_library()
    dev = [hip.put(fields[k]) for k in FIELDS]
    n = ctypes.sizeof(vofmod.VofDiag)
    out = hip.put(np.full(n, 0xAB, dtype=np.uint8))  # stale bytes: every field of the record must be written
    vofmod._check(lib.vof_diagnostics_noalloc(ctypes.byref(params), *dev, out, None))
    assert hip.L.hipDeviceSynchronize() == 0
    d = vofmod.VofDiag.from_buffer_copy(hip.get(out, n))
    return {k: getattr(d, k) for k, _ in vofmod.VofDiag._fields_}


@pytest.mark.parametrize("prec", [4, 8])
def test_noalloc_diagnostics_equal_vof_diagnostics(vofmod, prec):
    hip = Hip()
    try:
        for n, steps in ((40, 6), (22, 3)):  # second call: a smaller grid
            print(f"noalloc diagnostics, {n}^3, prec {prec}, {steps} steps", flush=True)
            p = vofmod.default_params(n=n, precision=prec, g=(0.0, -5.0, -9.81))
            s = vofmod.Solver(p)
            s.init_ic(1)
            s.step(1, steps, graph=True)
            fields = {k: s.download(k) for k in FIELDS}
            want = s.diagnostics()
            got = _noalloc_record(vofmod, hip, p, fields)
            assert got == want, (got, want)
            s.close()
    finally:
        hip.close()


def test_noalloc_diagnostics_refuses_bad_arguments(vofmod):
    lib = vofmod.load_library()
    p = vofmod.default_params(n=8)
    with pytest.raises(vofmod.VofError):
        vofmod._check(lib.vof_diagnostics_noalloc(ctypes.byref(p), None, None, None, None, None, None, None, None))
    # 5002 x 4002 stored cells per plane: 78,158 work-groups even with one chunk
    big = vofmod.default_params(nx=5000, ny=4000, nz=1, precision=4)
    one = ctypes.c_void_p(1)  # never dereferenced: the grid is refused before any launch
    with pytest.raises(vofmod.VofError, match="static"):
        vofmod._check(lib.vof_diagnostics_noalloc(ctypes.byref(big), one, one, one, one, one, one, one, None))
