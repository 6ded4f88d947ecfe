"""The no-allocation telemetry entry point against aether_fluid_compute_telemetry.

aether_fluid_compute_telemetry_noalloc reduces fields the caller holds on the
device with its partial sums in a library-owned static device array.  It runs
the same kernels on the same reduction tiling as the solver's own telemetry,
so the frames must be equal bit for bit.  Calls in a row share the array:
each, on a smaller grid than the one before, must not see the earlier calls'
partial sums.  The first grid fills the array to its stated maximum.

The test plays the caller: it allocates its device buffers itself, through
the HIP runtime the library is linked against (the same process-wide
libamdhip64.so.7), and hands the library plain device pointers.
"""
import ctypes as C

import numpy as np
import pytest

from pairs import random_state

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    import aether_fluid as af
    af.lib()  # the library's runtime is the one already in the process
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    return hip


def _ok(rc):
    assert rc == 0, f"HIP error {rc}"


def _frame_noalloc(hip, stream, precision, dims, fields, t):
    import aether_fluid as af
    nx, ny, nz = dims
    dtype = np.float32 if precision == "fp32" else np.float64
    bufs = []
    try:
        for n in ("u", "v", "w", "density"):
            host = np.ascontiguousarray(fields[n], dtype=dtype)
            p = C.c_void_p()
            _ok(hip.hipMalloc(C.byref(p), host.nbytes))
            bufs.append(p)
            _ok(hip.hipMemcpy(p, host.ctypes.data, host.nbytes, H2D))
        out = C.c_void_p()
        _ok(hip.hipMalloc(C.byref(out), 64))
        bufs.append(out)
        nan = np.full(8, np.nan)
        _ok(hip.hipMemcpy(out, nan.ctypes.data, 64, H2D))
        af._lib.check(af.lib().aether_fluid_compute_telemetry_noalloc(
            af.FP32 if precision == "fp32" else af.FP64, nx, ny, nz, 0.0, *bufs[:4], float(t),
            out, stream))
        _ok(hip.hipStreamSynchronize(stream))
        got = np.empty(8)
        _ok(hip.hipMemcpy(got.ctypes.data, out, 64, D2H))
        return got
    finally:
        for p in bufs:
            hip.hipFree(p)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_noalloc_telemetry_equals_solver_telemetry(precision):
    import aether_fluid as af
    hip = _hip()
    stream = C.c_void_p()
    _ok(hip.hipStreamCreate(C.byref(stream)))
    try:
        # each grid smaller than the one before: 1024 x 1024 uses all 4096
        # partial slots of the array (16 x 256 reduction columns), 128 x 2200
        # needs 1100 (more than the ~1024 of most grids), 20 x 18 x 16 few
        for dims, t in (((1024, 1024, 6), 0.5), ((128, 2200, 6), 0.25), ((20, 18, 16), 1.5)):
            print(f"[case] noalloc telemetry {precision} {dims}", flush=True)
            s = af.FluidSolver3D(*dims, precision=precision)
            st = random_state(*dims, seed=sum(dims))
            for n in ("u", "v", "w", "density"):
                s.upload(n, st[n])
            ref = s.compute_telemetry(t).as_array()
            held = {n: s.download(n) for n in ("u", "v", "w", "density")}
            s.close()
            got = _frame_noalloc(hip, stream, precision, dims, held, t)
            np.testing.assert_array_equal(got, ref, err_msg=str(dims))
    finally:
        hip.hipStreamDestroy(stream)
