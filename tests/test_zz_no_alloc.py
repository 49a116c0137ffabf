"""rtm_hip_image_postprocess_noalloc, called through the C ABI: its output is
bit-identical to rtm_hip_image_postprocess's (the same kernels, the partial
maxima in the library's static scratch instead of an allocated buffer) for
every flag combination, FP32 image and FP64 stack, on two calls in a row of
which the second post-processes a smaller image (the scratch keeps the first
call's partial maxima, which must not leak into the second)."""
import ctypes as C

import numpy as np
import pytest

from conftest import bits_equal, line_shots, small_model

pytestmark = pytest.mark.gpu


def _migrated(eng, nx, nz, stack, seed):
    import rtm_amd
    model = small_model(nx, nz)
    params = rtm_amd.RTMParams(nb=12, order=8, store_interval=2, f0=15.0)
    nt, dt = 160, 0.001
    rng = np.random.default_rng(seed)
    shot = line_shots(model, nt, dt, [0.4 * (nx - 1) * model.dx])[0]
    shot.traces = (rng.standard_normal((shot.nrec, nt)) * 1e-2).astype(np.float32)
    eng.setup(model, params, nt, dt, image_stack=stack)
    eng.migrate([shot])
    return nx * nz


class DeviceBuffer:
    """Caller-owned device memory for the output (the caller may allocate;
    the entry point under test must not), through the HIP runtime."""

    def __init__(self, nbytes):
        self.hip = C.CDLL("libamdhip64.so")
        self.nbytes = nbytes
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(nbytes)) == 0
        # a NaN pattern, so a cell the call does not write fails the comparison
        assert self.hip.hipMemset(self.ptr, 0xFF, C.c_size_t(nbytes)) == 0
        # the fill runs on the null stream, which the engine's non-blocking
        # stream does not wait for: finish it before the call
        assert self.hip.hipDeviceSynchronize() == 0

    def get(self, n):
        out = np.empty(n, np.float32)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(4 * n), 2) == 0  # D2H
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


@pytest.mark.parametrize("stack", [0, 1])
def test_noalloc_postprocess_equals_allocating(hip_engine_cls, stack):
    with hip_engine_cls() as eng:
        # the larger image first, then a smaller one on the same handle
        for nx, nz, seed in ((130, 90, 5), (47, 33, 6)):
            n = _migrated(eng, nx, nz, stack, seed)
            for flags in (1, 2, 3):
                print(f"case: noalloc stack={stack} grid={nx}x{nz} flags={flags}", flush=True)
                ref = eng.postprocess(illum=bool(flags & 1), laplacian=bool(flags & 2))
                buf = DeviceBuffer(4 * n)
                try:
                    eng._check(eng.lib.rtm_hip_image_postprocess_noalloc(eng.h, flags, 1e-4, buf.ptr))
                    got = buf.get(n)
                finally:
                    buf.free()
                assert bits_equal(got, ref), f"max diff {np.nanmax(np.abs(got - ref))}"
                assert np.abs(ref).max() > 0
