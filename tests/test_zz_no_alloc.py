"""The allocation-free transform entry point (dns_fft_execute_device_noalloc) against the
allocating device entry points (dns_fft_forward_device / dns_fft_backward_device), through the
C ABI on HBM pointers: bit for bit, on two calls in a row, the second with fewer planes.

Device memory comes from the HIP runtime the library itself links (libamdhip64.so.7), through
ctypes: a second HIP runtime in the process (torch bundles its own) would not share devices
reliably with the first."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FWD, BWD = -1, 1
H2D, D2H = 1, 2   # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class Hip:
    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        self.held = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.held.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def nan_filled(self, nbytes):
        p = self.alloc(nbytes)
        assert self.L.hipMemset(p, 0xFF, nbytes) == 0     # all-ones bytes: NaN doubles
        return p

    def download(self, p, shape):
        out = np.empty(shape)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free_all(self):
        self.L.hipDeviceSynchronize()
        for p in self.held:
            self.L.hipFree(p)
        self.held = []


@pytest.fixture(scope="module")
def env():
    import f90_dns_amd
    L = f90_dns_amd.lib()
    return f90_dns_amd, L, Hip()


def _plan(L, nx, ny, precision):
    h = ctypes.c_void_p()
    assert L.dns_fft_plan_create(nx, ny, precision, ctypes.byref(h)) == 0, L.dns_last_error(None)
    return h


# 192 x 192 and 256 x 128 FP64 take the two passes through the scratch (70 planes of 192 x 192
# are two chunks of it); 128 x 128 FP64 and 192 x 192 FP32 are LDS-resident (no scratch)
@pytest.mark.parametrize("nx,ny,precision,planes", [(192, 192, 0, (70, 9)), (256, 128, 0, (17, 3)),
                                                    (128, 128, 0, (12, 4)), (192, 192, 1, (6, 2))])
def test_noalloc_transforms_equal_allocating_ones(env, nx, ny, precision, planes):
    dns, L, hip = env
    _say(f"noalloc {nx}x{ny} precision {precision} planes {planes}")
    h = _plan(L, nx, ny, precision)
    nxh = nx // 2 + 1
    try:
        for call, npl in enumerate(planes):
            f = np.random.default_rng(17 + call).standard_normal((npl, ny, nx))
            df = hip.upload(f)
            nc = npl * ny * nxh * 2 * 8
            X1, X2 = hip.alloc(nc), hip.nan_filled(nc)
            assert L.dns_fft_forward_device(h, df, X1, npl, None) == 0
            assert L.dns_fft_execute_device_noalloc(h, FWD, df, X2, npl, None) == 0
            x1 = hip.download(X1, (npl, ny, nxh, 2))
            x2 = hip.download(X2, (npl, ny, nxh, 2))
            assert np.array_equal(x1, x2), (call, "forward")
            nr = npl * ny * nx * 8
            B1, B2 = hip.alloc(nr), hip.nan_filled(nr)
            assert L.dns_fft_backward_device(h, X1, B1, npl, None) == 0
            assert L.dns_fft_execute_device_noalloc(h, BWD, X1, B2, npl, None) == 0
            b1 = hip.download(B1, (npl, ny, nx))
            b2 = hip.download(B2, (npl, ny, nx))
            assert np.array_equal(b1, b2), (call, "backward")
            # and they are the transforms: numpy's rfft2, and the round trip
            ref = np.fft.rfft2(f)
            tol = 2e-6 if precision else 1e-11
            assert np.max(np.abs(x2[..., 0] + 1j * x2[..., 1] - ref)) < tol * np.abs(ref).max()
            assert np.max(np.abs(b2 - f)) < (1e-5 if precision else 1e-13) * np.abs(f).max()
            hip.free_all()
    finally:
        hip.free_all()
        L.dns_fft_plan_destroy(h)


def test_noalloc_rejects_bad_direction(env):
    dns, L, hip = env
    h = _plan(L, 192, 192, 0)
    try:
        x = hip.alloc(192 * 192 * 8)
        y = hip.alloc(192 * 97 * 16)
        assert L.dns_fft_execute_device_noalloc(h, 0, x, y, 1, None) == 1
        assert b"bad argument" in L.dns_last_error(None)
    finally:
        hip.free_all()
        L.dns_fft_plan_destroy(h)
