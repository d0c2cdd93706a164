"""vps_mlp_loss_and_grad_device, the entry point that allocates nothing,
against vps_mlp_loss_and_grad on the same batch, through the C ABI: two
calls in a row, the second on a smaller batch (the scratch keeps the first
call's partials; every partial must be rewritten).  The test's device
buffers come from hipMalloc of the HIP runtime libvps.so is linked against.
Same kernels, so the results must be bit-equal."""
import ctypes as C
import os

import numpy as np
import pytest

import mlp_oracle as M
from vps import Projection, scenes

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    for path in ("/opt/rocm/lib/libamdhip64.so.7", "libamdhip64.so.7"):
        if path.startswith("/") and not os.path.exists(path):
            continue
        lib = C.CDLL(path)
        lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        lib.hipFree.argtypes = [C.c_void_p]
        lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        return lib
    raise RuntimeError("HIP runtime not found")


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), nbytes) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.zeros(count, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    f = rng.normal(size=(n, 8)).astype(np.float32)
    t = rng.normal(size=n).astype(np.float32)
    return f, t


def test_device_loss_and_grad_matches_the_host_entry_point():
    print("no-alloc entry point: sizes 65536 then 5003", flush=True)
    g = Projection(scenes.small_3d(12, 8, 6))
    g.mlp_set_parameters(M.random_params(31).astype(np.float32))
    hip = _hip()
    nmax = 65536
    d_f, d_t = DevBuf(hip, nmax * 8 * 4), DevBuf(hip, nmax * 4)
    d_loss, d_grad = DevBuf(hip, 8), DevBuf(hip, 4801 * 4)
    for n, seed in ((nmax, 1), (5003, 2)):
        f, t = _batch(n, seed)
        loss_h, grad_h = g.mlp_loss_and_grad(f, t)
        d_f.put(f)
        d_t.put(t)
        g.mlp_loss_and_grad_device(n, d_f.p.value, d_t.p.value, d_loss.p.value, d_grad.p.value)
        g.synchronize()
        loss_d = float(d_loss.get(np.float64, 1)[0])
        grad_d = d_grad.get(np.float32, 4801)
        assert loss_d == loss_h, n
        assert np.array_equal(grad_d, grad_h), n
        assert np.linalg.norm(grad_h) > 0
    # sizes beyond the maximum and NULL pointers are refused before any launch
    with pytest.raises(Exception):
        g.mlp_loss_and_grad_device(67_108_865, d_f.p.value, d_t.p.value, d_loss.p.value,
                                   d_grad.p.value)
    with pytest.raises(Exception):
        g.mlp_loss_and_grad_device(16, 0, d_t.p.value, d_loss.p.value, d_grad.p.value)
    g.close()
    for b in (d_f, d_t, d_loss, d_grad):
        b.free()
