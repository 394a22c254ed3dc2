"""cpt_trace_closest_device_noalloc: the no-allocation variant of
cpt_trace_closest_device (traversal overflow stacks in a file-scope device
array) gives the same hits bit for bit, on two calls in a row, the second with
a smaller batch.  Device buffers come from hipMalloc through ctypes; the calls
under test get device pointers only."""
import ctypes as C

import numpy as np
import pytest

import cpt
from cpt import raybatch

pytestmark = pytest.mark.gpu


def _hip():
    try:
        return C.CDLL("libamdhip64.so")
    except OSError:
        return C.CDLL("/opt/rocm/lib/libamdhip64.so")


class _Dev:
    """A device buffer of nbytes (hipMalloc), freed on close."""

    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), C.c_size_t(nbytes)) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def get(self, shape, dtype):
        out = np.zeros(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(out.nbytes), 2) == 0
        return out

    def close(self):
        self.hip.hipFree(self.p)


def _device_rays(rays7):
    r = np.zeros((rays7.shape[0], 8), np.float32)
    r[:, 0:3] = rays7[:, 0:3]
    r[:, 3] = rays7[:, 6]
    r[:, 4:7] = rays7[:, 3:6]
    return r


@pytest.mark.parametrize("which", ["box", "terrain"])
def test_noalloc_trace_equals_allocating_entry(which):
    hip = _hip()
    scene = (cpt.scenes.box_room(res=(256, 256)) if which == "box"
             else cpt.scenes.terrain_scene(res=(320, 180), n_spheres=20, grid=(256, 128)))
    with cpt.Context(0) as ctx:
        ctx.upload(scene)
        for n, seed in ((1 << 18, 4), (3001, 9)):
            print(f"{which}: batch of {n} rays", flush=True)
            rays = raybatch.make_rays(scene, n, seed=seed)
            d_r, d_a, d_b = _Dev(hip, n * 32), _Dev(hip, n * 16), _Dev(hip, n * 16)
            try:
                d_r.put(_device_rays(rays))
                d_a.put(np.full((n, 4), -7, np.int32))
                d_b.put(np.full((n, 4), -9, np.int32))
                ctx._check(ctx.lib.cpt_trace_closest_device(ctx.h, d_r.p, n, d_a.p, None), "cpt_trace_closest_device")
                ctx._check(ctx.lib.cpt_trace_closest_device_noalloc(ctx.h, d_r.p, n, d_b.p, None),
                           "cpt_trace_closest_device_noalloc")
                assert hip.hipDeviceSynchronize() == 0
                a = d_a.get((n, 4), np.int32)
                b = d_b.get((n, 4), np.int32)
            finally:
                for d in (d_r, d_a, d_b):
                    d.close()
            assert np.array_equal(a, b)
            tri, t = ctx.trace_closest(rays)
            assert np.array_equal(b[:, 3], tri)
            hit = tri >= 0
            assert np.array_equal(b[hit, 0].view(np.float32), t[hit])
            print(f"{which}: {n} rays, {hit.mean() * 100:.1f}% hit, identical", flush=True)


def test_noalloc_rejects_bad_arguments():
    with cpt.Context(0) as ctx:
        assert ctx.lib.cpt_trace_closest_device_noalloc(ctx.h, None, 4, None, None) == cpt.CPT_E_INVALID
        ctx.upload(cpt.scenes.box_room(res=(16, 16)))
        assert ctx.lib.cpt_trace_closest_device_noalloc(ctx.h, C.c_void_p(16), -1, C.c_void_p(16), None) == \
            cpt.CPT_E_INVALID
        assert ctx.lib.cpt_trace_closest_device_noalloc(ctx.h, C.c_void_p(16), 0, C.c_void_p(16), None) == 0  # CPT_OK
