"""amps_sf_step_fp64_noalloc, the FP64 snowflake step that allocates no
device memory, against amps_step of a PURE_FP64 context of the same config:
bit for bit, on two calls in a row (two grids, the second smaller), three
steps each, through the C ABI.  The caller's buffers are allocated here with
the HIP runtime the library itself loaded."""
import ctypes

import numpy as np
import pytest

from tests.helpers import rel_err, smooth_fields

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libamps.so loaded."""

    def __init__(self, amps):
        amps.lib()
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
        self.bufs.append(p)
        return p.value

    def upload(self, dst, a):
        a = np.ascontiguousarray(a, np.float64)
        assert self.h.hipMemcpy(ctypes.c_void_p(dst), a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0

    def download(self, src, shape):
        out = np.empty(shape, np.float64)
        assert self.h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(src), ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def config(amps, O, n, tile, stencil, bc):
    c = amps.Config.default(amps.SNOWFLAKE)
    c.nx = c.ny = n
    c.tile, c.stencil, c.bc, c.mode = tile, stencil, bc, amps.PURE_FP64
    p = O.sf_params_config(n, n)
    for i in range(17):
        c.params[i] = p[i]
    c.seed = 9
    return c, p


def test_noalloc_step_equals_context_step(amps, O):
    hip = Hip(amps)
    try:
        for n, tile, stencil, bc in ((256, 32, 1, 1), (128, 16, 0, 0)):  # the second call is smaller
            c, p = config(amps, O, n, tile, stencil, bc)
            phi, T = smooth_fields(n, n, seed=n, nbump=3)
            ref = []
            with amps.Context(c) as ctx:
                ctx.init()
                ctx.set_field(0, phi)
                ctx.set_field(1, T)
                for _ in range(3):
                    ctx.step(1)
                    ref.append(np.stack([ctx.get_field(0), ctx.get_field(1)]))
            a, b = hip.alloc(2 * n * n * 8), hip.alloc(2 * n * n * 8)
            hip.upload(a, np.stack([phi, T]))
            po, to = phi, T
            for s in range(3):
                amps.sf_step_fp64_noalloc(c, a, b, s)
                hip.sync()
                got = hip.download(b, (2, n, n))
                assert np.array_equal(got, ref[s]), (n, s)
                po, to = O.sf_step(po, to, p, stencil=stencil, bc=bc, seed=c.seed, step=s)
                assert rel_err(got[0], po) <= 1e-12 and rel_err(got[1], to) <= 1e-12
                po, to = got[0], got[1]
                a, b = b, a
    finally:
        hip.free()


def test_noalloc_step_rejects_bad_arguments(amps, O):
    hip = Hip(amps)
    try:
        c, _ = config(amps, O, 64, 16, 1, 1)
        a = hip.alloc(2 * 64 * 64 * 8)
        with pytest.raises(amps.AmpsError):
            amps.sf_step_fp64_noalloc(c, a, a + 8, 0)  # overlapping buffers
        b = hip.alloc(2 * 64 * 64 * 8)
        g = c.copy()
        g.model = amps.GRAINGROWTH
        with pytest.raises(amps.AmpsError):
            amps.sf_step_fp64_noalloc(g, a, b, 0)
        g = c.copy()
        g.mode = amps.ADAPTIVE
        with pytest.raises(amps.AmpsError):
            amps.sf_step_fp64_noalloc(g, a, b, 0)
        g = c.copy()
        g.nx = 72  # not a multiple of the tile
        with pytest.raises(amps.AmpsError):
            amps.sf_step_fp64_noalloc(g, a, b, 0)
        with pytest.raises(amps.AmpsError):
            amps.sf_step_fp64_noalloc(c, a, b, -1)
    finally:
        hip.free()
