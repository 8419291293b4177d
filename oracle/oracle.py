"""ctypes bindings of the CPU oracle (oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker of the HIP engine.  The product
(w3914777_amd) never imports this module.

Arrays are float64 numpy arrays, row-major [ny][nx] (grain growth: [q][ny][nx]).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

# parameter slots, mirrored from include/amps.h
SF = dict(DX=0, DY=1, DT=2, TAU=3, EPSB=4, KAPPA=5, DELTA=6, ANISO=7, ALPHA=8, GAMMA=9,
          TEQ=10, THETA0=11, PI=12, NOISE=13, R2=14, CX=15, CY=16)
GG = dict(DX=0, DY=1, DT=2, L=3, KAPPA=4, ALPHA=5, BETA=6, GAMMA=7, R=8, CX=9, CY=10, AMP=11)
SN = dict(DX=0, DY=1, DT=2, A=3, B=4, KC=5, KE=6, L=7, DVOL=8, DVAP=9, DSURF=10, DGB=11,
          R1=12, R2=13, X1=14, Y1=15, X2=16, Y2=17, WIDTH=18, NOISE=19)
NPARAMS = 32
STENCIL_REF, STENCIL_LAP5 = 0, 1
BC_PERIODIC, BC_NEUMANN = 0, 1
INIT_REF, INIT_CONFIG = 0, 1
MON_SPATIAL, MON_TEMPORAL, MON_GRADIENT = 1, 2, 3
GRAD_SF = dict(DELTA=1, KAPPA=2, EPSB=3, TAU=4)
GRAD_SN = dict(DSURF=101, DGB=102, KC=103, KE=104, L=105, DVOL=106)


def build(force=False):
    """Compile liboracle.so with the committed Makefile (g++, seconds)."""
    srcs = [os.path.join(_HERE, f) for f in ("oracle.cpp", "low_emulation.cpp")]
    if force or not os.path.exists(_LIB) or os.path.getmtime(_LIB) < max(os.path.getmtime(f) for f in srcs):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB)
        D = ctypes.POINTER(ctypes.c_double)
        I = ctypes.c_int
        I32P = ctypes.POINTER(ctypes.c_int32)
        U64 = ctypes.c_uint64
        I64 = ctypes.c_int64
        L.oracle_rng_u01.argtypes = [U64, U64, U64]
        L.oracle_rng_u01.restype = ctypes.c_double
        L.oracle_noise_u.argtypes = [U64, U64, ctypes.c_uint32]
        L.oracle_noise_u.restype = ctypes.c_double
        L.oracle_bmap.argtypes = [I, I, I]
        L.oracle_f64_to_f16.argtypes = [ctypes.c_double]
        L.oracle_f64_to_f16.restype = ctypes.c_uint16
        L.oracle_f16_to_f64.argtypes = [ctypes.c_uint16]
        L.oracle_f16_to_f64.restype = ctypes.c_double
        L.oracle_sf_init.argtypes = [I, I, D, D, D]
        L.oracle_gg_init.argtypes = [I, I, I, D, I, U64, D]
        L.oracle_sn_init.argtypes = [I, I, D, I, U64, D, D, D]
        L.oracle_sf_step.argtypes = [I, I, D, I, I, U64, I64, D, D, D, D, I]
        L.oracle_gg_step.argtypes = [I, I, I, D, I, I, D, D, I]
        L.oracle_sn_step.argtypes = [I, I, D, I, D, D, D, D, D, D, I]
        L.oracle_sn_step_conservative.argtypes = [I, I, D, I, D, D, D, D, D, D, I]
        L.oracle_sf_run.argtypes = [I, I, D, I, I, U64, I64, I, D, D, I]
        L.oracle_sf_run.restype = ctypes.c_double
        L.oracle_gg_run.argtypes = [I, I, I, D, I, I, I, D, I]
        L.oracle_gg_run.restype = ctypes.c_double
        L.oracle_sn_run.argtypes = [I, I, D, I, I, D, D, D, I]
        L.oracle_sn_run.restype = ctypes.c_double
        L.oracle_sf_tangent.argtypes = [I, I, D, I, I, U64, I64, I, I, D, D, D, D, I]
        L.oracle_sn_tangent.argtypes = [I, I, D, I, I, I, D, D, D, D, D, D, I]
        L.oracle_sample_offsets.argtypes = [I, I, ctypes.POINTER(ctypes.c_int)]
        L.oracle_monitor_spatial.argtypes = [I, I, I, D, D]
        L.oracle_monitor_temporal.argtypes = [I, I, I, D, D, D]
        L.oracle_monitor_gradient.argtypes = [I, I, I, I, ctypes.c_double, ctypes.c_double, D, D]
        L.oracle_get_type.argtypes = [I, I, D, ctypes.c_double, I32P, I32P]
        U16P = ctypes.POINTER(ctypes.c_uint16)
        FP = ctypes.POINTER(ctypes.c_float)
        L.oracle_sf_low_step.argtypes = [I, I, I, D, I, I, U64, I64, I32P, D, D, U16P, D, FP, ctypes.c_float,
                                         U16P, D, FP, I]
        L.oracle_rh.argtypes = [ctypes.c_double]
        L.oracle_rh.restype = ctypes.c_double
        L.oracle_atan_h2.argtypes = [FP, FP, I]
        L.oracle_aniso6_h2.argtypes = [FP, FP, ctypes.c_float, ctypes.c_float, FP, FP, I]
        L.oracle_h2d.argtypes = [ctypes.c_uint16]
        L.oracle_h2d.restype = ctypes.c_double
        L.oracle_f2h.argtypes = [ctypes.c_float]
        L.oracle_f2h.restype = ctypes.c_uint16
        L.oracle_gg_low_step.argtypes = [I, I, I, I, D, I, I, I32P, D, U16P, D, FP, U16P, D, FP]
        _lib = L
    return _lib


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i32(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def params(values):
    p = np.zeros(NPARAMS, np.float64)
    for k, v in values.items():
        p[k] = v
    return p


# ------------------------------------------------------------------ defaults
def sf_params_ref(k=1, nx=512, ny=512):
    """snowflake/paras/para{k}.h values (k=1..5) with the reference's nucleus
    r^2 < 10 at (nx/2+8, ny/2+8) and pi = 3.1415926 (total.h:47)."""
    aniso = {1: 4.0, 2: 6.0, 3: 5.0, 4: 6.0, 5: 4.0}[k]
    theta0 = {1: 0.0, 2: 0.0, 3: 0.0, 4: 0.4, 5: 0.25}[k]
    v = {SF["DX"]: 0.03, SF["DY"]: 0.03, SF["DT"]: 1.0e-4, SF["TAU"]: 0.0003, SF["EPSB"]: 0.01,
         SF["KAPPA"]: 1.8, SF["DELTA"]: 0.02, SF["ANISO"]: aniso, SF["ALPHA"]: 0.9, SF["GAMMA"]: 10.0,
         SF["TEQ"]: 1.0, SF["THETA0"]: theta0, SF["PI"]: 3.1415926, SF["NOISE"]: 0.0, SF["R2"]: 10.0,
         SF["CX"]: nx // 2 + 8, SF["CY"]: ny // 2 + 8}
    return params(v)


def sf_params_config(nx=4096, ny=4096):
    """Snowflake workload parameters: eps = 0.01(1+0.05cos(6(theta-pi/2))),
    tau=3e-4, alpha=0.9, gamma=10, K=1.6, Teq=1, dx=0.03, dt=1e-4, radius-5
    disc at the centre, noise a=0.01."""
    v = {SF["DX"]: 0.03, SF["DY"]: 0.03, SF["DT"]: 1.0e-4, SF["TAU"]: 0.0003, SF["EPSB"]: 0.01,
         SF["KAPPA"]: 1.6, SF["DELTA"]: 0.05, SF["ANISO"]: 6.0, SF["ALPHA"]: 0.9, SF["GAMMA"]: 10.0,
         SF["TEQ"]: 1.0, SF["THETA0"]: 3.1415926 / 2, SF["PI"]: 3.1415926, SF["NOISE"]: 0.01,
         SF["R2"]: 25.0, SF["CX"]: nx // 2, SF["CY"]: ny // 2}
    return params(v)


def gg_params_ref(k=1, nx=512, ny=512):
    """graingrowth/paras/para{k}.h values."""
    L = {1: 5.0, 2: 10.0, 3: 5.0, 4: 5.0, 5: 20.0}[k]
    kap = {1: 0.1, 2: 0.1, 3: 0.2, 4: 0.1, 5: 0.1}[k]
    R = {1: 50.0, 2: 50.0, 3: 50.0, 4: 80.0, 5: 50.0}[k]
    v = {GG["DX"]: 0.5, GG["DY"]: 0.5, GG["DT"]: 0.005, GG["L"]: L, GG["KAPPA"]: kap, GG["ALPHA"]: 1.0,
         GG["BETA"]: 1.0, GG["GAMMA"]: 1.0, GG["R"]: R, GG["CX"]: nx // 2, GG["CY"]: ny // 2, GG["AMP"]: 0.0}
    return params(v)


def gg_params_config():
    v = {GG["DX"]: 0.5, GG["DY"]: 0.5, GG["DT"]: 0.005, GG["L"]: 5.0, GG["KAPPA"]: 0.5, GG["ALPHA"]: 1.0,
         GG["BETA"]: 1.0, GG["GAMMA"]: 1.0, GG["R"]: 0.0, GG["CX"]: 0.0, GG["CY"]: 0.0, GG["AMP"]: 1e-3}
    return params(v)


def sn_params_ref(k=1, nx=512, ny=512):
    """sintering/paras/para{k}.h values (coefm=kappa_c, coefk=kappa_eta,
    coefl=L; A=16, B=1 from sintering/function.h:72)."""
    coefm = {1: 5.0, 2: 5.0, 3: 20.0, 4: 20.0, 5: 20.0}[k]
    dsur = {1: 32.0, 2: 64.0, 3: 16.0, 4: 16.0, 5: 64.0}[k]
    R1 = {1: 50.0, 2: 50.0, 3: 50.0, 4: 20.0, 5: 20.0}[k]
    R2 = R1 / 2
    v = {SN["DX"]: 0.5, SN["DY"]: 0.5, SN["DT"]: 1.0e-4, SN["A"]: 16.0, SN["B"]: 1.0, SN["KC"]: coefm,
         SN["KE"]: 2.0, SN["L"]: 5.0, SN["DVOL"]: 0.040, SN["DVAP"]: 0.002, SN["DSURF"]: dsur,
         SN["DGB"]: 1.6, SN["R1"]: R1, SN["R2"]: R2, SN["X1"]: nx // 2 + 8, SN["Y1"]: 240.0,
         SN["X2"]: nx // 2 + 8, SN["Y2"]: 240.0 + R1 + R2, SN["WIDTH"]: 0.0, SN["NOISE"]: 0.0}
    return params(v)


def sn_params_config(nx=1024, ny=1024, R=150.0):
    v = {SN["DX"]: 0.5, SN["DY"]: 0.5, SN["DT"]: 1.0e-4, SN["A"]: 16.0, SN["B"]: 1.0, SN["KC"]: 5.0,
         SN["KE"]: 2.0, SN["L"]: 10.0, SN["DVOL"]: 0.04, SN["DVAP"]: 0.002, SN["DSURF"]: 16.0,
         SN["DGB"]: 1.6, SN["R1"]: R, SN["R2"]: R, SN["X1"]: nx / 2 - R, SN["Y1"]: ny / 2,
         SN["X2"]: nx / 2 + R, SN["Y2"]: ny / 2, SN["WIDTH"]: 4.0, SN["NOISE"]: 1e-3}
    return params(v)


# ------------------------------------------------------------------ wrappers
def rng_u01(seed, stream, idx):
    return lib().oracle_rng_u01(seed, stream, idx)


def noise_u(seed, step, cell):
    return lib().oracle_noise_u(seed, step, cell)


def f64_to_f16_bits(x):
    return lib().oracle_f64_to_f16(float(x))


def f16_bits_to_f64(h):
    return lib().oracle_f16_to_f64(int(h))


def sf_init(nx, ny, p):
    phi = np.zeros((ny, nx)); T = np.zeros((ny, nx))
    lib().oracle_sf_init(nx, ny, _d(p), _d(phi), _d(T))
    return phi, T


def gg_init(nx, ny, q, p, init=INIT_REF, seed=1):
    eta = np.zeros((q, ny, nx))
    lib().oracle_gg_init(nx, ny, q, _d(p), init, seed, _d(eta))
    return eta


def sn_init(nx, ny, p, init=INIT_REF, seed=1):
    c = np.zeros((ny, nx)); e1 = np.zeros((ny, nx)); e2 = np.zeros((ny, nx))
    lib().oracle_sn_init(nx, ny, _d(p), init, seed, _d(c), _d(e1), _d(e2))
    return c, e1, e2


def sf_step(phi, T, p, stencil=STENCIL_REF, bc=BC_PERIODIC, seed=1, step=0, threads=0):
    ny, nx = phi.shape
    phi = np.ascontiguousarray(phi, np.float64); T = np.ascontiguousarray(T, np.float64)
    po = np.empty_like(phi); to = np.empty_like(T)
    lib().oracle_sf_step(nx, ny, _d(p), stencil, bc, seed, step, _d(phi), _d(T), _d(po), _d(to), threads)
    return po, to


def sf_run(phi, T, p, nsteps, stencil=STENCIL_REF, bc=BC_PERIODIC, seed=1, step0=0, threads=0):
    """Returns (phi, T, seconds)."""
    ny, nx = phi.shape
    phi = np.array(phi, np.float64, order="C"); T = np.array(T, np.float64, order="C")
    sec = lib().oracle_sf_run(nx, ny, _d(p), stencil, bc, seed, step0, nsteps, _d(phi), _d(T), threads)
    return phi, T, sec


def gg_step(eta, p, stencil=STENCIL_REF, bc=BC_PERIODIC, threads=0):
    q, ny, nx = eta.shape
    eta = np.ascontiguousarray(eta, np.float64)
    out = np.empty_like(eta)
    lib().oracle_gg_step(nx, ny, q, _d(p), stencil, bc, _d(eta), _d(out), threads)
    return out


def gg_run(eta, p, nsteps, stencil=STENCIL_REF, bc=BC_PERIODIC, threads=0):
    q, ny, nx = eta.shape
    eta = np.array(eta, np.float64, order="C")
    sec = lib().oracle_gg_run(nx, ny, q, _d(p), stencil, bc, nsteps, _d(eta), threads)
    return eta, sec


def sn_step(c, e1, e2, p, bc=BC_PERIODIC, threads=0, conservative=False):
    """One sintering step; conservative: the Cahn-Hilliard update in
    divergence form (face mobilities 0.5 (M_left + M_right))."""
    ny, nx = c.shape
    c, e1, e2 = (np.ascontiguousarray(a, np.float64) for a in (c, e1, e2))
    co, ao, bo = np.empty_like(c), np.empty_like(c), np.empty_like(c)
    f = lib().oracle_sn_step_conservative if conservative else lib().oracle_sn_step
    f(nx, ny, _d(p), bc, _d(c), _d(e1), _d(e2), _d(co), _d(ao), _d(bo), threads)
    return co, ao, bo


def sn_run(c, e1, e2, p, nsteps, bc=BC_PERIODIC, threads=0):
    ny, nx = c.shape
    c, e1, e2 = (np.array(a, np.float64, order="C") for a in (c, e1, e2))
    sec = lib().oracle_sn_run(nx, ny, _d(p), bc, nsteps, _d(c), _d(e1), _d(e2), threads)
    return c, e1, e2, sec


def sf_tangent(phi, T, p, nsteps, param_id, dphi=None, dT=None, stencil=STENCIL_REF, bc=BC_PERIODIC,
               seed=1, step0=0, threads=0):
    """Forward-mode: final (phi, T) and d(phi, T)/d(param) after nsteps."""
    ny, nx = phi.shape
    phi = np.array(phi, np.float64, order="C"); T = np.array(T, np.float64, order="C")
    dphi = np.zeros_like(phi) if dphi is None else np.array(dphi, np.float64, order="C")
    dT = np.zeros_like(T) if dT is None else np.array(dT, np.float64, order="C")
    rc = lib().oracle_sf_tangent(nx, ny, _d(p), stencil, bc, seed, step0, nsteps, param_id,
                                 _d(phi), _d(T), _d(dphi), _d(dT), threads)
    assert rc == 0
    return phi, T, dphi, dT


def sn_tangent(c, e1, e2, p, nsteps, param_id, dc=None, de1=None, de2=None, bc=BC_PERIODIC, threads=0):
    ny, nx = c.shape
    c, e1, e2 = (np.array(a, np.float64, order="C") for a in (c, e1, e2))
    z = lambda a, ref: np.zeros_like(ref) if a is None else np.array(a, np.float64, order="C")
    dc, de1, de2 = z(dc, c), z(de1, c), z(de2, c)
    rc = lib().oracle_sn_tangent(nx, ny, _d(p), bc, nsteps, param_id, _d(c), _d(e1), _d(e2),
                                 _d(dc), _d(de1), _d(de2), threads)
    assert rc == 0
    return c, e1, e2, dc, de1, de2


def sample_offsets(tile, kind):
    out = (ctypes.c_int * 4)()
    lib().oracle_sample_offsets(tile, kind, out)
    return list(out)


def monitor_spatial(f, tile):
    ny, nx = f.shape
    md = np.zeros((ny // tile, nx // tile))
    lib().oracle_monitor_spatial(nx, ny, tile, _d(np.ascontiguousarray(f)), _d(md))
    return md


def monitor_temporal(f_last, f, tile):
    ny, nx = f.shape
    md = np.zeros((ny // tile, nx // tile))
    lib().oracle_monitor_temporal(nx, ny, tile, _d(np.ascontiguousarray(f_last)), _d(np.ascontiguousarray(f)), _d(md))
    return md


def monitor_gradient(f, tile, dx=1.0, dy=1.0, bc=BC_PERIODIC):
    ny, nx = f.shape
    md = np.zeros((ny // tile, nx // tile))
    lib().oracle_monitor_gradient(nx, ny, tile, bc, dx, dy, _d(np.ascontiguousarray(f)), _d(md))
    return md


def get_type(md, thr, types):
    """Returns (type_old, type_new) after one get_type pass."""
    ty, tx = md.shape
    t = np.array(types, np.int32, order="C").reshape(ty, tx)
    old = np.zeros_like(t)
    lib().oracle_get_type(tx, ty, _d(np.ascontiguousarray(md)), thr, _i32(old), _i32(t))
    return old, t


# --------------------------------------------------------- low-tile emulation
def _u16(a):
    assert a.dtype == np.uint16 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))


def _f32(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def sf_low_step(types, phi, T, words, off, scl, p, tile, stencil=STENCIL_REF, bc=BC_PERIODIC, seed=1, step=0,
                noise_skip=0.0, half2=False):
    """One step of the low tiles (type 1) of a snowflake grid, emulating the
    engine's fp16-storage / FP32-arithmetic scheme (low_emulation.cpp);
    half2: the AMPS_FLAG_HALF2 form (packed-half2 anisotropy and atan).
    words [2, ny, nx] uint16, off [2, ty, tx] float64, scl [2, ty, tx]
    float32; phi, T hold the FP64 tiles' values.  Returns (words, off, scl)."""
    ny, nx = phi.shape
    types = np.ascontiguousarray(types, np.int32)
    words = np.ascontiguousarray(words, np.uint16)
    off = np.ascontiguousarray(off, np.float64)
    scl = np.ascontiguousarray(scl, np.float32)
    wo, oo, so = np.empty_like(words), np.empty_like(off), np.empty_like(scl)
    lib().oracle_sf_low_step(nx, ny, tile, _d(p), stencil, bc, seed, step, _i32(types),
                             _d(np.ascontiguousarray(phi, np.float64)), _d(np.ascontiguousarray(T, np.float64)),
                             _u16(words), _d(off), _f32(scl), float(noise_skip), _u16(wo), _d(oo), _f32(so),
                             int(bool(half2)))
    return wo, oo, so


def fp16_round(x):
    """The emulator's rounding of a double to fp16 (nearest even), as a double."""
    return lib().oracle_rh(float(x))


def atan_h2(x):
    """The engine's packed-half2 atan (AMPS_FLAG_HALF2) of float32 arguments."""
    x = np.ascontiguousarray(x, np.float32)
    r = np.empty_like(x)
    lib().oracle_atan_h2(_f32(x), _f32(r), x.size)
    return r


def aniso6_h2(u, v, cth0, sth0):
    """The engine's packed-half2 (cos 6(theta-theta0), sin 6(theta-theta0)) of
    rescaled gradients (u, v) (max(|u|,|v|) in [0.5, 1), or (1, 0))."""
    u = np.ascontiguousarray(u, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    ca, sa = np.empty_like(u), np.empty_like(u)
    lib().oracle_aniso6_h2(_f32(u), _f32(v), float(cth0), float(sth0), _f32(ca), _f32(sa), u.size)
    return ca, sa


_H2D = None


def f16_words_to_f64(words):
    """Exact values of an array of fp16 words."""
    global _H2D
    if _H2D is None:
        _H2D = np.array([lib().oracle_h2d(i) for i in range(65536)], np.float64)
    return _H2D[np.asarray(words, np.uint16)]


def decode_tiles(words, off, scl, tile):
    """off + h * scl per cell of a [ny, nx] word array with [ty, tx] codes."""
    h = f16_words_to_f64(words)
    return np.repeat(np.repeat(off, tile, 0), tile, 1) + h * np.repeat(np.repeat(scl.astype(np.float64), tile, 0), tile, 1)


def f16_ulp_value(words, scl, tile):
    """Spacing of the fp16 grid at each word's magnitude (the gap to the next
    larger fp16 value), times its tile's scale: the resolution of a low
    tile's stored value at that cell."""
    h = np.abs(f16_words_to_f64(words)).astype(np.float16)
    gap = np.spacing(h).astype(np.float64)
    return gap * np.repeat(np.repeat(np.asarray(scl, np.float64), tile, 0), tile, 1)


def _f32_down(x):
    f = np.float32(x)
    return np.nextafter(f, np.float32(-np.inf)) if float(f) > x else f


def _f32_up(x):
    f = np.float32(x)
    return np.nextafter(f, np.float32(np.inf)) if float(f) < x else f


def sn_encode_low(v, types, tile):
    """The engine's re-encoding of a sintering eta field on its low tiles
    (types == 1), a restatement of its low-tile arithmetic (no reference
    counterpart: the reference's type-1 tiles compute in half2): the tile's
    bounds are the minimum rounded down and the maximum rounded up to float;
    off = (lo + hi) / 2 and hw = (hi - lo) / 2 in double; e is frexp's
    exponent of hw (-100 for hw = 0, clamped at -100); each word is the RNE
    fp16 of float(v - off) * 2^(14 - e); the scale is 2^(e - 14) (0 for
    hw = 0).  Returns (words [ny, nx], off [ty, tx], scl [ty, tx]); FP64
    tiles are left zero."""
    L = lib()
    ny, nx = v.shape
    ty, tx = ny // tile, nx // tile
    words = np.zeros((ny, nx), np.uint16)
    off = np.zeros((ty, tx))
    scl = np.zeros((ty, tx), np.float32)
    for j in range(ty):
        for i in range(tx):
            if types[j, i] != 1:
                continue
            b = v[j * tile:(j + 1) * tile, i * tile:(i + 1) * tile]
            lo, hi = float(_f32_down(b.min())), float(_f32_up(b.max()))
            o, hw = 0.5 * (lo + hi), 0.5 * (hi - lo)
            e = np.frexp(hw)[1] if hw > 0.0 else -100
            e = max(int(e), -100)
            invs = np.float32(2.0 ** (14 - e))
            x = (b - o).astype(np.float32) * invs
            words[j * tile:(j + 1) * tile, i * tile:(i + 1) * tile] = np.array(
                [L.oracle_f2h(float(t)) for t in x.ravel()], np.uint16).reshape(b.shape)
            off[j, i] = o
            scl[j, i] = 0.0 if hw == 0.0 else np.float32(2.0 ** (e - 14))
    return words, off, scl


def gg_low_step(types, v, words, off, scl, p, tile, stencil, bc):
    """One step of the grain-growth low tiles (low_emulation.cpp): types
    [q, ty, tx] int32, v decoded fields [q, ny, nx], words [q, ny, nx] uint16,
    off [q, ty, tx], scl [q, ty, tx] float32.  Returns (words, off, scl) of
    the next state on the low tiles (FP64 tiles copied through)."""
    q, ny, nx = v.shape
    types = np.ascontiguousarray(types, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    words = np.ascontiguousarray(words, np.uint16)
    off = np.ascontiguousarray(off, np.float64)
    scl = np.ascontiguousarray(scl, np.float32)
    wo, oo, so = words.copy(), off.copy(), scl.copy()
    lib().oracle_gg_low_step(nx, ny, q, tile, _d(np.ascontiguousarray(p, np.float64)), stencil, bc, _i32(types),
                             _d(v), _u16(words), _d(off), _f32(scl), _u16(wo), _d(oo), _f32(so))
    return wo, oo, so
