"""CPU checks of the low-tile emulation (oracle/low_emulation.cpp) itself,
before it is used as the checker of the engine's low tiles:

- its fp16 codec equals the exhaustively pinned double -> fp16 RNE encoder of
  the oracle on float inputs (random magnitudes, ties, overflow and subnormal
  edges) and decodes every one of the 65,536 words exactly;
- on a state held entirely in low tiles, one emulated step decodes to the
  FP64 oracle's step of the same decoded state within the fp16 re-encoding
  (half an fp16 ulp) plus the FP32 arithmetic's rounding, with interface
  updates hundreds of tolerances large, so a dropped or wrong term fails;
- a tile that is uniform at 0 with uniform neighbours stays exactly uniform
  (scale 0), and an FP64-tile neighbour is read from the FP64 values."""
import ctypes
import math

import numpy as np
import pytest

from tests.helpers import smooth_fields


def encode_grid(O, v, tile):
    """Per-tile codes of the engine's encoder: off = mid-range, scl = 2^(e-14)
    for a half-range below 2^e (0 for a uniform tile), words = fp16_rne((v -
    off) 2^(14-e))."""
    ny, nx = v.shape
    ty, tx = ny // tile, nx // tile
    words = np.zeros((ny, nx), np.uint16)
    off = np.zeros((ty, tx))
    scl = np.zeros((ty, tx), np.float32)
    for j in range(ty):
        for i in range(tx):
            s = np.s_[j * tile:(j + 1) * tile, i * tile:(i + 1) * tile]
            b = v[s]
            mn, mx = b.min(), b.max()
            o, hw = 0.5 * (mn + mx), 0.5 * (mx - mn)
            e = max(math.frexp(hw)[1], -100) if hw > 0 else -100
            off[j, i] = o
            scl[j, i] = 0.0 if hw == 0 else 2.0 ** (e - 14)
            words[s] = np.array([O.f64_to_f16_bits(x) for x in ((b - o) * 2.0 ** (14 - e)).ravel()],
                                np.uint16).reshape(b.shape)
    return words, off, scl


def test_codec_matches_pinned_encoder(O):
    L = O.lib()
    L.oracle_f2h.argtypes = [ctypes.c_float]
    L.oracle_f2h.restype = ctypes.c_uint16
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.standard_normal(20000) * np.exp(rng.uniform(-30, 12, 20000)),
                         (np.arange(2049) + 0.5) * 2.0 ** -24,  # subnormal ties
                         (2049 + 2 * np.arange(500)) / 2.0 * 2.0 ** -10,  # ties of normal numbers
                         [0.0, -0.0, 65504.0, 65519.0, 65520.0, -65520.0, 2.0 ** -14, 2.0 ** -25]]).astype(np.float32)
    for x in xs:
        assert L.oracle_f2h(float(x)) == O.f64_to_f16_bits(float(x)), float(x)
    h = np.arange(65536, dtype=np.uint16)
    v = O.f16_words_to_f64(h)
    ref = np.array([O.f16_bits_to_f64(int(i)) for i in h])
    fin = np.isfinite(ref)
    assert np.array_equal(v[fin], ref[fin]) and np.all(np.isnan(v[np.isnan(ref)]))


def box_scale(v, off, tile):
    """Per cell: the largest |value - own tile offset| over the cell's tile
    and its 8 neighbour tiles (periodic), the magnitude of the deviations the
    tile's FP32 arithmetic rounds."""
    ty, tx = off.shape
    out = np.zeros_like(v)
    for j in range(ty):
        for i in range(tx):
            box = np.concatenate([v[((j + dj) % ty) * tile:((j + dj) % ty + 1) * tile,
                                    ((i + di) % tx) * tile:((i + di) % tx + 1) * tile].ravel()
                                  for dj in (-1, 0, 1) for di in (-1, 0, 1)])
            out[j * tile:(j + 1) * tile, i * tile:(i + 1) * tile] = np.max(np.abs(box - off[j, i]))
    return out


@pytest.mark.parametrize("stencil,bc,tile", [(1, 1, 32), (0, 0, 16)])
def test_all_low_step_close_to_fp64_step(O, stencil, bc, tile):
    n = 128
    p = O.sf_params_config(n, n)
    phi, T = smooth_fields(n, n, seed=2, nbump=4)
    wp, op, sp = encode_grid(O, phi, tile)
    wt, ot, st = encode_grid(O, T, tile)
    dphi, dT = O.decode_tiles(wp, op, sp, tile), O.decode_tiles(wt, ot, st, tile)
    types = np.ones((n // tile, n // tile), np.int32)
    ew, eo, es = O.sf_low_step(types, dphi, dT, np.stack([wp, wt]), np.stack([op, ot]), np.stack([sp, st]), p, tile,
                               stencil, bc, seed=3, step=5)
    rp, rt = O.sf_step(dphi, dT, p, stencil=stencil, bc=bc, seed=3, step=5)
    # FP32 rounds each deviation (own and halo values, relative to the tile's
    # offset) by 2^-24 of its magnitude; through the Laplacian, flux and T
    # coupling that moves a cell by well under 16 such roundings
    s9 = box_scale(dphi, op, tile) + box_scale(dT, ot, tile)
    for f, (ref, old) in enumerate(((rp, dphi), (rt, dT))):
        e = O.decode_tiles(ew[f], eo[f], es[f], tile)
        ulp = O.f16_ulp_value(ew[f], es[f], tile)
        upd = np.abs(ref - old)
        # half an fp16 ulp of re-encoding plus the FP32 rounding
        d = np.abs(e - ref)
        tol = 0.5 * ulp + 16 * 2.0 ** -24 * s9
        assert np.all(d <= tol), (f, float(np.max(d / tol)))
        assert np.max(upd) > 100 * np.max(tol)  # the update is far larger than the tolerance
    # dropping the noise term (a = 0) moves the result by far more than the tolerance
    p0 = p.copy()
    p0[O.SF["NOISE"]] = 0.0
    nw, no, ns = O.sf_low_step(types, dphi, dT, np.stack([wp, wt]), np.stack([op, ot]), np.stack([sp, st]), p0, tile,
                               stencil, bc, seed=3, step=5)
    tol0 = 0.5 * O.f16_ulp_value(ew[0], es[0], tile) + 16 * 2.0 ** -24 * s9
    assert np.max(np.abs(O.decode_tiles(nw[0], no[0], ns[0], tile) - rp) / tol0) > 20


def test_uniform_tiles_and_fp64_neighbours(O):
    n, tile = 64, 16
    p = O.sf_params_config(n, n)
    p[O.SF["NOISE"]] = 0.0
    phi = np.zeros((n, n))
    T = np.full((n, n), 0.25)
    phi[40:, 40:] = 0.7  # tile (3, 3) and its FP64 neighbourhood carry a step in phi
    types = np.ones((4, 4), np.int32)
    types[2:, 2:] = 2
    wp, op, sp = encode_grid(O, phi, tile)
    wt, ot, st = encode_grid(O, T, tile)
    ew, eo, es = O.sf_low_step(types, phi, T, np.stack([wp, wt]), np.stack([op, ot]), np.stack([sp, st]), p, tile,
                               1, 1, seed=3, step=0)
    # tile (0, 0): phi uniform 0, T uniform, neighbours uniform -> unchanged, scale 0
    assert es[0][0, 0] == 0.0 and eo[0][0, 0] == 0.0 and not np.any(ew[0][:16, :16])
    assert es[1][0, 0] == 0.0 and eo[1][0, 0] == 0.25
    # tile (2, 1) sees the FP64 tile (2, 2)'s phi = 0.7 through its halo only
    # after (2, 2) holds 0.7 at its left edge: here phi[32:40, 32:48] = 0, so
    # the low tile (1, 2) above the step block must react to phi[40:, 40:]
    # only through FP64 values two rows below its own rows: unchanged
    assert es[0][1, 2] == 0.0
    # FP64 tiles are left to the FP64 path (words and codes untouched)
    assert np.array_equal(ew[0][32:, 32:], wp[32:, 32:]) and eo[0][3, 3] == op[3, 3]


def test_sintering_eta_encode_restatement(O):
    """The restated sintering low-eta encoding: per low tile the scale is a
    power of two, every |h| <= 2^14 (off is the mid-range of float bounds
    that cover the values), a uniform tile at a float-exact value gets scale
    0, and decoding gives
    the value back within half an fp16 ulp (RNE) plus the float rounding of
    v - off; FP64 tiles are untouched."""
    rng = np.random.default_rng(4)
    n, tile = 64, 16
    v = np.zeros((n, n))
    for j in range(4):
        for i in range(4):
            lo = float(np.float32(rng.uniform(-0.2, 1.0)))  # float-exact: a uniform tile's bounds coincide
            span = 10.0 ** rng.uniform(-12, 0) if (i + j) % 5 else 0.0
            v[j * tile:(j + 1) * tile, i * tile:(i + 1) * tile] = lo + span * rng.random((tile, tile))
    types = np.ones((4, 4), np.int32)
    types[1, 2] = 2
    w, off, scl = O.sn_encode_low(v, types, tile)
    assert scl[1, 2] == 0.0 and off[1, 2] == 0.0 and not np.any(w[16:32, 32:48])
    h = O.f16_words_to_f64(w)
    for j in range(4):
        for i in range(4):
            if types[j, i] != 1:
                continue
            s = np.s_[j * tile:(j + 1) * tile, i * tile:(i + 1) * tile]
            if np.ptp(v[s]) == 0.0:
                assert scl[j, i] == 0.0 and not np.any(w[s])
                continue
            assert math.frexp(float(scl[j, i]))[0] == 0.5  # a power of two
            assert np.max(np.abs(h[s])) <= 2.0 ** 14
    dec = O.decode_tiles(w, off, scl, tile)
    ulp = O.f16_ulp_value(w, scl, tile)
    dev = np.abs(v - np.repeat(np.repeat(off, tile, 0), tile, 1))
    low = np.repeat(np.repeat(types == 1, tile, 0), tile, 1)
    assert np.all(np.abs(dec - v)[low] <= (0.5 * ulp + 2.0 ** -24 * dev + 1e-300)[low])


@pytest.mark.parametrize("stencil,bc", [(0, 0), (1, 1)])
def test_gg_all_low_step_close_to_fp64_step(O, stencil, bc):
    """The grain-growth low-tile emulation on a state held entirely in low
    tiles: one step equals the FP64 oracle step of the decoded state within
    half an fp16 ulp plus the FP32 rounding of the deviations and of the
    reaction terms, the update being far larger than that; with the coupling
    2 gamma eta sum_{j!=i} eta_j^2 dropped (gamma = 0 in the emulator only)
    the result moves by far more than the tolerance."""
    n, tile, q = 64, 16, 3
    p = O.gg_params_config()
    a, b = smooth_fields(n, n, seed=5, nbump=4)
    fields = [0.9 * a + 0.05, 0.8 * (1.0 - a) + 0.1 * b, 0.3 * b]
    enc = [encode_grid(O, f, tile) for f in fields]
    words = np.stack([e[0] for e in enc])
    off = np.stack([e[1] for e in enc])
    scl = np.stack([e[2] for e in enc])
    dec = np.stack([O.decode_tiles(words[f], off[f], scl[f], tile) for f in range(q)])
    types = np.ones((q, n // tile, n // tile), np.int32)
    ew, eo, es = O.gg_low_step(types, dec, words, off, scl, p, tile, stencil, bc)
    ref = O.gg_step(dec, p, stencil=stencil, bc=bc)
    S = np.sum(dec * dec, axis=0)
    inv = 1.0 / (p[O.GG["DX"]] * p[O.GG["DY"]])
    dtl = p[O.GG["DT"]] * p[O.GG["L"]]
    for f in range(q):
        e = O.decode_tiles(ew[f], eo[f], es[f], tile)
        ulp = O.f16_ulp_value(ew[f], es[f], tile)
        eta = np.abs(dec[f])
        react = 2 * p[O.GG["GAMMA"]] * eta * (S + eta * eta) + p[O.GG["BETA"]] * eta ** 3 + p[O.GG["ALPHA"]] * eta
        s9 = box_scale(dec[f], off[f], tile)
        tol = 0.5 * ulp + 32 * 2.0 ** -24 * (s9 * (1 + dtl * p[O.GG["KAPPA"]] * 12 * inv) + dtl * react + eta)
        d = np.abs(e - ref[f])
        assert np.all(d <= tol), (f, float(np.max(d / tol)))
        assert np.max(np.abs(ref[f] - dec[f])) > 100 * np.median(tol)
    p0 = p.copy()
    p0[O.GG["GAMMA"]] = 0.0
    nw, no, ns = O.gg_low_step(types, dec, words, off, scl, p0, tile, stencil, bc)
    moved = max(np.max(np.abs(O.decode_tiles(nw[f], no[f], ns[f], tile) - ref[f]) / (0.5 * O.f16_ulp_value(ew[f], es[f], tile) + 1e-300))
                for f in range(q))
    assert moved > 20
