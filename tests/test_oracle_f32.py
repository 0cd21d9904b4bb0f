"""Pins of the oracle's FP32-bulk restatement.

* encode/decode round trip: positions within one FP32 ulp of a cell-relative
  offset, velocities on the 2^-21 grid within half a grid step.
* One FP32-bulk step against the FP64 collision+stream on the same injected
  state: velocities within 1e-5 (norm-wise, thermal scale), positions within
  1e-6 modulo the box, cell ids equal except for particles on a face.
* Momentum conservation is exact: the integer momentum (units of 2^-21) of
  the fluid is identical before and after every step; with beads, fluid +
  bead momentum is conserved to FP64 roundoff.
* The collision arithmetic brute-forced on a tiny box in Python, with exact
  fused multiply-adds (rationals), the exact integer sums and an explicit
  lexicographic search for the residual owner over coarse velocities that
  tie: the oracle must agree bit for bit.
"""
import struct
from fractions import Fraction

import numpy as np

import oracle as O
from cases import faithful_case, periodic_diff, random_draws

H = 0.1


def test_encode_decode_roundtrip():
    L = (12, 10, 9)
    case = faithful_case(L, seed=3, steps=0)
    a = case["a0"]
    cell, off, w = O.f32_encode(L, a, case["pos0"], case["vel0"])
    assert np.all((off >= 0) & (off < 1))
    pos, vel = O.f32_decode(L, a, cell, off, w)
    assert np.max(np.abs(periodic_diff(pos, case["pos0"], L))) < 1.2e-7
    assert np.max(np.abs(vel[:, :3] - case["vel0"][:, :3])) <= 2.0 ** -22
    q = vel[:, :3] * 2 ** 21
    assert np.array_equal(q, np.round(q))
    assert np.array_equal(cell, O.box(L, a, pos))


def test_f32_step_close_to_fp64_and_exact_momentum():
    L = (16, 12, 8)
    rng = np.random.default_rng(8)
    case = faithful_case(L, seed=21, steps=0)
    a = case["a0"]
    ncells = int(np.prod(L))
    cell, off, w = O.f32_encode(L, a, case["pos0"], case["vel0"])
    # FP64 reference on the decoded (identical) state
    p64, v64 = O.f32_decode(L, a, cell, off, w)
    q0 = np.round(w.astype(np.float64).sum(0) * 2 ** 21).astype(np.int64)
    for s in range(5):
        g, ax = random_draws(ncells, rng)
        a_next = rng.uniform(0, 1, 3)
        cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, gamma=g, axis=ax)
        O.collide(cfg, a, p64, v64)
        O.stream(p64, v64, H)
        cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, gamma=g, axis=ax)
        O.f32_step(cfg, a, a_next, H, cell, off, w)
        pos, vel = O.f32_decode(L, a_next, cell, off, w)
        q = np.round(w.astype(np.float64).sum(0) * 2 ** 21).astype(np.int64)
        assert np.array_equal(q, q0), "momentum must be exact"
        same = cell == O.box(L, a_next, p64)
        assert same.mean() > 0.999
        err = np.linalg.norm(vel[:, :3] - v64[:, :3], axis=1) / np.maximum(np.linalg.norm(v64[:, :3], axis=1), 1)
        if s == 0:
            # grid rounding (2^-22) plus the momentum residual one particle
            # per cell absorbs (a few 2^-21 steps): the 1e-5 FP32 bar
            assert np.quantile(err, 0.99) < 1e-6
            assert err.max() < 1e-5
        assert np.max(np.abs(periodic_diff(pos, p64, L))) < 1e-5
        # resync the FP64 state to the FP32 one (drift over steps is chaotic)
        p64, v64 = pos.copy(), vel.copy()
        a = a_next


def test_f32_beads_total_momentum():
    L = (8, 8, 8)
    rng = np.random.default_rng(9)
    case = faithful_case(L, seed=5, steps=0)
    a = case["a0"]
    cell, off, w = O.f32_encode(L, a, case["pos0"], case["vel0"])
    nb = 150
    bpos = rng.uniform(0, 8, (nb, 3))
    bvel = np.zeros((nb, 4))
    bvel[:, :3] = rng.normal(size=(nb, 3))
    bvel[:, 3] = 5.0
    p0 = w.astype(np.float64).sum(0) + (bvel[:, :3] * 5).sum(0)
    g, ax = random_draws(512, rng)
    cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, gamma=g, axis=ax)
    O.f32_step(cfg, a, rng.uniform(0, 1, 3), H, cell, off, w, bpos, bvel)
    p1 = w.astype(np.float64).sum(0) + (bvel[:, :3] * 5).sum(0)
    assert np.allclose(p0, p1, atol=1e-11)


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _f32(x):
    """Exact rational -> nearest float32, ties to even (no double rounding)."""
    f = np.float32(float(x))
    cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
    return min(cands, key=lambda c: (abs(Fraction(float(c)) - x), int(c.view(np.uint32)) & 1))


def _fmaf(a, b, c):
    return _f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def _bits(f):
    return struct.unpack("<I", struct.pack("<f", float(f)))[0]


def _rotmat(ax, c, s):
    x, y, z = (float(t) for t in ax)
    return [x * x + (1.0 - x * x) * c, x * y * (1.0 - c) - z * s, x * z * (1.0 - c) + y * s,
            x * y * (1.0 - c) + z * s, y * y + (1.0 - y * y) * c, y * z * (1.0 - c) - x * s,
            x * z * (1.0 - c) - y * s, y * z * (1.0 - c) + x * s, z * z + (1.0 - z * z) * c]


def _quant(vn):
    r = float(np.rint(float(vn) * 2.0 ** 21))
    if abs(r) < 2.0 ** 24:
        return np.float32(r / 2.0 ** 21)
    return np.float32(vn)


def _q(w):
    return int(float(w) * 2.0 ** 21)


def _brute_collision(L, cell, off, w, g, ax, ids, kT=1.0, nkeys=5, with_stream=False):
    """Expected velocities after one FP32-bulk collision (no beads): the
    per-cell map in FP64, rounded to FP32 for the particle update; the
    residual owner compares the first nkeys keys (then the first index).
    with_stream: also the velocities the stream uses (before the residual)."""
    Q = 2.0 ** -21
    out = w.copy()
    stream_w = w.copy()
    for c in range(int(np.prod(L))):
        idx = np.flatnonzero(cell == c)
        if len(idx) == 0:
            continue
        q = [[_q(w[i, k]) for k in range(3)] for i in idx]
        S1 = [sum(t[k] for t in q) for k in range(3)]
        lo = sum(sum(x * x for x in t) & ((1 << 40) - 1) for t in q)
        hi = sum(sum(x * x for x in t) >> 40 for t in q)
        s2 = float(hi) * 2.0 ** 40 + float(lo)
        m = len(idx)
        u = [(float(S1[k]) * Q + 0.0) / m for k in range(3)]
        rel = (s2 * (Q * Q) + 0.0) - (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * m
        sc = float(np.sqrt(kT * g[c] / (rel * 0.5))) if (rel > 0 and m > 1) else 1.0
        R = _rotmat(ax[3 * c:3 * c + 3], O.CALPHA, O.SALPHA)
        M = [r * sc for r in R]
        b = [u[i] - _fma(M[3 * i], u[0], _fma(M[3 * i + 1], u[1], M[3 * i + 2] * u[2])) for i in range(3)]
        Mf = [np.float32(x) for x in M]
        bf = [np.float32(x) for x in b]
        D = [0, 0, 0]
        keys = []
        for j, i in enumerate(idx):
            v = [w[i, k] for k in range(3)]
            for k in range(3):
                vn = _fmaf(Mf[3 * k], v[0], _fmaf(Mf[3 * k + 1], v[1], _fmaf(Mf[3 * k + 2], v[2], bf[k])))
                out[i, k] = _quant(vn)
                D[k] += q[j][k] - _q(out[i, k])
            keys.append(((_bits(w[i, 0]) << 32) | _bits(w[i, 1]),
                         (_bits(out[i, 0]) << 32) | _bits(out[i, 1]),
                         (_bits(out[i, 2]) << 32) | _bits(off[i, 0]),
                         (_bits(off[i, 1]) << 32) | _bits(off[i, 2]),
                         int(ids[i])))
        own = idx[max(range(len(idx)), key=lambda j: (keys[j][:nkeys], -j))]
        stream_w[idx] = out[idx]
        for k in range(3):
            out[own, k] = np.float32((_q(out[own, k]) + D[k]) * Q)
    return (out, stream_w) if with_stream else out


def _brute_stream(L, a, a_next, h, cell, off, ws):
    """FP32 stream + rebin: per axis o + w h + (a_next - a) in FP32, the
    floor moves the cell (periodic), an offset rounding up to 1 moves it one
    more."""
    L = np.asarray(L)
    ck = np.stack([cell % L[0], (cell // L[0]) % L[1], cell // (L[0] * L[1])], axis=1)
    da = (np.asarray(a_next, dtype=np.float64) - np.asarray(a, dtype=np.float64)).astype(np.float32)
    pk = (off + ws * np.float32(h)) + da
    fl = np.floor(pk)
    o = (pk - fl).astype(np.float32)
    n = fl.astype(np.int64) + (o >= 1.0)
    o[o >= 1.0] = 0.0
    ck = (ck + n) % L
    return (ck[:, 0] + L[0] * (ck[:, 1] + L[1] * ck[:, 2])).astype(np.int32), o


def test_f32_collision_matches_brute_force():
    L = (3, 2, 2)
    rng = np.random.default_rng(4)
    n = 150
    pos = rng.uniform(0, 1, (n, 3)) * np.array(L)
    vel = np.ones((n, 4))
    vel[:, :3] = rng.normal(size=(n, 3))
    coarse = rng.random(n) < 0.6  # equal velocity bits: owner ties
    vel[coarse, :3] = rng.integers(-2, 3, (coarse.sum(), 3)) * 0.5
    vel[:5, :3] = [9.5, -1.0, 0.25]  # beyond the 2^-21 grid's exact range
    dup = rng.choice(n, 10, replace=False)
    pos[dup[:5]] = pos[dup[5:]]
    vel[dup[:5]] = vel[dup[5:]]
    ids = rng.permutation(n).astype(np.int32)
    a = rng.uniform(0, 1, 3)
    cell, off, w = O.f32_encode(L, a, pos, vel)
    ncells = int(np.prod(L))
    g, ax = random_draws(ncells, rng)
    expect, ws = _brute_collision(L, cell, off, w, g, ax, ids, with_stream=True)
    w1 = w.copy()
    c1, o1 = cell.copy(), off.copy()
    a_next = rng.uniform(0, 1, 3)
    cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, gamma=g, axis=ax)
    O.f32_step(cfg, a, a_next, H, c1, o1, w1, ids=ids)
    assert np.array_equal(w1, expect)
    # the stream: every particle moves with its velocity before the residual
    ec, eo = _brute_stream(L, a, a_next, H, cell, off, ws)
    assert np.array_equal(c1, ec) and np.array_equal(o1, eo)
    assert not np.array_equal(ws, expect)
    # the ties matter: an owner picked by the old velocity alone differs
    assert not np.array_equal(w1, _brute_collision(L, cell, off, w, g, ax, ids, nkeys=1))
