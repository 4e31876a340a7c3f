"""The host libm's log2f, exp2f and powf as the shaders call them (ShaderSoft.h:132, PbrSoft.h:201-235,313,
BlinnPhongSoft.h:227), reproduced from librhip's tables (rhip_libm.h; log2f and exp2f by glibc's own algorithms,
with empty tables), and glibc's atan2f (SkyboxSoft.h:84) restated step for step.

CPU: the table lookup run on the host (rhip_libm_eval_host) against glibc itself, called through ctypes, on
dense runs of consecutive floats and on random inputs of every function's range; the runs hold inputs where
glibc is not correctly rounded, so the lookup is exercised. GPU: the device evaluation (rhip_libm_eval) against
glibc on the same inputs and against the host evaluation on millions more."""
import ctypes

import numpy as np
import pytest

from softglrender_amd import _lib

LIBM = ctypes.CDLL("libm.so.6")
for _f in ("log2f", "exp2f", "powf"):
    getattr(LIBM, _f).restype = ctypes.c_float
LIBM.log2f.argtypes = [ctypes.c_float]
LIBM.exp2f.argtypes = [ctypes.c_float]
LIBM.powf.argtypes = [ctypes.c_float, ctypes.c_float]
INV_GAMMA = float(np.float32(1.0) / np.float32(2.2))

# fn: (name, glibc call as the shaders make it, input range (inclusive bits), dense runs (start, count))
FNS = {
    0: ("log2f(x)", lambda x: LIBM.log2f(x), (0x00000001, 0x7F7FFFFF),
        [(0x3F780000, 1 << 18), (0x4B3A0000, 1 << 17), (0x00000001, 1 << 12), (0x00800000 - 4096, 8192)]),
    1: ("exp2f(-x)", lambda x: LIBM.exp2f(-x), (0x32000000, 0x41800000), [(0x3F000000, 1 << 17), (0x41000000, 1 << 16)]),
    2: ("powf(x, 5)", lambda x: LIBM.powf(x, 5.0), (0x30000000, 0x3F802000), [(0x3F400000, 1 << 17), (0x3E000000, 1 << 16)]),
    3: ("powf(x, 128)", lambda x: LIBM.powf(x, 128.0), (0x3ECCCCCD, 0x3F802000), [(0x3F7E0000, 1 << 17)]),
    4: ("powf(x, 1/2.2)", lambda x: LIBM.powf(x, INV_GAMMA), (0x00000001, 0x3F800000),
        [(0x3F200000, 1 << 17), (0x3B000000, 1 << 16), (0x00000001, 1 << 12)]),
}


def inputs(fn, seed=0, n_random=1 << 16):
    lo, hi = FNS[fn][2]
    runs = [np.arange(s, min(s + c, hi + 1), dtype=np.uint32) for s, c in FNS[fn][3]]
    rng = np.random.default_rng(seed + fn)
    rnd = rng.integers(lo, hi + 1, n_random, dtype=np.uint64).astype(np.uint32)
    edges = np.array([lo, lo + 1, hi - 1, hi], np.uint32)
    return np.concatenate(runs + [rnd, edges]).view(np.float32)


def glibc(fn, x):
    f = FNS[fn][1]
    return np.array([f(float(v)) for v in x], np.float32)


def host_eval(fn, x):
    lib = _lib.load()
    out = np.empty_like(x)
    _lib.check(lib.rhip_libm_eval_host(fn, x.ctypes.data, x.size, out.ctypes.data))
    return out


@pytest.mark.parametrize("fn", sorted(FNS))
def test_table_lookup_reproduces_glibc(fn):
    lib = _lib.load()
    entries, gate = ctypes.c_longlong(), ctypes.c_double()
    _lib.check(lib.rhip_libm_info(fn, ctypes.byref(entries), ctypes.byref(gate)))
    if fn in (0, 1):  # log2f, exp2f: glibc's own algorithms, restated, equal the host's on every float
        assert entries.value == 0 and gate.value < 0.0
    else:
        assert entries.value > 1000 and 0.0 <= gate.value < 0.5
    x = inputs(fn)
    want = glibc(fn, x)
    got = host_eval(fn, x)
    bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, (FNS[fn][0], x[bad[:5]], got[bad[:5]], want[bad[:5]])
    # the inputs include some where glibc is not the correctly rounded value (so the table was consulted)
    xd = x.astype(np.float64)
    cr = [lambda: np.log2(xd), lambda: np.exp2(-xd), lambda: xd ** 5, lambda: xd ** 128,
          lambda: xd ** np.float64(INV_GAMMA)][fn]()
    assert int((cr.astype(np.float32).view(np.uint32) != want.view(np.uint32)).sum()) > 0


def test_outside_the_ranges_glibc_equals_the_correctly_rounded_value():
    # the device falls back to correctly rounded values there: spot checks of the ranges' tails
    tiny = np.array([1e-30, 2.0 ** -28, 7e-9], np.float32)
    assert all(LIBM.exp2f(-float(u)) == 1.0 for u in tiny)
    assert all(LIBM.powf(float(x), 5.0) == 0.0 for x in np.array([2.0 ** -32, 1e-12, 3e-10], np.float32))
    assert all(LIBM.powf(float(x), 128.0) == 0.0 for x in np.array([0.39, 0.2, 0.01], np.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("fn", sorted(FNS))
def test_device_reproduces_glibc(renderer, fn):
    lib = _lib.load()
    print(f"[gpu] libm fn {fn}", flush=True)
    x = inputs(fn, seed=7)
    out = np.empty_like(x)
    _lib.check(lib.rhip_libm_eval(renderer.dev, fn, x.ctypes.data, x.size, out.ctypes.data))
    want = glibc(fn, x)
    bad = np.flatnonzero(out.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, (FNS[fn][0], x[bad[:5]], out[bad[:5]], want[bad[:5]])
    # millions more, against the host's evaluation of the same tables
    lo, hi = FNS[fn][2]
    y = np.random.default_rng(100 + fn).integers(lo, hi + 1, 1 << 24, dtype=np.uint64).astype(np.uint32).view(np.float32)
    dev = np.empty_like(y)
    _lib.check(lib.rhip_libm_eval(renderer.dev, fn, y.ctypes.data, y.size, dev.ctypes.data))
    assert np.array_equal(dev.view(np.uint32), host_eval(fn, y).view(np.uint32)), FNS[fn][0]


# ---------------------------------------------------------------- atan2f (SkyboxSoft.h:84), fn 5, no table

LIBM.atan2f.restype = ctypes.c_float
LIBM.atan2f.argtypes = [ctypes.c_float, ctypes.c_float]
ATAN2 = 5


def atan2_pairs(seed, n):
    """(y, x) pairs: unit-vector components as the skybox makes them, arbitrary bit patterns (every exponent,
    subnormals, infinities, NaNs) and the special cases fdlibm branches on."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    unit = d[:, [2, 0]].astype(np.float32)
    bits = rng.integers(0, 2 ** 32, (n, 2), dtype=np.uint64).astype(np.uint32).view(np.float32)
    vals = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 1e-30, -1e30, 2.0 ** -149, 0.4375, 1.1875,
                     2.4375, 2.0 ** 25, 2.0 ** -29], np.float32)
    special = np.array([(a, b) for a in vals for b in vals], np.float32)
    return np.ascontiguousarray(np.concatenate([unit, bits, special]))


def glibc_atan2(yx):
    return np.array([LIBM.atan2f(float(a), float(b)) for a, b in yx], np.float32)


def same_float(a, b):
    return (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))


def test_atan2f_restatement_reproduces_glibc():
    lib = _lib.load()
    yx = atan2_pairs(3, 100_000)
    got = np.empty(len(yx), np.float32)
    _lib.check(lib.rhip_libm_eval_host(ATAN2, yx.ctypes.data, len(yx), got.ctypes.data))
    want = glibc_atan2(yx)
    bad = np.flatnonzero(~same_float(got, want))
    assert bad.size == 0, (yx[bad[:5]], got[bad[:5]], want[bad[:5]])
    # glibc's atan2f is not correctly rounded: the restatement is checked where it matters
    with np.errstate(invalid="ignore"):
        cr = np.arctan2(yx[:, 0].astype(np.float64), yx[:, 1].astype(np.float64)).astype(np.float32)
    assert int((~same_float(cr, want)).sum()) > 1000


@pytest.mark.gpu
def test_device_atan2f_reproduces_glibc(renderer):
    lib = _lib.load()
    print("[gpu] libm atan2f", flush=True)
    yx = atan2_pairs(11, 100_000)
    out = np.empty(len(yx), np.float32)
    _lib.check(lib.rhip_libm_eval(renderer.dev, ATAN2, yx.ctypes.data, len(yx), out.ctypes.data))
    want = glibc_atan2(yx)
    bad = np.flatnonzero(~same_float(out, want))
    assert bad.size == 0, (yx[bad[:5]], out[bad[:5]], want[bad[:5]])
