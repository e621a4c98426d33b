"""The allocation-free counter launch (gbh_launch_count_noalloc, include/gbh.h)
against the launch it varies (gbh_launch_count with a caller workspace),
through the device-level C ABI: the same outputs on two calls in a row, the
second smaller (the library's static workspace keeps the first call's
contents), and both pinned by numpy (prime counts per class, twin pairs,
every gap between consecutive primes: histogram, largest gap and its place,
the tail gaps). Out-of-contract calls are refused before any launch."""
import ctypes
import re

import numpy as np
import pytest

import gbamd
from conftest import ROOT

pytestmark = pytest.mark.gpu

U32, U64, P = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p
GBH_E_ARG = 2
NOALLOC_WORDS = 4 << 20
# pieces per map word of the counter's gap walk (include/gbh.h)
COUNT_SUB = int(re.search(r"#define GBH_COUNT_SUB (\d+)", (ROOT / "include" / "gbh.h").read_text()).group(1))
WBITS = 4096


class Batch(ctypes.Structure):
    _fields_ = [("geoms", P), ("nseg", U32), ("cw", U32), ("wbits", U32), ("nwin", U32), ("map", P)]


class CountSeg(ctypes.Structure):
    _fields_ = [("cls", U32 * 8), ("pos0", U32 * 8), ("lo", U32 * 8), ("hi", U32 * 8), ("tk", U32 * 3),
                ("tk2", U32 * 3), ("td", U32 * 3), ("tlo", U32 * 3), ("thi", U32 * 3), ("nwords", U32), ("on", U32)]


class CountArgs(ctypes.Structure):
    _fields_ = [("b", Batch), ("cs", P), ("nwords", U32), ("swords", U32), ("scratch", P), ("out", P)]


class SegOut(ctypes.Structure):
    _fields_ = [("acct", U32 * 2), ("pending_n", U32), ("left_n", U32), ("err", U32), ("cnt", U32 * 8),
                ("twins", U32 * 3), ("pfirst", U32), ("plast", U32), ("gtail_n", U32), ("vdiff", U32),
                ("gmax", U64), ("ghist", U32 * 1025), ("pad1", U32), ("gtail", U32 * 8192),
                ("hist", U32 * (15 * 1205)), ("first", U32 * 1205), ("ridge_n", U32), ("pad2", U32),
                ("ridge", U64 * 2048), ("left", U32 * 65536)]


# a segment whose q_first = 1 mod 30: class k holds q = q_first + pos0[k] + 30 t
POS0 = [0, 6, 10, 12, 16, 18, 22, 28]
# twin pairs (11, 13), (17, 19), (29, 31): lower class, partner class, partner one bit on
TWINS = [(2, 3, 0), (4, 5, 0), (7, 0, 1)]


@pytest.fixture(scope="module")
def dev():
    lib = gbamd.load()
    lib.gbh_last_error.restype = ctypes.c_char_p
    for f in ("gbh_alloc", "gbh_free", "gbh_zero", "gbh_copy_h2d", "gbh_copy_d2h"):
        getattr(lib, f).restype = ctypes.c_int
    lib.gbh_alloc.argtypes = [P, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(P)]
    lib.gbh_free.argtypes = [P, P, ctypes.c_int]
    lib.gbh_zero.argtypes = [P, ctypes.c_int, P, ctypes.c_size_t]
    lib.gbh_copy_h2d.argtypes = [P, ctypes.c_int, P, P, ctypes.c_size_t]
    lib.gbh_copy_d2h.argtypes = [P, ctypes.c_int, P, P, ctypes.c_size_t]
    lib.gbh_reset_outputs.argtypes = [P, ctypes.c_int, P, U32]
    lib.gbh_launch_count.argtypes = [P, ctypes.c_int, ctypes.POINTER(CountArgs)]
    lib.gbh_launch_count_noalloc.argtypes = [P, ctypes.c_int, ctypes.POINTER(CountArgs)]
    ctx = P()
    assert lib.gbh_open(0, ctypes.byref(ctx)) == 0, lib.gbh_last_error()
    bufs = []

    def alloc(nbytes):
        p = P()
        assert lib.gbh_alloc(ctx, nbytes, 0, ctypes.byref(p)) == 0, lib.gbh_last_error()
        bufs.append(p)
        return p.value

    yield lib, ctx, alloc
    assert lib.gbh_stream_sync(ctx, 0) == 0
    for p in bufs:
        lib.gbh_free(ctx, p, 0)
    lib.gbh_close(ctx)


def _segments(rng, nseg, nwords):
    """Random class maps (prime density about 1/6 per class bit) with a run
    of empty words, so there are gaps of every size up to several thousand."""
    nwin = -(-nwords // (WBITS // 64))
    cw = nwin * (WBITS // 64) + 16
    maps = []
    for s in range(nseg):
        m = np.zeros((8, cw), dtype=np.uint64)
        for _ in range(3):   # AND of three random words: about 1/8 of the bits; OR in a sparse fourth
            m_r = rng.integers(0, 2**63, size=(8, cw), dtype=np.uint64) * 2 + rng.integers(0, 2, size=(8, cw), dtype=np.uint64)
            m = m_r if _ == 0 else m & m_r
        m[:, 100 + 7 * s:103 + 7 * s] = 0
        m[:, 400 + s] &= np.uint64(1)
        maps.append(m)
    return np.stack(maps), nwin, cw


def _countseg(nwords):
    c = CountSeg()
    nbits = nwords * 64
    for k in range(8):
        c.cls[k], c.pos0[k], c.lo[k], c.hi[k] = k, POS0[k], 0, nbits
    for j, (a, b, d) in enumerate(TWINS):
        c.tk[j], c.tk2[j], c.td[j], c.tlo[j], c.thi[j] = a, b, d, 0, nbits - d
    c.nwords, c.on = nwords, 1
    return c


def _popcount(a):
    return int(np.unpackbits(a.view(np.uint8)).sum())


def _expected(m, nwords):
    """Counts, twins and the gap walk of one segment, by numpy."""
    words = m[:, :nwords]
    cnt = [_popcount(words[k]) for k in range(8)]
    twins = []
    for a, b, d in TWINS:
        lw = words[a].copy()
        pw = m[b, :nwords + 1]
        if d:
            lw[-1] &= np.uint64((1 << 63) - 1)
            pw = (pw[:-1] >> np.uint64(1)) | (pw[1:] << np.uint64(63))
        else:
            pw = pw[:-1]
        twins.append(_popcount(lw & pw))
    pos = []
    for k in range(8):
        bits = np.unpackbits(words[k].view(np.uint8), bitorder="little")
        t = np.nonzero(bits)[0].astype(np.int64)
        pos.append(30 * t + POS0[k])
    pos = np.sort(np.concatenate(pos))
    g = np.diff(pos)
    hist = np.bincount(np.minimum(g // 2, 1024), minlength=1025)
    gmax = int(g.max())
    up = int(pos[1:][g == gmax].min())
    tail = sorted((int(u), int(x)) for u, x in zip(pos[1:][g >= 700], g[g >= 700]))
    return dict(cnt=cnt, twins=twins, pfirst=int(pos[0]), plast=int(pos[-1]), hist=hist.tolist(), gmax=gmax,
                gmax_up=up, tail=tail)


def _summary(o: SegOut):
    n = o.gtail_n
    return dict(cnt=list(o.cnt), twins=list(o.twins), pfirst=o.pfirst, plast=o.plast, hist=list(o.ghist),
                gmax=o.gmax >> 32, gmax_up=(~o.gmax) & 0xFFFFFFFF,
                tail=sorted((o.gtail[2 * i], o.gtail[2 * i + 1]) for i in range(n)))


def _call(dev, maps, nwin, cw, noalloc):
    lib, ctx, alloc = dev
    nseg, _, _ = maps.shape
    nwords = NWORDS[nseg]
    d_map = alloc(maps.nbytes)
    assert lib.gbh_copy_h2d(ctx, 0, d_map, maps.ctypes.data, maps.nbytes) == 0
    d_geo = alloc(4096 * nseg)
    assert lib.gbh_zero(ctx, 0, d_geo, 4096 * nseg) == 0     # geometry: not read by the counter
    cs = (CountSeg * nseg)(*[_countseg(nwords) for _ in range(nseg)])
    d_cs = alloc(ctypes.sizeof(cs))
    assert lib.gbh_copy_h2d(ctx, 0, d_cs, ctypes.addressof(cs), ctypes.sizeof(cs)) == 0
    d_out = alloc(ctypes.sizeof(SegOut) * nseg)
    assert lib.gbh_reset_outputs(ctx, 0, d_out, nseg) == 0
    swords = 2 * COUNT_SUB * nwords
    a = CountArgs(b=Batch(geoms=d_geo, nseg=nseg, cw=cw, wbits=WBITS, nwin=nwin, map=d_map), cs=d_cs,
                  nwords=nwords, swords=swords, out=d_out)
    if noalloc:
        a.scratch = None
        rc = lib.gbh_launch_count_noalloc(ctx, 0, ctypes.byref(a))
    else:
        a.scratch = alloc(4 * swords * nseg)
        rc = lib.gbh_launch_count(ctx, 0, ctypes.byref(a))
    assert rc == 0, lib.gbh_last_error()
    outs = (SegOut * nseg)()
    assert lib.gbh_copy_d2h(ctx, 0, ctypes.addressof(outs), d_out, ctypes.sizeof(outs)) == 0
    assert lib.gbh_stream_sync(ctx, 0) == 0
    return [_summary(o) for o in outs]


# words per class array of each call shape (nseg -> nwords): the second call is smaller
NWORDS = {3: 6000, 1: 1500}


@pytest.mark.parametrize("nseg", [3, 1], ids=["first-3x6000-words", "second-1x1500-words"])
def test_noalloc_count_equals_workspace_count(dev, nseg):
    rng = np.random.default_rng(nseg)
    maps, nwin, cw = _segments(rng, nseg, NWORDS[nseg])
    print(f"noalloc count: {nseg} segments x {NWORDS[nseg]} words", flush=True)
    want = _call(dev, maps, nwin, cw, noalloc=False)
    got = _call(dev, maps, nwin, cw, noalloc=True)
    assert got == want
    for s in range(nseg):
        assert got[s] == _expected(maps[s], NWORDS[nseg])
        assert got[s]["tail"]                      # the empty words made tail gaps


def test_noalloc_refuses_out_of_contract(dev):
    lib, ctx, alloc = dev
    d = alloc(1 << 20)
    a = CountArgs(b=Batch(geoms=d, nseg=2, cw=WBITS // 64 + 16, wbits=WBITS, nwin=1, map=d), cs=d, nwords=64,
                  swords=NOALLOC_WORDS // 2 + 1, scratch=None, out=d)
    assert lib.gbh_launch_count_noalloc(ctx, 0, ctypes.byref(a)) == GBH_E_ARG
    assert b"GBH_NOALLOC_SCRATCH_WORDS" in lib.gbh_last_error()
    a.swords, a.scratch = 128, d
    assert lib.gbh_launch_count_noalloc(ctx, 0, ctypes.byref(a)) == GBH_E_ARG
    assert b"NULL" in lib.gbh_last_error()
