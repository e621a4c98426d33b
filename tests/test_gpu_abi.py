"""The device-level C ABI (include/gbh.h) refuses arguments outside a
stage's contract before it launches anything: a non-zero GBH_E_ARG status
and a message in gbh_last_error(), the reference's fail-closed ZE_CHECK
behaviour (include/l0_context.hpp:14-20) as a status code. Every call
below fails its checks before it launches; the pointers it passes still
point into one real, zeroed 4 MiB device buffer, so a check that stopped
refusing would launch on mapped memory of a harmless size, not on a wild
address."""
import ctypes

import pytest

import gbamd

pytestmark = pytest.mark.gpu

GBH_E_ARG = 2
P, U32 = ctypes.c_void_p, ctypes.c_uint32
DUMMY = None             # the device buffer of the `dev` fixture (set there)


class Batch(ctypes.Structure):
    _fields_ = [("geoms", P), ("nseg", U32), ("cw", U32), ("wbits", U32), ("nwin", U32), ("map", P)]


class CheckArgs(ctypes.Structure):
    _fields_ = [("b", Batch), ("ka", U32), ("pcap", U32), ("coff", P), ("coffa", P), ("pending", P), ("out", P),
                ("kinds", U32), ("class_groups", U32)]


class PartitionArgs(ctypes.Structure):
    _fields_ = [("b", Batch), ("rdesc", P), ("region", P), ("fills", P), ("groups", P), ("ngroups", U32),
                ("cap", U32), ("hits", P), ("hit_n", P), ("err", P), ("part_cap", U32), ("pad0", U32)]


class CheckBArgs(ctypes.Structure):
    _fields_ = [("b", Batch), ("ka", U32), ("tail_t", U32), ("coff", P), ("pending", P), ("pcap", U32),
                ("kend", U32), ("out", P)]


class WindowArgs(ctypes.Structure):
    _fields_ = [("b", Batch), ("primes", P), ("nprimes", U32), ("ncoop", U32), ("i0", P), ("hits", P),
                ("hit_n", P), ("cap", U32), ("phases", U32)]


@pytest.fixture(scope="module")
def dev():
    lib = gbamd.load()
    lib.gbh_last_error.restype = ctypes.c_char_p
    ctx = P()
    assert lib.gbh_open(0, ctypes.byref(ctx)) == 0, lib.gbh_last_error()
    global DUMMY
    buf = P()
    nbytes = 4 << 20
    assert lib.gbh_alloc(ctx, ctypes.c_size_t(nbytes), 0, ctypes.byref(buf)) == 0, lib.gbh_last_error()
    assert lib.gbh_zero(ctx, 0, buf, ctypes.c_size_t(nbytes)) == 0
    assert lib.gbh_stream_sync(ctx, 0) == 0
    DUMMY = buf.value
    yield lib, ctx
    assert lib.gbh_stream_sync(ctx, 0) == 0
    lib.gbh_free(ctx, buf, 0)
    lib.gbh_close(ctx)


def _batch(**kw):
    b = Batch(geoms=DUMMY, nseg=1, cw=4096 // 64 + 16, wbits=4096, nwin=1, map=DUMMY)
    for k, v in kw.items():
        setattr(b, k, v)
    return b


def _refused(lib, rc, word):
    assert rc == GBH_E_ARG
    assert word in lib.gbh_last_error().decode()


def test_batch_contract(dev):
    lib, ctx = dev
    w = WindowArgs(b=_batch(wbits=3000), primes=DUMMY, nprimes=1, ncoop=0, i0=DUMMY, phases=1)
    _refused(lib, lib.gbh_launch_window(ctx, 0, ctypes.byref(w)), "power of two")
    w = WindowArgs(b=_batch(nseg=0), primes=DUMMY, nprimes=1, ncoop=0, i0=DUMMY, phases=1)
    _refused(lib, lib.gbh_launch_window(ctx, 0, ctypes.byref(w)), "batch size")
    w = WindowArgs(b=_batch(cw=8), primes=DUMMY, nprimes=1, ncoop=0, i0=DUMMY, phases=1)
    _refused(lib, lib.gbh_launch_window(ctx, 0, ctypes.byref(w)), "class stride")


def test_window_phases(dev):
    lib, ctx = dev
    w = WindowArgs(b=_batch(), primes=DUMMY, nprimes=4, ncoop=0, i0=DUMMY, phases=5)
    _refused(lib, lib.gbh_launch_window(ctx, 0, ctypes.byref(w)), "phases")
    w = WindowArgs(b=_batch(), primes=DUMMY, nprimes=4, ncoop=5, i0=DUMMY, phases=1)
    _refused(lib, lib.gbh_launch_window(ctx, 0, ctypes.byref(w)), "ncoop")


def test_check_class_groups(dev):
    lib, ctx = dev
    a = CheckArgs(b=_batch(), ka=96, pcap=1024, coff=DUMMY, coffa=DUMMY, pending=DUMMY, out=DUMMY, kinds=1,
                  class_groups=4)
    _refused(lib, lib.gbh_launch_check_a(ctx, 0, ctypes.byref(a)), "class_groups")
    a.class_groups, a.ka = 3, 100
    _refused(lib, lib.gbh_launch_check_a(ctx, 0, ctypes.byref(a)), "ka")


def test_partition_size(dev):
    lib, ctx = dev
    a = PartitionArgs(b=_batch(), rdesc=DUMMY, region=DUMMY, fills=DUMMY, groups=DUMMY, ngroups=1, cap=64,
                      hits=DUMMY, hit_n=DUMMY, err=DUMMY, part_cap=1000)
    _refused(lib, lib.gbh_launch_partition(ctx, 0, ctypes.byref(a)), "part_cap")
    a.part_cap, a.hits = 8192, None
    _refused(lib, lib.gbh_launch_partition(ctx, 0, ctypes.byref(a)), "null")


def test_check_b_slot_limit(dev):
    lib, ctx = dev
    a = CheckBArgs(b=_batch(), ka=96, tail_t=3000, coff=DUMMY, pending=DUMMY, pcap=1024, kend=96, out=DUMMY)
    _refused(lib, lib.gbh_launch_check_b(ctx, 0, ctypes.byref(a)), "kend")
