"""The partition step of the direct harvest (gbh_launch_partition) through
the C ABI: every (segment, class) region's entries land in the window lists
of their window, and a fill count above its region's size (a chunk that
failed before storing its counts) is taken at the region's size, so the
kernel never reads past a region. The step has no counterpart in the
reference, whose harvest (hl_harvest_mt, kernels/goldbach.cl:3251) appends
each hit to its bucket list directly for hl_apply (kernels/goldbach.cl:3460):
the window lists built here play that role, and the parity tests check the
maps made from them."""
import ctypes

import numpy as np
import pytest

import gbamd

pytestmark = pytest.mark.gpu

P, U32 = ctypes.c_void_p, ctypes.c_uint32
DUMMY = 0x10000          # geometry and map pointers: the partition reads neither
WBITS, NWIN, RCAP, CAP = 4096, 2, 8, 16


class Batch(ctypes.Structure):
    _fields_ = [("geoms", P), ("nseg", U32), ("cw", U32), ("wbits", U32), ("nwin", U32), ("map", P)]


class PartitionArgs(ctypes.Structure):
    _fields_ = [("b", Batch), ("rdesc", P), ("region", P), ("fills", P), ("groups", P), ("ngroups", U32),
                ("cap", U32), ("hits", P), ("hit_n", P), ("err", P), ("part_cap", U32), ("pad0", U32)]


def _run(fill_value, seed):
    torch = pytest.importorskip("torch")
    lib = gbamd.load()
    lib.gbh_last_error.restype = ctypes.c_char_p
    ctx = P()
    assert lib.gbh_open(0, ctypes.byref(ctx)) == 0, lib.gbh_last_error()
    try:
        rng = np.random.default_rng(seed)
        nbb = 8                                           # one segment: 8 classes
        region = rng.integers(0, NWIN * WBITS, size=nbb * RCAP, dtype=np.uint32)
        dev = lambda a: torch.from_numpy(np.asarray(a, dtype=np.uint32).view(np.int32)).cuda()
        d_region = dev(region)
        d_rdesc = dev([0, RCAP])                          # chunk 0: base 0, RCAP entries per list
        d_fills = dev(np.full(nbb, fill_value))
        d_groups = dev([0, 1])
        d_hits = torch.zeros(nbb * NWIN * CAP, dtype=torch.int32, device="cuda")
        d_hit_n = torch.zeros(nbb * NWIN, dtype=torch.int32, device="cuda")
        d_err = torch.zeros(1, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        b = Batch(geoms=DUMMY, nseg=1, cw=WBITS * NWIN // 64 + 16, wbits=WBITS, nwin=NWIN, map=DUMMY)
        a = PartitionArgs(b=b, rdesc=d_rdesc.data_ptr(), region=d_region.data_ptr(), fills=d_fills.data_ptr(),
                          groups=d_groups.data_ptr(), ngroups=1, cap=CAP, hits=d_hits.data_ptr(),
                          hit_n=d_hit_n.data_ptr(), err=d_err.data_ptr(), part_cap=8192)
        assert lib.gbh_launch_partition(ctx, 0, ctypes.byref(a)) == 0, lib.gbh_last_error()
        assert lib.gbh_stream_sync(ctx, 0) == 0, lib.gbh_last_error()
        hits = d_hits.cpu().numpy().view(np.uint32).reshape(nbb, NWIN, CAP)
        hit_n = d_hit_n.cpu().numpy().reshape(nbb, NWIN)
        err = int(d_err.cpu()[0])
    finally:
        lib.gbh_close(ctx)
    return region.reshape(nbb, RCAP), hits, hit_n, err


@pytest.mark.parametrize("fill_value,seed", [(RCAP, 1), (5, 2), (1000, 3), (0xFFFFFFFF, 4)])
def test_partition_lists(fill_value, seed):
    region, hits, hit_n, err = _run(fill_value, seed)
    assert err == 0
    used = min(fill_value, RCAP)                          # counts beyond the region are taken at its size
    for bk in range(8):
        entries = region[bk, :used]
        for w in range(NWIN):
            want = sorted(int(t) & (WBITS - 1) for t in entries if int(t) // WBITS == w)
            assert hit_n[bk, w] == len(want)
            assert sorted(hits[bk, w, :hit_n[bk, w]].tolist()) == want
