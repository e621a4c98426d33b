"""Python binding of the MI355X verifier (ctypes over the run-level C ABI,
include/goldbach_amd.h).

This is the host-side mirror of the reference's entry point
``gb::run_gpu`` (include/verifier.hpp:103-110): same arguments, same range
contract, the GBSTATS statistics files written as the reference writes them,
and failures raised (never reported as PASS). The compute path is the HIP
library ``libgoldbach_amd.so`` built next to this file; there is no fallback:
without it, or without a gfx950 GPU, every call raises.
"""
from __future__ import annotations

import ctypes
import os
import pathlib

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libgoldbach_amd.so"
CLI_PATH = HERE / "goldbach_amd"


class GbaRunResult(ctypes.Structure):
    _fields_ = [
        ("limit", ctypes.c_uint64), ("start", ctypes.c_uint64), ("segments", ctypes.c_uint64),
        ("leftovers", ctypes.c_uint64),
        ("init_seconds", ctypes.c_double), ("gpu_seconds", ctypes.c_double), ("total_seconds", ctypes.c_double),
        ("evens_per_second", ctypes.c_double),
        ("sieve_seconds", ctypes.c_double), ("large_seconds", ctypes.c_double), ("verify_seconds", ctypes.c_double),
        ("count_seconds", ctypes.c_double), ("sync_seconds", ctypes.c_double),
        ("total_tiles", ctypes.c_uint64), ("total_large_items", ctypes.c_uint64),
        ("split_used", ctypes.c_uint32), ("stats_on", ctypes.c_int32),
        ("stats_pmax", ctypes.c_uint64), ("stats_primes", ctypes.c_uint64), ("stats_twins", ctypes.c_uint64),
        ("stats_sampled", ctypes.c_uint64), ("stats_tail", ctypes.c_uint64),
        ("harvest_kernel_seconds", ctypes.c_double), ("window_kernel_seconds", ctypes.c_double),
        ("harvest_launches", ctypes.c_uint64), ("window_launches", ctypes.c_uint64),
        ("harvest_bytes", ctypes.c_double), ("window_bytes", ctypes.c_double),
        ("host_plan_seconds", ctypes.c_double), ("harvest_reruns", ctypes.c_uint64),
        ("graph_launches", ctypes.c_uint64), ("harvest_words", ctypes.c_double),
    ]


class GbaSegmentStats(ctypes.Structure):
    _fields_ = [
        ("seg_start", ctypes.c_uint64), ("seg_end", ctypes.c_uint64), ("evens", ctypes.c_uint64),
        ("tiles", ctypes.c_uint64), ("large_items", ctypes.c_uint64),
        ("sieve_split", ctypes.c_uint32), ("mask_count", ctypes.c_uint32), ("leftovers", ctypes.c_uint32),
        ("pad0", ctypes.c_uint32),
        ("sieve_seconds", ctypes.c_double), ("large_seconds", ctypes.c_double), ("check_seconds", ctypes.c_double),
        ("offset_seconds", ctypes.c_double), ("wall_seconds", ctypes.c_double),
    ]


class VerificationError(RuntimeError):
    """A run that did not pass (device error, coverage, counterexample, ...)."""


_lib = None


def load() -> ctypes.CDLL:
    """Loads the HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        lib = ctypes.CDLL(str(LIB_PATH))
        lib.gba_run_gpu.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_char_p, ctypes.c_int, ctypes.c_uint64,
                                    ctypes.c_uint32, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(GbaRunResult),
                                    ctypes.POINTER(GbaSegmentStats), ctypes.c_size_t,
                                    ctypes.POINTER(ctypes.c_size_t), ctypes.c_char_p, ctypes.c_size_t,
                                    ctypes.c_char_p, ctypes.c_size_t]
        lib.gba_run_gpu.restype = ctypes.c_int
        lib.gba_segment_map.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p,
                                        ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint32),
                                        ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint32),
                                        ctypes.c_char_p, ctypes.c_size_t]
        lib.gba_segment_map.restype = ctypes.c_int
        lib.gba_version.restype = ctypes.c_char_p
        lib.gba_engine_open.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p),
                                        ctypes.POINTER(ctypes.c_double), ctypes.c_char_p, ctypes.c_size_t]
        lib.gba_engine_open.restype = ctypes.c_int
        lib.gba_engine_run.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                       ctypes.POINTER(GbaRunResult), ctypes.c_char_p, ctypes.c_size_t,
                                       ctypes.c_char_p, ctypes.c_size_t]
        lib.gba_engine_run.restype = ctypes.c_int
        lib.gba_engine_close.argtypes = [ctypes.c_void_p]
        lib.gba_engine_close.restype = None
        lib.gba_engine_prepare.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                           ctypes.POINTER(ctypes.c_double), ctypes.c_char_p, ctypes.c_size_t]
        lib.gba_engine_prepare.restype = ctypes.c_int
        vp, u64, u32, cp, sz = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_size_t
        lib.gba_batcher_open.argtypes = [vp, u64, u64, u64, u64, ctypes.POINTER(vp), ctypes.POINTER(u32), cp, sz]
        lib.gba_batcher_open.restype = ctypes.c_int
        lib.gba_batcher_submit.argtypes = [vp, u64, u64, ctypes.POINTER(u64), cp, sz]
        lib.gba_batcher_submit.restype = ctypes.c_int
        lib.gba_batcher_step.argtypes = [vp, ctypes.POINTER(u64), u32, ctypes.POINTER(u32), ctypes.POINTER(ctypes.c_int),
                                         cp, sz]
        lib.gba_batcher_step.restype = ctypes.c_int
        lib.gba_batcher_result.argtypes = [vp, u64, ctypes.POINTER(GbaRunResult), cp, sz, cp, sz, cp, sz]
        lib.gba_batcher_result.restype = ctypes.c_int
        lib.gba_batcher_totals.argtypes = [vp, ctypes.POINTER(GbaRunResult)]
        lib.gba_batcher_totals.restype = ctypes.c_int
        lib.gba_batcher_memory.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u32)]
        lib.gba_batcher_memory.restype = ctypes.c_int
        lib.gba_batcher_segments.argtypes = [vp, u64, u64, ctypes.POINTER(u64), cp, sz]
        lib.gba_batcher_segments.restype = ctypes.c_int
        lib.gba_batcher_close.argtypes = [vp]
        lib.gba_batcher_close.restype = None
        lib.gba_tune_load.argtypes = [cp, u64, cp, sz, cp, sz]
        lib.gba_tune_load.restype = ctypes.c_int
        _lib = lib
    return _lib


def version() -> str:
    return load().gba_version().decode()


def run_gpu(limit: int, segment_even: int = 467_500_000, spv_path: str = "", verbose: bool = False,
            start_even: int = 4, sieve_split: int = 0, prime_cache_path: str | None = None,
            collect_segments: bool = False) -> dict:
    """Verifies every even n in [start_even, limit] on the GPU; the
    reference's eight arguments in its order (spv_path is ignored: the
    device code is linked into the library).

    Returns the run's figures, its GBSTATS line ("" when statistics are off)
    and, with collect_segments, "per_segment": one dict per segment. Raises
    VerificationError when the run does not pass.
    """
    lib = load()
    res = GbaRunResult()
    line = ctypes.create_string_buffer(1 << 20)
    err = ctypes.create_string_buffer(4096)
    cap = 1 << 16 if collect_segments else 0
    segs = (GbaSegmentStats * max(cap, 1))()
    nseg = ctypes.c_size_t(0)
    cache = prime_cache_path.encode() if prime_cache_path else None
    rc = lib.gba_run_gpu(int(limit), int(segment_even), spv_path.encode(), int(verbose), int(start_even),
                         int(sieve_split), cache, int(bool(collect_segments)), ctypes.byref(res), segs, cap,
                         ctypes.byref(nseg), line, len(line), err, len(err))
    if rc:
        raise VerificationError(err.value.decode())
    out = {name: getattr(res, name) for name, _ in GbaRunResult._fields_}
    out["stats_line"] = line.value.decode()
    if collect_segments:
        if nseg.value > cap:
            raise VerificationError(f"{nseg.value} segments: more than the {cap} collected")
        out["per_segment"] = [{n: getattr(segs[k], n) for n, _ in GbaSegmentStats._fields_ if n != "pad0"}
                              for k in range(nseg.value)]
    return out


class Engine:
    """Device state resident between runs (base primes up to sqrt(max_limit),
    prime tables in HBM, buffers for spans up to max_span); run() has the
    contract of run_gpu for limit <= max_limit."""

    def __init__(self, max_limit: int, max_span: int = 467_500_000):
        self._lib = load()
        self._h = ctypes.c_void_p()
        init = ctypes.c_double(0)
        err = ctypes.create_string_buffer(4096)
        if self._lib.gba_engine_open(int(max_limit), int(max_span), ctypes.byref(self._h), ctypes.byref(init), err,
                                     len(err)):
            raise VerificationError(err.value.decode())
        self.init_seconds = init.value
        # the result text buffers, allocated once (a run writes a terminated string)
        self._line = ctypes.create_string_buffer(1 << 20)
        self._err = ctypes.create_string_buffer(4096)

    def run(self, limit: int, segment_even: int = 467_500_000, start_even: int = 4) -> dict:
        res = GbaRunResult()
        line, err = self._line, self._err
        if self._lib.gba_engine_run(self._h, int(limit), int(segment_even), int(start_even), ctypes.byref(res), line,
                                    len(line), err, len(err)):
            raise VerificationError(err.value.decode())
        out = {name: getattr(res, name) for name, _ in GbaRunResult._fields_}
        out["stats_line"] = line.value.decode()
        return out

    def prepare(self, limit: int, start_even: int = 4) -> float:
        """Builds the range-dependent resident tables for runs over
        [start_even, limit] ahead of the first run; returns the build seconds."""
        sec = ctypes.c_double(0)
        err = ctypes.create_string_buffer(4096)
        if self._lib.gba_engine_prepare(self._h, int(limit), int(start_even), ctypes.byref(sec), err, len(err)):
            raise VerificationError(err.value.decode())
        return sec.value

    def close(self):
        if self._h:
            self._lib.gba_engine_close(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batcher:
    """Many ranges through one engine's pipeline (batched and served
    verification): submit() queues a range, step() issues the next batch of
    waiting segments and collects the one before it, returning the ids of
    ranges it completed; result() gives a range's figures, GBSTATS line and
    the state text of a fresh campaign after it alone."""

    def __init__(self, engine: Engine, max_limit: int, min_start: int, segment_even: int = 467_500_000,
                 device_budget: int = 0):
        self._lib = load()
        self._engine = engine            # the engine must outlive the batcher
        self._h = ctypes.c_void_p()
        kb = ctypes.c_uint32(0)
        err = ctypes.create_string_buffer(4096)
        if self._lib.gba_batcher_open(engine._h, int(max_limit), int(min_start), int(segment_even),
                                      int(device_budget), ctypes.byref(self._h), ctypes.byref(kb), err, len(err)):
            raise VerificationError(err.value.decode())
        self.batch_segments = kb.value
        self._ids = (ctypes.c_uint64 * 1024)()
        self._line = ctypes.create_string_buffer(1 << 20)
        self._state = ctypes.create_string_buffer(1 << 20)
        self._err = ctypes.create_string_buffer(4096)

    def submit(self, start_even: int, limit: int) -> int:
        rid = ctypes.c_uint64(0)
        err = ctypes.create_string_buffer(4096)
        if self._lib.gba_batcher_submit(self._h, int(start_even), int(limit), ctypes.byref(rid), err, len(err)):
            raise VerificationError(err.value.decode())
        return rid.value

    def step(self):
        """One pipeline step: (ids of ranges completed, whether work remains)."""
        n = ctypes.c_uint32(0)
        more = ctypes.c_int(0)
        err = self._err
        if self._lib.gba_batcher_step(self._h, self._ids, len(self._ids), ctypes.byref(n), ctypes.byref(more), err,
                                      len(err)):
            raise VerificationError(err.value.decode())
        return [self._ids[k] for k in range(n.value)], bool(more.value)

    def result(self, rid: int) -> dict:
        res = GbaRunResult()
        line, state, err = self._line, self._state, self._err
        if self._lib.gba_batcher_result(self._h, int(rid), ctypes.byref(res), line, len(line), state, len(state),
                                        err, len(err)):
            raise VerificationError(err.value.decode())
        out = {name: getattr(res, name) for name, _ in GbaRunResult._fields_}
        out["stats_line"] = line.value.decode()
        out["state"] = state.value.decode()
        return out

    def totals(self) -> dict:
        res = GbaRunResult()
        self._lib.gba_batcher_totals(self._h, ctypes.byref(res))
        return {name: getattr(res, name) for name, _ in GbaRunResult._fields_}

    def memory(self) -> tuple:
        """(device bytes the batch buffers hold now, segments in flight at once)."""
        nbytes, inflight = ctypes.c_uint64(0), ctypes.c_uint32(0)
        self._lib.gba_batcher_memory(self._h, ctypes.byref(nbytes), ctypes.byref(inflight))
        return nbytes.value, inflight.value

    def segments_of(self, start_even: int, limit: int) -> int:
        """The segments [start_even, limit] splits into."""
        n = ctypes.c_uint64(0)
        err = ctypes.create_string_buffer(4096)
        if self._lib.gba_batcher_segments(self._h, int(start_even), int(limit), ctypes.byref(n), err, len(err)):
            raise VerificationError(err.value.decode())
        return n.value

    def run_all(self) -> list:
        """Steps until every submitted range is done; returns the completion order."""
        order = []
        more = True
        while more:
            done, more = self.step()
            order += done
        return order

    def close(self):
        if self._h:
            self._lib.gba_batcher_close(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tune_load(path, limit: int) -> str:
    """Installs the tuned knob defaults of `path`'s band holding `limit`
    (path None: drops them); returns the installed "KNOB=value ..." text."""
    lib = load()
    out = ctypes.create_string_buffer(4096)
    err = ctypes.create_string_buffer(4096)
    if lib.gba_tune_load(None if path is None else str(path).encode(), int(limit), out, len(out), err, len(err)):
        raise VerificationError(err.value.decode())
    return out.value.decode()


def segment_map(seg_start: int, limit: int, span: int):
    """One segment's class map from the GPU: (words [8, cw] uint64, Q[8], nc[8])."""
    import numpy as np

    lib = load()
    # upper bound of the stride: the merged last segment spans below 2 x span
    cap = 8 * ((span * 2 + 20000) // 30 // 64 + (1 << 20) // 64 + 64)
    words = np.zeros(cap, dtype=np.uint64)
    cw = ctypes.c_uint32(0)
    Q = (ctypes.c_uint64 * 8)()
    nc = (ctypes.c_uint32 * 8)()
    err = ctypes.create_string_buffer(4096)
    rc = lib.gba_segment_map(int(seg_start), int(limit), int(span), words.ctypes.data, cap, ctypes.byref(cw), Q, nc,
                             err, len(err))
    if rc:
        raise VerificationError(err.value.decode())
    n = 8 * cw.value
    if n > cap:
        raise VerificationError("segment map larger than the buffer")
    return words[:n].reshape(8, cw.value), list(Q), list(nc)


def env_scope(**kv):
    """Context manager that sets GB_* knobs for the calls inside it."""
    import contextlib

    @contextlib.contextmanager
    def scope():
        old = {k: os.environ.get(k) for k in kv}
        try:
            for k, v in kv.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = str(v)
            yield
        finally:
            for k, v in old.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v

    return scope()
