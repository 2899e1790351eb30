"""Python view of the streaming server (include/mdsm_c.h, "Streaming server").

    srv = Server(hbm_budget=0)
    b = srv.open_beam(obs_kwargs, ranges, clip_nsigma=5.0, snr_thr=6.0,
                      widths=(1, 2, 4, 8, 16, 32, 64))
    srv.push(b, block)                      # [nchans][N] host block
    res, cands = srv.poll(wait_ms=-1)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import (CAND_DTYPE, DTYPES, NP_DTYPE, Cand, MdsmError, Obs, PlanInfo, Range, check,
               lib)


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_inflight", C.c_uint32),
                ("max_batch", C.c_uint32), ("persistent", C.c_uint32),
                ("queue_depth", C.c_uint32)]


class DetectConfig(C.Structure):
    _fields_ = [("clip_nsigma", C.c_float), ("snr_thr", C.c_double),
                ("widths_mask", C.c_uint32), ("cand_cap", C.c_uint32),
                ("retain_output", C.c_uint32), ("pad", C.c_uint32)]


class Result(C.Structure):
    _fields_ = [("beam", C.c_uint32), ("status", C.c_int32), ("seq", C.c_uint64),
                ("ncand", C.c_uint32), ("nreturned", C.c_uint32), ("nclipped", C.c_uint32),
                ("batch", C.c_uint32), ("flags", C.c_uint32), ("pad", C.c_uint32),
                ("t_arrival", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]


# Result.flags
OVERFLOW, ZERO_FILLED, SKIPPED = 1, 2, 4
CANCELLED = -9


class ServerStats(C.Structure):
    _fields_ = [("pushed", C.c_uint64), ("admitted", C.c_uint64), ("completed", C.c_uint64),
                ("deferred_budget", C.c_uint64), ("deferred_inflight", C.c_uint64),
                ("launches", C.c_uint64), ("bytes_in_use", C.c_uint64),
                ("bytes_peak", C.c_uint64), ("budget", C.c_uint64), ("beams", C.c_uint32),
                ("inflight", C.c_uint32)]


_ready = False


def _sig():
    global _ready
    L = lib()
    if _ready:
        return L
    vp = C.c_void_p
    spec = {
        "mdsm_server_create": [C.POINTER(ServerConfig), C.POINTER(vp)],
        "mdsm_beam_open": [vp, C.POINTER(Obs), C.POINTER(Range), C.c_uint32,
                           C.POINTER(DetectConfig), C.POINTER(C.c_uint32)],
        "mdsm_beam_close": [vp, C.c_uint32],
        "mdsm_beam_cancel": [vp, C.c_uint32],
        "mdsm_beam_push_seq": [vp, C.c_uint32, vp, C.c_int64, C.c_uint64, C.c_double],
        "mdsm_beam_push_device": [vp, C.c_uint32, vp, C.c_int64, C.c_double,
                                  C.POINTER(C.c_uint64)],
        "mdsm_beam_export": [vp, C.c_uint32, vp, C.c_uint64, C.POINTER(C.c_uint64)],
        "mdsm_beam_import": [vp, C.c_uint32, vp, C.c_uint64],
        "mdsm_beam_info": [vp, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(PlanInfo)],
        "mdsm_beam_push": [vp, C.c_uint32, vp, C.c_int64, C.c_double, C.POINTER(C.c_uint64)],
        "mdsm_beam_acquire": [vp, C.c_uint32, C.POINTER(vp)],
        "mdsm_beam_commit": [vp, C.c_uint32, vp, C.c_double, C.POINTER(C.c_uint64)],
        "mdsm_server_poll": [vp, C.POINTER(Result), vp, C.c_uint32, C.c_int32],
        "mdsm_server_output": [vp, C.c_uint32, C.c_uint64, C.POINTER(vp)],
        "mdsm_server_release": [vp, C.c_uint32, C.c_uint64],
        "mdsm_server_hold": [vp, C.c_int],
        "mdsm_server_drain": [vp],
        "mdsm_server_get_stats": [vp, C.POINTER(ServerStats)],
    }
    for n, a in spec.items():
        f = getattr(L, n)
        f.argtypes = a
        f.restype = C.c_int
    L.mdsm_server_destroy.argtypes = [vp]
    L.mdsm_server_destroy.restype = None
    L.mdsm_now.restype = C.c_double
    L.mdsm_now.argtypes = []
    _ready = True
    return L


def now() -> float:
    return _sig().mdsm_now()


class Beam:
    def __init__(self, srv, bid, obs, cap, overlap, info):
        self.srv, self.id, self.obs, self.cap = srv, bid, obs, cap
        self.overlap, self.info = overlap, info


class Server:
    def __init__(self, hbm_budget=0, max_inflight=0, max_batch=0, persistent=False,
                 queue_depth=0):
        # persistent: False/0 off, True/1 default grid, n > 1: n workgroups
        cfg = ServerConfig(hbm_budget, max_inflight, max_batch, int(persistent), queue_depth)
        h = C.c_void_p()
        check(_sig().mdsm_server_create(C.byref(cfg), C.byref(h)))
        self.h = h.value
        self.beams = {}

    def open_beam(self, nchans, fch1, foff, tsamp, nsamp_block, dtype="u8", ranges=(),
                  shift_rule=0, clip_nsigma=0.0, snr_thr=6.0, widths=(1, 2, 4, 8, 16, 32, 64),
                  cand_cap=65536, retain_output=False) -> Beam:
        dt = DTYPES[dtype] if isinstance(dtype, str) else int(dtype)
        obs = Obs(nchans, fch1, foff, tsamp, nsamp_block, dt, shift_rule)
        rr = (Range * len(ranges))(*[Range(float(a), float(b), int(n), int(k))
                                     for a, b, n, k in ranges])
        mask = 0
        for w in widths:
            mask |= int(w)
        det = DetectConfig(clip_nsigma, snr_thr, mask, cand_cap, 1 if retain_output else 0, 0)
        bid = C.c_uint32()
        check(_sig().mdsm_beam_open(self.h, C.byref(obs), rr, len(ranges), C.byref(det),
                                    C.byref(bid)))
        ov = C.c_uint32()
        info = PlanInfo()
        check(_sig().mdsm_beam_info(self.h, bid.value, C.byref(ov), C.byref(info)))
        b = Beam(self, bid.value, obs, cand_cap, ov.value, info)
        self.beams[b.id] = b
        return b

    def close_beam(self, b: Beam):
        """Process every block already pushed, then free the beam."""
        check(_sig().mdsm_beam_close(self.h, b.id))
        self.beams.pop(b.id, None)

    def cancel_beam(self, b: Beam):
        """Drop the beam's queued blocks (reported with status CANCELLED) and
        free it once its block on the device has completed."""
        check(_sig().mdsm_beam_cancel(self.h, b.id))
        self.beams.pop(b.id, None)

    def push_seq(self, b: Beam, block: np.ndarray, seq: int, arrival=-1.0):
        """Push block number `seq`: gaps are zero-filled, duplicates raise
        (MdsmError -6); returns False on back-pressure."""
        block = np.ascontiguousarray(block, NP_DTYPE[b.obs.dtype])
        rc = _sig().mdsm_beam_push_seq(self.h, b.id, block.ctypes.data, block.shape[1], seq,
                                       arrival)
        if rc == -8:
            return False
        check(rc)
        return True

    def push_device(self, b: Beam, d_block, stride: int, arrival=-1.0):
        """Queue a block already in device memory (kept unchanged by the
        caller until its result is polled); None on back-pressure."""
        seq = C.c_uint64()
        ptr = d_block.ptr if hasattr(d_block, "ptr") else d_block
        rc = _sig().mdsm_beam_push_device(self.h, b.id, ptr, stride, arrival, C.byref(seq))
        if rc == -8:
            return None
        check(rc)
        return seq.value

    def export_state(self, b: Beam) -> bytes:
        n = C.c_uint64()
        rc = _sig().mdsm_beam_export(self.h, b.id, None, 0, C.byref(n))
        if rc not in (0, -4):
            check(rc)
        buf = (C.c_char * n.value)()
        check(_sig().mdsm_beam_export(self.h, b.id, buf, n.value, C.byref(n)))
        return bytes(buf)

    def import_state(self, b: Beam, state: bytes):
        buf = (C.c_char * len(state)).from_buffer_copy(state)
        check(_sig().mdsm_beam_import(self.h, b.id, buf, len(state)))

    def push(self, b: Beam, block: np.ndarray, arrival=-1.0) -> int:
        block = np.ascontiguousarray(block, NP_DTYPE[b.obs.dtype])
        seq = C.c_uint64()
        check(_sig().mdsm_beam_push(self.h, b.id, block.ctypes.data, block.shape[1], arrival,
                                    C.byref(seq)))
        return seq.value

    def acquire(self, b: Beam):
        """Lend the beam's next free pinned staging block as a [nchans][N]
        array to fill in place, or None when none is free."""
        ptr = C.c_void_p()
        rc = _sig().mdsm_beam_acquire(self.h, b.id, C.byref(ptr))
        if rc == -8:
            return None
        check(rc)
        dt = np.dtype(NP_DTYPE[b.obs.dtype])
        n = b.obs.nchans * b.obs.nsamp_in
        buf = (C.c_char * (n * dt.itemsize)).from_address(ptr.value)
        return np.frombuffer(buf, dt, n).reshape(b.obs.nchans, b.obs.nsamp_in)

    def commit(self, b: Beam, block: np.ndarray, arrival=-1.0) -> int:
        seq = C.c_uint64()
        check(_sig().mdsm_beam_commit(self.h, b.id, block.ctypes.data, arrival, C.byref(seq)))
        return seq.value

    def try_push(self, b: Beam, block: np.ndarray, arrival=-1.0):
        """push, or None when the beam's queue is full (back-pressure)."""
        try:
            return self.push(b, block, arrival)
        except MdsmError as e:
            if e.code == -8:
                return None
            raise

    def poll(self, wait_ms=-1, cap=1 << 20):
        res = Result()
        buf = (Cand * cap)()
        rc = _sig().mdsm_server_poll(self.h, C.byref(res), C.cast(buf, C.c_void_p), cap,
                                     wait_ms)
        if rc == -8:
            return None, None
        check(rc)
        n = res.nreturned
        arr = np.zeros(n, CAND_DTYPE)
        if n:
            raw = np.ctypeslib.as_array(buf)[:n]
            for f in ("dm", "t", "width", "snr"):
                arr[f] = raw[f]
        return res, arr

    def output(self, b: Beam, seq: int) -> np.ndarray:
        """Host copy of a retained block output (packed plan layout)."""
        from . import D2H
        p = C.c_void_p()
        check(_sig().mdsm_server_output(self.h, b.id, seq, C.byref(p)))
        out = np.empty(b.info.out_elems, np.float32)
        check(lib().mdsm_memcpy(out.ctypes.data, p.value, out.nbytes, D2H, None))
        return out

    def release(self, b: Beam, seq: int):
        check(_sig().mdsm_server_release(self.h, b.id, seq))

    def hold(self, flag=True):
        """Pause (True) or resume (False) admissions: blocks pushed while held
        are admitted together on resume."""
        check(_sig().mdsm_server_hold(self.h, int(bool(flag))))

    def drain(self):
        check(_sig().mdsm_server_drain(self.h))

    def stats(self) -> dict:
        st = ServerStats()
        check(_sig().mdsm_server_get_stats(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            _sig().mdsm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
