"""MI355X-native incoherent-dedispersion path of MDSM (host bindings).

Python bindings over the C ABI of ``libmdsm_hip.so`` (include/mdsm_c.h).  The
compute path is the gfx950 HIP library; nothing here computes samples on the
CPU.  If the library has not been built, every entry point raises: there is
no fallback.

    from mdsm_amd import Plan, DeviceBuffer, synth_filterbank
    plan = Plan.brute(nchans=256, fch1=1500.0, foff=-0.390625, tsamp=64e-6,
                      nsamp_in=65536, dtype="f32",
                      ranges=[(0.0, 0.2, 512, 1)])
    outs = plan.dedisperse_host(x)      # list of [ndms_r][nout_r] arrays
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MDSM_LIB", os.path.join(HERE, "libmdsm_hip.so"))

U8, F16, F32 = 0, 1, 2
SHIFT_REF, SHIFT_NS = 0, 1
H2D, D2H, D2D = 1, 2, 3
DTYPES = {"u8": U8, "f16": F16, "f32": F32}
NP_DTYPE = {U8: np.uint8, F16: np.float16, F32: np.float32}
STATUS = {0: "OK", -1: "INVALID", -2: "NOMEM", -3: "DEVICE", -4: "CAPACITY", -5: "BUDGET",
          -6: "STATE", -7: "PARSE", -8: "AGAIN"}


class MdsmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mdsm error {code} ({STATUS.get(code, '?')}): {msg}")
        self.code = code


# ---------------------------------------------------------------- C structs
class Obs(C.Structure):
    _fields_ = [("nchans", C.c_uint32), ("fch1", C.c_float), ("foff", C.c_float),
                ("tsamp", C.c_float), ("nsamp_in", C.c_uint32), ("dtype", C.c_int32),
                ("shift_rule", C.c_int32)]


class Range(C.Structure):
    _fields_ = [("lowdm", C.c_float), ("dmstep", C.c_float), ("ndms", C.c_uint32),
                ("downsample", C.c_uint32)]


class Pass(C.Structure):
    _fields_ = [("lowdm", C.c_float), ("highdm", C.c_float), ("dmstep", C.c_float),
                ("sub_dmstep", C.c_float), ("binsize", C.c_uint32), ("ndms", C.c_uint32),
                ("ncalls", C.c_uint32), ("calldms", C.c_uint32)]


class PlanInfo(C.Structure):
    _fields_ = [("nranges", C.c_uint32), ("ndms_total", C.c_uint32), ("maxshift", C.c_uint32),
                ("nchans", C.c_uint32), ("out_elems", C.c_uint64), ("work_bytes", C.c_uint64),
                ("dm_chan_samples", C.c_double), ("adds", C.c_double), ("bytes", C.c_double),
                ("is_subband", C.c_uint32), ("nblocks", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("chunk_rows", C.c_uint32), ("lds_budget", C.c_uint32),
                ("persistent", C.c_uint32), ("grid", C.c_uint32), ("tile_group", C.c_uint32)]

    def as_dict(self) -> dict:
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


class SigprocHeader(C.Structure):
    _fields_ = [("source_name", C.c_char * 80), ("rawdatafile", C.c_char * 80),
                ("machine_id", C.c_int32), ("telescope_id", C.c_int32), ("data_type", C.c_int32),
                ("nchans", C.c_int32), ("nbits", C.c_int32), ("nifs", C.c_int32),
                ("nbeams", C.c_int32), ("ibeam", C.c_int32), ("barycentric", C.c_int32),
                ("pulsarcentric", C.c_int32), ("nbins", C.c_int32), ("nsamples", C.c_int32),
                ("tstart", C.c_double), ("tsamp", C.c_double), ("fch1", C.c_double),
                ("foff", C.c_double), ("refdm", C.c_double), ("az_start", C.c_double),
                ("za_start", C.c_double), ("src_raj", C.c_double), ("src_dej", C.c_double),
                ("period", C.c_double), ("npuls", C.c_int64), ("header_bytes", C.c_uint64),
                ("nspectra", C.c_uint64)]

    def as_dict(self) -> dict:
        out = {}
        for f, _ in self._fields_:
            v = getattr(self, f)
            out[f] = v.decode() if isinstance(v, bytes) else v
        return out


class DDPlanParams(C.Structure):
    _fields_ = [("lodm", C.c_double), ("hidm", C.c_double), ("f_ctr", C.c_double),
                ("bw", C.c_double), ("dt", C.c_double), ("nchans", C.c_uint32),
                ("nsubs", C.c_uint32), ("ok_smearing", C.c_double)]


class RangeInfo(C.Structure):
    _fields_ = [("lowdm", C.c_float), ("dmstep", C.c_float), ("ndms", C.c_uint32),
                ("downsample", C.c_uint32), ("nout", C.c_uint32), ("maxshift", C.c_uint32),
                ("out_offset", C.c_uint64), ("first_dm", C.c_uint32), ("ncalls", C.c_uint32)]


class Cand(C.Structure):
    _fields_ = [("dm", C.c_uint32), ("t", C.c_int32), ("width", C.c_uint32),
                ("snr", C.c_float)]


CAND_DTYPE = np.dtype([("dm", np.uint32), ("t", np.int32), ("width", np.uint32),
                       ("snr", np.float32)])


class Occupancy(C.Structure):
    _fields_ = [("workgroups_per_cu", C.c_int32), ("waves_per_simd", C.c_int32),
                ("vgprs", C.c_int32), ("static_lds", C.c_int32), ("dynamic_lds", C.c_int32),
                ("scratch_bytes", C.c_int32)]


class Pulse(C.Structure):
    _fields_ = [("dm", C.c_double), ("t0", C.c_uint32), ("width", C.c_uint32),
                ("amp", C.c_float), ("pad", C.c_uint32)]


class SynthSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("dtype", C.c_int32), ("nchans", C.c_uint32),
                ("nsamp", C.c_uint32), ("fch1", C.c_float), ("foff", C.c_float),
                ("tsamp", C.c_float), ("mean", C.c_float), ("sigma", C.c_float),
                ("gain_lo", C.c_float), ("gain_hi", C.c_float), ("npulses", C.c_uint32),
                ("pulses", C.POINTER(Pulse)), ("nthreads", C.c_uint32)]


class DeviceProps(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("total_mem", C.c_uint64),
                ("free_mem", C.c_uint64), ("cus", C.c_int32), ("clock_khz", C.c_int32),
                ("mem_clock_khz", C.c_int32), ("mem_bus_bits", C.c_int32),
                ("lds_per_block", C.c_uint32), ("pad", C.c_uint32)]


# ---------------------------------------------------------------- loading
_lib = None


def lib():
    """The loaded libmdsm_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with python -m mdsm_amd.build "
                           "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, i32, u32, u64, i64, f32, f64 = (C.c_void_p, C.c_int32, C.c_uint32, C.c_uint64,
                                        C.c_int64, C.c_float, C.c_double)
    sig = {
        "mdsm_plan_create": [C.POINTER(Obs), C.POINTER(Range), u32, C.POINTER(vp)],
        "mdsm_plan_create_subband": [C.POINTER(Obs), u32, C.POINTER(Pass), u32, C.POINTER(vp)],
        "mdsm_plan_get_info": [vp, C.POINTER(PlanInfo)],
        "mdsm_plan_get_range": [vp, u32, C.POINTER(RangeInfo)],
        "mdsm_plan_shift_table": [vp, u32, vp],
        "mdsm_plan_fine_table": [vp, u32, vp],
        "mdsm_plan_dm_values": [vp, vp],
        "mdsm_dedisperse": [vp, vp, i64, vp, vp, vp],
        "mdsm_decimate": [vp, i32, u32, i64, u32, u32, vp, i64, vp],
        "mdsm_clip_u8": [vp, i64, u32, i64, u32, f32, vp, i64, vp, vp, vp, vp],
        "mdsm_detect_boxcar": [vp, u32, u32, i64, vp, u32, f64, vp, u32, vp, vp, vp, vp],
        "mdsm_detect_boxcar_stream": [vp, u32, u32, i64, vp, i64, i32, vp, u32, f64, vp, u32, vp,
                                      vp, vp, vp],
        "mdsm_detect_median5": [vp, u32, u32, i64, vp, f64, i32, vp, vp, u32, vp, u32, vp, vp,
                                vp],
        "mdsm_synth_filterbank": [C.POINTER(SynthSpec), vp, i64],
        "mdsm_synth_filterbank_device": [C.POINTER(SynthSpec), vp, i64, vp],
        "mdsm_timing_enable": [C.c_int],
        "mdsm_timing_reset": [],
        "mdsm_timing_get": [C.c_char_p, C.POINTER(f64), C.POINTER(u64)],
        "mdsm_dedisperse_plain": [vp, vp, i64, vp, vp],
        "mdsm_sigproc_read_header": [C.c_char_p, C.POINTER(SigprocHeader)],
        "mdsm_ddplan": [C.POINTER(DDPlanParams), vp, u32, C.POINTER(u32)],
        "mdsm_sigproc_open": [C.c_char_p, C.POINTER(vp), C.POINTER(SigprocHeader)],
        "mdsm_sigproc_read": [vp, u64, vp, C.POINTER(u64)],
        "mdsm_sigproc_seek": [vp, u64],
        "mdsm_unpack_spectra": [vp, i32, u32, u32, i64, i32, vp, i32, i64, i64, vp],
        "mdsm_series_writer_open": [C.c_char_p, vp, u32, f64, f64, f64, f64, C.POINTER(vp)],
        "mdsm_series_writer_write": [vp, vp, vp, u32, u32, i64],
        "mdsm_corner_turn": [vp, i32, u32, u32, i64, i32, vp, i32, i64, i64, vp],
        "mdsm_detrend": [vp, u32, u32, i64, u32, vp],
        "mdsm_batch_create": [C.POINTER(vp), u32, C.POINTER(vp)],
        "mdsm_batch_run": [vp, C.POINTER(vp), C.POINTER(i64), C.POINTER(vp), C.POINTER(vp), vp],
        "mdsm_batch_capture": [vp, C.POINTER(vp), C.POINTER(i64), C.POINTER(vp), C.POINTER(vp),
                               vp],
        "mdsm_batch_replay": [vp, vp],
        "mdsm_batch_autotune": [vp, C.POINTER(vp), C.POINTER(i64), C.POINTER(vp), C.POINTER(vp),
                                vp, u32, C.POINTER(Tuning)],
        "mdsm_plan_set_tuning": [vp, C.POINTER(Tuning)],
        "mdsm_plan_occupancy": [vp, u32, C.POINTER(Occupancy)],
        "mdsm_plan_lds_bytes": [vp, u32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
        "mdsm_plan_upload": [vp],
        "mdsm_probe_lds_read": [u32, u32, C.POINTER(f64)],
        "mdsm_probe_lds_write": [u32, u32, C.POINTER(f64)],
        "mdsm_probe_lds_read_form": [u32, u32, C.POINTER(f64)],
        "mdsm_dedisperse_noalloc": [vp, vp, i64, vp, vp],
        "mdsm_plan_get_tuning": [vp, C.POINTER(Tuning)],
        "mdsm_plan_autotune": [vp, vp, i64, vp, vp, vp, u32, C.POINTER(Tuning)],
        "mdsm_device_count": [C.POINTER(C.c_int)],
        "mdsm_device_set": [C.c_int],
        "mdsm_device_query": [C.c_int, C.POINTER(DeviceProps)],
        "mdsm_device_malloc": [C.POINTER(vp), u64],
        "mdsm_device_free": [vp],
        "mdsm_host_malloc_pinned": [C.POINTER(vp), u64],
        "mdsm_host_free_pinned": [vp],
        "mdsm_memcpy": [vp, vp, u64, C.c_int, vp],
        "mdsm_memcpy_2d": [vp, u64, vp, u64, u64, u64, C.c_int, vp],
        "mdsm_fill_u32": [vp, u32, u64, vp],
        "mdsm_stream_create": [C.POINTER(vp)],
        "mdsm_stream_destroy": [vp],
        "mdsm_stream_synchronize": [vp],
        "mdsm_device_synchronize": [],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.mdsm_plan_destroy.argtypes = [vp]
    L.mdsm_plan_destroy.restype = None
    L.mdsm_batch_destroy.argtypes = [vp]
    L.mdsm_batch_destroy.restype = None
    for name in ("mdsm_sigproc_close", "mdsm_series_writer_close"):
        getattr(L, name).argtypes = [vp]
        getattr(L, name).restype = None
    L.mdsm_noalloc_max_work.restype = C.c_uint64
    L.mdsm_noalloc_max_work.argtypes = []
    L.mdsm_median5_stats_len.restype = C.c_uint64
    L.mdsm_median5_stats_len.argtypes = [u32, u32]
    L.mdsm_boxcar_carry_len.restype = C.c_uint32
    L.mdsm_boxcar_carry_len.argtypes = [u32]
    L.mdsm_last_error.restype = C.c_char_p
    L.mdsm_version.restype = C.c_char_p
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise MdsmError(rc, lib().mdsm_last_error().decode(errors="replace"))
    return rc


def version() -> str:
    return lib().mdsm_version().decode()


# ---------------------------------------------------------------- device plumbing
def device_count() -> int:
    n = C.c_int(0)
    rc = lib().mdsm_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def device_props(dev: int = 0) -> dict:
    p = DeviceProps()
    check(lib().mdsm_device_query(dev, C.byref(p)))
    return {"name": p.name.decode(), "arch": p.arch.decode(), "total_mem": p.total_mem,
            "free_mem": p.free_mem, "cus": p.cus, "clock_khz": p.clock_khz,
            "mem_clock_khz": p.mem_clock_khz, "mem_bus_bits": p.mem_bus_bits,
            "lds_per_block": p.lds_per_block}


def synchronize():
    check(lib().mdsm_device_synchronize())


class Stream:
    def __init__(self):
        h = C.c_void_p()
        check(lib().mdsm_stream_create(C.byref(h)))
        self.handle = h.value

    def synchronize(self):
        check(lib().mdsm_stream_synchronize(self.handle))

    def close(self):
        if self.handle:
            lib().mdsm_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _sh(stream):
    if stream is None:
        return None
    return stream.handle if isinstance(stream, Stream) else stream


class DeviceBuffer:
    """A device allocation owned by this object (freed on close/__del__)."""

    def __init__(self, nbytes: int):
        h = C.c_void_p()
        check(lib().mdsm_device_malloc(C.byref(h), int(nbytes)))
        self.ptr = h.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_host(cls, arr: np.ndarray, stream=None) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes)
        b.upload(arr, stream)
        return b

    def upload(self, arr: np.ndarray, stream=None, offset: int = 0):
        arr = np.ascontiguousarray(arr)
        if offset + arr.nbytes > self.nbytes:
            raise ValueError("upload exceeds buffer")
        check(lib().mdsm_memcpy(self.ptr + offset, arr.ctypes.data, arr.nbytes, H2D, _sh(stream)))

    def download_rows(self, dtype, nrows: int, row_elems: int, pitch_elems: int,
                      offset: int = 0) -> np.ndarray:
        """Download nrows rows of row_elems elements spaced pitch_elems apart."""
        out = np.empty((nrows, row_elems), dtype)
        es = out.itemsize
        if offset + ((nrows - 1) * pitch_elems + row_elems) * es > self.nbytes:
            raise ValueError("download exceeds buffer")
        check(lib().mdsm_memcpy_2d(out.ctypes.data, row_elems * es, self.ptr + offset,
                                   pitch_elems * es, row_elems * es, nrows, D2H, None))
        return out

    def download(self, dtype, shape, offset: int = 0) -> np.ndarray:
        out = np.empty(shape, dtype)
        if offset + out.nbytes > self.nbytes:
            raise ValueError("download exceeds buffer")
        check(lib().mdsm_memcpy(out.ctypes.data, self.ptr + offset, out.nbytes, D2H, None))
        return out

    def fill_u32(self, value: int = 0, count: int | None = None, stream=None):
        n = self.nbytes // 4 if count is None else count
        check(lib().mdsm_fill_u32(self.ptr, value, n, _sh(stream)))

    def close(self):
        if self.ptr:
            lib().mdsm_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- timing
def timing_enable(on: bool = True):
    check(lib().mdsm_timing_enable(1 if on else 0))


def timing_reset():
    check(lib().mdsm_timing_reset())


def timing_get(kernel: str):
    ms, n = C.c_double(), C.c_uint64()
    check(lib().mdsm_timing_get(kernel.encode(), C.byref(ms), C.byref(n)))
    return ms.value, n.value


# ---------------------------------------------------------------- plans
@dataclass
class RangeLayout:
    lowdm: float
    dmstep: float
    ndms: int
    downsample: int
    nout: int
    maxshift: int
    out_offset: int
    first_dm: int
    ncalls: int


class Plan:
    """A dedispersion plan: brute-force multi-range or two-stage subband."""

    def __init__(self, handle, obs: Obs):
        self._h = handle
        self.obs = obs
        info = PlanInfo()
        check(lib().mdsm_plan_get_info(self._h, C.byref(info)))
        self.info = info
        self.ranges: list[RangeLayout] = []
        for r in range(info.nranges):
            ri = RangeInfo()
            check(lib().mdsm_plan_get_range(self._h, r, C.byref(ri)))
            self.ranges.append(RangeLayout(ri.lowdm, ri.dmstep, ri.ndms, ri.downsample, ri.nout,
                                           ri.maxshift, ri.out_offset, ri.first_dm, ri.ncalls))
        self._work = None

    @staticmethod
    def make_obs(nchans, fch1, foff, tsamp, nsamp_in, dtype="f32", shift_rule=SHIFT_REF) -> Obs:
        dt = DTYPES[dtype] if isinstance(dtype, str) else int(dtype)
        return Obs(nchans, fch1, foff, tsamp, nsamp_in, dt, shift_rule)

    @classmethod
    def brute(cls, nchans, fch1, foff, tsamp, nsamp_in, dtype="f32",
              ranges: Sequence[tuple] = ((0.0, 1.0, 1, 1),), shift_rule=SHIFT_REF) -> "Plan":
        obs = cls.make_obs(nchans, fch1, foff, tsamp, nsamp_in, dtype, shift_rule)
        arr = (Range * len(ranges))(*[Range(float(a), float(b), int(n), int(k))
                                      for (a, b, n, k) in ranges])
        h = C.c_void_p()
        check(lib().mdsm_plan_create(C.byref(obs), arr, len(ranges), C.byref(h)))
        return cls(h.value, obs)

    @classmethod
    def subband(cls, nchans, fch1, foff, tsamp, nsamp_in, nsubs, passes: Sequence[dict],
                dtype="f32", shift_rule=SHIFT_REF) -> "Plan":
        obs = cls.make_obs(nchans, fch1, foff, tsamp, nsamp_in, dtype, shift_rule)
        arr = (Pass * len(passes))()
        for i, p in enumerate(passes):
            arr[i] = Pass(p.get("lowdm", 0.0), p.get("highdm", 0.0), p["dmstep"],
                          p["sub_dmstep"], p.get("binsize", 1), p["ncalls"] * p["calldms"],
                          p["ncalls"], p["calldms"])
        h = C.c_void_p()
        check(lib().mdsm_plan_create_subband(C.byref(obs), nsubs, arr, len(passes), C.byref(h)))
        return cls(h.value, obs)

    @property
    def handle(self):
        return self._h

    def shift_table(self, r: int) -> np.ndarray:
        R = self.ranges[r]
        out = np.empty((R.ncalls, self.obs.nchans), np.int32)
        check(lib().mdsm_plan_shift_table(self._h, r, out.ctypes.data))
        return out

    def fine_table(self, r: int, nsubs: int) -> np.ndarray:
        R = self.ranges[r]
        out = np.empty((R.ndms, nsubs), np.int32)
        check(lib().mdsm_plan_fine_table(self._h, r, out.ctypes.data))
        return out

    def dm_values(self) -> np.ndarray:
        out = np.empty(self.info.ndms_total, np.float32)
        check(lib().mdsm_plan_dm_values(self._h, out.ctypes.data))
        return out

    def dedisperse(self, d_in, in_stride: int, d_out, d_work=None, stream=None):
        pin = d_in.ptr if isinstance(d_in, DeviceBuffer) else d_in
        pout = d_out.ptr if isinstance(d_out, DeviceBuffer) else d_out
        pw = d_work.ptr if isinstance(d_work, DeviceBuffer) else d_work
        check(lib().mdsm_dedisperse(self._h, pin, int(in_stride), pout, pw, _sh(stream)))

    def upload(self):
        """Put the plan's device tables on the device now (set-up)."""
        check(lib().mdsm_plan_upload(self._h))

    def dedisperse_noalloc(self, d_in, in_stride: int, d_out, stream=None):
        """mdsm_dedisperse_noalloc: workspace and counters in the library's
        static per-GPU array; needs upload() first."""
        pin = d_in.ptr if isinstance(d_in, DeviceBuffer) else d_in
        pout = d_out.ptr if isinstance(d_out, DeviceBuffer) else d_out
        check(lib().mdsm_dedisperse_noalloc(self._h, pin, int(in_stride), pout, _sh(stream)))

    def dedisperse_plain(self, d_in, in_stride: int, d_out, stream=None):
        """The reference-shaped plain float32 kernel (a performance baseline)."""
        pin = d_in.ptr if isinstance(d_in, DeviceBuffer) else d_in
        pout = d_out.ptr if isinstance(d_out, DeviceBuffer) else d_out
        check(lib().mdsm_dedisperse_plain(self._h, pin, int(in_stride), pout, _sh(stream)))

    def set_tuning(self, chunk_rows=0, lds_budget=0, persistent=0, grid=0, tile_group=0):
        t = Tuning(chunk_rows, lds_budget, persistent, grid, tile_group)
        check(lib().mdsm_plan_set_tuning(self._h, C.byref(t)))

    def tuning(self) -> dict:
        t = Tuning()
        check(lib().mdsm_plan_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def autotune(self, d_in, in_stride: int, d_out, d_work=None, stream=None, reps: int = 3) -> dict:
        pin = d_in.ptr if isinstance(d_in, DeviceBuffer) else d_in
        pout = d_out.ptr if isinstance(d_out, DeviceBuffer) else d_out
        pw = d_work.ptr if isinstance(d_work, DeviceBuffer) else d_work
        t = Tuning()
        check(lib().mdsm_plan_autotune(self._h, pin, int(in_stride), pout, pw, _sh(stream),
                                       int(reps), C.byref(t)))
        return t.as_dict()

    def occupancy(self, stage: int = 0) -> dict:
        """Theoretical occupancy of the dedispersion launch of `stage`."""
        o = Occupancy()
        check(lib().mdsm_plan_occupancy(self._h, int(stage), C.byref(o)))
        return {f: getattr(o, f) for f, _ in Occupancy._fields_}

    def lds_bytes(self, stage: int = 0) -> tuple:
        """(read, write) LDS bytes of one launch of dedispersion `stage`: the
        trial reads and the staging stores."""
        r, w = C.c_uint64(), C.c_uint64()
        check(lib().mdsm_plan_lds_bytes(self._h, int(stage), C.byref(r), C.byref(w)))
        return r.value, w.value

    def split_output(self, flat: np.ndarray) -> list:
        return [flat[R.out_offset:R.out_offset + R.ndms * R.nout].reshape(R.ndms, R.nout)
                for R in self.ranges]

    def dedisperse_host(self, x: np.ndarray) -> list:
        """Upload a channel-major block, dedisperse on the GPU, download."""
        x = np.ascontiguousarray(x, NP_DTYPE[self.obs.dtype])
        assert x.shape[0] == self.obs.nchans and x.shape[1] >= self.obs.nsamp_in
        din = DeviceBuffer.from_host(x)
        dout = DeviceBuffer(self.info.out_elems * 4)
        try:
            self.dedisperse(din, x.shape[1], dout)
            synchronize()
            flat = dout.download(np.float32, (self.info.out_elems,))
        finally:
            din.close()
            dout.close()
        return self.split_output(flat)

    def close(self):
        if self._h:
            lib().mdsm_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- clipper / detectors
def clip_u8(d_in, in_stride, nchans, t0, n, nsigma, d_out, out_stride, d_med=None, d_mad=None,
            d_ncl=None, stream=None):
    g = lambda b: b.ptr if isinstance(b, DeviceBuffer) else b
    check(lib().mdsm_clip_u8(g(d_in), in_stride, nchans, t0, n, nsigma, g(d_out), out_stride,
                             g(d_med), g(d_mad), g(d_ncl), _sh(stream)))


class Batch:
    """Up to 64 plans run as one step (one dedispersion launch per stage)."""

    def __init__(self, plans: Sequence["Plan"]):
        self.plans = list(plans)
        arr = (C.c_void_p * len(self.plans))(*[p.handle for p in self.plans])
        h = C.c_void_p()
        check(lib().mdsm_batch_create(arr, len(self.plans), C.byref(h)))
        self._h = h.value

    def _bufs(self, ins, strides, outs, works):
        g = lambda b: (b.ptr if isinstance(b, DeviceBuffer) else b) or None
        n = len(self.plans)
        a_in = (C.c_void_p * n)(*[g(b) for b in ins])
        a_st = (C.c_int64 * n)(*[int(x) for x in strides])
        a_out = (C.c_void_p * n)(*[g(b) for b in outs])
        a_w = (C.c_void_p * n)(*[g(b) for b in works]) if works is not None else None
        return a_in, a_st, a_out, a_w

    def run(self, ins, strides, outs, works=None, stream=None):
        a = self._bufs(ins, strides, outs, works)
        check(lib().mdsm_batch_run(self._h, *a, _sh(stream)))

    def capture(self, ins, strides, outs, works=None, stream=None):
        a = self._bufs(ins, strides, outs, works)
        check(lib().mdsm_batch_capture(self._h, *a, _sh(stream)))

    def replay(self, stream=None):
        check(lib().mdsm_batch_replay(self._h, _sh(stream)))

    def autotune(self, ins, strides, outs, works=None, stream=None, reps: int = 2) -> dict:
        a = self._bufs(ins, strides, outs, works)
        t = Tuning()
        check(lib().mdsm_batch_autotune(self._h, *a, _sh(stream), int(reps), C.byref(t)))
        return t.as_dict()

    def close(self):
        if self._h:
            lib().mdsm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- DM plans
def ddplan(lodm, hidm, f_ctr, bw, nchans, dt, nsubs=0, ok_smearing=0.0) -> list:
    """Passes (DM range, step, downsampling; subband step and calls) from the
    smearing budget."""
    prm = DDPlanParams(lodm, hidm, f_ctr, bw, dt, nchans, nsubs, ok_smearing)
    n = C.c_uint32()
    buf = (Pass * 64)()
    check(lib().mdsm_ddplan(C.byref(prm), buf, 64, C.byref(n)))
    return [dict(lowdm=p.lowdm, highdm=p.highdm, dmstep=p.dmstep, sub_dmstep=p.sub_dmstep,
                 binsize=p.binsize, ndms=p.ndms, ncalls=p.ncalls, calldms=p.calldms)
            for p in buf[:n.value]]


# ---------------------------------------------------------------- SIGPROC files
def sigproc_read_header(path: str) -> dict:
    h = SigprocHeader()
    check(lib().mdsm_sigproc_read_header(path.encode(), C.byref(h)))
    return h.as_dict()


class SigprocReader:
    """Sequential reader of raw packed spectra (time-major)."""

    def __init__(self, path: str):
        h = SigprocHeader()
        r = C.c_void_p()
        check(lib().mdsm_sigproc_open(path.encode(), C.byref(r), C.byref(h)))
        self._h = r.value
        self.header = h.as_dict()
        nifs = max(self.header["nifs"], 1)
        self.spectrum_bytes = self.header["nchans"] * nifs * self.header["nbits"] // 8

    def read(self, nspectra: int) -> np.ndarray:
        buf = np.empty(nspectra * self.spectrum_bytes, np.uint8)
        got = C.c_uint64()
        check(lib().mdsm_sigproc_read(self._h, nspectra, buf.ctypes.data, C.byref(got)))
        return buf[:got.value * self.spectrum_bytes].reshape(got.value, self.spectrum_bytes)

    def seek(self, spectrum: int):
        check(lib().mdsm_sigproc_seek(self._h, spectrum))

    def close(self):
        if self._h:
            lib().mdsm_sigproc_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def unpack_spectra(d_raw, nbits, nsamp, nchans, in_stride_bytes, d_out, out_stride, out_col0=0,
                   invert=False, dtype_out="f32", stream=None):
    g = lambda b: b.ptr if isinstance(b, DeviceBuffer) else b
    do = DTYPES[dtype_out] if isinstance(dtype_out, str) else int(dtype_out)
    check(lib().mdsm_unpack_spectra(g(d_raw), int(nbits), nsamp, nchans, in_stride_bytes,
                                    1 if invert else 0, g(d_out), do, out_stride, out_col0,
                                    _sh(stream)))


class SeriesWriter:
    """One float32 SIGPROC file <prefix>_<dm>.fil per selected DM value."""

    def __init__(self, prefix: str, dm_values, fch1, foff, tsamp, tstart=0.0):
        dms = np.ascontiguousarray(dm_values, np.float32)
        h = C.c_void_p()
        check(lib().mdsm_series_writer_open(prefix.encode(), dms.ctypes.data, dms.size, fch1,
                                            foff, tsamp, tstart, C.byref(h)))
        self._h = h.value

    def write(self, rows: np.ndarray, row_dm_values):
        rows = np.ascontiguousarray(rows, np.float32)
        dmv = np.ascontiguousarray(row_dm_values, np.float32)
        check(lib().mdsm_series_writer_write(self._h, rows.ctypes.data, dmv.ctypes.data,
                                             rows.shape[0], rows.shape[1], rows.shape[1]))

    def close(self):
        if self._h:
            lib().mdsm_series_writer_close(self._h)
            self._h = None


def corner_turn(d_in, dtype_in, nsamp, nchans, in_stride, d_out, out_stride, out_col0=0,
                invert=False, dtype_out=None, stream=None):
    """Time-major [nsamp][in_stride] -> channel-major rows out[c][out_col0 + t]."""
    g = lambda b: b.ptr if isinstance(b, DeviceBuffer) else b
    di = DTYPES[dtype_in] if isinstance(dtype_in, str) else int(dtype_in)
    do = di if dtype_out is None else (DTYPES[dtype_out] if isinstance(dtype_out, str)
                                       else int(dtype_out))
    check(lib().mdsm_corner_turn(g(d_in), di, nsamp, nchans, in_stride, 1 if invert else 0,
                                 g(d_out), do, out_stride, out_col0, _sh(stream)))


def detrend(d_x, nrows, n, stride, seg_len, stream=None):
    """In-place per-segment detrend and normalise of float rows."""
    g = lambda b: b.ptr if isinstance(b, DeviceBuffer) else b
    check(lib().mdsm_detrend(g(d_x), nrows, n, stride, seg_len, _sh(stream)))


def detect_boxcar(d_x, ndms, n, stride, widths, thr, d_cands, cap, d_count, d_mean, d_sigma,
                  stream=None):
    g = lambda b: b.ptr if isinstance(b, DeviceBuffer) else b
    w = (C.c_uint32 * len(widths))(*widths)
    check(lib().mdsm_detect_boxcar(g(d_x), ndms, n, stride, w, len(widths), thr, g(d_cands), cap,
                                   g(d_count), g(d_mean), g(d_sigma), _sh(stream)))


def probe_lds_read(bytes_per_lane: int, waves_per_simd: int) -> float:
    """Measured LDS read bandwidth (TB/s) of the current GPU."""
    v = C.c_double()
    check(lib().mdsm_probe_lds_read(int(bytes_per_lane), int(waves_per_simd), C.byref(v)))
    return v.value


def probe_lds_read_form(form: int, waves_per_simd: int) -> float:
    """Measured LDS read bandwidth (TB/s) of read form 1..4 (ds_read2st64_b32,
    ds_read2st64_b64, ds_read_b64 at a 4-byte-aligned address, ds_read2_b32
    of adjacent dwords)."""
    v = C.c_double()
    check(lib().mdsm_probe_lds_read_form(int(form), int(waves_per_simd), C.byref(v)))
    return v.value


def probe_lds_write(bytes_per_lane: int, waves_per_simd: int) -> float:
    """Measured LDS write bandwidth (TB/s) of the current GPU."""
    v = C.c_double()
    check(lib().mdsm_probe_lds_write(int(bytes_per_lane), int(waves_per_simd), C.byref(v)))
    return v.value


def boxcar_carry_len(widths) -> int:
    """Carried samples per DM row of a streaming search with these widths."""
    mask = 0
    for w in widths:
        mask |= int(w)
    return int(lib().mdsm_boxcar_carry_len(mask))


def detect_boxcar_stream(d_x, ndms, n, stride, d_carry, carry_stride, carry_valid, widths, thr,
                         d_cands, cap, d_count, d_mean, d_sigma, stream=None):
    """Streaming boxcar search of one block (mdsm_detect_boxcar_stream): the
    carry [ndms][carry_stride] holds the previous block's last samples and is
    updated with this block's."""
    g = lambda b: b.ptr if isinstance(b, DeviceBuffer) else b
    w = (C.c_uint32 * len(widths))(*widths)
    check(lib().mdsm_detect_boxcar_stream(g(d_x), ndms, n, stride, g(d_carry), carry_stride,
                                          1 if carry_valid else 0, w, len(widths), thr,
                                          g(d_cands), cap, g(d_count), g(d_mean), g(d_sigma),
                                          _sh(stream)))


def median5_stats_len(ndms: int, n: int) -> int:
    """Doubles of device scratch detect_median5 needs for ndms rows of n samples."""
    return int(lib().mdsm_median5_stats_len(int(ndms), int(n)))


def detect_median5(d_x, ndms, n, stride, d_dmv, thr, d_cands, cap, d_count, d_stats,
                   per_row=False, d_chunk_mean=None, d_chunk_rms=None, samples_per_chunk=0,
                   stream=None):
    g = lambda b: b.ptr if isinstance(b, DeviceBuffer) else b
    check(lib().mdsm_detect_median5(g(d_x), ndms, n, stride, g(d_dmv), thr, 1 if per_row else 0,
                                    g(d_chunk_mean), g(d_chunk_rms), samples_per_chunk,
                                    g(d_cands), cap, g(d_count), g(d_stats), _sh(stream)))


def read_candidates(d_cands, d_count, cap) -> np.ndarray:
    """Download and sort (dm, width, t) a candidate list written by a detector."""
    cnt = int(d_count.download(np.uint32, (1,))[0])
    m = min(cnt, cap)
    arr = d_cands.download(CAND_DTYPE, (m,)) if m else np.zeros(0, CAND_DTYPE)
    arr = np.sort(arr, order=["dm", "width", "t"])
    if cnt > cap:
        raise MdsmError(-4, f"candidate overflow: {cnt} found, capacity {cap}")
    return arr


# ---------------------------------------------------------------- synthetic data
def synth_filterbank_device(d_out, nchans, nsamp, dtype="u8", seed=1, fch1=1500.0,
                            foff=-0.390625, tsamp=64e-6, mean=0.0, sigma=1.0, gain=(1.0, 1.0),
                            pulses=(), stride=None, stream=None):
    """synth_filterbank into the device buffer d_out ([nchans][stride])."""
    dt = DTYPES[dtype] if isinstance(dtype, str) else int(dtype)
    stride = nsamp if stride is None else stride
    P = (Pulse * max(1, len(pulses)))()
    for i, (dm, t0, w, a) in enumerate(pulses):
        P[i] = Pulse(float(dm), int(t0), int(w), float(a), 0)
    spec = SynthSpec(int(seed), dt, nchans, nsamp, fch1, foff, tsamp, mean, sigma, gain[0],
                     gain[1], len(pulses), P, 0)
    ptr = d_out.ptr if isinstance(d_out, DeviceBuffer) else d_out
    check(lib().mdsm_synth_filterbank_device(C.byref(spec), ptr, stride, _sh(stream)))


def synth_filterbank(nchans, nsamp, dtype="u8", seed=1, fch1=1500.0, foff=-0.390625,
                     tsamp=64e-6, mean=0.0, sigma=1.0, gain=(1.0, 1.0), pulses=(),
                     stride=None, nthreads=0) -> np.ndarray:
    """Seeded channel-major filterbank [nchans][stride] (host).

    pulses: sequence of (dm, t0, width, amp); amp in units of sigma.  Pulses
    follow the nearest-sample FP64 delay track (shift rule NS)."""
    dt = DTYPES[dtype] if isinstance(dtype, str) else int(dtype)
    stride = nsamp if stride is None else stride
    out = np.empty((nchans, stride), NP_DTYPE[dt])
    P = (Pulse * max(1, len(pulses)))()
    for i, (dm, t0, w, a) in enumerate(pulses):
        P[i] = Pulse(float(dm), int(t0), int(w), float(a), 0)
    spec = SynthSpec(int(seed), dt, nchans, nsamp, fch1, foff, tsamp, mean, sigma, gain[0],
                     gain[1], len(pulses), P, nthreads)
    check(lib().mdsm_synth_filterbank(C.byref(spec), out.ctypes.data, stride))
    return out
