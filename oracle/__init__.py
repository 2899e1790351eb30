"""ctypes bindings for the oracle (CPU restatement of MDSM's dedispersion path).

TEST INFRASTRUCTURE ONLY: the parity checker.  Only tests/, the smoke()
entry point and bench.py's cpu_baseline leg import this package; the
product path (mdsm_amd) never does.

``lib()`` loads oracle/liboracle.so (built by ``make -C oracle``).
``ref()`` loads oracle/_ref/libmdsm_ref.so, the reference's own
cpu_dedisp.c / dedispersion_output.cpp / file_handler.cpp compiled by
``make -C oracle ref``; it returns None when that build is absent.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
U8, F16, F32 = 0, 1, 2
_DT = {np.dtype(np.uint8): U8, np.dtype(np.float16): F16, np.dtype(np.float32): F32}

_lib = None
_ref = None


class Cand(C.Structure):
    _fields_ = [("dm", C.c_uint32), ("t", C.c_int32), ("width", C.c_uint32),
                ("snr", C.c_float)]


class SCand(C.Structure):
    _fields_ = [("dm", C.c_uint32), ("t", C.c_int32), ("width", C.c_uint32),
                ("snr", C.c_float), ("blk", C.c_uint32)]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so missing: run make -C oracle")
        L = C.CDLL(path)
        vp, i32, i64, f32, f64 = C.c_void_p, C.c_int, C.c_int64, C.c_float, C.c_double
        L.orc_dmdelay.restype = f32
        L.orc_dmdelay.argtypes = [f32, f32]
        L.orc_delay_table.argtypes = [f32, f32, i32, vp]
        L.orc_shift_table_ref.argtypes = [vp, i32, f32, f32, i32, f32, vp]
        L.orc_shift_table_ns.argtypes = [f64, f64, i32, f64, f64, i32, f64, vp]
        L.orc_half_to_float.restype = f32
        L.orc_half_to_float.argtypes = [C.c_uint16]
        L.orc_dedisperse_ref_tm.argtypes = [vp, i32, i32, i32, vp, vp]
        L.orc_dedisperse_ref_tm_mt.argtypes = [vp, i32, i32, i32, vp, vp, i32]
        L.orc_dedisperse_f32.argtypes = [vp, i32, i32, i64, i32, i32, vp, vp, i64]
        L.orc_dedisperse_f64.argtypes = [vp, i32, i32, i64, i32, i32, vp, vp, i64]
        L.orc_decimate_f64.argtypes = [vp, i32, i32, i64, i32, i32, vp]
        L.orc_range_f32.argtypes = [vp, i32, i32, i64, i32, i32, i32, i32, vp, vp, i64]
        L.orc_range_f64.argtypes = [vp, i32, i32, i64, i32, i32, i32, i32, vp, vp, i64]
        L.orc_subband_tables.argtypes = [vp, i32, i32, f32, f32, f32, i32, i32, f32, vp, vp]
        L.orc_subband_f64.argtypes = [vp, i32, i32, i64, i32, i32, i32, i32, i32, vp, vp,
                                      i32, i32, vp, i64]
        L.orc_clip_u8.argtypes = [vp, i32, i64, i64, i32, f32, vp, i64, vp, vp, vp]
        L.orc_row_stats.argtypes = [vp, i32, C.POINTER(f64), C.POINTER(f64)]
        L.orc_boxcar_detect.restype = i32
        L.orc_boxcar_detect.argtypes = [vp, i32, i32, i64, vp, i32, f64, vp, i32, vp, vp]
        L.orc_boxcar_stream.restype = i32
        L.orc_boxcar_stream.argtypes = [vp, i32, i32, i32, i64, vp, i32, f64, vp, i32]
        L.orc_median_of_five.restype = f32
        L.orc_median_of_five.argtypes = [f32] * 5
        L.orc_brute_stats.restype = None
        L.orc_brute_stats.argtypes = [vp, i32, i32, i64, C.POINTER(f32), C.POINTER(f32)]
        L.orc_corner_turn_f32.restype = None
        L.orc_corner_turn_f32.argtypes = [vp, i32, i32, i64, i32, vp, i64]
        L.orc_unpack_bits.restype = None
        L.orc_unpack_bits.argtypes = [vp, i32, i64, vp]
        L.orc_detrend_f64.restype = None
        L.orc_detrend_f64.argtypes = [vp, i32, i32, i64, i32, vp]
        _lib = L
    return _lib


# ---------------------------------------------------------------- shifts
def delay_table(fch1, foff, nchans):
    d = np.empty(nchans, np.float32)
    lib().orc_delay_table(fch1, foff, nchans, _p(d))
    return d


def shift_table_ref(fch1, foff, nchans, startdm, dmstep, ndms, tsamp):
    d = delay_table(fch1, foff, nchans)
    s = np.empty((ndms, nchans), np.int32)
    lib().orc_shift_table_ref(_p(d), nchans, startdm, dmstep, ndms, tsamp, _p(s))
    return s


def shift_table_ns(fch1, foff, nchans, startdm, dmstep, ndms, tsamp):
    s = np.empty((ndms, nchans), np.int32)
    lib().orc_shift_table_ns(fch1, foff, nchans, startdm, dmstep, ndms, tsamp, _p(s))
    return s


# ---------------------------------------------------------------- brute force
def _cm(x):
    x = np.ascontiguousarray(x)
    return x, _DT[x.dtype], x.shape[0], x.shape[1]


def dedisperse_ref_tm(x_tm, shift, nsamp):
    """cpu_dedisp.c loop on time-major input [S][C]; returns time-major [nsamp][D]."""
    x_tm = np.ascontiguousarray(x_tm, np.float32)
    shift = np.ascontiguousarray(shift, np.int32)
    ndms, nchans = shift.shape
    out = np.empty((nsamp, ndms), np.float32)
    lib().orc_dedisperse_ref_tm(_p(x_tm), nchans, nsamp, ndms, _p(shift), _p(out))
    return out


def dedisperse_ref_tm_mt(x_tm, shift, nsamp, nthreads):
    """dedisperse_ref_tm on nthreads OpenMP threads (same float sums)."""
    x_tm = np.ascontiguousarray(x_tm, np.float32)
    shift = np.ascontiguousarray(shift, np.int32)
    ndms, nchans = shift.shape
    out = np.empty((nsamp, ndms), np.float32)
    lib().orc_dedisperse_ref_tm_mt(_p(x_tm), nchans, nsamp, ndms, _p(shift), _p(out), nthreads)
    return out


def dedisperse_f32(x, shift, nout):
    x, dt, nchans, stride = _cm(x)
    shift = np.ascontiguousarray(shift, np.int32)
    ndms = shift.shape[0]
    out = np.empty((ndms, nout), np.float32)
    lib().orc_dedisperse_f32(_p(x), dt, nchans, stride, nout, ndms, _p(shift), _p(out), nout)
    return out


def dedisperse_f64(x, shift, nout):
    x, dt, nchans, stride = _cm(x)
    shift = np.ascontiguousarray(shift, np.int32)
    ndms = shift.shape[0]
    out = np.empty((ndms, nout), np.float64)
    lib().orc_dedisperse_f64(_p(x), dt, nchans, stride, nout, ndms, _p(shift), _p(out), nout)
    return out


def decimate_f64(x, k):
    x, dt, nchans, stride = _cm(x)
    out = np.empty((nchans, stride // k), np.float64)
    lib().orc_decimate_f64(_p(x), dt, nchans, stride, stride, k, _p(out))
    return out


def range_f32(x, k, shift, nout):
    x, dt, nchans, stride = _cm(x)
    shift = np.ascontiguousarray(shift, np.int32)
    ndms = shift.shape[0]
    out = np.empty((ndms, nout), np.float32)
    lib().orc_range_f32(_p(x), dt, nchans, stride, stride, k, nout, ndms, _p(shift), _p(out), nout)
    return out


def range_f64(x, k, shift, nout):
    x, dt, nchans, stride = _cm(x)
    shift = np.ascontiguousarray(shift, np.int32)
    ndms = shift.shape[0]
    out = np.empty((ndms, nout), np.float64)
    lib().orc_range_f64(_p(x), dt, nchans, stride, stride, k, nout, ndms, _p(shift), _p(out), nout)
    return out


# ---------------------------------------------------------------- subband
def subband_tables(fch1, foff, nchans, nsubs, lowdm, sub_dmstep, dmstep, ncalls,
                   calldms, tsamp_k):
    d = delay_table(fch1, foff, nchans)
    coarse = np.empty((ncalls, nchans), np.int32)
    fine = np.empty((ncalls * calldms, nsubs), np.int32)
    lib().orc_subband_tables(_p(d), nchans, nsubs, lowdm, sub_dmstep, dmstep, ncalls,
                             calldms, tsamp_k, _p(coarse), _p(fine))
    return coarse, fine


def subband_f64(x, k, nsubs, ncalls, calldms, coarse, fine, len1, nout):
    x, dt, nchans, stride = _cm(x)
    coarse = np.ascontiguousarray(coarse, np.int32)
    fine = np.ascontiguousarray(fine, np.int32)
    out = np.empty((ncalls * calldms, nout), np.float64)
    lib().orc_subband_f64(_p(x), dt, nchans, stride, stride, k, nsubs, ncalls, calldms,
                          _p(coarse), _p(fine), len1, nout, _p(out), nout)
    return out


# ---------------------------------------------------------------- clipper
def clip_u8(x, t0, n, nsigma):
    x = np.ascontiguousarray(x, np.uint8)
    nchans, stride = x.shape
    out = np.empty((nchans, n), np.uint8)
    med = np.empty(nchans, np.int32)
    mad = np.empty(nchans, np.int32)
    ncl = np.empty(nchans, np.int32)
    lib().orc_clip_u8(_p(x), nchans, stride, t0, n, nsigma, _p(out), n, _p(med), _p(mad),
                      _p(ncl))
    return out, med, mad, ncl


# ---------------------------------------------------------------- detector
def row_stats(row):
    row = np.ascontiguousarray(row, np.float32)
    m, s = C.c_double(), C.c_double()
    lib().orc_row_stats(_p(row), row.size, C.byref(m), C.byref(s))
    return m.value, s.value


def boxcar_detect(x, widths, thr, cap=1 << 20):
    x = np.ascontiguousarray(x, np.float32)
    ndms, n = x.shape
    w = np.ascontiguousarray(widths, np.int32)
    while True:
        cands = (Cand * cap)()
        means = np.empty(ndms, np.float64)
        sig = np.empty(ndms, np.float64)
        cnt = lib().orc_boxcar_detect(_p(x), ndms, n, n, _p(w), w.size, thr,
                                      C.cast(cands, C.c_void_p), cap, _p(means), _p(sig))
        if cnt <= cap:
            break
        cap = cnt
    arr = np.zeros(cnt, dtype=[("dm", np.uint32), ("t", np.int32), ("width", np.uint32),
                               ("snr", np.float32)])
    if cnt:
        raw = np.ctypeslib.as_array(cands)[:cnt]
        for f in ("dm", "t", "width", "snr"):
            arr[f] = raw[f]
    return arr, means, sig


def boxcar_stream(x, n, widths, thr, cap=1 << 20):
    """Streaming boxcar search over rows x [ndms][nblk*n] taken as nblk
    consecutive blocks of n samples (orc_boxcar_stream): per-block
    statistics, each window searched once by the block its end falls in.
    Returns candidates with fields dm, t (relative to the owning block's
    first sample), width, snr, blk."""
    x = np.ascontiguousarray(x, np.float32)
    ndms, T = x.shape
    if T % n:
        raise ValueError("row length must be a whole number of blocks")
    w = np.ascontiguousarray(widths, np.int32)
    while True:
        cands = (SCand * cap)()
        cnt = lib().orc_boxcar_stream(_p(x), ndms, n, T // n, T, _p(w), w.size, thr,
                                      C.cast(cands, C.c_void_p), cap)
        if cnt <= cap:
            break
        cap = cnt
    arr = np.zeros(cnt, dtype=[("dm", np.uint32), ("t", np.int32), ("width", np.uint32),
                               ("snr", np.float32), ("blk", np.uint32)])
    if cnt:
        raw = np.ctypeslib.as_array(cands)[:cnt]
        for f in ("dm", "t", "width", "snr", "blk"):
            arr[f] = raw[f]
    return arr


def median_of_five(a, b, c, d, e):
    return lib().orc_median_of_five(a, b, c, d, e)


def brute_stats(x):
    """orc_brute_stats: the reference output thread's chunked whole-block
    mean and rms of rows x [ndms][n] (float32 values)."""
    x = np.ascontiguousarray(x, np.float32)
    m, r = C.c_float(), C.c_float()
    lib().orc_brute_stats(_p(x), x.shape[0], x.shape[1], x.shape[1], C.byref(m), C.byref(r))
    return np.float32(m.value), np.float32(r.value)


def process_brute_lines(x, lowdm, dmstep, thr, timestamp, block_rate):
    """Candidate lines of the reference's brute-force output thread without
    noise arrays (src/mdsm/dedispersion_output.cpp:143-195) on rows x
    [tdms][nsamp]: statistics from brute_stats, the median of each 5-sample
    window (l < nsamp - 4), a line "%lf, %f, %f" (time, DM, median/rms)
    where median - mean >= rms * thr and the row's DM > 1, all in float."""
    x = np.ascontiguousarray(x, np.float32)
    lm, lr = brute_stats(x)
    thr = np.float32(thr)
    out = []
    for k in range(x.shape[0]):
        dm = np.float32(np.float32(lowdm) + np.float32(0)) + np.float32(k) * np.float32(dmstep)
        if not dm > 1.0:
            continue
        w = np.lib.stride_tricks.sliding_window_view(x[k], 5)
        med = np.sort(w, axis=1)[:, 2]
        hit = np.nonzero((med - lm) >= lr * thr)[0]
        for l in hit:
            out.append("%f, %f, %f" % (timestamp + float(l) * block_rate, float(dm),
                                       float(np.float32(med[l] / lr))))
    return out


def corner_turn_f32(x_tm, invert=False):
    """Time-major [nsamp][nchans] -> channel-major [nchans][nsamp] float32."""
    x = np.ascontiguousarray(x_tm, np.float32)
    T, Cn = x.shape
    out = np.empty((Cn, T), np.float32)
    lib().orc_corner_turn_f32(_p(x), T, Cn, Cn, 1 if invert else 0, _p(out), T)
    return out


def unpack_bits(raw, nbits, nvals):
    raw = np.ascontiguousarray(np.frombuffer(bytes(raw), np.uint8))
    out = np.empty(int(nvals), np.float32)
    lib().orc_unpack_bits(_p(raw), int(nbits), int(nvals), _p(out))
    return out


def detrend_f64(x, seg_len):
    x = np.ascontiguousarray(x, np.float32)
    R, N = x.shape
    out = np.empty((R, N), np.float64)
    lib().orc_detrend_f64(_p(x), R, N, N, int(seg_len), _p(out))
    return out


# ---------------------------------------------------------------- DM plan (DDplan)
_ALLOW_DDM = [0.0001, 0.0002, 0.0005, 0.001, 0.002, 0.005, 0.01, 0.02, 0.03, 0.05, 0.1, 0.2, 0.3,
              0.5, 1.0, 2.0, 3.0, 5.0, 10.0, 20.0, 30.0, 50.0, 100.0, 200.0, 300.0, 500.0, 1000.0]


def _dm_smear(dm, bw, f):            # python/ddplan.py:126-132
    return 1000.0 * dm * bw / (0.0001205 * f ** 3.0)


def _bw_smear(step, bw, f):          # :134-141
    return _dm_smear(0.5 * step, bw, f)


def _sub_smear(substep, nsub, bw, f):  # :151-161
    return 0.0 if nsub == 0 else _dm_smear(0.5 * substep, bw / nsub, f)


def _method(dt, f, bw, chw, downsamp, lodm, hidm, ddm, nsub):  # dedisp_method, :23-63
    bws = _bw_smear(ddm, bw, f)
    per = 0
    if nsub:
        per = 2
        while True:
            if _sub_smear((per + 2) * ddm, nsub, bw, f) > 0.8 * min(bws, 1000.0 * dt * downsamp):
                dsub = per * ddm
                subs = _sub_smear(dsub, nsub, bw, f)
                break
            per += 2
    else:
        dsub, subs = ddm, 0.0
    other = math.sqrt((1000.0 * dt) ** 2.0 + (1000.0 * dt * downsamp) ** 2.0 + bws ** 2.0 +
                      subs ** 2.0)
    cross = min(2.0 * 0.001 * other / chw * 0.0001205 * f ** 3.0, hidm)   # :84-94, :51-52
    n = int(math.ceil((cross - lodm) / ddm))
    calls = 1
    if nsub:
        calls = int(math.ceil(n * ddm / dsub))
        n = calls * per
    return dict(lowdm=lodm, highdm=lodm + n * ddm, dmstep=ddm, sub_dmstep=dsub,
                binsize=downsamp, ndms=n, ncalls=calls, calldms=per if nsub else n)


def ddplan(lodm, hidm, f_ctr, bw, nchans, dt, nsubs=0, ok_smearing=0.0):
    """dm_steps of python/ddplan.py:163-253, in double precision."""
    chw = bw / nchans
    dtms = 1000.0 * dt
    ff = 1.2
    ids = idm = 0
    downsamp = 1
    min_chan = _dm_smear(lodm + 0.5 * _ALLOW_DDM[0], chw, f_ctr)
    min_bw = _bw_smear(_ALLOW_DDM[0], bw, f_ctr)
    ok = max(ok_smearing, min_chan, min_bw, dtms)
    if ff * min_chan > dtms or ok > dtms:
        okval = ok if ok > ff * min_chan else ff * min_chan
        while dtms * 2.0 ** (ids + 1) < okval:
            ids += 1
        downsamp = 2 ** ids
    ddm = dt * downsamp * 0.0001205 * f_ctr ** 3.0 / (0.5 * bw)      # guess_DMstep, :143-149
    while _ALLOW_DDM[idm + 1] < ff * ddm:
        idm += 1
    ms = [_method(dt, f_ctr, bw, chw, downsamp, lodm, hidm, _ALLOW_DDM[idm], nsubs)]
    while ms[-1]["highdm"] < hidm:
        ids += 1
        downsamp = 2 ** ids
        eff = dtms * downsamp
        while _bw_smear(_ALLOW_DDM[idm + 1], bw, f_ctr) < ff * eff:
            idm += 1
        ms.append(_method(dt, f_ctr, bw, chw, downsamp, ms[-1]["highdm"], hidm,
                          _ALLOW_DDM[idm], nsubs))
    return ms


# ---------------------------------------------------------------- reference SIGPROC reader
class RefFileHeader(C.Structure):
    """FILE_HEADER of src/mdsm/file_handler.h:6-28 (x86-64 layout)."""
    _fields_ = [("rawdatafile", C.c_char * 80), ("source_name", C.c_char * 80),
                ("machine_id", C.c_int), ("telescope_id", C.c_int), ("data_type", C.c_int),
                ("nchans", C.c_int), ("nbits", C.c_int), ("nifs", C.c_int),
                ("scan_number", C.c_int), ("barycentric", C.c_int), ("pulsarcentric", C.c_int),
                ("tstart", C.c_double), ("mjdobs", C.c_double), ("tsamp", C.c_double),
                ("fch1", C.c_double), ("foff", C.c_double), ("refdm", C.c_double),
                ("az_start", C.c_double), ("za_start", C.c_double), ("src_raj", C.c_double),
                ("src_dej", C.c_double), ("gal_l", C.c_double), ("gal_b", C.c_double),
                ("header_tobs", C.c_double), ("raw_fch1", C.c_double), ("raw_foff", C.c_double),
                ("nbeams", C.c_int), ("ibeam", C.c_int), ("srcl", C.c_double),
                ("srcb", C.c_double), ("ast0", C.c_double), ("lst0", C.c_double),
                ("wapp_scan_number", C.c_long), ("project", C.c_char * 8),
                ("culprits", C.c_char * 24), ("analog_power", C.c_double * 2),
                ("frequency_table", C.c_double * 4096), ("npuls", C.c_long),
                ("period", C.c_double), ("nbins", C.c_int), ("itmp", C.c_int),
                ("total_bytes", C.c_int)]


def _libc():
    c = C.CDLL(None)
    c.fopen.restype = C.c_void_p
    c.fopen.argtypes = [C.c_char_p, C.c_char_p]
    c.fclose.argtypes = [C.c_void_p]
    c.fseek.argtypes = [C.c_void_p, C.c_long, C.c_int]
    return c


def ref_read_header(path):
    """The reference's read_header on a file: (FILE_HEADER copy, total_bytes) or None
    when oracle/_ref is not built."""
    R = ref()
    if R is None:
        return None
    fn = getattr(R, "_Z11read_headerP8_IO_FILE")
    fn.restype = C.POINTER(RefFileHeader)
    fn.argtypes = [C.c_void_p]
    libc = _libc()
    f = libc.fopen(path.encode(), b"rb")
    try:
        h = fn(f)
        return RefFileHeader.from_buffer_copy(h.contents) if h else None
    finally:
        libc.fclose(f)


def ref_read_block(path, offset, nbits, nvals):
    """The reference's read_block: nvals samples from byte offset."""
    R = ref()
    if R is None:
        return None
    fn = getattr(R, "_Z10read_blockP8_IO_FILEiPfm")
    fn.restype = C.c_ulong
    fn.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_ulong]
    libc = _libc()
    f = libc.fopen(path.encode(), b"rb")
    try:
        libc.fseek(f, int(offset), 0)
        out = np.zeros(int(nvals), np.float32)
        n = fn(f, int(nbits), out.ctypes.data, int(nvals))
        return out[:n]
    finally:
        libc.fclose(f)


# ---------------------------------------------------------------- reference build
def ref():
    """The reference's own code (oracle/_ref/libmdsm_ref.so) or None."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libmdsm_ref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.dmdelay.restype = C.c_float
        R.dmdelay.argtypes = [C.c_float, C.c_float]
        R.mdsm_ref_cpu_dedisp_main.restype = C.c_int
        R.mdsm_ref_cpu_dedisp_main.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
        R.mdsm_ref_hooks_reset.restype = None
        R.mdsm_ref_hooks_set_input.argtypes = [C.c_void_p, C.c_size_t]
        R.mdsm_ref_hooks_count.restype = C.c_int
        R.mdsm_ref_hooks_ptr.restype = C.c_void_p
        R.mdsm_ref_hooks_ptr.argtypes = [C.c_int]
        R.mdsm_ref_hooks_size.restype = C.c_size_t
        R.mdsm_ref_hooks_size.argtypes = [C.c_int]
        med5 = getattr(R, "_Z12medianOfFivefffff")
        med5.restype = C.c_float
        med5.argtypes = [C.c_float] * 5
        R.median_of_five = med5
        if hasattr(R, "mdsm_ref_process_brute"):
            R.mdsm_ref_process_brute.restype = C.c_int
            R.mdsm_ref_process_brute.argtypes = [C.c_void_p, C.c_uint, C.c_uint, C.c_int,
                                                 C.c_float, C.c_float, C.c_float, C.c_double,
                                                 C.c_double, C.c_char_p]
        _ref = R
    return _ref


def ref_process_brute(buf, read_nsamp, lowdm, dmstep, thr, timestamp, block_rate, path):
    """The reference's own process_brute (oracle/_ref) on a [tdms][nsamp]
    float buffer, one thread, no noise arrays; its candidate lines go to
    `path`.  Returns False when the reference build is absent."""
    R = ref()
    if R is None or not hasattr(R, "mdsm_ref_process_brute"):
        return False
    b = np.ascontiguousarray(buf, np.float32).copy()
    tdms, nsamp = b.shape
    rc = R.mdsm_ref_process_brute(_p(b), nsamp, tdms, int(read_nsamp), lowdm, dmstep, thr,
                                  timestamp, block_rate, str(path).encode())
    if rc != 0:
        raise RuntimeError("reference process_brute could not open its output file")
    return True


def ref_cpu_dedisp(fch1, foff, tsamp, nchans, nsamp, startdm, dmstep, tdms,
                   data_tm=None):
    """Run the reference's cpu_dedisp.c main() in-process.

    The reference hard-codes fch1/foff/tsamp as globals and then rescales
    foff /= nchans, tsamp *= nchans (src/prototypes/cpu_dedisp.c:9,36-37);
    the globals are pre-set so that the rescaled values equal the requested
    ones exactly (the factor nchans must be a power of two).  When data_tm
    (time-major float32 [nsamp+maxshift][nchans]) is given it replaces the
    reference's constant 0.1 input.  Returns (shift[tdms][nchans],
    output_tm[nsamp][tdms], input_tm).
    """
    R = ref()
    if R is None:
        raise RuntimeError("oracle/_ref not built")
    assert nchans & (nchans - 1) == 0, "nchans must be a power of two"
    g = lambda name, ct: ct.in_dll(R, name)
    g("fch1", C.c_float).value = fch1
    g("foff", C.c_float).value = np.float32(foff) * np.float32(nchans)
    g("tsamp", C.c_float).value = np.float32(tsamp) / np.float32(nchans)
    R.mdsm_ref_hooks_reset()
    if data_tm is not None:
        data_tm = np.ascontiguousarray(data_tm, np.float32)
        R.mdsm_ref_hooks_set_input(_p(data_tm), data_tm.nbytes)
    args = ["cpu", "-nchans", str(nchans), "-nsamp", str(nsamp), "-startdm", repr(float(startdm)),
            "-dmstep", repr(float(dmstep)), "-tdms", str(tdms)]
    argv = (C.c_char_p * len(args))(*[a.encode() for a in args])
    R.mdsm_ref_cpu_dedisp_main(len(args), argv)
    assert np.float32(g("foff", C.c_float).value) == np.float32(foff)
    assert np.float32(g("tsamp", C.c_float).value) == np.float32(tsamp)
    n = R.mdsm_ref_hooks_count()
    assert n == tdms + 3, n
    shift = np.empty((tdms, nchans), np.int32)
    for i in range(tdms):
        p = R.mdsm_ref_hooks_ptr(1 + i)
        shift[i] = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_int32)), (nchans,))
    maxshift = int(shift[-1, -1])
    pin = R.mdsm_ref_hooks_ptr(tdms + 1)
    pout = R.mdsm_ref_hooks_ptr(tdms + 2)
    inp = np.ctypeslib.as_array(C.cast(pin, C.POINTER(C.c_float)),
                                ((nsamp + maxshift), nchans)).copy()
    out = np.ctypeslib.as_array(C.cast(pout, C.POINTER(C.c_float)), (nsamp, tdms)).copy()
    R.mdsm_ref_hooks_reset()
    return shift, out, inp
