"""Python view of the drop-in manager API (include/mdsm_compat.h).

Mirrors src/mdsm/dedispersion_manager.h: processSurveyParameters,
initialiseMDSM, next_chunk, start_processing, tearDownMDSM, with SURVEY and
SUBBAND_PASSES laid out field for field as src/mdsm/survey.h.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import check, lib


class SUBBAND_PASSES(C.Structure):
    _fields_ = [("lowdm", C.c_float), ("highdm", C.c_float), ("dmstep", C.c_float),
                ("sub_dmstep", C.c_float), ("binsize", C.c_int), ("ndms", C.c_int),
                ("ncalls", C.c_int), ("calldms", C.c_int), ("mean", C.c_int),
                ("stddev", C.c_int)]


class SURVEY(C.Structure):
    _fields_ = [("nsamp", C.c_uint), ("nchans", C.c_uint), ("tdms", C.c_uint),
                ("maxshift", C.c_uint), ("nbits", C.c_uint), ("tsamp", C.c_float),
                ("foff", C.c_float), ("fch1", C.c_float), ("useBruteForce", C.c_bool),
                ("useL1Cache", C.c_bool), ("lowdm", C.c_float), ("dmstep", C.c_float),
                ("pass_parameters", C.POINTER(SUBBAND_PASSES)), ("num_passes", C.c_uint),
                ("nsubs", C.c_uint), ("timestamp", C.c_double), ("blockRate", C.c_double),
                ("noiseMean", C.POINTER(C.c_float)), ("noiseRMS", C.POINTER(C.c_float)),
                ("samplesPerChunk", C.c_uint), ("fp", C.c_void_p),
                ("fileprefix", C.c_char * 80), ("basedir", C.c_char * 120),
                ("secs_per_file", C.c_uint), ("use_pc_time", C.c_char),
                ("single_file_mode", C.c_char), ("detection_threshold", C.c_float),
                ("num_threads", C.c_uint), ("gpu_ids", C.POINTER(C.c_uint)),
                ("num_gpus", C.c_uint)]


_ready = False


def _sig():
    global _ready
    if _ready:
        return lib()
    L = lib()
    L.processSurveyParameters.restype = C.POINTER(SURVEY)
    L.processSurveyParameters.argtypes = [C.c_char_p]
    L.mdsm_parse_survey_xml.restype = C.POINTER(SURVEY)
    L.mdsm_parse_survey_xml.argtypes = [C.c_char_p, C.c_size_t]
    L.mdsm_free_survey.restype = None
    L.mdsm_free_survey.argtypes = [C.POINTER(SURVEY)]
    L.initialiseMDSM.restype = C.POINTER(C.c_float)
    L.initialiseMDSM.argtypes = [C.POINTER(SURVEY)]
    L.next_chunk.restype = C.POINTER(C.c_float)
    L.next_chunk.argtypes = [C.c_uint, C.POINTER(C.c_uint), C.c_double, C.c_double,
                             C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.start_processing.restype = C.c_int
    L.start_processing.argtypes = [C.c_uint]
    L.tearDownMDSM.restype = None
    L.tearDownMDSM.argtypes = []
    L.mdsm_compat_device_output.restype = C.c_void_p
    _ready = True
    return L


def _err():
    return lib().mdsm_last_error().decode(errors="replace")


def parse_survey_xml(xml: str):
    b = xml.encode()
    p = _sig().mdsm_parse_survey_xml(b, len(b))
    if not p:
        raise ValueError(_err())
    return p


def processSurveyParameters(path: str):
    p = _sig().processSurveyParameters(path.encode())
    if not p:
        raise ValueError(_err())
    return p


def free_survey(p):
    _sig().mdsm_free_survey(p)


def passes(p):
    s = p.contents
    return [s.pass_parameters[i] for i in range(s.num_passes)]


def initialiseMDSM(p):
    buf = _sig().initialiseMDSM(p)
    if not buf:
        raise RuntimeError(_err())
    s = p.contents
    return np.ctypeslib.as_array(buf, (s.nchans, s.nsamp + s.maxshift))


def next_chunk(data_read, timestamp=0.0, block_rate=0.0, mean=None, rms=None):
    """Returns (output or None, samples)."""
    n = C.c_uint(0)
    fp = C.POINTER(C.c_float)
    m = mean.ctypes.data_as(fp) if mean is not None else None
    r = rms.ctypes.data_as(fp) if rms is not None else None
    out = _sig().next_chunk(data_read, C.byref(n), timestamp, block_rate, m, r)
    return out, n.value


def start_processing(data_read):
    return _sig().start_processing(data_read)


def tearDownMDSM():
    _sig().tearDownMDSM()


# ------------------------------------------------ host/device ABI (mdsm_wrapper.h)
class DEVICE_INFO(C.Structure):
    """src/mdsm/dedispersion_thread.h:9-20."""
    _fields_ = [("multiprocessor_count", C.c_int), ("constant_memory", C.c_int),
                ("shared_memory", C.c_int), ("register_count", C.c_int),
                ("thread_count", C.c_int), ("clock_rate", C.c_int), ("device_id", C.c_int)]


class DEVICES(C.Structure):
    """src/mdsm/dedispersion_thread.h:22-29 (minTotalGlobalMem in KiB)."""
    _fields_ = [("num_devices", C.c_int), ("minTotalGlobalMem", C.c_ulong),
                ("devices", C.POINTER(DEVICE_INFO))]


class THREAD_PARAMS(C.Structure):
    """src/mdsm/dedispersion_thread.h:31-57; the lock and barriers are
    pointers to glibc pthread_rwlock_t / pthread_barrier_t."""
    _fields_ = [("iterations", C.c_int), ("maxiters", C.c_int), ("maxshift", C.c_int),
                ("binsize", C.c_int), ("dmshifts", C.POINTER(C.c_float)),
                ("inputsize", C.c_size_t), ("outputsize", C.c_size_t),
                ("dm_output", C.c_size_t), ("dedispersed_size", C.c_size_t),
                ("output", C.POINTER(C.c_float)), ("input", C.POINTER(C.c_float)),
                ("device_id", C.c_int), ("thread_num", C.c_ushort),
                ("num_threads", C.c_ushort), ("stop", C.c_ushort),
                ("rw_lock", C.c_void_p), ("input_barrier", C.c_void_p),
                ("output_barrier", C.c_void_p), ("start", C.c_long),
                ("survey", C.POINTER(SURVEY))]


def wrapper_lib():
    """The library with call_initialise_devices / call_dedisperse typed."""
    L = lib()
    L.call_initialise_devices.restype = C.POINTER(DEVICES)
    L.call_initialise_devices.argtypes = [C.POINTER(SURVEY)]
    L.call_dedisperse.restype = C.c_void_p
    L.call_dedisperse.argtypes = [C.c_void_p]
    return L


def call_initialise_devices(survey=None):
    """List of DEVICE_INFO dicts and minTotalGlobalMem (KiB); ([], 0) when no
    GPU qualifies.  The DEVICES block is freed here."""
    p = wrapper_lib().call_initialise_devices(survey)
    if not p:
        return [], 0
    d = p.contents
    out = [{f: getattr(d.devices[i], f) for f, _ in DEVICE_INFO._fields_}
           for i in range(d.num_devices)]
    mem = d.minTotalGlobalMem
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    libc.free(C.cast(d.devices, C.c_void_p))
    libc.free(C.cast(p, C.c_void_p))
    return out, mem


# ------------------------------------------------ MdsmModule adapter (mdsm_module.h)
class MODULE_OUTPUT(C.Structure):
    _fields_ = [("nseries", C.c_uint32), ("dm_values", C.POINTER(C.c_float)),
                ("nsamp", C.POINTER(C.c_uint32)),
                ("data", C.POINTER(C.POINTER(C.c_float)))]


def _module_sig():
    L = lib()
    L.mdsm_module_create.restype = C.c_int
    L.mdsm_module_create.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    L.mdsm_module_run.restype = C.c_int
    L.mdsm_module_run.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.c_int64, C.c_int64, C.c_int, C.c_float, C.c_float,
                                  C.c_double, C.c_double, C.POINTER(MODULE_OUTPUT)]
    L.mdsm_module_survey.restype = C.POINTER(SURVEY)
    L.mdsm_module_survey.argtypes = [C.c_void_p]
    L.mdsm_module_destroy.restype = None
    L.mdsm_module_destroy.argtypes = [C.c_void_p]
    return L


class MdsmModule:
    """MdsmModule (src/pelican-mdsm/modules/MdsmModule.h:16-45) over the C ABI:
    run() takes one blob of spectra [ntime][nsubbands][nchannels] float32 (or
    None / ntime 0 for the end of the stream) and returns the dedispersed
    series of a completed block as a list of (dm, samples) -- copies -- or []."""

    def __init__(self, observation_file, create_output_blob=True, invert_channels=True):
        self._L = _module_sig()
        h = C.c_void_p()
        rc = self._L.mdsm_module_create(str(observation_file).encode(), int(create_output_blob),
                                        int(invert_channels), C.byref(h))
        if rc:
            raise RuntimeError(f"mdsm_module_create: {_err()}")
        self._h = h

    @property
    def survey(self):
        return self._L.mdsm_module_survey(self._h).contents

    def run(self, spectra=None, weighted=False, mean=0.0, rms=0.0, timestamp=0.0,
            block_rate=0.0):
        out = MODULE_OUTPUT()
        if spectra is None or spectra.shape[0] == 0:
            rc = self._L.mdsm_module_run(self._h, None, 0, 0, 0, 0, 0, int(weighted), mean, rms,
                                         timestamp, block_rate, C.byref(out))
        else:
            x = np.asarray(spectra, np.float32)
            assert x.ndim == 3 and x.strides[2] == 4
            rc = self._L.mdsm_module_run(self._h, x.ctypes.data, x.shape[0], x.shape[1],
                                         x.shape[2], x.strides[0] // 4, x.strides[1] // 4,
                                         int(weighted), mean, rms, timestamp, block_rate,
                                         C.byref(out))
        if rc:
            raise RuntimeError(f"mdsm_module_run: {_err()}")
        return [(out.dm_values[i],
                 np.ctypeslib.as_array(out.data[i], (out.nsamp[i],)).copy())
                for i in range(out.nseries)]

    def close(self):
        if self._h:
            self._L.mdsm_module_destroy(self._h)
            self._h = None
