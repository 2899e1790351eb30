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
