"""The manager's side of the worker protocol of include/mdsm_wrapper.h, for
tests: glibc pthread barriers and rw lock, workers started on Python threads
running call_dedisperse, and the calls of src/mdsm/dedispersion_manager.cpp
(initialiseMDSM :320-381, next_chunk :418-483, start_processing :486-500)
with no output thread (barriers of num_threads + 1 parties).

Every manager step runs on a daemon thread joined with a time limit, so a
protocol mistake fails the test instead of blocking the test run."""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from mdsm_amd import compat

_SERIAL = -1  # glibc PTHREAD_BARRIER_SERIAL_THREAD


class Manager:
    def __init__(self, survey, nthreads, maxshift, dedispersed_size, device_id=0,
                 iterations=1, maxiters=2, limit=120.0):
        self.libc = C.CDLL("libc.so.6")
        self.limit = limit
        self.lock = C.create_string_buffer(64)   # >= sizeof(pthread_rwlock_t)
        self.inb = C.create_string_buffer(64)    # >= sizeof(pthread_barrier_t)
        self.outb = C.create_string_buffer(64)
        assert self.libc.pthread_rwlock_init(self.lock, None) == 0
        assert self.libc.pthread_barrier_init(self.inb, None, nthreads + 1) == 0
        assert self.libc.pthread_barrier_init(self.outb, None, nthreads + 1) == 0
        self.survey = survey
        s = survey.contents
        self.nchans, self.maxshift = s.nchans, maxshift
        self.stride = s.nsamp + maxshift
        self.input = np.zeros((s.nchans, self.stride), np.float32)
        self.output = np.zeros((nthreads, dedispersed_size), np.float32)
        self.params = (compat.THREAD_PARAMS * nthreads)()
        fp = C.POINTER(C.c_float)
        for k in range(nthreads):
            p = self.params[k]
            p.iterations, p.maxiters, p.maxshift, p.binsize = iterations, maxiters, maxshift, 1
            p.dedispersed_size = dedispersed_size
            p.output = self.output[k].ctypes.data_as(fp)
            p.input = self.input.ctypes.data_as(fp)
            p.device_id = device_id
            p.thread_num, p.num_threads, p.stop = k, nthreads, 0
            p.rw_lock = C.addressof(self.lock)
            p.input_barrier = C.addressof(self.inb)
            p.output_barrier = C.addressof(self.outb)
            p.survey = survey
        L = compat.wrapper_lib()
        self.threads = [threading.Thread(target=L.call_dedisperse,
                                         args=(C.addressof(self.params[k]),), daemon=True)
                        for k in range(nthreads)]
        for t in self.threads:
            t.start()
        self.blocks = 0
        self._run(lambda: self._wait(self.inb))     # initialiseMDSM's first input barrier

    def _run(self, fn):
        err = []

        def body():
            try:
                fn()
            except BaseException as e:  # noqa: BLE001 - reported below
                err.append(e)
        t = threading.Thread(target=body, daemon=True)
        t.start()
        t.join(self.limit)
        assert not t.is_alive(), "manager step did not finish: protocol deadlock"
        if err:
            raise err[0]

    def _wait(self, b):
        r = self.libc.pthread_barrier_wait(b)
        assert r in (0, _SERIAL), r

    def _next_chunk(self, nsamp=None, stop=False):
        assert self.libc.pthread_rwlock_wrlock(self.lock) == 0
        self._wait(self.outb)
        if stop:
            for p in self.params:
                p.stop = 1
        elif nsamp is not None:
            self.survey.contents.nsamp = nsamp
        assert self.libc.pthread_rwlock_unlock(self.lock) == 0

    def block(self, data, nsamp=None):
        """Write one block (the first: [nchans][nsamp + maxshift]; later ones:
        the new samples, placed after the carried maxshift), then next_chunk and
        start_processing.  nsamp lowers survey->nsamp for this block, as
        next_chunk does for a short subband block.  Returns the output of the
        previous block (a copy, [threads][dedispersed_size]) or None."""
        def step():
            if self.blocks == 0:
                self.input[:, :data.shape[1]] = data
            else:
                self.input[:, self.maxshift:self.maxshift + data.shape[1]] = data
            self._next_chunk(nsamp)
            self._wait(self.inb)
        self._run(step)
        self.blocks += 1
        return self.output.copy() if self.blocks >= 2 else None

    def finish(self):
        """next_chunk(0), start_processing(0), next_chunk(0): returns the last
        block's output, joins the workers and destroys the objects."""
        res = []

        def step():
            self._next_chunk(stop=True)
            self._wait(self.inb)
            res.append(self.output.copy())
            assert self.libc.pthread_rwlock_wrlock(self.lock) == 0
            self._wait(self.outb)
            assert self.libc.pthread_rwlock_unlock(self.lock) == 0
            for t in self.threads:
                t.join(self.limit)
                assert not t.is_alive(), "worker did not exit"
        self._run(step)
        self.libc.pthread_barrier_destroy(self.inb)
        self.libc.pthread_barrier_destroy(self.outb)
        self.libc.pthread_rwlock_destroy(self.lock)
        return res[0]


def brute_survey(nchans, nsamp, fch1, foff, tsamp, lowdm, dmstep, tdms, nthreads):
    s = compat.SURVEY()
    s.nsamp, s.nchans, s.tdms = nsamp, nchans, tdms
    s.tsamp, s.foff, s.fch1 = tsamp, foff, fch1
    s.useBruteForce = True
    s.lowdm, s.dmstep = lowdm, dmstep
    s.num_threads = nthreads
    return s


def subband_survey(nchans, nsamp, fch1, foff, tsamp, nsubs, passes, nthreads):
    s = compat.SURVEY()
    s.nsamp, s.nchans = nsamp, nchans
    s.tsamp, s.foff, s.fch1 = tsamp, foff, fch1
    s.useBruteForce = False
    arr = (compat.SUBBAND_PASSES * len(passes))()
    for i, q in enumerate(passes):
        for f, v in q.items():
            setattr(arr[i], f, v)
    s.pass_parameters = C.cast(arr, C.POINTER(compat.SUBBAND_PASSES))
    s.num_passes, s.nsubs = len(passes), nsubs
    s.num_threads = nthreads
    return s, arr


def ref_maxshift(fch1, foff, nchans, tsamp, hidm):
    """maxshift = (int)(dmshifts[nchans-1] * hiDM / tsamp) in float32, as
    src/mdsm/dedispersion_manager.cpp:291-297 computes it."""
    import oracle
    d = oracle.delay_table(fch1, foff, nchans)[-1]
    return int(np.float32(np.float32(d * np.float32(hidm)) / np.float32(tsamp)))
