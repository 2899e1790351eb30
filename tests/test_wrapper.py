"""The reference's host/device ABI (include/mdsm_wrapper.h): the manager's
side of the worker protocol (tests/wrapper_harness.py) drives
call_dedisperse workers on threads, and every block they return must equal
the oracle over the continuous stream, for each thread's share of the trials
(src/mdsm/dedispersion_thread.cu:161-163, 225)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import mdsm_amd as M
from mdsm_amd import compat
import wrapper_harness as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FCH1, TSAMP = 1500.0, 64e-6


def test_struct_layout_matches_the_header(tmp_path):
    """compat's ctypes mirror has the C header's offsets (gcc, offsetof)."""
    fields = {"THREAD_PARAMS": compat.THREAD_PARAMS, "DEVICES": compat.DEVICES,
              "DEVICE_INFO": compat.DEVICE_INFO, "SURVEY": compat.SURVEY}
    lines = ['#include "mdsm_wrapper.h"', "#include <stdio.h>", "#include <stddef.h>",
             "int main(void) {"]
    for name, cls in fields.items():
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-lpthread"], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                   check=True).stdout.splitlines())
    for name, cls in fields.items():
        assert int(got[name]) == C.sizeof(cls), name
        for f, _ in cls._fields_:
            assert int(got[f"{name}.{f}"]) == getattr(cls, f).offset, f"{name}.{f}"


def test_initialise_devices_outside_the_gpu_list_returns_none():
    """No device qualifies: NULL, and the process goes on (the reference
    exits, src/mdsm/dedispersion_thread.cu:11-12)."""
    s = compat.SURVEY()
    ids = (C.c_uint * 1)(4000)
    s.gpu_ids, s.num_gpus = C.cast(ids, C.POINTER(C.c_uint)), 1
    devs, mem = compat.call_initialise_devices(C.pointer(s))
    assert devs == [] and mem == 0


@pytest.mark.parametrize("nthreads", [1, 3])
def test_worker_that_cannot_set_up_keeps_the_protocol(nthreads):
    """A worker whose device does not exist still meets every barrier of the
    manager's sequence, writes zeros for every block and exits on stop."""
    C_, nsamp = 16, 1024
    s = H.brute_survey(C_, nsamp, FCH1, -1.0, TSAMP, 0.0, 1.0, 6 * nthreads, nthreads)
    Mx = H.ref_maxshift(FCH1, -1.0, C_, TSAMP, 6.0 * nthreads - 1)
    size = nsamp * 6
    m = H.Manager(C.pointer(s), nthreads, Mx, size, device_id=1 << 20, limit=30.0)
    m.output[:] = 7.0
    outs = [m.block(np.ones((C_, nsamp + Mx), np.float32))]
    for _ in range(2):
        outs.append(m.block(np.ones((C_, nsamp), np.float32)))
    outs.append(m.finish())
    assert outs[0] is None
    for o in outs[1:]:
        assert o.shape == (nthreads, size) and not o.any()


# ------------------------------------------------------------------- GPU
def _brute_case(nthreads, C_, nsamp, nblocks, foff, lowdm, dmstep, tdms, seed):
    s = H.brute_survey(C_, nsamp, FCH1, foff, TSAMP, lowdm, dmstep, tdms, nthreads)
    hidm = np.float32(np.float32(lowdm) + np.float32(dmstep) * np.float32(tdms - 1))
    Mx = H.ref_maxshift(FCH1, foff, C_, TSAMP, hidm)
    per = tdms // nthreads
    stream = M.synth_filterbank(C_, nsamp * nblocks + Mx, "u8", seed=seed, mean=128, sigma=16,
                                fch1=FCH1, foff=foff,
                                pulses=[(lowdm + dmstep * tdms * 0.6, nsamp + 500, 6, 4.0)]
                                ).astype(np.float32)
    m = H.Manager(C.pointer(s), nthreads, Mx, per * nsamp)
    got = []
    for b in range(nblocks):
        data = stream[:, :nsamp + Mx] if b == 0 else stream[:, Mx + b * nsamp:Mx + (b + 1) * nsamp]
        o = m.block(data)
        if o is not None:
            got.append(o)
    got.append(m.finish())
    return stream, Mx, per, got


@pytest.mark.gpu
@pytest.mark.parametrize("nthreads", [1, 2])
def test_brute_force_workers_match_oracle(gpu, nthreads):
    import oracle
    C_, nsamp, nblocks, foff, lowdm, dmstep, tdms = 64, 4096, 3, -300.0 / 64, 1.0, 0.75, 96
    print(f"case wrapper brute nthreads={nthreads}", flush=True)
    stream, Mx, per, got = _brute_case(nthreads, C_, nsamp, nblocks, foff, lowdm, dmstep,
                                       tdms, seed=41 + nthreads)
    assert len(got) == nblocks
    for t in range(nthreads):
        f32 = np.float32
        start = f32(f32(lowdm) + f32(f32(f32(dmstep) * f32(tdms)) / f32(nthreads)) * f32(t))
        sh = oracle.shift_table_ref(FCH1, foff, C_, start, dmstep, per, TSAMP)
        assert int(sh.max()) <= Mx
        for b in range(nblocks):
            x = stream[:, b * nsamp:(b + 1) * nsamp + Mx]
            ref = oracle.dedisperse_f32(x, sh, nsamp)
            assert np.array_equal(got[b][t].reshape(per, nsamp), ref), (t, b)


def _subband_expected(x, nsamp, passes, nsubs, foff, nthreads, t):
    import oracle
    f32 = np.float32
    rows = []
    for q in passes:
        k = q["binsize"]
        nc = q["ncalls"] // nthreads
        lo = f32(f32(q["lowdm"]) + f32(f32(q["sub_dmstep"]) * f32(nc)) * f32(t))
        co, fi = oracle.subband_tables(FCH1, foff, x.shape[0], nsubs, lo, q["sub_dmstep"],
                                       q["dmstep"], nc, q["calldms"], f32(TSAMP) * f32(k))
        n = nsamp // k
        M2 = int(fi.max())
        cps = x.shape[0] // nsubs
        M1 = int((co - co[:, (np.arange(x.shape[0]) // cps) * cps]).max())
        assert n + M1 + M2 <= x.shape[1] // k
        r = oracle.subband_f64(x, k, nsubs, nc, q["calldms"], co, fi, n + M2, n)
        rows.append(r.astype(np.float32).ravel())
    return np.concatenate(rows)


@pytest.mark.gpu
@pytest.mark.parametrize("nthreads", [1, 2])
def test_subband_workers_match_oracle_with_short_last_block(gpu, nthreads):
    """Two passes (binsize 1 and 2), ncalls split over the threads, and a last
    block shorter than nsamp (survey->nsamp lowered to a multiple of the
    largest binsize, as next_chunk does)."""
    C_, nsubs, nsamp, foff = 128, 16, 6000, -300.0 / 128
    passes = [dict(lowdm=0.0, highdm=24.0, dmstep=0.5, sub_dmstep=6.0, binsize=1, ndms=48,
                   ncalls=4, calldms=12),
              dict(lowdm=24.0, highdm=60.0, dmstep=1.0, sub_dmstep=9.0, binsize=2, ndms=36,
                   ncalls=4, calldms=9)]
    print(f"case wrapper subband nthreads={nthreads}", flush=True)
    s, keep = H.subband_survey(C_, nsamp, FCH1, foff, TSAMP, nsubs, passes, nthreads)
    Mx = H.ref_maxshift(FCH1, foff, C_, TSAMP, passes[-1]["highdm"] + 6.0)
    size = sum(q["ncalls"] // nthreads * q["calldms"] * (nsamp // q["binsize"]) for q in passes)
    short = 4001 - 4001 % 2
    lens = [nsamp, nsamp, short]
    stream = M.synth_filterbank(C_, sum(lens) + Mx, "u8", seed=51 + nthreads, mean=128,
                                sigma=16, fch1=FCH1, foff=foff,
                                pulses=[(30.0, 7000, 6, 4.0), (50.0, 13000, 8, 4.0)]
                                ).astype(np.float32)
    m = H.Manager(C.pointer(s), nthreads, Mx, size)
    got, pos = [], 0
    for b, n in enumerate(lens):
        data = stream[:, :n + Mx] if b == 0 else stream[:, Mx + pos:Mx + pos + n]
        o = m.block(data, nsamp=n if n != nsamp else None)
        if o is not None:
            got.append(o)
        pos += n
    got.append(m.finish())
    assert len(got) == len(lens)
    pos = 0
    for b, n in enumerate(lens):
        x = stream[:, pos:pos + nsamp + Mx]
        for t in range(nthreads):
            exp = _subband_expected(x, n, passes, nsubs, foff, nthreads, t)
            assert np.array_equal(got[b][t][:exp.size], exp), (b, t)
        pos += n
    del keep


@pytest.mark.gpu
def test_initialise_devices_lists_the_gpu(gpu):
    devs, mem = compat.call_initialise_devices(None)
    assert len(devs) == M.device_count()
    d = devs[0]
    assert d["device_id"] == 0 and d["multiprocessor_count"] >= 1
    assert d["thread_count"] >= 256 and d["shared_memory"] >= 65536
    assert mem > 2.5 * 1024 * 1024


@pytest.mark.gpu
def test_worker_with_short_maxshift_writes_zeros(gpu):
    """A manager whose maxshift is shorter than the trials' delays: the
    worker reports it, keeps the protocol and writes zeros (the reference
    would read past its input rows)."""
    C_, nsamp = 64, 2048
    s = H.brute_survey(C_, nsamp, FCH1, -300.0 / 64, TSAMP, 0.0, 2.0, 32, 1)
    m = H.Manager(C.pointer(s), 1, 8, 32 * nsamp)
    m.output[:] = 5.0
    x = np.ones((C_, nsamp + 8), np.float32)
    o1 = m.block(x)
    o2 = m.block(np.ones((C_, nsamp), np.float32))
    last = m.finish()
    assert o1 is None and not o2.any() and not last.any()
