"""The MdsmModule adapter (include/mdsm_module.h): spectra blobs in, series
out.  GPU cases stream a synthetic observation through it blob by blob --
blobs that straddle blocks, channel inversion, a short last block, weighted
blobs whose per-chunk mean/rms must reach the median-of-five detector at
the right samples -- and check every block against the oracle over the
continuous stream."""
import math
import os
import subprocess

import numpy as np
import pytest

import mdsm_amd as M
from mdsm_amd import compat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RATE = 64e-6


def _brute_xml(tmp, C, nsamp, ndms, dmstep, thr):
    return f"""<observation>
  <frequencies top="1500" offset="{-300.0 / C}"/>
  <dm lowDM="0" numDMs="{ndms}" dmStep="{dmstep}" useBruteForce="1"/>
  <channels number="{C}" subbands="1"/>
  <timing tsamp="{RATE}"/>
  <samples number="{nsamp}"/>
  <detection threshold="{thr}"/>
  <output filePrefix="cands" baseDirectory="{tmp}" singleFileMode="1"/>
</observation>"""


def _write(tmp_path, xml):
    p = tmp_path / "obs.xml"
    p.write_text(xml)
    return p


def test_header_compiles_as_c(tmp_path):
    src = tmp_path / "m.c"
    src.write_text('#include "mdsm_module.h"\n#include "mdsm_wrapper.h"\nint main(void) '
                   '{ mdsm_module_output o; (void)o; return 0; }\n')
    subprocess.run(["gcc", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-o", str(tmp_path / "m")], check=True)


def test_missing_observation_file_is_an_error(tmp_path):
    with pytest.raises(RuntimeError):
        compat.MdsmModule(tmp_path / "none.xml")


def _feed(mod, stream_tm, ntime, nsub, ts_of=None, stats=None):
    """Blobs of ntime spectra (the last one shorter), then end-of-stream
    blobs; returns the list of non-empty outputs in order."""
    outs = []
    total = stream_tm.shape[0]
    j = 0
    for a in range(0, total, ntime):
        blob = stream_tm[a:a + ntime].reshape(-1, nsub, stream_tm.shape[1] // nsub)
        kw = {}
        if ts_of is not None:
            kw["timestamp"] = ts_of(a)
            kw["block_rate"] = RATE
        if stats is not None:
            kw.update(weighted=True, mean=stats[j][0], rms=stats[j][1])
        o = mod.run(blob, **kw)
        if o:
            outs.append(o)
        j += 1
    for _ in range(4):
        # the end-of-stream blob is the first blob of a last partial block:
        # its timestamp is that block's
        o = mod.run(None, **({} if ts_of is None else dict(timestamp=ts_of(total),
                                                           block_rate=RATE)))
        if o:
            outs.append(o)
    return outs


@pytest.mark.gpu
def test_module_blocks_match_oracle(gpu, tmp_path):
    """Inverted channels, 1000-spectrum blobs that straddle every block, four
    full blocks and a short last one."""
    import oracle
    C, nsub, nsamp, ndms, dmstep = 64, 4, 4096, 48, 2.0
    print("case module brute blocks", flush=True)
    mod = compat.MdsmModule(_write(tmp_path, _brute_xml(tmp_path, C, nsamp, ndms, dmstep, 50.0)),
                            create_output_blob=True, invert_channels=True)
    Mx = mod.survey.maxshift
    short = 1500
    total = Mx + 4 * nsamp + short
    x = M.synth_filterbank(C, total, "u8", seed=71, mean=128, sigma=16, fch1=1500.0,
                           foff=-300.0 / C, pulses=[(40.0, 6000, 6, 4.0)]).astype(np.float32)
    # channel c of the block is blob channel C-1-c: the blob carries x reversed
    stream_tm = np.ascontiguousarray(x[::-1].T)
    outs = _feed(mod, stream_tm, 1000, nsub)
    mod.close()
    assert len(outs) == 5
    sh = oracle.shift_table_ref(1500.0, -300.0 / C, C, 0.0, dmstep, ndms, RATE)
    for b, o in enumerate(outs):
        n = nsamp if b < 4 else short
        ref = oracle.dedisperse_f32(x[:, b * nsamp:b * nsamp + n + Mx], sh, n)
        assert len(o) == ndms
        for d, (dm, ser) in enumerate(o):
            assert dm == np.float32(np.float32(dmstep) * np.float32(d))
            assert ser.shape == (n,)
            assert np.array_equal(ser, ref[d]), (b, d)


def _expected_chunk_stats(nsamp, Mx, ntime, total, stats):
    """The weighted-blob bookkeeping of MdsmModule.cpp:57-96 and :219-237 for
    blobs of ntime spectra: the mean/rms arrays each block hands to
    next_chunk, in block order."""
    chunks = (nsamp + Mx) // ntime
    n = max(chunks + 1, -(-(nsamp + Mx) // ntime) + 1)
    means = [np.zeros(n, np.float32) for _ in range(3)]
    rmss = [np.zeros(n, np.float32) for _ in range(3)]
    counter, samples, res, j = 0, 0, [], 0
    for a in range(0, total, ntime):
        nt = min(ntime, total - a)
        idx = samples // ntime if counter == 0 else samples // ntime + math.floor(Mx / ntime)
        if idx < n:
            means[counter % 3][idx], rmss[counter % 3][idx] = stats[j]
        req = nsamp + Mx if counter == 0 else nsamp
        copy = min(nt, req - samples)
        samples += copy
        if samples == req:
            res.append((means[counter % 3].copy(), rmss[counter % 3].copy()))
            counter += 1
            samples = 0
            fr, k = nsamp // ntime, Mx // ntime
            means[counter % 3][:k] = means[(counter - 1) % 3][fr:fr + k]
            rmss[counter % 3][:k] = rmss[(counter - 1) % 3][fr:fr + k]
            samples += nt - copy
        j += 1
    if samples:
        res.append((means[counter % 3].copy(), rmss[counter % 3].copy()))
    return res


@pytest.mark.gpu
def test_module_weighted_blobs_reach_the_detector(gpu, tmp_path):
    """Every blob carries its own mean and rms; each candidate line must show
    the statistics of the chunk its sample falls in, the block's timestamp
    must be its first blob's minus maxshift samples, and the candidate set
    must be exactly the median-of-five threshold rule over those statistics
    (src/mdsm/dedispersion_output.cpp:176-208)."""
    C, nsub, nsamp, ndms, dmstep, thr, ntime = 64, 8, 4096, 24, 3.0, 1.0, 512
    print("case module weighted", flush=True)
    mod = compat.MdsmModule(_write(tmp_path, _brute_xml(tmp_path, C, nsamp, ndms, dmstep, thr)),
                            create_output_blob=True, invert_channels=False)
    Mx = mod.survey.maxshift
    nblk = 3
    # whole blobs only: the stream ends with a short block of the padding
    total = -(-(Mx + nblk * nsamp) // ntime) * ntime
    x = M.synth_filterbank(C, total, "u8", seed=72, mean=128, sigma=16, fch1=1500.0,
                           foff=-300.0 / C).astype(np.float32)
    stream_tm = np.ascontiguousarray(x.T)
    nblobs = -(-total // ntime)
    stats = [(np.float32(C * 128 + (j % 7) * 9.0 - 27.0), np.float32(90.0 + (j % 5) * 15.0))
             for j in range(nblobs)]

    def block_of(g):
        return 0 if g < nsamp + Mx else (g - Mx) // nsamp

    outs = _feed(mod, stream_tm, ntime, nsub, ts_of=lambda g: 1.0e4 * block_of(g), stats=stats)
    mod.close()
    nout = nblk + (total > Mx + nblk * nsamp)
    assert len(outs) == nout
    exp_stats = _expected_chunk_stats(nsamp, Mx, ntime, total, stats)
    assert len(exp_stats) == nout
    got = set()
    for line in (tmp_path / "cands.dat").read_text().splitlines():
        t, dm, snr, mean, rms = (float(v) for v in line.split(","))
        k = int(round((t + RATE * Mx) / 1.0e4)) if t > 1.0 else 0
        base = 0.0 if k == 0 else 1.0e4 * k - RATE * Mx
        l = int(round((t - base) / RATE))
        em, er = exp_stats[k][0][l // ntime], exp_stats[k][1][l // ntime]
        assert abs(mean - em) <= 1e-3 and abs(rms - er) <= 1e-3, (k, l, mean, em, rms, er)
        got.add((k, int(round(dm / dmstep)), l))
    exp = set()
    for k, o in enumerate(outs):
        em, er = exp_stats[k]
        for d, (dm, ser) in enumerate(o):
            if not dm > 1.0:
                continue
            w = np.lib.stride_tricks.sliding_window_view(ser, 5)
            med = np.sort(w, axis=1)[:, 2]
            l = np.arange(med.size)
            hit = (med - em[l // ntime]) >= er[l // ntime] * np.float32(thr)
            exp |= {(k, d, int(i)) for i in np.nonzero(hit)[0]}
    assert len(exp) > 100 and got == exp


SUBBAND_XML = """<observation>
  <frequencies top="1500" offset="-2.34375"/>
  <channels number="128" subbands="16"/>
  <timing tsamp="0.000064"/>
  <samples number="6000"/>
  <detection threshold="50"/>
  <output filePrefix="cands" baseDirectory="{tmp}" singleFileMode="1"/>
  <passes>
    <pass><lowDm>0</lowDm><highDm>24</highDm><deltaDm>0.5</deltaDm><downsample>1</downsample>
          <subDm>6</subDm><numDms>48</numDms><dmsPerCall>12</dmsPerCall><ncalls>4</ncalls></pass>
    <pass><lowDm>24</lowDm><highDm>60</highDm><deltaDm>1</deltaDm><downsample>2</downsample>
          <subDm>9</subDm><numDms>36</numDms><dmsPerCall>9</dmsPerCall><ncalls>4</ncalls></pass>
  </passes>
</observation>"""


@pytest.mark.gpu
def test_module_subband_series_and_labels(gpu, tmp_path):
    """Subband passes: series of samples/binsize samples, DM labels as
    MdsmModule.cpp:195-214 sets them, values equal to the oracle's two-stage
    subband dedispersion."""
    import oracle
    print("case module subband", flush=True)
    mod = compat.MdsmModule(_write(tmp_path, SUBBAND_XML.format(tmp=tmp_path)),
                            create_output_blob=True, invert_channels=False)
    s = mod.survey
    C, nsub, nsamp, Mx = 128, 16, s.nsamp, s.maxshift
    total = Mx + 2 * nsamp
    x = M.synth_filterbank(C, total, "u8", seed=73, mean=128, sigma=16, fch1=1500.0,
                           foff=-2.34375, pulses=[(30.0, 5000, 6, 4.0)]).astype(np.float32)
    outs = _feed(mod, np.ascontiguousarray(x.T), 700, 4)
    mod.close()
    assert len(outs) == 2
    f32 = np.float32
    pp = [dict(lowdm=0.0, dmstep=0.5, sub=6.0, k=1, ncalls=4, calldms=12),
          dict(lowdm=24.0, dmstep=1.0, sub=9.0, k=2, ncalls=4, calldms=9)]
    for b, o in enumerate(outs):
        xb = x[:, b * nsamp:b * nsamp + nsamp + Mx]
        i = 0
        for q in pp:
            co, fi = oracle.subband_tables(1500.0, -2.34375, C, nsub, q["lowdm"], q["sub"],
                                           q["dmstep"], q["ncalls"], q["calldms"],
                                           f32(RATE) * f32(q["k"]))
            n = nsamp // q["k"]
            ref = oracle.subband_f64(xb, q["k"], nsub, q["ncalls"], q["calldms"], co, fi,
                                     n + int(fi.max()), n)
            for d in range(q["ncalls"] * q["calldms"]):
                dm, ser = o[i]
                assert dm == f32(f32(q["lowdm"]) + f32(d) * f32(q["dmstep"]))
                assert ser.shape == (n,)
                assert np.array_equal(ser.astype(np.float64), ref[d].astype(np.float32)
                                      .astype(np.float64)), (b, d)
                i += 1
        assert i == len(o)


@pytest.mark.gpu
def test_module_rejects_bad_blobs(gpu, tmp_path):
    """Channel counts that do not match the observation, and blobs longer
    than a block, are errors (the module stays usable)."""
    C, nsamp = 64, 4096
    mod = compat.MdsmModule(_write(tmp_path, _brute_xml(tmp_path, C, nsamp, 16, 1.0, 50.0)))
    with pytest.raises(RuntimeError):
        mod.run(np.zeros((100, 4, 8), np.float32))          # 32 channels, not 64
    with pytest.raises(RuntimeError):
        mod.run(np.zeros((nsamp + 1, 4, 16), np.float32))  # longer than a block
    assert mod.run(np.zeros((100, 4, 16), np.float32)) == []
    mod.close()
