"""The drop-in manager API: observation XML parsing (CPU) and the
next_chunk/start_processing pipeline against the oracle (GPU)."""
import os

import numpy as np
import pytest

import mdsm_amd as M
from mdsm_amd import compat

OBS_XML = """<?xml version="1.0"?>
<!-- observation file as written by the reference pipelines -->
<observation>
    <frequencies top="1198.4375" offset="-0.390625" />
    <dm lowDM="0" numDMs="1024" dmStep="0.1" useBruteForce="1" useL1Cache="1" />
    <samples number="131040" />
    <channels number="256" subbands="32" />
    <timing tsamp="0.00000256" />
    <detection threshold="5" />
    <output filePrefix="GUPPI" baseDirectory="." secondsPerFile="180"
            usePCTime="0" singleFileMode="1" />
    <gpus ids="0,2" />
</observation>
"""

SUBBAND_XML = """<observation>
  <frequencies top="1500" offset="-0.78125"/>
  <DM lowDM="0" numDMs="0" dmStep="0"/>
  <channels number="128" subbands="16"/>
  <timing tsamp="0.000064"/>
  <passes>
    <pass><lowDm>6</lowDm><highDm>126</highDm><deltaDm>0.6</deltaDm><downsample>1</downsample>
          <subDm>12</subDm><numDms>120</numDms><dmsPerCall>20</dmsPerCall><ncalls>6</ncalls></pass>
    <pass><lowDm>126</lowDm><highDm>246</highDm><deltaDm>1.2</deltaDm><downsample>2</downsample>
          <subDm>24</subDm><numDms>100</numDms><dmsPerCall>20</dmsPerCall><ncalls>5</ncalls></pass>
  </passes>
</observation>"""


def test_parse_reference_obs_xml():
    p = compat.parse_survey_xml(OBS_XML)
    s = p.contents
    assert s.fch1 == np.float32(1198.4375) and s.foff == np.float32(-0.390625)
    assert s.lowdm == 0 and s.tdms == 1024 and s.dmstep == np.float32(0.1)
    assert s.useBruteForce and s.useL1Cache
    assert s.nsamp == 131040 and s.nchans == 256 and s.nsubs == 32
    assert s.tsamp == np.float32(0.00000256) and s.detection_threshold == 5.0
    assert s.fileprefix == b"GUPPI" and s.basedir == b"." and s.secs_per_file == 180
    assert s.use_pc_time == b"\x00" and s.single_file_mode == b"\x01"
    assert s.num_gpus == 2 and [s.gpu_ids[i] for i in range(2)] == [0, 2]
    compat.free_survey(p)


def test_parse_defaults_and_subband_passes():
    p = compat.parse_survey_xml(SUBBAND_XML)
    s = p.contents
    assert not s.useBruteForce           # absent attribute reads 0 (dedispersion_manager.cpp:112)
    assert not s.useL1Cache
    assert s.fileprefix == b"output" and s.basedir == b"." and s.secs_per_file == 600
    assert s.use_pc_time == b"\x01" and s.single_file_mode == b"\x00"
    assert s.detection_threshold == 5.0 and s.num_gpus == 0
    ps = compat.passes(p)
    assert len(ps) == 2 and s.tdms == 220     # tdms = sum of numDms for subband surveys
    assert ps[1].lowdm == 126 and ps[1].dmstep == np.float32(1.2) and ps[1].binsize == 2
    assert ps[1].sub_dmstep == 24 and ps[1].calldms == 20 and ps[1].ncalls == 5
    compat.free_survey(p)


@pytest.mark.parametrize("bad", ["<observation><dm lowDM='1'></observation>", "<obs/>",
                                 "<observation", ""])
def test_parse_errors_are_reported(bad):
    with pytest.raises(ValueError):
        compat.parse_survey_xml(bad)


def test_process_survey_parameters_file(tmp_path):
    f = tmp_path / "obs.xml"
    f.write_text(OBS_XML)
    p = compat.processSurveyParameters(str(f))
    assert p.contents.tdms == 1024
    compat.free_survey(p)
    with pytest.raises(ValueError):
        compat.processSurveyParameters(str(tmp_path / "missing.xml"))


# ---------------------------------------------------------------- GPU pipeline
def _brute_xml(tmp, C, nsamp, ndms, dmstep, thr=5.0):
    return f"""<observation>
  <frequencies top="1500" offset="{-300.0 / C}"/>
  <dm lowDM="0" numDMs="{ndms}" dmStep="{dmstep}" useBruteForce="1" useL1Cache="1"/>
  <channels number="{C}" subbands="1"/>
  <timing tsamp="0.000064"/>
  <samples number="{nsamp}"/>
  <detection threshold="{thr}"/>
  <output filePrefix="cands" baseDirectory="{tmp}" singleFileMode="1"/>
</observation>"""


@pytest.mark.gpu
def test_manager_pipeline_matches_oracle(gpu, tmp_path):
    """Stream 4 blocks through initialiseMDSM/next_chunk/start_processing the
    way src/MDSM.cpp:139-166 does; every returned block equals the oracle
    brute force over the concatenated stream (continuity across blocks)."""
    import oracle
    C, nsamp, ndms, dmstep = 64, 4096, 48, 2.0
    p = compat.parse_survey_xml(_brute_xml(tmp_path, C, nsamp, ndms, dmstep))
    inbuf = compat.initialiseMDSM(p)
    s = p.contents
    Mx = s.maxshift
    assert inbuf.shape == (C, nsamp + Mx)
    nblocks = 4
    total = nsamp * nblocks + Mx
    stream = M.synth_filterbank(C, total, "u8", seed=77, mean=128, sigma=16,
                                fch1=1500.0, foff=-300.0 / C,
                                pulses=[(60.0, 5000, 6, 4.0)]).astype(np.float32)
    got = []
    k = 0
    pos = 0
    while True:
        if k < nblocks:
            if k == 0:
                inbuf[:, :] = stream[:, :nsamp + Mx]
                pos = nsamp + Mx
            else:
                inbuf[:, Mx:] = stream[:, pos:pos + nsamp]
                pos += nsamp
            data_read = nsamp
        else:
            data_read = 0
        out, n = compat.next_chunk(data_read, 0.0, 64e-6)
        if out:
            got.append(np.ctypeslib.as_array(out, (ndms, nsamp)).copy())
        if not compat.start_processing(data_read):
            break
        k += 1
    compat.tearDownMDSM()
    assert len(got) == nblocks
    sh = oracle.shift_table_ref(1500.0, -300.0 / C, C, 0.0, dmstep, ndms, 64e-6)
    ref = oracle.dedisperse_f32(stream, sh, nsamp * nblocks)
    for b in range(nblocks):
        assert np.array_equal(got[b], ref[:, b * nsamp:(b + 1) * nsamp]), f"block {b}"
    # the candidate file is the reference output thread's: process_brute
    # without noise arrays (chunked block statistics) on every block, line
    # for line -- against the reference's own process_brute where
    # oracle/_ref is built, and against the oracle's restatement of it
    txt = (tmp_path / "cands.dat").read_text().strip().splitlines()
    exp = []
    for b in range(nblocks):
        exp += oracle.process_brute_lines(got[b], 0.0, dmstep, 5.0, 0.0, 64e-6)
    assert txt == exp and len(txt) > 0
    if oracle.ref() is not None:
        ref = []
        for b in range(nblocks):
            path = tmp_path / f"ref{b}.dat"
            assert oracle.ref_process_brute(got[b], nsamp, 0.0, dmstep, 5.0, 0.0, 64e-6, path)
            ref += path.read_text().strip().splitlines()
        assert txt == ref
