"""The standalone file pipeline as a compiled C++ program (mdsm_amd/mdsm_file,
the shape of the reference's src/MDSM.cpp:100-170): a SIGPROC file is read,
unpacked on the GPU and driven through processSurveyParameters /
initialiseMDSM / next_chunk(data_read, unsigned&) / start_processing; its
candidate file must equal, line for line, the reference output thread's
process_brute (restated by the oracle and pinned to the reference's own
build) on the oracle's dedispersion of the same stream."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from test_sigproc import _s

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "mdsm_amd", "mdsm_file")


def _write_u8_fil(path, x_cm, tsamp, fch1, foff):
    """8-bit SIGPROC file of channel-major samples x_cm [C][T] (time-major on disk)."""
    import struct
    C = x_cm.shape[0]
    h = _s(b"HEADER_START") + _s(b"source_name") + _s(b"synthetic")
    for k, v in dict(machine_id=0, telescope_id=0, data_type=1, nchans=C, nbits=8,
                     nifs=1).items():
        h += _s(k.encode()) + struct.pack("<i", v)
    for k, v in dict(tstart=60000.0, tsamp=tsamp, fch1=fch1, foff=foff).items():
        h += _s(k.encode()) + struct.pack("<d", v)
    h += _s(b"HEADER_END")
    with open(path, "wb") as f:
        f.write(h + np.ascontiguousarray(x_cm.T).tobytes())


def _obs_xml(tmp, C, nsamp, ndms, dmstep, thr):
    return f"""<observation>
  <frequencies top="1500" offset="{-300.0 / C}"/>
  <dm lowDM="0" numDMs="{ndms}" dmStep="{dmstep}" useBruteForce="1" useL1Cache="0"/>
  <channels number="{C}" subbands="1"/>
  <timing tsamp="0.000064"/>
  <samples number="{nsamp}"/>
  <detection threshold="{thr}"/>
  <output filePrefix="cands" baseDirectory="{tmp}" singleFileMode="1"/>
</observation>"""


def test_driver_binary_is_built():
    assert os.access(DRIVER, os.X_OK), "mdsm_amd/mdsm_file missing: run mdsm_amd.build"


@pytest.mark.gpu
def test_file_pipeline_matches_oracle(gpu, tmp_path):
    M = gpu
    print("case C++ file pipeline", flush=True)
    C, nsamp, ndms, dmstep, thr, nblocks = 64, 4096, 48, 2.0, 5.0, 3
    obs = tmp_path / "obs.xml"
    obs.write_text(_obs_xml(tmp_path, C, nsamp, ndms, dmstep, thr))
    tsamp = float(np.float32(64e-6))
    sh = oracle.shift_table_ref(1500.0, -300.0 / C, C, 0.0, dmstep, ndms, np.float32(64e-6))
    # the manager's overlap: the plan's largest shift
    Mx = int(sh.max())
    total = nblocks * nsamp + Mx
    x = M.synth_filterbank(C, total, "u8", seed=91, mean=128, sigma=16, fch1=1500.0,
                           foff=-300.0 / C, pulses=[(40.0, 3000, 6, 4.0), (70.0, 9000, 3, 5.0)])
    fil = tmp_path / "data.fil"
    _write_u8_fil(str(fil), x, 64e-6, 1500.0, -300.0 / C)
    r = subprocess.run([DRIVER, str(fil), "-obs", str(obs)], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert f"maxshift {Mx} nsamp {nsamp}" in r.stdout
    # every block read in full; two calls with data_read = 0 drain the pipeline
    assert f"samples {nblocks * nsamp} " in r.stdout
    assert r.stdout.count("data_read 0") == 2
    ref = oracle.dedisperse_f32(x.astype(np.float32), sh, nblocks * nsamp)
    exp = []
    for k in range(nblocks):
        blk = np.ascontiguousarray(ref[:, k * nsamp:(k + 1) * nsamp])
        exp += oracle.process_brute_lines(blk, 0.0, dmstep, thr, tsamp * nsamp * k, tsamp)
    got = (tmp_path / "cands.dat").read_text().strip().splitlines()
    assert got == exp and len(got) > 0
