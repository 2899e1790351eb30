"""SIGPROC filterbank files: header parsing and raw reads (host), sample
unpacking (oracle restatement, pinned to the reference's own read_header and
read_block built under oracle/_ref, and to numpy), and the dedispersed-series
writer.  The GPU unpack is in test_gpu_prep.py."""
import struct

import numpy as np
import pytest

import oracle


def _s(b):
    return struct.pack("<i", len(b)) + b


def write_fil(path, nbits, nchans, nspec, seed=0, extra=True):
    """A SIGPROC file with the keywords the reference parses, and random data."""
    rng = np.random.default_rng(seed)
    h = _s(b"HEADER_START")
    if extra:
        h += _s(b"rawdatafile") + _s(b"raw_input.dat")
        h += _s(b"source_name") + _s(b"J0000+0000")
    ints = dict(machine_id=7, telescope_id=4, data_type=1, nchans=nchans, nbits=nbits, nifs=1,
                nbeams=13, ibeam=2, barycentric=0, pulsarcentric=0)
    dbls = dict(tstart=56789.125, tsamp=64e-6, fch1=1500.0, foff=-300.0 / nchans, refdm=0.0,
                az_start=12.5, za_start=33.25, src_raj=123456.78, src_dej=-234512.5)
    for k, v in ints.items():
        h += _s(k.encode()) + struct.pack("<i", v)
    for k, v in dbls.items():
        h += _s(k.encode()) + struct.pack("<d", v)
    h += _s(b"HEADER_END")
    nbytes = nspec * nchans * nbits // 8
    if nbits == 32:
        data = rng.standard_normal(nspec * nchans).astype(np.float32).tobytes()
    elif nbits == 16:
        data = rng.integers(0, 65536, nspec * nchans, dtype=np.uint16).tobytes()
    else:
        data = rng.integers(0, 256, nbytes, dtype=np.uint8).tobytes()
    with open(path, "wb") as f:
        f.write(h + data)
    return len(h), data, ints, dbls


@pytest.mark.parametrize("nbits", [1, 2, 4, 8, 16, 32])
def test_header_and_raw_reads(tmp_path, nbits):
    import mdsm_amd as M
    path = str(tmp_path / f"x{nbits}.fil")
    C, T = 64, 300
    hb, data, ints, dbls = write_fil(path, nbits, C, T, seed=nbits)
    h = M.sigproc_read_header(path)
    assert h["header_bytes"] == hb and h["nspectra"] == T
    assert h["source_name"] == "J0000+0000" and h["rawdatafile"] == "raw_input.dat"
    for k, v in ints.items():
        assert h[k] == v, k
    for k, v in dbls.items():
        assert h[k] == v, k
    r = M.SigprocReader(path)
    a = r.read(100)
    r.seek(250)
    b = r.read(100)                       # only 50 left
    r.close()
    sb = C * nbits // 8
    assert a.shape == (100, sb) and b.shape == (50, sb)
    assert a.tobytes() == data[:100 * sb] and b.tobytes() == data[250 * sb:]


@pytest.mark.parametrize("nbits", [1, 2, 4, 8, 16, 32])
def test_unpack_matches_reference_read_block(tmp_path, nbits):
    path = str(tmp_path / f"u{nbits}.fil")
    C, T = 64, 40
    hb, data, _, _ = write_fil(path, nbits, C, T, seed=10 + nbits)
    got = oracle.unpack_bits(data, nbits, C * T)
    # independent: numpy
    raw = np.frombuffer(data, np.uint8)
    if nbits in (1, 2, 4):
        bits = np.unpackbits(raw, bitorder="little").reshape(-1, nbits)
        exp = (bits * (1 << np.arange(nbits))).sum(axis=1).astype(np.float32)
    elif nbits == 8:
        exp = raw.astype(np.float32)
    elif nbits == 16:
        exp = np.frombuffer(data, np.uint16).astype(np.float32)
    else:
        exp = np.frombuffer(data, np.float32)
    assert np.array_equal(got, exp)
    if oracle.ref() is None:
        pytest.skip("oracle/_ref not built (no reference tree)")
    ref = oracle.ref_read_block(path, hb, nbits, C * T)
    assert ref.size == C * T and np.array_equal(got, ref)


def test_header_matches_reference_read_header(tmp_path):
    import mdsm_amd as M
    if oracle.ref() is None:
        pytest.skip("oracle/_ref not built (no reference tree)")
    path = str(tmp_path / "h.fil")
    write_fil(path, 8, 256, 10)
    ours = M.sigproc_read_header(path)
    ref = oracle.ref_read_header(path)
    assert ref.total_bytes == ours["header_bytes"]
    for k in ("machine_id", "telescope_id", "data_type", "nchans", "nbits", "nifs", "nbeams",
              "ibeam", "tstart", "tsamp", "fch1", "foff", "az_start", "za_start", "src_raj",
              "src_dej"):
        assert getattr(ref, k) == ours[k], k
    assert ref.source_name.decode() == ours["source_name"]
    assert ref.rawdatafile.decode() == ours["rawdatafile"]


def test_unknown_keyword_is_an_error(tmp_path):
    import mdsm_amd as M
    path = str(tmp_path / "bad.fil")
    with open(path, "wb") as f:
        f.write(_s(b"HEADER_START") + _s(b"no_such_key") + struct.pack("<i", 1) + _s(b"HEADER_END"))
    with pytest.raises(M.MdsmError) as e:
        M.sigproc_read_header(path)
    assert e.value.code == -7
    with open(path, "wb") as f:
        f.write(b"\x00" * 64)
    with pytest.raises(M.MdsmError):
        M.sigproc_read_header(path)


def test_series_writer_roundtrip(tmp_path):
    import mdsm_amd as M
    prefix = str(tmp_path / "MDSM")
    dms = np.float32([1.5, 30.25])
    w = M.SeriesWriter(prefix, dms, 1500.0, -0.29296875, 64e-6, 0.0)
    rows = np.arange(4 * 100, dtype=np.float32).reshape(4, 100)
    row_dms = np.float32([0.0, 1.5, 10.0, 30.25])
    w.write(rows, row_dms)
    w.write(rows + 1000, row_dms)
    w.close()
    for dm, r in ((1.5, 1), (30.25, 3)):
        path = f"{prefix}_{dm:g}.fil"
        h = M.sigproc_read_header(path)
        assert h["nchans"] == 1 and h["nbits"] == 32 and h["nifs"] == 1 and h["data_type"] == 1
        assert h["tsamp"] == 64e-6 and h["fch1"] == 1500.0 and h["nspectra"] == 200
        data = np.fromfile(path, np.float32, offset=h["header_bytes"])
        assert np.array_equal(data, np.concatenate([rows[r], rows[r] + 1000]))
