"""Corner turn (time-major spectra -> channel-major rows) and detrend/normalise
of dedispersed rows, on the GPU, against the oracle."""
import numpy as np
import pytest

import oracle
from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,out", [("u8", "u8"), ("u8", "f32"), ("f16", "f16"),
                                       ("f16", "f32"), ("f32", "f32")])
@pytest.mark.parametrize("invert", [False, True])
def test_corner_turn_matches_oracle(gpu, dtype, out, invert):
    M = gpu
    print(f"case corner_turn {dtype}->{out} invert={invert}", flush=True)
    rng = np.random.default_rng(11)
    T, Cn, in_stride = 1000, 300, 307          # ragged against the 64x64 tiles
    npd = {"u8": np.uint8, "f16": np.float16, "f32": np.float32}
    if dtype == "u8":
        x = rng.integers(0, 256, (T, in_stride), dtype=np.uint8)
    else:
        x = rng.standard_normal((T, in_stride)).astype(npd[dtype])
    col0, out_stride = 13, 1100
    din = M.DeviceBuffer.from_host(x)
    dout = M.DeviceBuffer(Cn * out_stride * np.dtype(npd[out]).itemsize)
    dout.fill_u32(0xFFFFFFFF)
    M.corner_turn(din, dtype, T, Cn, in_stride, dout, out_stride, col0, invert, out)
    M.synchronize()
    got = dout.download(npd[out], (Cn, out_stride))
    exp = oracle.corner_turn_f32(x[:, :Cn].astype(np.float32), invert)
    assert np.array_equal(got[:, col0:col0 + T].astype(np.float32), exp)
    # nothing outside the target window is touched
    assert np.all(got[:, :col0].view(np.uint8) == 0xFF)


@pytest.mark.parametrize("invert", [False, True])
@pytest.mark.parametrize("T,Cn,pad", [(1000, 300, 8), (4096, 1024, 8), (17, 4, 8),
                                      (1000, 320, 16), (4096, 1024, 0), (33, 16, 16)])
def test_corner_turn_u8_aligned_ragged(gpu, T, Cn, pad, invert):
    """uint8 rows with aligned input and 16-byte-aligned output (the
    register-transpose kernels: dword rows with nchans % 4 == 0, 16-byte rows
    with nchans % 16 == 0): spectra counts off the 16-sample and workgroup
    grids, channel counts off the channel blocks; bytes outside the window
    untouched."""
    M = gpu
    print(f"case corner_turn_u8_aligned T={T} C={Cn} pad={pad} invert={invert}", flush=True)
    rng = np.random.default_rng(14)
    in_stride, col0 = Cn + pad, 16
    out_stride = (col0 + T + 15) // 16 * 16 + 32
    x = rng.integers(0, 256, (T, in_stride), dtype=np.uint8)
    din = M.DeviceBuffer.from_host(x)
    dout = M.DeviceBuffer(Cn * out_stride)
    dout.fill_u32(0xFFFFFFFF)
    M.corner_turn(din, "u8", T, Cn, in_stride, dout, out_stride, col0, invert, "u8")
    M.synchronize()
    got = dout.download(np.uint8, (Cn, out_stride))
    exp = x[:, :Cn][:, ::-1].T if invert else x[:, :Cn].T
    assert np.array_equal(got[:, col0:col0 + T], exp)
    assert np.all(got[:, :col0] == 0xFF) and np.all(got[:, col0 + T:] == 0xFF)


def test_corner_turn_full_config2_block(gpu):
    """1024 channels x 2^20 uint8 samples: bytes round-trip exactly (checksum of
    the transposed block equals the oracle's on a slice and in total)."""
    M = gpu
    print("case corner_turn 1024x2^20 u8", flush=True)
    Cn, T = 1024, 1 << 20
    rng = np.random.default_rng(12)
    x = rng.integers(0, 256, (T, Cn), dtype=np.uint8)
    din = M.DeviceBuffer.from_host(x)
    dout = M.DeviceBuffer(Cn * T)
    M.corner_turn(din, "u8", T, Cn, Cn, dout, T, 0, True, "u8")
    M.synchronize()
    got = dout.download(np.uint8, (Cn, T))
    assert np.array_equal(got[:, 5000:5100], x[5000:5100, ::-1].T)
    assert np.array_equal(got.sum(axis=1, dtype=np.uint64), x.sum(axis=0, dtype=np.uint64)[::-1])


@pytest.mark.parametrize("n,seg,pad", [(4096, 1024, 5), (5000, 1024, 5), (777, 64, 5),
                                       (3001, 1000, 5), (5000, 2000, 5),
                                       (8192, 1024, 0), (5000, 1024, 4)])
def test_detrend_matches_oracle(gpu, n, seg, pad):
    """pad 0 or 4: 16-byte aligned rows, whose whole 1024-sample segments
    take the kernel's 16-consecutive-samples-per-lane form."""
    M = gpu
    print(f"case detrend n={n} seg={seg} pad={pad}", flush=True)
    rng = np.random.default_rng(13)
    R, stride = 9, n + pad
    t = np.arange(n)
    x = np.zeros((R, stride), np.float32)
    for r in range(R):
        x[r, :n] = 50 + 0.02 * r * t + 3 * rng.standard_normal(n)
    x[3, :n] = 7.0                               # constant row: zero residual -> zeros
    d = M.DeviceBuffer.from_host(x)
    M.detrend(d, R, n, stride, seg)
    M.synchronize()
    got = d.download(np.float32, (R, stride))
    exp = oracle.detrend_f64(np.ascontiguousarray(x[:, :n]), seg)
    err, ab = rel_err(got[:, :n], exp)
    print(f"detrend max-abs {ab:.3e} rel {err:.3e}")
    assert err <= 1e-6
    assert np.all(got[3, :n] == 0.0)
    assert np.array_equal(got[:, n:], x[:, n:])   # padding untouched (if any)


@pytest.mark.parametrize("aligned", [False, True])
@pytest.mark.parametrize("nbits,out", [(1, "u8"), (2, "u8"), (4, "u8"), (8, "u8"), (1, "f32"),
                                       (4, "f32"), (8, "f32"), (16, "f32"), (32, "f32")])
def test_unpack_spectra_matches_oracle(gpu, nbits, out, aligned):
    """aligned: 16-byte-aligned output rows (the register-transpose kernel
    for uint8 output); otherwise the tiled kernel."""
    M = gpu
    print(f"case unpack nbits={nbits} -> {out} aligned={aligned}", flush=True)
    rng = np.random.default_rng(20 + nbits)
    T, Cn = 700, 192
    sb = Cn * nbits // 8
    stride = sb + 16                       # padded spectra
    raw = rng.integers(0, 256, (T, stride), dtype=np.uint8)
    if nbits == 32:                        # finite floats
        raw[:, :sb] = np.frombuffer(rng.standard_normal(T * Cn).astype(np.float32).tobytes(),
                                    np.uint8).reshape(T, sb)
    exp = np.stack([oracle.unpack_bits(raw[t, :sb].tobytes(), nbits, Cn) for t in range(T)])
    for invert in (False, True):
        d_raw = M.DeviceBuffer.from_host(raw)
        col0, ostride = (16, 736) if aligned else (5, T + 9)
        npd = np.uint8 if out == "u8" else np.float32
        d_out = M.DeviceBuffer(Cn * ostride * np.dtype(npd).itemsize)
        M.unpack_spectra(d_raw, nbits, T, Cn, stride, d_out, ostride, col0, invert, out)
        M.synchronize()
        got = d_out.download(npd, (Cn, ostride))[:, col0:col0 + T].astype(np.float32)
        ref = exp.T[::-1] if invert else exp.T
        assert np.array_equal(got, ref), f"nbits {nbits} invert {invert}"
