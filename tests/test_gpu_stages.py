"""GPU parity of the other stages: two-stage subband dedispersion, the uint8
RFI clipper, the boxcar detector and the median-of-five detector, each
through the C ABI against the oracle."""
import numpy as np
import pytest

import oracle
from conftest import rel_err

pytestmark = pytest.mark.gpu
WIDTHS = [1, 2, 4, 8, 16, 32, 64]


def _cand_key(a):
    return sorted(zip(a["dm"].tolist(), a["t"].tolist(), a["width"].tolist()))


# ---------------------------------------------------------------- subband
@pytest.mark.parametrize("dtype,k", [("u8", 1), ("u8", 2), ("f16", 1), ("f32", 3)])
def test_subband_matches_oracle(gpu, dtype, k):
    M = gpu
    C, S, nsubs, ncalls, calldms, dmstep = 128, 12000, 16, 6, 20, 0.6
    print(f"case subband dtype={dtype} k={k}", flush=True)
    if dtype == "u8":
        x = M.synth_filterbank(C, S, "u8", seed=31, mean=128, sigma=16,
                               pulses=[(40.0, 4000, 6, 3.0)])
    else:
        x = M.synth_filterbank(C, S, dtype, seed=32, gain=(0.5, 2.0),
                               pulses=[(40.0, 4000, 6, 3.0)])
    p = M.Plan.subband(C, 1500.0, -0.78125, 64e-6, S, nsubs,
                       [dict(lowdm=calldms * dmstep / 2, dmstep=dmstep,
                             sub_dmstep=dmstep * calldms, ncalls=ncalls, calldms=calldms,
                             binsize=k)], dtype=dtype)
    out = p.dedisperse_host(x)[0]
    co = p.shift_table(0)
    fi = p.fine_table(0, nsubs)
    R = p.ranges[0]
    cps = C // nsubs
    M2 = int(fi.max())
    ref = oracle.subband_f64(x, k, nsubs, ncalls, calldms, co, fi, R.nout + M2, R.nout)
    if dtype == "u8":
        assert np.array_equal(out.astype(np.float64), ref.astype(np.float32).astype(np.float64))
    else:
        e, a = rel_err(out, ref)
        print(f"subband {dtype} k={k}: rel {e:.3e} maxabs {a:.3e}")
        assert e <= 1e-5


def test_subband_one_channel_per_subband_equals_brute(gpu):
    M = gpu
    C, S, calldms, ncalls, dmstep = 32, 5000, 8, 4, 1.5
    print("case subband_reduces_to_brute", flush=True)
    x = M.synth_filterbank(C, S, "u8", seed=33, mean=128, sigma=16)
    p = M.Plan.subband(C, 1500.0, -0.78125, 64e-6, S, C,
                       [dict(lowdm=calldms * dmstep / 2, dmstep=dmstep, sub_dmstep=dmstep * calldms,
                             ncalls=ncalls, calldms=calldms)], dtype="u8")
    out = p.dedisperse_host(x)[0]
    fi = p.fine_table(0, C)
    assert np.array_equal(out, oracle.dedisperse_f32(x, fi, p.ranges[0].nout))


# ---------------------------------------------------------------- clipper
@pytest.mark.parametrize("S,t0,n", [(70000, 1234, 65536), (65568, 16, 65531), (65536, 0, 65536)])
def test_clip_u8_matches_oracle(gpu, S, t0, n):
    """Unaligned rows (per-sample passes) and 16-byte-aligned rows (vector
    passes, with and without a ragged tail)."""
    M = gpu
    C = 64
    print(f"case clip_u8 S={S} t0={t0} n={n}", flush=True)
    x = M.synth_filterbank(C, S, "u8", seed=41, mean=128, sigma=16)
    rng = np.random.default_rng(41)
    x[3, rng.integers(t0, t0 + n, 200)] = 255          # impulsive RFI
    x[7, t0:t0 + n] = 90                                # constant channel
    x[9, t0 + 100:t0 + 5000] = 250                      # long burst
    din = M.DeviceBuffer.from_host(x)
    dout = M.DeviceBuffer.from_host(x)
    dmed, dmad, dncl = (M.DeviceBuffer(C * 4) for _ in range(3))
    M.clip_u8(din, S, C, t0, n, 5.0, dout, S, dmed, dmad, dncl)
    M.synchronize()
    got = dout.download(np.uint8, (C, S))
    out, med, mad, ncl = oracle.clip_u8(x, t0, n, 5.0)
    assert np.array_equal(got[:, t0:t0 + n], out)
    assert np.array_equal(got[:, :t0], x[:, :t0]) and np.array_equal(got[:, t0 + n:], x[:, t0 + n:])
    assert np.array_equal(dmed.download(np.int32, (C,)), med)
    assert np.array_equal(dmad.download(np.int32, (C,)), mad)
    assert np.array_equal(dncl.download(np.int32, (C,)), ncl)
    assert ncl[3] >= 200 and ncl[9] > 4000


def test_clip_in_place_odd_length(gpu):
    M = gpu
    C, S = 5, 1001
    print("case clip_in_place", flush=True)
    x = np.random.default_rng(3).integers(0, 256, (C, S), dtype=np.uint8)
    d = M.DeviceBuffer.from_host(x)
    M.clip_u8(d, S, C, 0, S, 2.0, d, S)
    M.synchronize()
    out, *_ = oracle.clip_u8(x, 0, S, 2.0)
    assert np.array_equal(d.download(np.uint8, (C, S)), out)


# ---------------------------------------------------------------- boxcar
def _run_boxcar(M, x, thr, widths=WIDTHS, cap=1 << 16):
    D, n = x.shape
    dx = M.DeviceBuffer.from_host(x)
    dc = M.DeviceBuffer(cap * 16)
    dn = M.DeviceBuffer(4)
    dm = M.DeviceBuffer(D * 8)
    ds = M.DeviceBuffer(D * 8)
    M.detect_boxcar(dx, D, n, n, widths, thr, dc, cap, dn, dm, ds)
    M.synchronize()
    return (M.read_candidates(dc, dn, cap), dm.download(np.float64, (D,)),
            ds.download(np.float64, (D,)))


def test_boxcar_integer_data_exact(gpu):
    """Integer-valued rows (as uint8 dedispersion produces): candidate lists
    and statistics identical to the oracle."""
    M = gpu
    print("case boxcar_integer", flush=True)
    rng = np.random.default_rng(51)
    D, n = 96, 20000
    x = rng.integers(0, 4000, (D, n)).astype(np.float32) + 100000
    for d, t, w, a in [(3, 500, 1, 9000), (10, 7000, 8, 3000), (50, 19990, 4, 5000),
                       (60, 0, 64, 900), (95, 12345, 32, 1400)]:
        x[d, t:t + w] += a
    got, mean, sig = _run_boxcar(M, x, 6.0)
    ref, rm, rs = oracle.boxcar_detect(x, WIDTHS, 6.0)
    assert np.array_equal(mean, rm) and np.array_equal(sig, rs)
    assert _cand_key(got) == _cand_key(ref)
    assert len(ref) >= 5
    np.testing.assert_array_equal(np.sort(got, order=["dm", "width", "t"])["snr"],
                                  np.sort(ref, order=["dm", "width", "t"])["snr"])


def test_boxcar_float_data_within_tie_band(gpu):
    M = gpu
    print("case boxcar_float", flush=True)
    rng = np.random.default_rng(52)
    D, n = 40, 30000
    x = rng.standard_normal((D, n)).astype(np.float32)
    x[5, 1000:1016] += 2.0
    x[30, 29000:29002] += 6.0
    got, mean, sig = _run_boxcar(M, x, 5.0)
    ref, rm, rs = oracle.boxcar_detect(x, WIDTHS, 5.0)
    assert np.allclose(mean, rm, rtol=0, atol=1e-12) and np.allclose(sig, rs, rtol=1e-12)
    band = 1e-6
    keep = lambda a: a[np.abs(a["snr"] - 5.0) > band]
    assert _cand_key(keep(got)) == _cand_key(keep(ref))


def test_boxcar_capacity_overflow_reported(gpu):
    M = gpu
    print("case boxcar_overflow", flush=True)
    x = np.zeros((4, 5000), np.float32)
    x[:, ::50] = 100.0
    with pytest.raises(M.MdsmError):
        _run_boxcar(M, x, 3.0, cap=10)


# ---------------------------------------------------------------- median of five
def _np_median5(x, dmv, thr):
    """Whole-block statistics by the reference output thread's chunked rule
    (oracle.brute_stats, pinned to the reference's process_brute), then the
    median-of-five threshold in float."""
    import oracle
    lm, lr = oracle.brute_stats(x)
    res = []
    for d in range(x.shape[0]):
        if not dmv[d] > 1.0:
            continue
        w = np.lib.stride_tricks.sliding_window_view(x[d], 5)
        med = np.sort(w, axis=1)[:, 2]
        hit = (med - lm) >= lr * np.float32(thr)
        for l in np.nonzero(hit)[0]:
            res.append((d, int(l), 5))
    return res, lm, lr


@pytest.mark.parametrize("n,stride", [(4000, 4000), (4001, 4003), (4003, 4004), (4005, 4008)])
def test_median5_detector(gpu, n, stride):
    """Aligned rows (16-byte loads) and unaligned or ragged ones."""
    M = gpu
    print(f"case median5 n={n} stride={stride}", flush=True)
    rng = np.random.default_rng(61)
    D = 30
    x = rng.integers(0, 100, (D, n)).astype(np.float32)
    x[4, 100:110] += 400
    x[5, n - 7:n] += 400        # a burst in the last window
    x[0, 50:60] += 400          # DM 0 row: excluded by dm > 1
    dmv = np.arange(D, dtype=np.float32) * 0.5
    xs = np.zeros((D, stride), np.float32)
    xs[:, :n] = x
    dx = M.DeviceBuffer.from_host(xs)
    ddm = M.DeviceBuffer.from_host(dmv)
    cap = 4096
    dc, dn = M.DeviceBuffer(cap * 16), M.DeviceBuffer(4)
    dst = M.DeviceBuffer(M.median5_stats_len(D, n) * 8)      # 4 chunks of 32768: 3 read
    M.detect_median5(dx, D, n, stride, ddm, 3.0, dc, cap, dn, dst)
    M.synchronize()
    got = M.read_candidates(dc, dn, cap)
    exp, lm, lr = _np_median5(x, dmv, 3.0)
    assert _cand_key(got) == sorted(exp) and len(exp) > 0
    st = dst.download(np.float64, (2,))
    assert np.float32(st[0]) == lm and np.float32(st[1]) == lr
