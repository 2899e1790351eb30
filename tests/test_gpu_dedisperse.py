"""GPU parity: dedispersion through the C ABI against the oracle.

Bars: uint8 input -> bit-exact (integer sums below 2^24 are exact in float32
in any order); float16/float32 input -> normwise relative error
max|gpu - fp64| / max|fp64| <= 1e-5 (the path's stated tolerance), with the
max-abs error printed per case.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _plan(M, C, S, dtype, ranges, fch1=1500.0, foff=-0.390625, tsamp=64e-6, rule=0):
    return M.Plan.brute(C, fch1, foff, tsamp, S, dtype, ranges, shift_rule=rule)


def _check_range(plan, r, x, got, dtype):
    R = plan.ranges[r]
    sh = plan.shift_table(r)
    ref = oracle.range_f64(x, R.downsample, sh, R.nout)
    if dtype == "u8":
        exact = oracle.range_f32(x, R.downsample, sh, R.nout)
        assert np.array_equal(got, exact), f"range {r}: uint8 path not bit-exact"
        assert np.array_equal(got.astype(np.float64), ref.astype(np.float32).astype(np.float64))
    else:
        e, a = rel_err(got, ref)
        print(f"range {r} dtype {dtype}: rel {e:.3e} maxabs {a:.3e}")
        assert e <= TOL


@pytest.mark.parametrize("dtype", ["u8", "f16", "f32"])
def test_single_range_small(gpu, dtype):
    M = gpu
    C, S = 64, 3000
    print(f"case single_range_small dtype={dtype} C={C} S={S}", flush=True)
    if dtype == "u8":
        x = M.synth_filterbank(C, S, "u8", seed=11, mean=128, sigma=16)
    else:
        x = M.synth_filterbank(C, S, dtype, seed=12, mean=0.0, sigma=1.0)
    plan = _plan(M, C, S, dtype, [(0.0, 1.5, 70, 1)])
    outs = plan.dedisperse_host(x)
    _check_range(plan, 0, x, outs[0], dtype)


@pytest.mark.parametrize("C,S,dtype,ranges", [
    (1, 17, "u8", [(0.0, 1.0, 1, 1)]),                     # one channel, one trial, one short tile
    (2, 300, "f32", [(0.0, 3.0, 1, 1)]),
    (3, 257, "f16", [(0.0, 2.0, 5, 2)]),                   # decimated, 126 outputs
    (7, 1000, "u8", [(0.0, 1.0, 65, 1)]),                  # a second block holding one trial
    (1, 5000, "f32", [(0.0, 1.0, 3, 4)]),
    (5, 700, "u8", [(0.0, 0.5, 9, 1), (4.5, 1.0, 3, 16)]),  # 16x decimated uint16 rows, 43 outputs
])
def test_tiny_shapes(gpu, C, S, dtype, ranges):
    """Shapes below one tile, one channel chunk and one trial block."""
    M = gpu
    print(f"case tiny C={C} S={S} dtype={dtype} ranges={ranges}", flush=True)
    rng = np.random.default_rng(C * 1000 + S)
    if dtype == "u8":
        x = rng.integers(0, 256, (C, S), dtype=np.uint8)
    else:
        x = rng.standard_normal((C, S)).astype(np.float16 if dtype == "f16" else np.float32)
    foff = -25.0 / C
    plan = _plan(M, C, S, dtype, ranges, foff=foff)
    outs = plan.dedisperse_host(x)
    for r in range(len(ranges)):
        assert outs[r].shape == (ranges[r][2], plan.ranges[r].nout)
        _check_range(plan, r, x, outs[r], dtype)


def test_const_per_channel_kat(gpu):
    """x[c][t] = c gives every output C(C-1)/2 (src/prototypes/brute_dedisp_imp_tex.cu:112-114,167-175)."""
    M = gpu
    C, S = 128, 5000
    print("case const_kat", flush=True)
    x = np.repeat(np.arange(C, dtype=np.float32)[:, None], S, axis=1)
    plan = _plan(M, C, S, "f32", [(0.0, 2.0, 40, 1)])
    out = plan.dedisperse_host(x)[0]
    assert np.all(out == C * (C - 1) / 2)


def test_impulse_reads_every_shift(gpu):
    """x = delta(c0, t0): out[d][t] = 1 iff t = t0 - shift[d][c0] (end-to-end table check)."""
    M = gpu
    C, S = 96, 4096
    print("case impulse", flush=True)
    plan = _plan(M, C, S, "f32", [(0.0, 3.0, 64, 1)])
    sh = plan.shift_table(0)
    N = plan.ranges[0].nout
    for c0 in (0, 37, C - 1):
        x = np.zeros((C, S), np.float32)
        t0 = 2000
        x[c0, t0] = 1.0
        out = plan.dedisperse_host(x)[0]
        exp = np.zeros_like(out)
        for d in range(sh.shape[0]):
            t = t0 - sh[d, c0]
            if 0 <= t < N:
                exp[d, t] = 1.0
        assert np.array_equal(out, exp), f"impulse at channel {c0}"


def test_multi_range_decimation_u8(gpu):
    M = gpu
    C, S = 64, 8192
    print("case multi_range_decimation_u8", flush=True)
    x = M.synth_filterbank(C, S, "u8", seed=21, mean=128, sigma=16,
                           pulses=[(120.0, 3000, 8, 4.0)])
    plan = _plan(M, C, S, "u8", [(0.0, 1.0, 40, 1), (40.0, 2.0, 30, 2), (100.0, 4.0, 20, 4),
                                 (180.0, 5.0, 9, 3)])
    outs = plan.dedisperse_host(x)
    for r in range(4):
        _check_range(plan, r, x, outs[r], "u8")


@pytest.mark.parametrize("chunk_rows", [2, 3, 5, 7, 8])
def test_u8_chunk_shapes_exact(gpu, chunk_rows):
    """uint8 and decimated uint16 rows with every chunk shape the offset
    registers meet: odd chunks (a lone last channel), a channel count that
    is no multiple of the chunk, partial trial blocks; bit-exact."""
    M = gpu
    C, S = 75, 9000
    print(f"case u8_chunk_shapes chunk_rows={chunk_rows}", flush=True)
    x = M.synth_filterbank(C, S, "u8", seed=23, mean=128, sigma=16, foff=-300.0 / C,
                           pulses=[(60.0, 4000, 4, 4.0)])
    plan = _plan(M, C, S, "u8", [(0.0, 0.75, 100, 1), (75.0, 1.5, 41, 2)], foff=-300.0 / C)
    t = plan.tuning()
    t["chunk_rows"] = chunk_rows
    plan.set_tuning(**t)
    outs = plan.dedisperse_host(x)
    for r in range(2):
        _check_range(plan, r, x, outs[r], "u8")


@pytest.mark.parametrize("pad", [0, 5])
def test_decimation_pyramid_u8(gpu, pad):
    """uint8 ranges at k = 2, 4, 8, 16 (one fused decimation pass) and k = 3
    (its own pass), a block length that is no multiple of 16, and rows that
    are (pad 0: stride 10016) or are not (pad 5) 16-byte aligned."""
    M = gpu
    C, S = 40, 10011
    print(f"case decimation_pyramid_u8 pad={pad}", flush=True)
    x = M.synth_filterbank(C, 10016 + pad, "u8", seed=23, mean=128, sigma=40)
    plan = _plan(M, C, S, "u8", [(0.0, 2.0, 20, 1), (40.0, 3.0, 12, 2), (80.0, 4.0, 10, 4),
                                 (120.0, 6.0, 9, 8), (180.0, 8.0, 7, 16), (240.0, 9.0, 5, 3)])
    outs = plan.dedisperse_host(x)
    for r in range(6):
        _check_range(plan, r, x[:, :S], outs[r], "u8")


@pytest.mark.parametrize("dtype", ["f16", "f32"])
def test_multi_range_decimation_float(gpu, dtype):
    M = gpu
    C, S = 48, 6000
    print(f"case multi_range_decimation_float {dtype}", flush=True)
    x = M.synth_filterbank(C, S, dtype, seed=22, gain=(0.5, 2.0))
    plan = _plan(M, C, S, dtype, [(0.0, 2.0, 33, 1), (66.0, 4.0, 17, 2), (130.0, 6.0, 11, 8)])
    outs = plan.dedisperse_host(x)
    for r in range(3):
        _check_range(plan, r, x, outs[r], dtype)


def test_ragged_stride_and_ns_rule(gpu):
    """Row stride > nsamp, odd sizes, the NS shift rule."""
    M = gpu
    C, S, stride = 33, 3001, 3333
    print("case ragged_stride_ns", flush=True)
    x = M.synth_filterbank(C, S, "u8", seed=5, mean=100, sigma=20, stride=stride)
    plan = _plan(M, C, S, "u8", [(3.0, 1.7, 45, 1)], rule=1)
    sh = plan.shift_table(0)
    ns = oracle.shift_table_ns(1500.0, -0.390625, C, np.float32(3.0), np.float32(1.7), 45, np.float32(64e-6))
    assert np.array_equal(sh, ns)
    din = M.DeviceBuffer.from_host(x)
    dout = M.DeviceBuffer(plan.info.out_elems * 4)
    plan.dedisperse(din, stride, dout)
    M.synchronize()
    got = dout.download(np.float32, (plan.info.out_elems,)).reshape(45, -1)
    exact = oracle.range_f32(np.ascontiguousarray(x[:, :S]), 1, sh, plan.ranges[0].nout)
    assert np.array_equal(got, exact)


def test_ragged_stride_float16(gpu):
    """float16 rows whose starts are 2 bytes off a dword (odd row stride) and
    an odd row length: the two-plane staging's per-sample fetch at row starts
    and ends, odd and even shifts, windows wider than one tile."""
    M = gpu
    C, S, stride = 33, 3001, 3333
    print("case ragged_stride_float16", flush=True)
    x = M.synth_filterbank(C, S, "f16", seed=6, gain=(0.5, 2.0), stride=stride,
                           foff=-300.0 / C, pulses=[(40.0, 500, 3, 2.0)])
    plan = _plan(M, C, S, "f16", [(3.0, 1.7, 45, 1)], foff=-300.0 / C, rule=1)
    sh = plan.shift_table(0)
    assert (sh % 2 == 1).any() and (sh % 2 == 0).any() and sh.max() > 256
    din = M.DeviceBuffer.from_host(x)
    dout = M.DeviceBuffer(plan.info.out_elems * 4)
    plan.dedisperse(din, stride, dout)
    M.synchronize()
    got = dout.download(np.float32, (plan.info.out_elems,)).reshape(45, -1)
    ref = oracle.range_f64(np.ascontiguousarray(x[:, :S]), 1, sh, plan.ranges[0].nout)
    e, a = rel_err(got, ref)
    print(f"ragged float16 rel {e:.3e} maxabs {a:.3e}")
    assert e <= TOL


def test_config1_full_size(gpu):
    """Config 1 at full size (256 ch, 65536 samples, 512 DMs, f32, pulse at DM 56.8):
    all DM rows compared to the FP64 oracle on a subset of 64 rows."""
    M = gpu
    C, S = 256, 65536
    print("case config1_full", flush=True)
    x = M.synth_filterbank(C, S, "f32", seed=1, pulses=[(56.8, 30000, 4, 0.5)])
    plan = _plan(M, C, S, "f32", [(0.0, 0.2, 512, 1)])
    assert plan.ranges[0].maxshift == 433 and plan.ranges[0].nout == 65103
    out = plan.dedisperse_host(x)[0]
    sh = plan.shift_table(0)
    rows = np.arange(0, 512, 8)
    ref = oracle.dedisperse_f64(x, sh[rows], plan.ranges[0].nout)
    e, a = rel_err(out[rows], ref)
    print(f"config1 rel {e:.3e} maxabs {a:.3e}")
    assert e <= TOL
    # the injected pulse is the brightest point near DM 56.8 (row 284)
    d, t = np.unravel_index(np.argmax(out), out.shape)
    assert abs(d - 284) <= 3 and abs(t - 30000) <= 4


def _check_integer_plan_full(M, plan, x, nchk=4096):
    """uint8 plan at full size: (a) every DM row's total over its N_r output
    samples equals, exactly, sum_c P_c[N_r + s_dc] - P_c[s_dc] with P_c the
    prefix sums of channel c of the (k-summed) input -- a checksum of every
    row that moves if any shift, window or row is wrong; (b) the first nchk
    samples of the first, middle and last row of each range are bit-exact
    against the oracle run on the window they depend on."""
    C, S = x.shape[0], plan.obs.nsamp_in
    din = M.DeviceBuffer.from_host(x)
    dout = M.DeviceBuffer(plan.info.out_elems * 4)
    dw = M.DeviceBuffer(max(plan.info.work_bytes, 256))
    plan.dedisperse(din, x.shape[1], dout, dw)
    M.synchronize()
    for r, R in enumerate(plan.ranges):
        k, n = R.downsample, S // R.downsample
        sh = plan.shift_table(r)
        tot = np.zeros(R.ndms, np.int64)
        for c in range(C):
            xc = x[c, :n * k].reshape(n, k).sum(axis=1, dtype=np.int64)
            P = np.concatenate(([0], np.cumsum(xc)))
            tot += P[R.nout + sh[:, c]] - P[sh[:, c]]
        got = dout.download(np.float32, (R.ndms, R.nout), offset=R.out_offset * 4)
        gsum = np.add.reduce(got, axis=1, dtype=np.float64) * k
        bad = np.flatnonzero(gsum != tot.astype(np.float64))
        print(f"range {r}: k {k} ndms {R.ndms} nout {R.nout} maxshift {R.maxshift} "
              f"row totals differing {bad.size}", flush=True)
        assert bad.size == 0, f"range {r}: rows {bad[:8]} differ"
        rows = [0, R.ndms // 2, R.ndms - 1]
        win = np.ascontiguousarray(x[:, :(R.maxshift + nchk) * k])
        ref = oracle.range_f32(win, k, sh[rows], nchk)
        assert np.array_equal(got[rows, :nchk], ref), f"range {r}: sampled rows not bit-exact"
        del got


def test_config2_full_size(gpu):
    """Config 2 at full size: 1024 ch x 2^20 uint8 with injected pulses,
    ranges 0-200/0.2 (k 1), 200-500/0.5 (k 2), 500-1000/1.0 (k 4)."""
    import bench
    M = gpu
    print("case config2 full size", flush=True)
    w = bench.WORKLOADS["cfg2"]
    x, _ = bench.make_input(M, w)
    _check_integer_plan_full(M, bench.make_plan(M, w), x)


def test_config5_block_full_size(gpu):
    """One config-5 block at full size: 1024 ch x (65536 + 16176) uint8,
    2048 DMs over 0-1000."""
    import bench
    M = gpu
    print("case config5 block full size", flush=True)
    w = bench.WORKLOADS["cfg5"]
    S = w["nsamp"] + 16176
    x = M.synth_filterbank(w["nchans"], S, "u8", seed=55, mean=128, sigma=16,
                           pulses=[(400.0, 30000, 8, 3.0)])
    plan = M.Plan.brute(w["nchans"], w["fch1"], w["foff"], w["tsamp"], S, "u8", w["ranges"])
    assert plan.ranges[0].maxshift == 16176 and plan.ranges[0].nout == w["nsamp"]
    _check_integer_plan_full(M, plan, x)


def test_plain_kernel_equals_staged(gpu):
    """The reference-shaped plain float32 kernel sums in the same channel
    order as the staged path: identical outputs."""
    M = gpu
    print("case plain kernel", flush=True)
    C, S = 256, 20000
    x = M.synth_filterbank(C, S, "f32", seed=71, pulses=[(30.0, 5000, 4, 1.0)])
    p = M.Plan.brute(C, 1500.0, -0.390625, 64e-6, S, "f32", [(0.0, 0.5, 70, 1), (35.0, 1.0, 50, 1)])
    din = M.DeviceBuffer.from_host(x)
    a = M.DeviceBuffer(p.info.out_elems * 4)
    b = M.DeviceBuffer(p.info.out_elems * 4)
    p.dedisperse(din, S, a)
    p.dedisperse_plain(din, S, b)
    M.synchronize()
    assert np.array_equal(a.download(np.float32, (p.info.out_elems,)),
                          b.download(np.float32, (p.info.out_elems,)))
