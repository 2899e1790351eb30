"""Pins the oracle (CPU) against the reference's own code, golden fixtures,
closed forms and invariants.  No GPU needed."""
import hashlib
import itertools
import json
import math
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
needs_ref = pytest.mark.skipif(oracle.ref() is None, reason="oracle/_ref not built here")


def _digests():
    with open(os.path.join(GOLD, "ref_shift_digests.json")) as f:
        return json.load(f)


# ---------------------------------------------------------------- shift tables
@pytest.mark.parametrize("name", sorted(_digests()))
def test_ref_shift_table_matches_golden(name):
    """Oracle REF table == the table the reference's cpu_dedisp.c main() computed."""
    d = _digests()[name]
    sh = oracle.shift_table_ref(d["fch1"], d["foff"], d["nchans"], d["startdm"], d["dmstep"],
                                d["ndms"], d["tsamp"])
    assert list(sh.shape) == d["shape"]
    assert int(sh[-1, -1]) == d["maxshift"]
    assert sh.reshape(-1)[d["sample_idx"]].tolist() == d["sample_val"]
    assert int(sh.sum(dtype=np.int64)) == d["sum"]
    assert hashlib.sha256(np.ascontiguousarray(sh, np.int32).tobytes()).hexdigest() == d["sha256"]


@needs_ref
@pytest.mark.parametrize("case", [(1500.0, -0.390625, 64e-6, 256, 0.0, 0.2, 512),
                                  (1400.0, -0.5, 100e-6, 128, 3.0, 0.7, 200),
                                  (1500.0, -300.0 / 1024, 256e-6, 1024, 500.0, 1.0, 500)])
def test_ref_shift_table_matches_reference_build(case):
    fch1, foff, tsamp, C, lo, step, nd = case
    shr, _, _ = oracle.ref_cpu_dedisp(fch1, foff, tsamp, C, 1, lo, step, nd)
    sho = oracle.shift_table_ref(fch1, foff, C, lo, step, nd, tsamp)
    assert np.array_equal(shr, sho)


def test_ref_rule_disagreement_counts():
    """Regression numbers from an independent probe of the same formulas:
    REF (float32, truncated) differs from FP64-truncated in 2 entries of the
    config-1 table, and the 4148.741601 vs 4148.808 constants move 194
    nearest-rounded entries."""
    C, nd = 256, 512
    sh = oracle.shift_table_ref(1500.0, -0.390625, C, 0.0, 0.2, nd, 64e-6)
    f = 1500.0 - 0.390625 * np.arange(C)
    dm = 0.2 * np.arange(nd)
    x = 4148.741601 * (1 / f**2 - 1 / 1500.0**2)[None, :] * dm[:, None] / 64e-6
    assert int((np.floor(x) != sh).sum()) == 2
    y = 4148.808 * (1 / f**2 - 1 / 1500.0**2)[None, :] * dm[:, None] / 64e-6
    assert int((np.round(x) != np.round(y)).sum()) == 194


def test_ns_rule_closed_form():
    C, nd = 64, 50
    sh = oracle.shift_table_ns(1500.0, -1.0, C, 1.0, 3.0, nd, 64e-6)
    f = 1500.0 - np.arange(C, dtype=np.float64)
    dm = 1.0 + 3.0 * np.arange(nd)
    exp = np.array([[round(4148.808 * d * (1 / fc**2 - 1 / 1500.0**2) / 64e-6) for fc in f]
                    for d in dm])
    assert np.array_equal(sh, exp)
    assert np.all(np.diff(sh, axis=0) >= 0) and np.all(np.diff(sh, axis=1) >= 0)


# ---------------------------------------------------------------- brute force
def test_dedisperse_matches_reference_loop_golden():
    """Oracle loop == the reference's own loop output (fixture)."""
    g = np.load(os.path.join(GOLD, "ref_cpu_dedisp_small.npz"))
    C, N, nd = (int(v) for v in g["dims"])
    out = oracle.dedisperse_ref_tm(g["x_tm"], g["shift"], N)
    assert np.array_equal(out, g["out_tm"])
    # channel-major restatement, same float order
    out_cm = oracle.dedisperse_f32(np.ascontiguousarray(g["x_tm"].T), g["shift"], N)
    assert np.array_equal(out_cm, g["out_tm"].T)
    # the multi-threaded loop (the bench's all-cores CPU figure), same sums
    for nthreads in (1, 3, 8):
        assert np.array_equal(oracle.dedisperse_ref_tm_mt(g["x_tm"], g["shift"], N, nthreads),
                              g["out_tm"])


@needs_ref
def test_dedisperse_matches_reference_build_fresh():
    C, N, nd = 32, 777, 19
    sh, _, _ = oracle.ref_cpu_dedisp(1450.0, -2.0, 64e-6, C, N, 1.0, 3.3, nd)
    x_tm = np.random.default_rng(5).standard_normal((N + int(sh[-1, -1]), C)).astype(np.float32)
    _, out_ref, _ = oracle.ref_cpu_dedisp(1450.0, -2.0, 64e-6, C, N, 1.0, 3.3, nd, x_tm)
    assert np.array_equal(oracle.dedisperse_ref_tm(x_tm, sh, N), out_ref)


def test_reference_constant_input_kat():
    """cpu_dedisp.c's own input is a 0.1 constant: every output is the float sum of C x 0.1."""
    g = np.load(os.path.join(GOLD, "ref_cpu_dedisp_small.npz"))
    C, N, nd = (int(v) for v in g["dims"])
    x = np.full((N + int(g["shift"][-1, -1]), C), 0.1, np.float32)
    out = oracle.dedisperse_ref_tm(x, g["shift"], N)
    exp = np.float32(0)
    for _ in range(C):
        exp = np.float32(exp + np.float32(0.1))
    assert np.all(out == exp)


def test_constant_per_channel_kat():
    """x[c][t] = c gives C(C-1)/2 everywhere (src/prototypes/brute_dedisp_imp_tex.cu:112-114,167-175)."""
    C, S = 48, 900
    sh = oracle.shift_table_ref(1500.0, -1.0, C, 0.0, 5.0, 30, 64e-6)
    x = np.repeat(np.arange(C, dtype=np.float32)[:, None], S, 1)
    N = S - int(sh.max())
    assert np.all(oracle.dedisperse_f32(x, sh, N) == C * (C - 1) / 2)
    assert np.all(oracle.dedisperse_f64(x, sh, N) == C * (C - 1) / 2)


def test_impulse_and_dm0_and_linearity():
    C, S = 40, 700
    sh = oracle.shift_table_ref(1500.0, -1.0, C, 0.0, 4.0, 25, 64e-6)
    N = S - int(sh.max())
    x = np.zeros((C, S), np.float32)
    x[17, 400] = 1.0
    out = oracle.dedisperse_f64(x, sh, N)
    for d in range(sh.shape[0]):
        row = np.zeros(N)
        t = 400 - sh[d, 17]
        if 0 <= t < N:
            row[t] = 1.0
        assert np.array_equal(out[d], row)
    rng = np.random.default_rng(3)
    a = rng.integers(-1000, 1000, (C, S)).astype(np.float32)
    b = rng.integers(-1000, 1000, (C, S)).astype(np.float32)
    oa, ob, oab = (oracle.dedisperse_f64(v, sh, N) for v in (a, b, 2 * a - 3 * b))
    assert np.array_equal(oab, 2 * oa - 3 * ob)
    # DM 0 row: shifts are all zero -> plain channel sum
    assert np.all(sh[0] == 0)
    assert np.array_equal(oa[0], a[:, :N].astype(np.float64).sum(0))


def test_uint8_exact_in_float32():
    C, S = 300, 800
    x = np.random.default_rng(8).integers(0, 256, (C, S), dtype=np.uint8)
    sh = oracle.shift_table_ref(1500.0, -0.5, C, 0.0, 3.0, 20, 64e-6)
    N = S - int(sh.max())
    assert np.array_equal(oracle.dedisperse_f32(x, sh, N).astype(np.float64),
                          oracle.dedisperse_f64(x, sh, N))


# ---------------------------------------------------------------- decimation / ranges
def test_decimation_closed_form():
    x = np.random.default_rng(2).integers(0, 256, (5, 103), dtype=np.uint8)
    for k in (1, 2, 3, 4, 7):
        d = oracle.decimate_f64(x, k)
        n = 103 // k
        exp = x[:, :n * k].reshape(5, n, k).sum(-1)
        assert np.array_equal(d, exp)


def test_range_k1_is_brute_force_and_k_scales():
    C, S = 32, 1200
    x = np.random.default_rng(4).integers(0, 256, (C, S), dtype=np.uint8)
    sh = oracle.shift_table_ref(1500.0, -1.0, C, 0.0, 3.0, 12, 64e-6)
    N = S - int(sh.max())
    assert np.array_equal(oracle.range_f32(x, 1, sh, N), oracle.dedisperse_f32(x, sh, N))
    # k = 2: brute force on the explicit 2-sample sums, divided by 2
    sh2 = oracle.shift_table_ref(1500.0, -1.0, C, 0.0, 3.0, 12, np.float32(64e-6) * 2)
    dec = x[:, :S // 2 * 2].reshape(C, S // 2, 2).sum(-1).astype(np.float32)
    N2 = S // 2 - int(sh2.max())
    exp = oracle.dedisperse_f64(dec, sh2, N2) / 2
    assert np.array_equal(oracle.range_f64(x, 2, sh2, N2), exp)
    assert np.array_equal(oracle.range_f32(x, 2, sh2, N2).astype(np.float64), exp)


# ---------------------------------------------------------------- subband
def _sub_setup(C, nsubs, calldms, ncalls, dmstep, k=1):
    fch1, foff, tsamp = 1500.0, -1.5, 64e-6
    co, fi = oracle.subband_tables(fch1, foff, C, nsubs, 0.0, dmstep * calldms, dmstep, ncalls,
                                   calldms, np.float32(tsamp) * np.float32(k))
    cps = C // nsubs
    rel = co - co[:, (np.arange(C) // cps) * cps]
    M1, M2 = int(rel.max()), int(fi.max())
    return co, fi, M1, M2


def test_subband_one_channel_per_subband_is_brute_force():
    C, ncalls, calldms = 24, 3, 5
    co, fi, M1, M2 = _sub_setup(C, C, calldms, ncalls, 2.0)
    assert M1 == 0
    S = 900
    N = S - M1 - M2
    x = np.random.default_rng(6).integers(0, 256, (C, S), dtype=np.uint8)
    out = oracle.subband_f64(x, 1, C, ncalls, calldms, co, fi, N + M2, N)
    # fine table is the full brute-force shift table at the fine DMs
    assert np.array_equal(out, oracle.dedisperse_f64(x, fi, N))


def test_subband_tiny_hand_case():
    """2 subbands x 2 channels, one call, one fine DM, explicit shifts."""
    x = np.arange(4 * 20, dtype=np.float32).reshape(4, 20)
    coarse = np.array([[0, 1, 3, 5]], np.int32)       # rel: [0,1 | 0,2]
    fine = np.array([[0, 2]], np.int32)
    len1, nout = 10, 6
    out = oracle.subband_f64(x, 1, 2, 1, 1, coarse, fine, len1, nout)
    sb0 = x[0, 0:10] + x[1, 1:11]
    sb1 = x[2, 0:10] + x[3, 2:12]
    exp = sb0[0:6] + sb1[2:8]
    assert np.array_equal(out[0], exp)


def test_subband_tables_follow_reference_rule():
    """Fine DMs start at max(coarse - calldms*dmstep/2, 0) (src/mdsm/dedispersion_thread.cu:197-198)."""
    C, nsubs, ncalls, calldms, dmstep = 64, 8, 4, 6, 0.5
    delay = oracle.delay_table(1500.0, -1.5, C)
    co, fi = oracle.subband_tables(1500.0, -1.5, C, nsubs, 1.0, 3.0, dmstep, ncalls, calldms,
                                   np.float32(64e-6))
    for j in range(ncalls):
        cdm = np.float32(np.float32(1.0) + np.float32(j) * np.float32(3.0))
        start = max(np.float32(cdm - np.float32(np.float32(calldms) * np.float32(dmstep)) / np.float32(2)),
                    np.float32(0))
        for i in range(calldms):
            fdm = np.float32(start + np.float32(np.float32(i) * np.float32(dmstep)))
            for b in range(nsubs):
                v = np.float32(np.float32(delay[b * 8] * fdm) / np.float32(64e-6))
                assert fi[j * calldms + i, b] == int(v)
        for c in range(C):
            v = np.float32(np.float32(delay[c] * cdm) / np.float32(64e-6))
            assert co[j, c] == int(v)


# ---------------------------------------------------------------- clipper
@pytest.mark.parametrize("n", [1, 2, 7, 64, 1001])
def test_clip_matches_sort_median(n):
    rng = np.random.default_rng(n)
    C = 6
    x = rng.integers(100, 140, (C, n + 10), dtype=np.uint8)
    x[0, :] = 77                                   # constant channel
    if n > 3:
        x[1, 5 + 3] = 255                          # spike
        x[2, 5:5 + n] = np.where(np.arange(n) % 2 == 0, 10, 12)  # ties, even/odd
    out, med, mad, ncl = oracle.clip_u8(x, 5, n, 5.0)
    for c in range(C):
        v = np.sort(x[c, 5:5 + n].astype(np.int64))
        m = int(v[(n - 1) // 2])
        dev = np.sort(np.abs(x[c, 5:5 + n].astype(np.int64) - m))
        a = int(dev[(n - 1) // 2])
        thr = np.float32(np.float32(5.0) * np.float32(1.4826)) * np.float32(a)
        clip = np.abs(x[c, 5:5 + n].astype(np.int64) - m).astype(np.float32) > thr
        assert med[c] == m and mad[c] == a and ncl[c] == int(clip.sum())
        assert np.array_equal(out[c], np.where(clip, m, x[c, 5:5 + n]))


# ---------------------------------------------------------------- boxcar detector
def _np_boxcar(x, widths, thr):
    """Brute-force numpy restatement (pairwise tree sums, same peak rule)."""
    res = []
    for d, row in enumerate(x.astype(np.float64)):
        n = row.size
        s1 = float(np.sum(row))
        m = s1 / n
        var = float(np.sum(row * row)) / n - m * m
        sig = math.sqrt(max(var, 0.0))
        for w in widths:
            if w > n or sig <= 0:
                continue
            S = row.copy()
            cw = 1
            while cw < w:
                S = S[:-cw] + S[cw:]
                cw *= 2
            snr = (S - w * m) / (sig * math.sqrt(w))
            last = n - w
            for t in range(last + 1):
                if not snr[t] > thr:
                    continue
                if t > 0 and not S[t] >= S[t - 1]:
                    continue
                if t < last and not S[t] > S[t + 1]:
                    continue
                res.append((d, t, w))
    return res


def test_boxcar_matches_numpy_bruteforce():
    rng = np.random.default_rng(11)
    x = rng.integers(0, 50, (5, 300)).astype(np.float32)
    x[2, 100:108] += 60
    x[4, 7] += 90
    widths = [1, 2, 4, 8, 16]
    got, means, sig = oracle.boxcar_detect(x, widths, 4.0)
    exp = _np_boxcar(x, widths, 4.0)
    assert sorted(zip(got["dm"].tolist(), got["t"].tolist(), got["width"].tolist())) == sorted(exp)
    assert len(exp) > 0


def test_boxcar_spike_closed_form():
    """A spike of height h on a constant row: snr_w = h(1 - w/n)/(sigma sqrt(w)),
    sigma = h sqrt(n-1)/n, candidates only at widths below the threshold bound."""
    n, h = 1024, 100.0
    x = np.full((1, n), 5.0, np.float32)
    x[0, 500] += h
    got, means, sig = oracle.boxcar_detect(x, [1, 2, 4, 8, 16, 32, 64], 6.0)
    s = h * math.sqrt(n - 1) / n
    assert abs(sig[0] - s) < 1e-9 * s
    for w in (1, 2, 4, 8, 16, 32, 64):
        snr = h * (1 - w / n) / (s * math.sqrt(w))
        present = [int(t) for t, ww in zip(got["t"], got["width"]) if ww == w]
        if snr > 6.0:
            assert present == [500]   # peak rule: the last t of the plateau of equal sums
            sn = got["snr"][(got["width"] == w)][0]
            assert abs(sn - snr) < 1e-5 * snr
        else:
            assert present == []


def test_boxcar_tophat_widths():
    """Top hat of width 8: snr grows as sqrt(w) up to w = 8, then as 8/sqrt(w)."""
    n = 4096
    x = np.zeros((1, n), np.float32)
    x[0, 1000:1008] = 10.0
    got, m, s = oracle.boxcar_detect(x, [1, 2, 4, 8, 16, 32], 0.0)
    best = {}
    for t, w, v in zip(got["t"], got["width"], got["snr"]):
        best[int(w)] = max(best.get(int(w), -1), float(v))
    mu, sg = m[0], s[0]
    for w in (1, 2, 4, 8, 16, 32):
        exp = (10.0 * min(w, 8) - w * mu) / (sg * math.sqrt(w))
        assert abs(best[w] - exp) < 1e-5 * abs(exp)


def _np_boxcar_stream(x, n, widths, thr):
    """Independent restatement of the streaming search on integer rows: window
    sums from an int64 prefix sum (exact), per-block statistics, window owned
    by the block of its end sample + 1, right neighbour required."""
    res = []
    for d, row in enumerate(x.astype(np.int64)):
        T = row.size
        P = np.concatenate([[0], np.cumsum(row)])
        blocks = row.reshape(-1, n).astype(np.float64)
        means = blocks.sum(1) / n
        sig = np.sqrt(np.maximum((blocks * blocks).sum(1) / n - means * means, 0.0))
        for w in widths:
            S = (P[w:] - P[:-w]).astype(np.float64)      # S[t] = sum row[t:t+w]
            for t in range(0, T - w):                     # t + 1 + w <= T
                k = (t + w) // n
                if not sig[k] > 0:
                    continue
                if not (S[t] - w * means[k]) / (sig[k] * math.sqrt(w)) > thr:
                    continue
                if t > 0 and not S[t] >= S[t - 1]:
                    continue
                if not S[t] > S[t + 1]:
                    continue
                res.append((d, t - k * n, w, k))
    return sorted(res)


def test_boxcar_stream_matches_numpy_bruteforce():
    rng = np.random.default_rng(12)
    n, nblk = 200, 4
    x = rng.integers(0, 40, (6, n * nblk)).astype(np.float32)
    x[1, n - 5:n + 7] += 50          # straddles the edge of blocks 0 and 1
    x[3, 3 * n - 1] += 120           # last sample of block 2
    x[5, 2 * n] += 120               # first sample of block 2
    widths = [1, 2, 4, 8, 16, 32]
    got = oracle.boxcar_stream(x, n, widths, 4.0)
    key = sorted(zip(got["dm"].tolist(), got["t"].tolist(), got["width"].tolist(),
                     got["blk"].tolist()))
    exp = _np_boxcar_stream(x, n, widths, 4.0)
    assert key == exp and len(exp) > 0
    # windows owned by a block but starting in the previous one are reported
    assert any(t < 0 for _, t, _, _ in key)


def test_boxcar_stream_one_block_is_block_search_less_the_last_window():
    rng = np.random.default_rng(13)
    x = rng.integers(0, 50, (5, 300)).astype(np.float32)
    x[2, 100:108] += 60
    x[4, 296:300] += 90              # pulse in the last windows of the row
    widths = [1, 2, 4, 8, 16]
    blk, _, _ = oracle.boxcar_detect(x, widths, 4.0)
    st = oracle.boxcar_stream(x, 300, widths, 4.0)
    a = {(d, t, w) for d, t, w in zip(blk["dm"].tolist(), blk["t"].tolist(),
                                      blk["width"].tolist()) if t < 300 - w}
    b = set(zip(st["dm"].tolist(), st["t"].tolist(), st["width"].tolist()))
    # the block search's last window (t = n - w) has no right neighbour and is
    # left to the next block in the stream; every other candidate agrees
    assert b == a
    assert any(t == 300 - w for t, w in zip(blk["t"].tolist(), blk["width"].tolist()))


def test_boxcar_stream_straddling_pulse_found_once():
    """A top hat of 16 samples centred on a block edge: the stream search finds
    it once at its full width and S/N; searching each block alone splits it."""
    n, nblk = 4096, 2
    rng = np.random.default_rng(14)
    x = rng.integers(100, 140, (1, n * nblk)).astype(np.float32)
    x[0, n - 8:n + 8] += 200.0
    st = oracle.boxcar_stream(x, n, [16], 6.0)
    hits = [(int(t), int(k)) for t, k in zip(st["t"], st["blk"])]
    assert (-8, 1) in hits and len([h for h in hits if h[0] == -8]) == 1
    # block by block: neither block sees a window holding the whole pulse
    for k in range(nblk):
        c, _, _ = oracle.boxcar_detect(np.ascontiguousarray(x[:, k * n:(k + 1) * n]), [16], 6.0)
        starts = [int(t) + k * n for t in c["t"]]
        assert n - 8 not in starts


# ---------------------------------------------------------------- process_brute statistics
def _process_brute_input(seed, tdms, nsamp):
    rng = np.random.default_rng(seed)
    x = (100.0 + 10.0 * rng.standard_normal((tdms, nsamp))).astype(np.float32)
    x[tdms // 2, nsamp // 3:nsamp // 3 + 12] += 150.0
    return x


def test_process_brute_matches_reference_golden():
    """The oracle's brute-force output-thread detector (chunked block
    statistics, median of five, float thresholds) reproduces the reference
    process_brute's candidate file line for line: a block smaller than one
    chunk, two whole chunks (the second never read) and a partial chunk."""
    with open(os.path.join(GOLD, "ref_process_brute.json")) as f:
        cases = json.load(f)
    for c in cases:
        x = _process_brute_input(c["seed"], c["tdms"], c["nsamp"])
        got = oracle.process_brute_lines(x, c["lowdm"], c["dmstep"], c["thr"], c["timestamp"],
                                         c["block_rate"])
        assert got == c["lines"] and len(got) > 0, (c["tdms"], c["nsamp"])


def test_brute_stats_skip_the_last_chunk():
    """Closed forms of the chunk rule: with two whole chunks only the first is
    read; with fewer values than a chunk all are read."""
    x = np.zeros((2, 32768), np.float32)
    x[0] = 3.0
    x[1] = 100.0                     # the second chunk: never read
    m, r = oracle.brute_stats(x)
    assert m == np.float32(3.0) and r == np.float32(0.0)
    y = np.array([[1, 2, 3, 4, 5, 6]], np.float32)
    m, r = oracle.brute_stats(y)
    assert m == np.float32(3.5) and abs(float(r) - np.sqrt(35.0 / 12.0)) < 1e-6


def test_process_brute_matches_reference_build(tmp_path):
    """Against the reference's own process_brute built under oracle/_ref, on
    blocks not in the golden file (skipped where that build is absent)."""
    if oracle.ref() is None:
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(77)
    for tdms, nsamp in ((7, 9000), (33, 3000), (128, 1024)):
        x = (50.0 + 5.0 * rng.standard_normal((tdms, nsamp))).astype(np.float32)
        x[5, 10:30] += 60.0          # row 5: DM 2.5 > 1
        path = tmp_path / f"c{tdms}.dat"
        assert oracle.ref_process_brute(x, nsamp, 0.0, 0.5, 2.0, 10.0, 1e-3, path)
        ref = path.read_text().strip().splitlines()
        assert oracle.process_brute_lines(x, 0.0, 0.5, 2.0, 10.0, 1e-3) == ref and ref


# ---------------------------------------------------------------- median of five
def test_median_of_five_matches_reference_golden():
    g = np.load(os.path.join(GOLD, "ref_median5.npz"))
    got = np.array([oracle.median_of_five(*map(float, t)) for t in g["tuples"]], np.float32)
    assert np.array_equal(got, g["median"])
    assert np.array_equal(np.sort(g["tuples"], 1)[:, 2], g["median"])


@needs_ref
def test_median_of_five_matches_reference_build():
    R = oracle.ref()
    for t in itertools.product(range(5), repeat=5):
        assert R.median_of_five(*map(float, t)) == oracle.median_of_five(*map(float, t))


# ---------------------------------------------------------------- corner turn, detrend
@pytest.mark.parametrize("invert", [False, True])
def test_corner_turn_is_numpy_transpose(invert):
    rng = np.random.default_rng(5)
    x = rng.standard_normal((37, 23)).astype(np.float32)   # [t][c]
    got = oracle.corner_turn_f32(x, invert)
    exp = x.T[::-1] if invert else x.T
    assert np.array_equal(got, exp)


def test_detrend_matches_polyfit_and_invariants():
    rng = np.random.default_rng(6)
    L = 64
    t = np.arange(3 * L + 17, dtype=np.float64)
    rows = np.stack([0.3 * t - 5 + rng.standard_normal(t.size),
                     100 + 0.01 * t ** 1.5 + rng.standard_normal(t.size)]).astype(np.float32)
    got = oracle.detrend_f64(rows, L)
    for r in range(rows.shape[0]):
        for s0 in range(0, t.size, L):
            seg = rows[r, s0:s0 + L].astype(np.float64)
            n = seg.size
            x = np.arange(n) - n / 2 + 0.5
            b, a = np.polyfit(x, seg, 1)            # independent least-squares fit
            res = seg - (a + b * x)
            exp = res / np.sqrt(np.mean(res ** 2))
            g = got[r, s0:s0 + n]
            assert np.allclose(g, exp, rtol=1e-9, atol=1e-9)
            assert abs(g.mean()) < 1e-9 and abs((g * x).sum()) < 1e-6
            assert abs(np.mean(g ** 2) - 1.0) < 1e-9


def test_detrend_pure_line_is_zero():
    x = (np.arange(256, dtype=np.float32) * 0.25 - 3.0)[None, :]
    got = oracle.detrend_f64(x, 128)
    assert np.all(got == 0.0)
