"""The benchmark configurations at their full shapes, checked over every
output sample where the oracle cannot run in full.

Row totals: the sum of every output sample of DM row d equals
sum_c P_c[s_dc + nout] - P_c[s_dc], with P_c the FP64 prefix sums of channel
c of the (k-summed) input and s_dc the row's effective shift of channel c --
a checksum of every sample of every row that moves if any shift, window,
channel or row is wrong.  uint8 input: exact (each sample is rounded back to
its integer k-sum first).  float16/float32 input: |gpu - fp64| <= 1e-5 times
the row's L1 mass sum_c sum_t |x[c][t + s_dc]| (the path's 1e-5 relative
bar, applied per row).  Sampled windows at the start, middle and end of rows
are compared sample by sample with the FP64 oracle (bit-exact for uint8)."""
import numpy as np
import pytest

import oracle
from conftest import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _row_totals(x, eff, nout, k=1, block=64):
    """FP64 row totals and L1 masses of out[d][t] = (1/k) sum_c xk[c][t + eff[d][c]]
    (xk: k-sample sums of x) from prefix sums over blocks of channels (the L1
    mass only for float input)."""
    C = x.shape[0]
    D = eff.shape[0]
    integer = x.dtype == np.uint8
    tot = np.zeros(D, np.float64)
    l1 = np.zeros(D, np.float64)
    for c0 in range(0, C, block):
        c1 = min(C, c0 + block)
        xb = x[c0:c1].astype(np.int64 if integer else np.float64)
        if k > 1:
            n = xb.shape[1] // k
            xb = xb[:, :n * k].reshape(c1 - c0, n, k).sum(axis=2)
        cols = np.arange(c1 - c0)[None, :]
        e = eff[:, c0:c1]
        P = np.zeros((c1 - c0, xb.shape[1] + 1), xb.dtype)
        np.cumsum(xb, axis=1, out=P[:, 1:])
        tot += (P[cols, e + nout] - P[cols, e]).sum(axis=1).astype(np.float64)
        if not integer:
            np.cumsum(np.abs(xb), axis=1, out=P[:, 1:])
            l1 += (P[cols, e + nout] - P[cols, e]).sum(axis=1)
        del P, xb
    return tot / k, (l1 if not integer else tot) / k


def _gpu_row_totals(dout, nrows, nout, offset_elems=0, k=1, integer=False, chunk=256):
    """FP64 totals of the GPU rows (uint8: each sample rounded back to its
    integer k-sum, so the total is exact)."""
    tot = np.zeros(nrows, np.float64)
    for r0 in range(0, nrows, chunk):
        r1 = min(nrows, r0 + chunk)
        g = dout.download(np.float32, (r1 - r0, nout), offset=(offset_elems + r0 * nout) * 4)
        if integer:
            tot[r0:r1] = np.rint(g.astype(np.float64) * k).sum(axis=1) / k
        else:
            tot[r0:r1] = np.add.reduce(g, axis=1, dtype=np.float64)
        del g
    return tot


def _check_totals(name, got, tot, l1, integer):
    if integer:
        bad = np.flatnonzero(got != tot)
        print(f"{name}: {got.size} row totals, {bad.size} differ", flush=True)
        assert bad.size == 0, f"{name}: rows {bad[:8]} differ"
    else:
        ratio = np.abs(got - tot) / np.maximum(l1, 1e-300)
        print(f"{name}: {got.size} row totals, worst |err|/L1 {ratio.max():.3e}", flush=True)
        assert ratio.max() <= TOL


def test_config1_every_row_and_sample(gpu):
    """Config 1 (256 ch x 65536 f32, 512 DMs 0-102.2): every row's total, and
    every fourth row over all its 65103 samples against the FP64 oracle."""
    import bench
    M = gpu
    print("case config1 every row", flush=True)
    w = bench.WORKLOADS["cfg1"]
    x, _ = bench.make_input(M, w)
    plan = bench.make_plan(M, w)
    R = plan.ranges[0]
    din = M.DeviceBuffer.from_host(x)
    dout = M.DeviceBuffer(plan.info.out_elems * 4)
    plan.dedisperse(din, w["nsamp"], dout)
    M.synchronize()
    sh = plan.shift_table(0)
    tot, l1 = _row_totals(x, sh, R.nout)
    _check_totals("config1", _gpu_row_totals(dout, R.ndms, R.nout), tot, l1, False)
    got = dout.download(np.float32, (R.ndms, R.nout))
    rows = np.arange(0, R.ndms, 4)
    ref = oracle.dedisperse_f64(x, sh[rows], R.nout)
    e, a = rel_err(got[rows], ref)
    print(f"config1 {rows.size} rows x {R.nout} samples: rel {e:.3e} maxabs {a:.3e}")
    assert e <= TOL


def test_config3_every_row_and_windows(gpu):
    """Config 3 (4096 ch x 2^19 f16, 64 subbands, 4096 DMs): the total of
    every one of the 4096 rows over all 480777 samples (effective shift of
    channel c in subband b for row d = j*64 + i: fine[d][b] + coarse[j][c] -
    coarse[j][top(b)]), and every row at the start, middle and end of its
    samples against the FP64 two-stage oracle."""
    import bench
    M = gpu
    print("case config3 every row", flush=True)
    w = bench.WORKLOADS["cfg3"]
    x, _ = bench.make_input(M, w)
    plan = bench.make_plan(M, w)
    R = plan.ranges[0]
    q = w["passes"][0]
    din = M.DeviceBuffer.from_host(x)
    dout = M.DeviceBuffer(plan.info.out_elems * 4)
    plan.dedisperse(din, w["nsamp"], dout)
    M.synchronize()
    co = plan.shift_table(0)                      # [ncalls][nchans]
    fi = plan.fine_table(0, w["nsubs"])           # [ncalls*calldms][nsubs]
    C, nsubs, cd = w["nchans"], w["nsubs"], q["calldms"]
    cps = C // nsubs
    j = np.arange(R.ndms) // cd
    b = np.arange(C) // cps
    top = co[:, b * cps]                          # [ncalls][C]: coarse shift of the subband top
    eff = fi[:, b] + (co - top)[j]                # [ndms][C]
    tot, l1 = _row_totals(x, eff, R.nout)
    _check_totals("config3", _gpu_row_totals(dout, R.ndms, R.nout), tot, l1, False)
    nchk = 2048
    M2 = int(fi.max())
    for t_a in (0, R.nout // 2, R.nout - nchk):
        win = np.ascontiguousarray(x[:, t_a:t_a + R.maxshift + nchk])
        ref = oracle.subband_f64(win, 1, nsubs, q["ncalls"], cd, co, fi, nchk + M2, nchk)
        got = dout.download_rows(np.float32, R.ndms, nchk, R.nout, offset=t_a * 4)
        e, a = rel_err(got, ref)
        print(f"config3 window at {t_a}: {R.ndms} rows x {nchk}: rel {e:.3e} maxabs {a:.3e}")
        assert e <= TOL


def _cfg4_plans(M, seed=4004):
    """The 64 plans of the config-4 benchmark (the same seeded draw)."""
    rng = np.random.default_rng(seed)
    fch1, bw, tsamp = 1500.0, 300.0, 64e-6
    plans, beams = [], []
    while len(plans) < 64:
        C = int(rng.choice([256, 1024, 4096]))
        S = int(rng.integers(1 << 16, (1 << 19) + 1))
        dt = "u8" if rng.random() < 0.5 else "f16"
        total = int(rng.integers(256, 4097))
        nr = int(rng.integers(1, 4))
        cuts = np.sort(rng.choice(np.arange(1, total), nr - 1, replace=False)) if nr > 1 else []
        counts = np.diff(np.concatenate([[0], cuts, [total]])).astype(int)
        ks = np.sort(rng.integers(1, 9, nr))
        ranges, lo = [], 0.0
        for nd, k in zip(counts, ks):
            ranges.append((lo, 0.1 * float(k), int(nd), int(k)))
            lo += nd * 0.1 * float(k)
        try:
            p = M.Plan.brute(C, fch1, -bw / C, tsamp, S, dt, ranges)
        except M.MdsmError:
            continue
        plans.append(p)
        beams.append(dict(C=C, S=S, dtype=dt, ranges=ranges))
    return plans, beams


def test_config4_batch_at_workload(gpu):
    """Config 4 as benchmarked: 64 plans of uneven shapes (C in {256, 1024,
    4096}, S in [2^16, 2^19], uint8 or float16, 1-3 ranges, 256-4096 trials,
    k 1-8) in one persistent batch launch over windows of two resident
    pools.  Every plan's output equals the plan run alone (bit for bit, every
    sample); for a set of plans covering both dtypes, every channel count and
    decimations up to 8, every row's total and the first, middle and last row
    of every range match the oracle."""
    M = gpu
    print("case config4 batch at workload", flush=True)
    plans, beams = _cfg4_plans(M)
    Cmax, Smax = 4096, 1 << 19
    pools = {
        "u8": M.synth_filterbank(Cmax, Smax, "u8", seed=401, fch1=1500.0, foff=-300.0 / Cmax,
                                 tsamp=64e-6, mean=128.0, sigma=16.0,
                                 pulses=[(300.0, 200_000, 8, 1.0)], nthreads=16),
        "f16": M.synth_filterbank(Cmax, Smax, "f16", seed=402, fch1=1500.0, foff=-300.0 / Cmax,
                                  tsamp=64e-6, gain=(0.5, 2.0),
                                  pulses=[(300.0, 200_000, 8, 0.5)], nthreads=16),
    }
    d_pool = {k: M.DeviceBuffer.from_host(v) for k, v in pools.items()}
    ins = [d_pool[b["dtype"]] for b in beams]
    outs = [M.DeviceBuffer(p.info.out_elems * 4) for p in plans]
    batch = M.Batch(plans)
    s = M.Stream()
    batch.run(ins, [Smax] * len(plans), outs, stream=s)
    s.synchronize()
    # every plan alone, compared bit for bit with its batch output
    alone = None
    for i, p in enumerate(plans):
        if alone is None or alone.nbytes < p.info.out_elems * 4:
            alone = M.DeviceBuffer(max(p.info.out_elems * 4, 1 << 30))
        p.dedisperse(ins[i], Smax, alone)
        M.synchronize()
        n = p.info.out_elems
        step = 1 << 26
        for o in range(0, n, step):
            m = min(step, n - o)
            a = outs[i].download(np.float32, (m,), offset=o * 4)
            b = alone.download(np.float32, (m,), offset=o * 4)
            assert np.array_equal(a, b), f"plan {i} {beams[i]}: batch output differs from alone"
    print(f"config4: all {len(plans)} plans equal alone", flush=True)
    # oracle checks on plans covering both dtypes, every channel count and k up to 8
    chosen, seen = [], set()
    for i, b in enumerate(beams):
        key = (b["dtype"], b["C"])
        if key not in seen and (b["C"] < 4096 or not any(bb["C"] == 4096 for bb in
                                                         (beams[j] for j in chosen))):
            seen.add(key)
            chosen.append(i)
    kmax = max(range(len(beams)), key=lambda i: max(r[3] for r in beams[i]["ranges"]))
    if kmax not in chosen:
        chosen.append(kmax)
    for i in chosen:
        b, p = beams[i], plans[i]
        x = pools[b["dtype"]][:b["C"]]
        integer = b["dtype"] == "u8"
        for r, R in enumerate(p.ranges):
            sh = p.shift_table(r)
            xs = np.ascontiguousarray(x[:, :b["S"]])
            tot, l1 = _row_totals(xs, sh, R.nout, R.downsample)
            got = _gpu_row_totals(outs[i], R.ndms, R.nout, R.out_offset, R.downsample, integer)
            _check_totals(f"config4 plan {i} ({b['dtype']}, C {b['C']}, S {b['S']}) range {r} "
                          f"k {R.downsample}", got, tot, l1, integer)
            rows = [0, R.ndms // 2, R.ndms - 1]
            nchk = min(4096, R.nout)
            win = np.ascontiguousarray(xs[:, :(R.maxshift + nchk) * R.downsample])
            g = outs[i].download_rows(np.float32, R.ndms, nchk, R.nout,
                                      offset=R.out_offset * 4)[rows]
            if integer:
                assert np.array_equal(g, oracle.range_f32(win, R.downsample, sh[rows], nchk))
            else:
                e, a = rel_err(g, oracle.range_f64(win, R.downsample, sh[rows], nchk))
                assert e <= TOL


def test_config5_served_two_beams_at_shape(gpu):
    """Config 5 at its shape: two beams of 1024-channel uint8 blocks of 2^16
    samples, 2048 DMs over 0-1000, clip at 5 sigma, boxcar w = 1..64 at
    S/N > 6, two blocks each, served in one batch.  Each beam's blocks match
    the oracle pipeline over its stream: every output row's total (exact),
    sampled rows (bit-exact) and the candidate set of the streaming search
    (which includes windows across the block edge); each beam alone gives
    the same outputs and candidates."""
    import bench
    from mdsm_amd.server import Server
    M = gpu
    print("case config5 served at shape", flush=True)
    w = bench.WORKLOADS["cfg5"]
    C, N = w["nchans"], w["nsamp"]
    widths = (1, 2, 4, 8, 16, 32, 64)
    det = dict(clip_nsigma=5.0, snr_thr=6.0, widths=widths, cand_cap=1 << 18,
               retain_output=True)
    srv = Server(max_batch=8)
    beams = [srv.open_beam(C, w["fch1"], w["foff"], w["tsamp"], N, "u8", w["ranges"], **det)
             for _ in range(2)]
    Mo = beams[0].overlap
    assert Mo == 16176
    xs = [M.synth_filterbank(C, 2 * N, "u8", seed=500 + i, fch1=w["fch1"], foff=w["foff"],
                             tsamp=w["tsamp"], mean=128.0, sigma=16.0,
                             pulses=[(200.0 + 300 * i, N - Mo - 6, 8, 1.5),
                                     (600.0, 20000 + 7000 * i, 2, 3.0)], nthreads=16)
          for i in range(2)]
    xs[1][17, 1000:1100] = 255                     # impulsive RFI for the clipper

    def serve(server, pairs):
        """Push two blocks of each (beam, stream); the first blocks of all
        beams are admitted together."""
        res = {}
        server.hold(True)
        for k in range(2):
            for b, x in pairs:
                server.push(b, np.ascontiguousarray(x[:, k * N:(k + 1) * N]))
            if k == 0:
                server.hold(False)
        server.drain()
        while True:
            r, c = server.poll(wait_ms=0)
            if r is None:
                break
            assert r.status == 0
            res[(r.beam, r.seq)] = (server.output(server.beams[r.beam], r.seq), c, r.batch)
            server.release(server.beams[r.beam], r.seq)
        return res

    mix = serve(srv, list(zip(beams, xs)))
    assert max(v[2] for v in mix.values()) == 2, "the two beams' blocks were not batched"
    srv.close()
    D = 2048
    sh = oracle.shift_table_ref(w["fch1"], w["foff"], C, 0.0, 1000.0 / 2048, D,
                                np.float32(w["tsamp"]))
    for i, b in enumerate(beams):
        x = xs[i]
        clipped = np.zeros((C, Mo + 2 * N), np.uint8)
        for k in range(2):
            o, *_ = oracle.clip_u8(np.ascontiguousarray(x[:, k * N:(k + 1) * N]), 0, N, 5.0)
            clipped[:, Mo + k * N:Mo + (k + 1) * N] = o
        outs = []
        for k in range(2):
            out, _, _ = mix[(b.id, k)]
            seg = np.ascontiguousarray(clipped[:, k * N:k * N + Mo + N])
            tot, _ = _row_totals(seg, sh, N)
            g = out.reshape(D, N)
            _check_totals(f"config5 beam {i} block {k}", np.rint(g.astype(np.float64)).sum(1),
                          tot, tot, True)
            rows = np.linspace(0, D - 1, 8).astype(int)
            assert np.array_equal(g[rows], oracle.range_f32(seg, 1, sh[rows], N))
            outs.append(g)
        ref = oracle.boxcar_stream(np.concatenate(outs, axis=1), N, list(widths), 6.0)
        for k in range(2):
            c = mix[(b.id, k)][1]
            got = sorted(zip(c["dm"].tolist(), c["t"].tolist(), c["width"].tolist(),
                             c["snr"].tolist()))
            sel = ref[ref["blk"] == k]
            exp = sorted(zip(sel["dm"].tolist(), sel["t"].tolist(), sel["width"].tolist(),
                             sel["snr"].tolist()))
            assert [g[:3] for g in got] == [e[:3] for e in exp], \
                f"beam {i} block {k}: candidates differ from the oracle"
            # S/N from the row statistics the dedispersion launch writes (exact)
            assert [g[3] for g in got] == [e[3] for e in exp], \
                f"beam {i} block {k}: candidate S/N differs from the oracle"
        assert any(t < 0 for t in mix[(b.id, 1)][1]["t"].tolist()), "edge pulse not found"
    # each beam alone: identical outputs and candidates
    for i in range(2):
        s1 = Server(max_batch=1)
        b1 = s1.open_beam(C, w["fch1"], w["foff"], w["tsamp"], N, "u8", w["ranges"], **det)
        alone = serve(s1, [(b1, xs[i])])
        for k in range(2):
            assert np.array_equal(alone[(b1.id, k)][0], mix[(beams[i].id, k)][0])
            ka = sorted(zip(*(alone[(b1.id, k)][1][f].tolist() for f in ("dm", "t", "width"))))
            km = sorted(zip(*(mix[(beams[i].id, k)][1][f].tolist() for f in ("dm", "t", "width"))))
            assert ka == km
        s1.close()
