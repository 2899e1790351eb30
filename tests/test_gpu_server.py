"""Serving loop: continuous batching across beams, device-resident overlap,
clip -> dedisperse -> boxcar per block, HBM-budget admission.

The key property: each block's result (DM x time output and candidate list)
is the same when it runs alone and when it runs batched with other beams'
blocks, and equals the oracle pipeline over the concatenated stream."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
WIDTHS = (1, 2, 4, 8, 16, 32, 64)


def _beam_kwargs(C=64, N=4096, ranges=((0.0, 2.0, 40, 1), (80.0, 4.0, 20, 2)), clip=5.0):
    return dict(nchans=C, fch1=1500.0, foff=-300.0 / C, tsamp=64e-6, nsamp_block=N, dtype="u8",
                ranges=ranges, clip_nsigma=clip, snr_thr=6.0, widths=WIDTHS,
                retain_output=True)


def _stream(M, C, N, nblocks, seed, pulses):
    return M.synth_filterbank(C, N * nblocks, "u8", seed=seed, mean=128, sigma=16,
                              fch1=1500.0, foff=-300.0 / C, pulses=pulses)


def _run(srv, beams_and_streams, N, nblocks, interleave=True, hold_first=False):
    """Push every block of every beam; returns {(beam, seq): (out, cands)}.
    hold_first: admissions are held while every beam's first block is pushed,
    so that those blocks are admitted as one batch."""
    res = {}
    pending = 0
    order = []
    for k in range(nblocks):
        for b, x in beams_and_streams:
            order.append((b, k, x))
    if not interleave:
        order.sort(key=lambda t: (t[0].id, t[1]))
    if hold_first:
        srv.hold(True)
        for b, k, x in order[:len(beams_and_streams)]:
            assert srv.try_push(b, np.ascontiguousarray(x[:, k * N:(k + 1) * N])) is not None
            pending += 1
        order = order[len(beams_and_streams):]
        srv.hold(False)
    for b, k, x in order:
        blk = np.ascontiguousarray(x[:, k * N:(k + 1) * N])
        while srv.try_push(b, blk) is None:
            r, c = srv.poll(wait_ms=-1)
            res[(r.beam, r.seq)] = (srv.output(srv.beams[r.beam], r.seq), c, r)
            srv.release(srv.beams[r.beam], r.seq)
            pending -= 1
        pending += 1
    srv.drain()
    while True:
        r, c = srv.poll(wait_ms=0)
        if r is None:
            break
        assert r.status == 0
        res[(r.beam, r.seq)] = (srv.output(srv.beams[r.beam], r.seq), c, r)
        srv.release(srv.beams[r.beam], r.seq)
    return res


def _oracle_beam(x, beam, N, nblocks, ranges, C, clip=5.0):
    """Oracle pipeline over the stream: clip each block's new samples,
    dedisperse the zero-history + stream series block by block, then search
    each range's series -- the first N/k samples of every block's rows,
    concatenated -- with the streaming detector (every window once, per-block
    statistics).  Returns per block (packed output, candidates)."""
    M = beam.overlap
    clipped = np.empty_like(x)
    for k in range(nblocks):
        if clip > 0:
            out, *_ = oracle.clip_u8(x[:, k * N:(k + 1) * N], 0, N, clip)
        else:
            out = x[:, k * N:(k + 1) * N]
        clipped[:, k * N:(k + 1) * N] = out
    full = np.concatenate([np.zeros((C, M), np.uint8), clipped], axis=1)
    outs = [[] for _ in range(nblocks)]
    series = [[] for _ in ranges]
    for k in range(nblocks):
        seg = np.ascontiguousarray(full[:, k * N:k * N + M + N])
        for r, (lo, step, nd, kk) in enumerate(ranges):
            ts = np.float32(64e-6) * np.float32(kk)
            sh = oracle.shift_table_ref(1500.0, -300.0 / C, C, lo, step, nd, ts)
            nout = (M + N) // kk - int(sh.max())
            o = oracle.range_f32(seg, kk, sh, nout)
            outs[k].append(o.reshape(-1))
            series[r].append(o[:, :N // kk])
    cands = [[] for _ in range(nblocks)]
    first = 0
    for r, (lo, step, nd, kk) in enumerate(ranges):
        c = oracle.boxcar_stream(np.concatenate(series[r], axis=1), N // kk, list(WIDTHS), 6.0)
        c["dm"] += first
        for k in range(nblocks):
            sel = c[c["blk"] == k]
            cands[k].append(np.array(sel[["dm", "t", "width", "snr"]].tolist(),
                                     dtype=[("dm", np.uint32), ("t", np.int32),
                                            ("width", np.uint32), ("snr", np.float32)]))
        first += nd
    return [(np.concatenate(outs[k]), np.concatenate(cands[k])) for k in range(nblocks)]


def _overlap(C, N, ranges):
    """The overlap M a beam of this shape keeps (mdsm_beam_info)."""
    from mdsm_amd.server import Server
    srv = Server()
    b = srv.open_beam(**_beam_kwargs(C, N, ranges))
    M = b.overlap
    srv.close()
    return M


def _key(c):
    return sorted(zip(c["dm"].tolist(), c["t"].tolist(), c["width"].tolist(), c["snr"].tolist()))


def test_alone_and_in_mix_identical_and_match_oracle(gpu):
    from mdsm_amd.server import Server
    M = gpu
    C, N, nblocks = 64, 4096, 3
    ranges = ((0.0, 2.0, 40, 1), (80.0, 4.0, 20, 2))
    print("case server alone/mix", flush=True)
    Mo = _overlap(C, N, ranges)
    # pulses inside blocks and across block edges: output sample t of block k
    # is stream sample k*N - M + t, so a pulse arriving (top of the band) at
    # k*N - M - 6 lies across the edge of blocks k-1 and k in the series
    streams = [_stream(M, C, N, nblocks, 100 + i, [(30.0 + 20 * i, 2000 + 3000 * i, 4, 3.0),
                                                   (25.0 + 10 * i, N - Mo - 6 + i, 12, 2.0),
                                                   (60.0, 2 * N - Mo - 3, 32, 1.5)])
               for i in range(4)]
    for x in streams:          # impulsive RFI for the clipper
        x[5, 100:110] = 255
    # alone: one server per beam, one beam at a time
    alone = []
    for i, x in enumerate(streams):
        srv = Server(max_batch=1)
        b = srv.open_beam(**_beam_kwargs(C, N, ranges))
        r = _run(srv, [(b, x)], N, nblocks)
        alone.append({k[1]: v for k, v in r.items()})
        srv.close()
    # mix: all beams on one server, interleaved pushes, batched launches
    srv = Server(max_batch=8)
    bs = [srv.open_beam(**_beam_kwargs(C, N, ranges)) for _ in streams]
    mix = _run(srv, list(zip(bs, streams)), N, nblocks, hold_first=True)
    batches = [v[2].batch for v in mix.values()]
    assert max(batches) > 1, "blocks were never batched together"
    for i, b in enumerate(bs):
        ref = _oracle_beam(streams[i], b, N, nblocks, ranges, C)
        for k in range(nblocks):
            out_m, c_m, _ = mix[(b.id, k)]
            out_a, c_a, _ = alone[i][k]
            assert np.array_equal(out_m, out_a), f"beam {i} block {k}: output differs in the mix"
            assert _key(c_m) == _key(c_a), f"beam {i} block {k}: candidates differ in the mix"
            assert np.array_equal(out_m, ref[k][0]), f"beam {i} block {k}: output != oracle"
            assert [t[:3] for t in _key(c_m)] == [t[:3] for t in _key(ref[k][1])], \
                f"beam {i} block {k}: candidates != oracle"
            # S/N uses the row statistics: from the dedispersion launch's
            # partials for the full-rate range, from a pass over the output
            # for the decimated one; both exact, so the S/N is the oracle's
            assert [t[3] for t in _key(c_m)] == [t[3] for t in _key(ref[k][1])], \
                f"beam {i} block {k}: candidate S/N != oracle"
    # windows that start in the previous block were searched (t < 0)
    assert any(int(t) < 0 for v in mix.values() for t in v[1]["t"])
    st = srv.stats()
    assert st["completed"] == len(streams) * nblocks and st["launches"] < st["completed"]
    srv.close()


def test_budget_admission(gpu):
    """A beam that can never fit is rejected; a budget for one block at a time
    defers admission but completes everything."""
    from mdsm_amd.server import Server
    M = gpu
    print("case server budget", flush=True)
    srv = Server(hbm_budget=1 << 20)
    with pytest.raises(M.MdsmError) as ei:
        srv.open_beam(**_beam_kwargs())
    assert ei.value.code == -5
    srv.close()
    srv = Server(hbm_budget=(1 << 30), max_inflight=1, max_batch=4)
    C, N = 64, 4096
    bs = [srv.open_beam(**_beam_kwargs(C, N)) for _ in range(3)]
    xs = [_stream(M, C, N, 2, 7 + i, []) for i in range(3)]
    out = _run(srv, list(zip(bs, xs)), N, 2, hold_first=True)
    assert len(out) == 6
    st = srv.stats()
    assert st["deferred_inflight"] > 0 and st["bytes_peak"] <= st["budget"]
    srv.close()


def test_persistent_launch_matches(gpu):
    from mdsm_amd.server import Server
    M = gpu
    print("case server persistent", flush=True)
    C, N = 64, 4096
    xs = [_stream(M, C, N, 2, 50 + i, [(40.0, 1500, 8, 4.0)]) for i in range(3)]
    outs = []
    for persistent in (0, 24):  # 24 workgroups: fewer than the items, so the queue is used
        srv = Server(max_batch=8, persistent=persistent)
        bs = [srv.open_beam(**_beam_kwargs(C, N)) for _ in xs]
        r = _run(srv, list(zip(bs, xs)), N, 2)
        outs.append({(i, k): r[(b.id, k)][0] for i, b in enumerate(bs) for k in range(2)})
        srv.close()
    for key in outs[0]:
        assert np.array_equal(outs[0][key], outs[1][key])


def test_zero_copy_acquire_commit_matches_push(gpu):
    """Blocks written in place into lent staging buffers give the same
    results as blocks pushed by copy."""
    from mdsm_amd.server import Server
    M = gpu
    print("case server acquire/commit", flush=True)
    C, N = 64, 4096
    xs = [_stream(M, C, N, 2, 60 + i, [(35.0, 1200, 4, 4.0)]) for i in range(2)]
    outs = []
    for mode in ("push", "acquire"):
        srv = Server(max_batch=4)
        bs = [srv.open_beam(**_beam_kwargs(C, N)) for _ in xs]
        res = {}
        for k in range(2):
            for b, x in zip(bs, xs):
                blk = np.ascontiguousarray(x[:, k * N:(k + 1) * N])
                if mode == "push":
                    while srv.try_push(b, blk) is None:
                        r, c = srv.poll(wait_ms=-1)
                        res[(r.beam, r.seq)] = (srv.output(srv.beams[r.beam], r.seq), c)
                        srv.release(srv.beams[r.beam], r.seq)
                else:
                    while True:
                        buf = srv.acquire(b)
                        if buf is not None:
                            break
                        r, c = srv.poll(wait_ms=-1)
                        res[(r.beam, r.seq)] = (srv.output(srv.beams[r.beam], r.seq), c)
                        srv.release(srv.beams[r.beam], r.seq)
                    buf[...] = blk
                    srv.commit(b, buf)
        srv.drain()
        while True:
            r, c = srv.poll(wait_ms=0)
            if r is None:
                break
            res[(r.beam, r.seq)] = (srv.output(srv.beams[r.beam], r.seq), c)
            srv.release(srv.beams[r.beam], r.seq)
        outs.append({(i, k): res[(b.id, k)] for i, b in enumerate(bs) for k in range(2)})
        srv.close()
    for key in outs[0]:
        assert np.array_equal(outs[0][key][0], outs[1][key][0])
        assert _key(outs[0][key][1]) == _key(outs[1][key][1])


@pytest.mark.parametrize("fdt", ["f16", "f32"])
def test_float_beam_in_a_mixed_batch(gpu, fdt):
    """A float16 or float32 beam (no clipper: it runs on uint8 beams) served beside a
    uint8 beam, their first blocks admitted as one batch (one launch of the
    kernel built for every dtype, row statistics on): each beam's outputs and
    candidates equal that beam served alone; the float outputs are within
    the path's 1e-5 normwise tolerance of the FP64 oracle over the
    concatenated stream; its candidates are the oracle's streaming search of
    the served series (the FP64 tree on the same float32 values; the row
    statistics' FP64 sums may differ in rounding order, so S/N agree to 1e-6
    and the sets agree away from a 1e-6 band around the threshold)."""
    from mdsm_amd.server import Server
    from conftest import rel_err
    M = gpu
    C, N, nblocks = 64, 4096, 3
    ranges = ((0.0, 2.0, 40, 1), (80.0, 4.0, 20, 2))
    print(f"case server {fdt} beam in a mixed batch", flush=True)
    kw16 = dict(_beam_kwargs(C, N, ranges, clip=0.0), dtype=fdt)
    kw8 = _beam_kwargs(C, N, ranges)
    x16 = M.synth_filterbank(C, N * nblocks, fdt, seed=311, gain=(0.5, 2.0), fch1=1500.0,
                             foff=-300.0 / C, pulses=[(30.0, 2500, 4, 3.0), (50.0, N + 1000, 8, 2.0),
                                                     (20.0, 2 * N - 200, 16, 2.0)])
    x8 = _stream(M, C, N, nblocks, 312, [(40.0, 3000, 4, 3.0)])
    alone = {}
    for name, kw, x in (("f16", kw16, x16), ("u8", kw8, x8)):
        srv = Server(max_batch=1)
        b = srv.open_beam(**kw)
        r = _run(srv, [(b, x)], N, nblocks)
        alone[name] = {k: r[(b.id, k)] for k in range(nblocks)}
        srv.close()
    srv = Server(max_batch=4)
    b16 = srv.open_beam(**kw16)
    b8 = srv.open_beam(**kw8)
    Mo = b16.overlap
    mix = _run(srv, [(b16, x16), (b8, x8)], N, nblocks, hold_first=True)
    srv.close()
    for k in range(nblocks):
        for name, b in (("f16", b16), ("u8", b8)):
            o_a, c_a, _ = alone[name][k]
            o_m, c_m, _ = mix[(b.id, k)]
            assert np.array_equal(o_a, o_m), f"{name} block {k}: output differs in the mix"
            assert _key(c_a) == _key(c_m), f"{name} block {k}: candidates differ in the mix"
    full = np.concatenate([np.zeros((C, Mo), x16.dtype), x16], axis=1)
    series = [[] for _ in ranges]
    for k in range(nblocks):
        seg = np.ascontiguousarray(full[:, k * N:k * N + Mo + N])
        got = alone["f16"][k][0]
        off = 0
        for r, (lo, step, nd, kk) in enumerate(ranges):
            ts = np.float32(64e-6) * np.float32(kk)
            sh = oracle.shift_table_ref(1500.0, -300.0 / C, C, lo, step, nd, ts)
            nout = (Mo + N) // kk - int(sh.max())
            g = got[off:off + nd * nout].reshape(nd, nout)
            off += nd * nout
            e, a = rel_err(g, oracle.range_f64(seg, kk, sh, nout))
            print(f"block {k} range {r}: rel {e:.3e} maxabs {a:.3e}")
            assert e <= 1e-5
            series[r].append(g[:, :N // kk])
    thr = 6.0
    first = 0
    nfound = 0
    for r, (lo, step, nd, kk) in enumerate(ranges):
        ref = oracle.boxcar_stream(np.concatenate(series[r], axis=1), N // kk, list(WIDTHS), thr)
        for k in range(nblocks):
            c = alone["f16"][k][1]
            sel = (c["dm"] >= first) & (c["dm"] < first + nd)
            gotd = {(int(d) - first, int(t), int(w)): float(s)
                    for d, t, w, s in zip(c["dm"][sel], c["t"][sel], c["width"][sel], c["snr"][sel])}
            rk = ref[ref["blk"] == k]
            expd = {(int(d), int(t), int(w)): float(s)
                    for d, t, w, s in zip(rk["dm"], rk["t"], rk["width"], rk["snr"])}
            for key in set(gotd) | set(expd):
                if key in gotd and key in expd:
                    assert abs(gotd[key] - expd[key]) <= 1e-6 * abs(expd[key]), key
                else:
                    s = gotd.get(key, expd.get(key))
                    assert abs(s - thr) <= 1e-6 * thr, f"range {r} block {k}: {key} {s}"
            nfound += len(expd)
        first += nd
    print(f"{fdt} beam: {nfound} candidates")
    assert nfound > 0
