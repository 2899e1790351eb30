"""Streaming semantics of the serving loop and the streaming detector:
boxcar windows across block edges, sequence gaps and duplicates, cancel and
close (also while admissions are held), pushes racing a close, export and
import of a beam's resident state, candidate overflow and device-resident
pushes.  Every block's output and candidates are compared with the oracle
pipeline over the whole stream or with the same stream run uninterrupted."""
import threading

import numpy as np
import pytest

import oracle
from test_gpu_server import WIDTHS, _beam_kwargs, _key, _oracle_beam, _overlap, _stream

pytestmark = pytest.mark.gpu

C, N = 64, 4096
RANGES = ((0.0, 2.0, 40, 1), (80.0, 4.0, 20, 2))


def _collect(srv, res, retain=True):
    """Poll everything completed into res[(beam, seq)] = (out, cands, result)."""
    while True:
        r, c = srv.poll(wait_ms=0)
        if r is None:
            return
        out = None
        if retain and r.status == 0 and not (r.flags & 4):
            out = srv.output(srv.beams[r.beam], r.seq)
            srv.release(srv.beams[r.beam], r.seq)
        res[(r.beam, r.seq)] = (out, c, r)


def _cand_key(c):
    return [t[:3] for t in _key(c)]


def test_detect_boxcar_stream_matches_oracle(gpu):
    """mdsm_detect_boxcar_stream block by block with its carry equals the
    oracle's streaming search over the concatenated rows: windows across the
    edges are found once, at negative t in the later block."""
    M = gpu
    print("case detect_boxcar_stream", flush=True)
    rng = np.random.default_rng(31)
    ndms, n, nblk = 24, 3000, 4
    x = rng.integers(90, 170, (ndms, n * nblk)).astype(np.float32)
    for d in range(ndms):            # pulses of every width across or near each edge
        e = n * (1 + d % (nblk - 1))
        w = 1 << (d % 7)
        x[d, e - w // 2 - (d % 3):e - w // 2 - (d % 3) + w] += 400.0 / np.sqrt(w)
    H = M.boxcar_carry_len(WIDTHS)
    assert H == 68
    d_carry = M.DeviceBuffer(ndms * H * 4)
    cap = 1 << 16
    d_c, d_n = M.DeviceBuffer(cap * 16), M.DeviceBuffer(4)
    d_mean, d_sig = M.DeviceBuffer(ndms * 8), M.DeviceBuffer(ndms * 8)
    got = []
    for k in range(nblk):
        d_x = M.DeviceBuffer.from_host(np.ascontiguousarray(x[:, k * n:(k + 1) * n]))
        M.detect_boxcar_stream(d_x, ndms, n, n, d_carry, H, k > 0, WIDTHS, 6.0, d_c, cap, d_n,
                               d_mean, d_sig)
        M.synchronize()
        c = M.read_candidates(d_c, d_n, cap)
        got += [(int(a), int(t), int(w), k) for a, t, w in zip(c["dm"], c["t"], c["width"])]
    ref = oracle.boxcar_stream(x, n, list(WIDTHS), 6.0)
    exp = sorted(zip(ref["dm"].tolist(), ref["t"].tolist(), ref["width"].tolist(),
                     ref["blk"].tolist()))
    assert sorted(got) == exp
    assert any(t < 0 for _, t, _, _ in got), "no window across a block edge was reported"


def test_sequence_gap_zero_filled_and_duplicate_rejected(gpu):
    """Blocks 2 and 3 never arrive: block 2 is processed as zeros, block 3
    (whose history is then all zeros: M <= N) is skipped, both flagged; block
    4 and everything else equal the oracle on the stream with those blocks
    zeroed.  A duplicate block number is rejected."""
    from mdsm_amd.server import SKIPPED, ZERO_FILLED, Server
    M = gpu
    print("case server sequence gap", flush=True)
    nblk = 5
    Mo = _overlap(C, N, RANGES)
    assert Mo <= N
    x = _stream(M, C, N, nblk, 77, [(40.0, 1000, 8, 3.0), (50.0, 4 * N - Mo + 500, 4, 4.0)])
    srv = Server(max_batch=4)
    b = srv.open_beam(**_beam_kwargs(C, N, RANGES))
    res = {}
    for k in (0, 1, 4):
        blk = np.ascontiguousarray(x[:, k * N:(k + 1) * N])
        while not srv.push_seq(b, blk, k):
            srv.drain()
            _collect(srv, res)
    with pytest.raises(M.MdsmError) as ei:
        srv.push_seq(b, np.ascontiguousarray(x[:, N:2 * N]), 1)
    assert ei.value.code == -6
    srv.drain()
    _collect(srv, res)
    assert sorted(k[1] for k in res) == list(range(nblk))
    flags = {k[1]: v[2].flags for k, v in res.items()}
    assert flags[2] == ZERO_FILLED and flags[3] == ZERO_FILLED | SKIPPED
    assert flags[0] == flags[1] == flags[4] == 0
    xz = x.copy()
    xz[:, 2 * N:4 * N] = 0
    ref = _oracle_beam(xz, b, N, nblk, RANGES, C)
    for k in range(nblk):
        out, c, r = res[(b.id, k)]
        assert r.status == 0
        if k == 3:
            assert np.all(ref[k][0] == 0) and r.ncand == 0
            continue
        assert np.array_equal(out, ref[k][0]), f"block {k}: output != oracle"
        assert _cand_key(c) == _cand_key(ref[k][1]), f"block {k}: candidates != oracle"
    srv.close()


def test_cancel_and_close_while_held(gpu):
    """With admissions held: cancel drops a beam's queued blocks (reported as
    cancelled) and close processes the other beam's queued blocks."""
    from mdsm_amd.server import CANCELLED, Server
    M = gpu
    print("case server cancel/close held", flush=True)
    xs = [_stream(M, C, N, 3, 90 + i, [(30.0, 1500, 4, 3.0)]) for i in range(2)]
    srv = Server(max_batch=4)
    ba, bb = (srv.open_beam(**_beam_kwargs(C, N, RANGES)) for _ in range(2))
    srv.hold(True)
    for b, x in ((ba, xs[0]), (bb, xs[1])):
        for k in range(3):
            assert srv.try_push(b, np.ascontiguousarray(x[:, k * N:(k + 1) * N])) is not None
    res = {}
    ida, idb = ba.id, bb.id
    srv.cancel_beam(ba)
    srv.close_beam(bb)       # must not wait for the hold to be lifted
    _collect(srv, res, retain=False)
    assert sorted(s for (bid, s) in res if bid == ida) == [0, 1, 2]
    assert all(res[(ida, s)][2].status == CANCELLED for s in range(3))
    assert all(res[(idb, s)][2].status == 0 for s in range(3))
    srv.close()


def test_push_racing_close(gpu):
    """Pushes from another thread while the beam closes: every push either
    fails cleanly or its block is processed and reported before the beam is
    freed."""
    from mdsm_amd.server import Server
    M = gpu
    print("case server push vs close", flush=True)
    x = _stream(M, C, N, 1, 5, [])
    srv = Server(max_batch=4, queue_depth=8)
    b = srv.open_beam(**dict(_beam_kwargs(C, N, RANGES), retain_output=False))
    ok, errs = [], []

    def pusher():
        for _ in range(1 << 20):     # until the beam is gone
            try:
                if srv.try_push(b, x) is not None:
                    ok.append(1)
            except M.MdsmError as e:
                errs.append(e.code)
                return

    th = threading.Thread(target=pusher)
    th.start()
    import time
    t0 = time.monotonic()
    while not ok and not errs and time.monotonic() - t0 < 30:
        time.sleep(0.001)
    assert ok, f"no push succeeded ({errs})"
    srv.close_beam(b)
    th.join()
    res = {}
    _collect(srv, res, retain=False)
    assert len(res) == len(ok) and all(v[2].status == 0 for v in res.values())
    assert errs and errs[0] in (-1, -6)
    srv.close()


def test_export_import_continues_the_stream(gpu):
    """Blocks 0-1 on one beam, its state exported and imported into a beam of
    another server, blocks 2-3 there: outputs and candidates equal one beam
    that ran all four blocks."""
    from mdsm_amd.server import Server
    M = gpu
    print("case server export/import", flush=True)
    nblk = 4
    Mo = _overlap(C, N, RANGES)
    x = _stream(M, C, N, nblk, 123, [(35.0, 2 * N - Mo - 5, 8, 3.0), (70.0, 1200, 4, 3.0)])
    blocks = [np.ascontiguousarray(x[:, k * N:(k + 1) * N]) for k in range(nblk)]
    srv = Server()
    b = srv.open_beam(**_beam_kwargs(C, N, RANGES))
    whole = {}
    for blk in blocks:
        srv.push(b, blk)
        srv.drain()
        _collect(srv, whole)
    srv.close()
    s1 = Server()
    b1 = s1.open_beam(**_beam_kwargs(C, N, RANGES))
    part = {}
    for blk in blocks[:2]:
        s1.push(b1, blk)
    state = s1.export_state(b1)
    _collect(s1, part)
    s1.close_beam(b1)
    s1.close()
    s2 = Server()
    b2 = s2.open_beam(**_beam_kwargs(C, N, RANGES))
    s2.import_state(b2, state)
    for blk in blocks[2:]:
        assert s2.push(b2, blk) in (2, 3)
    s2.drain()
    _collect(s2, part)
    for k in range(nblk):
        bid = b1.id if k < 2 else b2.id
        out, c, _ = part[(bid, k)]
        wo, wc, _ = whole[(b.id, k)]
        assert np.array_equal(out, wo), f"block {k}: output differs after re-admission"
        assert _key(c) == _key(wc), f"block {k}: candidates differ after re-admission"
    # a state of another plan is refused
    b3 = s2.open_beam(**_beam_kwargs(C, N, ((0.0, 1.0, 40, 1),)))
    with pytest.raises(M.MdsmError):
        s2.import_state(b3, state)
    s2.close()


def test_candidate_overflow_is_flagged(gpu):
    from mdsm_amd.server import OVERFLOW, Server
    M = gpu
    print("case server candidate overflow", flush=True)
    x = _stream(M, C, N, 1, 8, [])
    srv = Server()
    kw = _beam_kwargs(C, N, RANGES)
    kw.update(snr_thr=1.0, cand_cap=5)
    b = srv.open_beam(**kw)
    srv.push(b, x)
    r, c = srv.poll(wait_ms=-1)
    assert r.status == 0 and r.flags & OVERFLOW
    assert r.ncand > r.nreturned and r.nreturned == len(c) == 5 * len(RANGES)
    assert np.all(c["width"] > 0)
    srv.close()


def test_device_push_matches_host_push(gpu):
    from mdsm_amd.server import Server
    M = gpu
    print("case server device push", flush=True)
    x = _stream(M, C, N, 3, 41, [(45.0, 3000, 4, 3.0)])
    outs = []
    for mode in ("host", "device"):
        srv = Server()
        b = srv.open_beam(**_beam_kwargs(C, N, RANGES))
        res, keep = {}, []
        for k in range(3):
            blk = np.ascontiguousarray(x[:, k * N:(k + 1) * N])
            if mode == "host":
                srv.push(b, blk)
            else:
                d = M.DeviceBuffer.from_host(blk)
                keep.append(d)           # kept until its result has been polled
                assert srv.push_device(b, d, N) == k
        srv.drain()
        _collect(srv, res)
        outs.append([res[(b.id, k)] for k in range(3)])
        srv.close()
    for (o1, c1, _), (o2, c2, _) in zip(*outs):
        assert np.array_equal(o1, o2) and _key(c1) == _key(c2)


def test_bright_wide_beam_detector_exact_at_the_float32_bound(gpu):
    """1024 uint8 channels near full scale (mean 240): the window sums of the
    widest boxcar reach 1.6e7, just under 2^24, where the server runs its
    detector's sum tree in float32 (exact for these integer series).  Every
    block's output, and its candidates with their S/N, equal the oracle's
    FP64 streaming search."""
    from mdsm_amd.server import Server
    M = gpu
    print("case server bright 1024-channel beam", flush=True)
    C2, N2, nblk = 1024, 2048, 3
    ranges = ((0.0, 1.0, 32, 1),)
    assert 255 * C2 * 64 < 1 << 24
    x = M.synth_filterbank(C2, N2 * nblk, "u8", seed=4242, mean=240, sigma=8,
                           fch1=1500.0, foff=-300.0 / C2,
                           pulses=[(12.0, 900, 6, 1.0), (20.0, N2 + 300, 40, 0.6),
                                   (8.0, 2 * N2 - 200, 64, 0.5)])
    srv = Server(max_batch=2)
    b = srv.open_beam(**_beam_kwargs(C2, N2, ranges))
    res = {}
    for k in range(nblk):
        blk = np.ascontiguousarray(x[:, k * N2:(k + 1) * N2])
        while not srv.push_seq(b, blk, k):
            srv.drain()
            _collect(srv, res)
    srv.drain()
    _collect(srv, res)
    ref = _oracle_beam(x, b, N2, nblk, ranges, C2)
    ncand = 0
    for k in range(nblk):
        out, c, r = res[(b.id, k)]
        assert r.status == 0
        assert np.array_equal(out, ref[k][0]), f"block {k}: output != oracle"
        got, exp = _key(c), _key(ref[k][1])
        assert [t[:3] for t in got] == [t[:3] for t in exp], f"block {k}: candidates != oracle"
        for g, e in zip(got, exp):
            assert abs(g[3] - e[3]) <= 1e-6 * abs(e[3]), f"block {k}: S/N {g} != {e}"
        ncand += len(got)
    assert ncand > 0
    assert float(np.max(ref[0][0])) * 64 > 0.9 * (1 << 24), "sums do not approach the bound"
    srv.close()
