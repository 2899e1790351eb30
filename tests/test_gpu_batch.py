"""Batches: many plans of uneven shapes (channels, samples, dtype, DM ranges,
decimation, brute force and subband) in one launch per stage, eager and as a
replayed HIP graph.  Every plan's output must equal the plan run alone,
bit for bit, and the oracle."""
import numpy as np
import pytest

import oracle
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _plans(M, seed=7, n=9):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        C = int(rng.choice([64, 128, 256]))
        dtype = "u8" if i % 3 != 2 else "f16"
        S = int(rng.integers(9000, 30000))
        foff = -300.0 / C
        nr = int(rng.integers(1, 4))
        ranges, lo = [], 0.0
        for r in range(nr):
            nd = int(rng.integers(20, 140))
            k = int(rng.choice([1, 2, 3, 4, 8]))
            step = float(rng.choice([0.25, 0.5, 1.0]))
            ranges.append((lo, step, nd, k))
            lo += nd * step
        pulses = [(float(rng.uniform(5, lo)), int(rng.integers(2000, S - 3000)), 4, 2.0)
                  for _ in range(int(rng.integers(0, 5)))]
        if dtype == "u8":
            x = M.synth_filterbank(C, S, "u8", seed=100 + i, mean=128, sigma=16, foff=foff,
                                   pulses=pulses)
        else:
            x = M.synth_filterbank(C, S, "f16", seed=100 + i, gain=(0.5, 2.0), foff=foff,
                                   pulses=pulses)
        p = M.Plan.brute(C, 1500.0, foff, 64e-6, S, dtype, ranges)
        out.append((p, x))
    # one subband plan in the mix
    C, S = 128, 20000
    x = M.synth_filterbank(C, S, "u8", seed=99, mean=128, sigma=16)
    p = M.Plan.subband(C, 1500.0, -300.0 / C, 64e-6, S, 16,
                       [dict(lowdm=6.0, dmstep=0.6, sub_dmstep=12.0, ncalls=4, calldms=20)],
                       dtype="u8")
    out.append((p, x))
    return out


def _alone(M, plans):
    res = []
    for p, x in plans:
        din = M.DeviceBuffer.from_host(x)
        dout = M.DeviceBuffer(p.info.out_elems * 4)
        p.dedisperse(din, x.shape[1], dout)
        M.synchronize()
        res.append(dout.download(np.float32, (p.info.out_elems,)))
    return res


@pytest.mark.parametrize("persistent", [0, 1])
def test_batch_equals_plans_alone(gpu, persistent):
    M = gpu
    print(f"case batch persistent={persistent}", flush=True)
    plans = _plans(M)
    ref = _alone(M, plans)
    for p, _ in plans:
        # a grid smaller than the work queue, so that persistent mode really pulls items
        p.set_tuning(persistent=persistent, chunk_rows=4, grid=40 if persistent else 0)
    b = M.Batch([p for p, _ in plans])
    ins = [M.DeviceBuffer.from_host(x) for _, x in plans]
    outs = [M.DeviceBuffer(p.info.out_elems * 4) for p, _ in plans]
    b.run(ins, [x.shape[1] for _, x in plans], outs)
    M.synchronize()
    for i, ((p, x), o) in enumerate(zip(plans, outs)):
        got = o.download(np.float32, (p.info.out_elems,))
        assert np.array_equal(got, ref[i]), f"plan {i} differs in the batch"
    # and the oracle, for the brute-force plans
    for i, (p, x) in enumerate(plans[:-1]):
        for r, R in enumerate(p.ranges):
            sh = p.shift_table(r)
            exp = oracle.range_f64(x, R.downsample, sh, R.nout)
            got = ref[i][R.out_offset:R.out_offset + R.ndms * R.nout].reshape(R.ndms, R.nout)
            if x.dtype == np.uint8:
                assert np.array_equal(got.astype(np.float64), exp.astype(np.float32).astype(np.float64))
            else:
                e, _ = rel_err(got, exp)
                assert e <= 1e-5, (i, r, e)
    b.close()


def test_batch_graph_replay_matches_eager(gpu):
    """Captured step replayed several times, with eager work and host
    synchronisation between replays; every replay equals the eager result."""
    M = gpu
    print("case batch graph replay", flush=True)
    plans = _plans(M, seed=8, n=7)
    for p, _ in plans:
        p.set_tuning(persistent=1, chunk_rows=6, grid=40)
    b = M.Batch([p for p, _ in plans])
    ins = [M.DeviceBuffer.from_host(x) for _, x in plans]
    strides = [x.shape[1] for _, x in plans]
    outs = [M.DeviceBuffer(p.info.out_elems * 4) for p, _ in plans]
    s = M.Stream()
    b.run(ins, strides, outs, stream=s)
    s.synchronize()
    eager = [o.download(np.float32, (p.info.out_elems,)) for (p, _), o in zip(plans, outs)]
    b.capture(ins, strides, outs, stream=s)
    for rep in range(3):
        for o in outs:
            o.fill_u32(0x7fc00000)                   # poison with NaN between replays
        M.synchronize()
        b.replay(s)
        s.synchronize()
        for i, ((p, _), o) in enumerate(zip(plans, outs)):
            got = o.download(np.float32, (p.info.out_elems,))
            assert np.array_equal(got, eager[i]), f"replay {rep}, plan {i}"
        # eager work in between
        plans[0][0].dedisperse(ins[0], strides[0], outs[0])
        M.synchronize()
    b.close()
    s.close()


def test_batch_autotune_keeps_results(gpu):
    M = gpu
    print("case batch autotune", flush=True)
    plans = _plans(M, seed=9, n=5)
    ref = _alone(M, plans)
    b = M.Batch([p for p, _ in plans])
    ins = [M.DeviceBuffer.from_host(x) for _, x in plans]
    strides = [x.shape[1] for _, x in plans]
    outs = [M.DeviceBuffer(p.info.out_elems * 4) for p, _ in plans]
    t = b.autotune(ins, strides, outs, reps=1)
    assert t["chunk_rows"] in (2, 3, 4, 6, 8, 12, 16) and t["persistent"] in (0, 1)
    for p, _ in plans:
        assert p.tuning()["chunk_rows"] == t["chunk_rows"]
    b.run(ins, strides, outs)
    M.synchronize()
    for i, ((p, _), o) in enumerate(zip(plans, outs)):
        assert np.array_equal(o.download(np.float32, (p.info.out_elems,)), ref[i])
    b.close()
