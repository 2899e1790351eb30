#!/usr/bin/env python3
"""Benchmark of the dedispersion hot path on one MI355X.

Workload (default, "cfg5"): the served pipeline -- 32 beams of 1024-channel
uint8 blocks of 2^16 samples (tsamp 64 us, 1200-1500 MHz), each beam with a
device-resident max-delay overlap and detector carry; per block: median/MAD
clip at 5 sigma, 2048-trial brute-force dedispersion over DM 0-1000 in FP32,
boxcar detection w = 1..64 at S/N > 6; pulses in 1% of blocks.  A step is one
block of every beam, the blocks already resident in HBM.  The line carries
the real-time factor, p50/p99 block latency at the real Poisson arrival rate,
the HBM roofline of the dedispersion kernel, the ratio to this build's
single-beam float32 kernel, the reference CPU dedisperser on a bounded
sample, and the parity leg (row checksums, FP64 error, candidates against
the oracle).

Other workloads (--workload): cfg1 (plain float32 brute force), cfg2 (3-range
uint8 plan with decimation), cfg3 (two-stage subband, float16), cfg4 (64
uneven beams in one batch, graph replay), stages (the other kernels).

    python bench.py [--steps K] [--warmup W] [--workload cfg5|cfg1|cfg2|cfg3|cfg4|stages]
                    [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
VALU_ADD_PEAK_T = 39.3         # unpacked fp32 adds/s (256 CU x 2.4 GHz x 64), T/s
# LDS read roof, measured on MI355X (profiles/lds_probe.md, mdsm_probe_lds_read):
# ds_read_b128 at 8 waves per SIMD 135.1 TB/s; ds_read_b64, the integer
# kernel's read, 97.4 TB/s at 4 waves per SIMD and 108 TB/s at 8.
LDS_PEAK_TBS = 135.1
LDS_PEAK_SOURCE = "profiles/lds_probe.md: ds_read_b128, 8 waves per SIMD"
LDS_B64_AT_4_WAVES_TBS = 97.4
# LDS write roof (mdsm_probe_lds_write): ds_write_b64 at 4 waves per SIMD
LDS_WRITE_B64_AT_4_WAVES_TBS = 50.8


def probe_lds_rates(M):
    """ds_read_b64 and ds_write_b64 rates (TB/s) at 4 waves per SIMD on this
    GPU, best of three probe launches each (a few ms); None if the probe
    fails."""
    try:
        return (max(M.probe_lds_read(8, 4) for _ in range(3)),
                max(M.probe_lds_write(8, 4) for _ in range(3)))
    except Exception as e:  # the roof then falls back to the profile's rates
        print(f"[bench] LDS probe failed: {e}", file=sys.stderr, flush=True)
        return None


def lds_roof(reads, writes, seconds, live=None):
    """The dedispersion kernel's LDS roof at its occupancy (4 waves per SIMD):
    the trial reads at the ds_read_b64 rate plus the staging stores at the
    ds_write_b64 rate, against the measured time.  The rates are those
    probed on this GPU in this run (live), else the profile's."""
    def roof(rd, wr):
        return reads / (rd * 1e12) + writes / (wr * 1e12)
    prof_s = roof(LDS_B64_AT_4_WAVES_TBS, LDS_WRITE_B64_AT_4_WAVES_TBS)
    rd, wr = live if live else (LDS_B64_AT_4_WAVES_TBS, LDS_WRITE_B64_AT_4_WAVES_TBS)
    roof_s = roof(rd, wr)
    return {"read_bytes": reads, "write_bytes": writes, "roof_ms": roof_s * 1e3,
            "rw_frac": roof_s / seconds,
            "read_b64_tbs": rd, "write_b64_tbs": wr,
            "rates_source": ("probed on this GPU in this run (mdsm_probe_lds_read/_write, 8 bytes "
                             "per lane, 4 waves per SIMD, best of 3)") if live
                            else "profiles/lds_probe.md",
            "rw_frac_at_profile_rates": prof_s / seconds,
            "rw_note": "LDS time of the trial reads (ds_read_b64) and staging stores "
                       "(ds_write_b64) at 4 waves per SIMD, each at its probed rate, over the "
                       "measured kernel time; rw_frac_at_profile_rates prices them at "
                       "profiles/lds_probe.md's 97.4 and 50.8 TB/s"}
# BASELINE's headline metric: throughput, its fraction of the HBM roofline,
# real-time factor and p99 block latency, against the plain float32 kernel
BASELINE_METRIC = ("DM\u00b7chan\u00b7samples/s, %HBM roofline, RT factor, p99 block latency "
                   "vs plain")
PROFILE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")

WORKLOADS = {
    "cfg2": dict(nchans=1024, nsamp=1 << 20, fch1=1500.0, foff=-300.0 / 1024, tsamp=64e-6,
                 dtype="u8", ranges=[(0.0, 0.2, 1000, 1), (200.0, 0.5, 600, 2),
                                     (500.0, 1.0, 500, 4)]),
    "cfg1": dict(nchans=256, nsamp=65536, fch1=1500.0, foff=-0.390625, tsamp=64e-6,
                 dtype="f32", ranges=[(0.0, 0.2, 512, 1)]),
    # two-stage subband: 4096 ch, 1100-1500 MHz, 2^19 f16 samples, 4096 DMs over
    # 0-1500 as 64 coarse DMs x 64 fine, 64 subbands (lowdm = half a call's span,
    # so the calls tile 0-1500 without overlap)
    "cfg3": dict(nchans=4096, nsamp=1 << 19, fch1=1500.0, foff=-400.0 / 4096, tsamp=54.6e-6,
                 dtype="f16", nsubs=64,
                 passes=[dict(lowdm=64 * (1500.0 / 4096) / 2, dmstep=1500.0 / 4096,
                              sub_dmstep=64 * (1500.0 / 4096), ncalls=64, calldms=64)]),
    # batched: 64 beams of uneven shapes, one launch per step, graph replay
    "cfg4": dict(beams=64, seed=4004),
    # the other stages of the path, each at the size it has in configs 2 and 5
    "stages": dict(),
    # serving: 32 beams of 1024-ch uint8 blocks of 2^16 samples, 2048 DMs over 0-1000
    "cfg5": dict(nchans=1024, nsamp=1 << 16, fch1=1500.0, foff=-300.0 / 1024, tsamp=64e-6,
                 dtype="u8", ranges=[(0.0, 1000.0 / 2048, 2048, 1)], beams=32,
                 mean_interval_s=4.194304),
}


def serve(M, w, args):
    """Config 5, the served pipeline: `beams` beams of 1024-channel uint8
    blocks of 2^16 samples (tsamp 64 us, 4.19 s of data each), each beam with
    a device-resident 16176-sample overlap and detector carry, per block: clip
    at 5 sigma (median/MAD), 2048-trial dedispersion over DM 0-1000, boxcar
    detection w = 1..64 at S/N > 6.  Pulses in 1% of blocks.

    A step is one block of every beam.  Timed phase (`value`): steps blocks
    per beam pushed back to back from HBM-resident blocks (mdsm_beam_push_device)
    -- throughput and real-time factor.  Latency phases: Poisson arrivals at
    the real rate (one block per 4.19 s per beam, host pushes through PCIe)
    and time-compressed by --speedup (device pushes; continuous batching under
    load); latency = candidates on the host - arrival.  Beside it: this
    build's staged float32 kernel and the reference-shaped plain float32
    kernel on one beam's block, the reference CPU dedisperser on a bounded
    sample, and a parity leg (one beam, three blocks with a pulse across a
    block edge: row checksums of every output row, FP64 error of sampled rows,
    candidates equal to the oracle's streaming search)."""
    from mdsm_amd.server import Server, now
    C, N, B = w["nchans"], w["nsamp"], w["beams"]
    widths = (1, 2, 4, 8, 16, 32, 64)
    rng = np.random.default_rng(5)
    synth = lambda n, seed, pulses: M.synth_filterbank(
        C, n, "u8", seed=seed, fch1=w["fch1"], foff=w["foff"], tsamp=w["tsamp"], mean=128.0,
        sigma=16.0, pulses=pulses, nthreads=16)
    npool = 8
    print("[bench] cfg5: generating blocks", file=sys.stderr, flush=True)
    pool = [synth(N, 900 + i, []) for i in range(npool - 1)]
    pool.append(synth(N, 990, [(float(rng.uniform(50, 900)), int(rng.integers(1000, N - 20000)),
                                8, 1.5), (float(rng.uniform(50, 900)), 40000, 2, 2.0)]))
    d_pool = [M.DeviceBuffer.from_host(b) for b in pool]
    pulse_rate = 0.01

    npulse = [0]

    def pick(r):  # pulses in 1% of blocks
        if r.random() < pulse_rate:
            npulse[0] += 1
            return npool - 1
        return int(r.integers(0, npool - 1))

    # the autotuner picks the launch shape for this block shape once; every
    # beam plan of the same shape adopts it
    tp = M.Plan.brute(C, w["fch1"], w["foff"], w["tsamp"], N, "u8", w["ranges"])
    S = N + tp.info.maxshift
    tp = M.Plan.brute(C, w["fch1"], w["foff"], w["tsamp"], S, "u8", w["ranges"])
    tx = np.concatenate([pool[1][:, :S - N], pool[0]], axis=1)
    d_tx = M.DeviceBuffer.from_host(tx)
    d_to = M.DeviceBuffer(tp.info.out_elems * 4)
    tuning = tp.autotune(d_tx, S, d_to, reps=2)
    del d_tx, d_to, tx
    srv = Server(max_inflight=getattr(args, "max_inflight", 16),
                 max_batch=getattr(args, "max_batch", 8),
                 persistent=getattr(args, "persistent", 0))
    det = dict(clip_nsigma=5.0, snr_thr=6.0, widths=widths, cand_cap=1 << 16)
    beams = [srv.open_beam(C, w["fch1"], w["foff"], w["tsamp"], N, "u8", w["ranges"], **det)
             for _ in range(B)]
    Mo = beams[0].overlap
    units_blk = float(beams[0].info.dm_chan_samples)
    D = int(beams[0].info.ndms_total)
    block_s = N * w["tsamp"]
    results = []

    def collect(wait_ms):
        r, c = srv.poll(wait_ms=wait_ms)
        if r is not None:
            assert r.status == 0, r.status
            results.append((r.beam, r.seq, r.t_arrival, r.t_admit, r.t_done, r.batch, r.ncand,
                            r.flags))
        return r

    def push_dev(b, i):
        while srv.push_device(b, d_pool[i], N) is None:
            collect(-1)

    def push_host(b, i, arr):
        while srv.try_push(b, pool[i], arrival=arr) is None:
            collect(-1)

    sched = np.random.default_rng(55)
    # warm-up rounds, then the timed rounds: every beam's block from HBM
    for _ in range(args.warmup):
        for b in beams:
            push_dev(b, pick(sched))
    srv.drain()
    while collect(0) is not None:
        pass
    results.clear()
    M.timing_reset()
    M.timing_enable(True)
    npulse[0] = 0
    t0 = now()
    for _ in range(args.steps):
        for b in beams:
            push_dev(b, pick(sched))
    srv.drain()
    wall = now() - t0
    pulses_thr = npulse[0]
    M.timing_enable(False)
    while collect(0) is not None:
        pass
    st_thr = srv.stats()
    nblk = args.steps * B
    assert len(results) == nblk
    if getattr(args, "throughput_only", False):   # the profiler's program: timed phase only
        srv.close()
        return {"blocks": nblk + args.warmup * B, "seconds": wall}
    kern = {k: M.timing_get(k) for k in ("dedisperse", "clip", "detect", "decimate")}
    value = nblk * units_blk / wall
    rt_factor = nblk * block_s / wall
    mean_batch_thr = float(np.mean([r[5] for r in results]))
    ded_ms, ded_n = kern["dedisperse"]
    # algorithmic HBM bytes of one block's dedispersion: the ring input
    # [C][M+N] u8 read once, the [D][N] float32 output written once
    alg_blk = C * (N + Mo) * 1.0 + D * N * 4.0
    ded_s_blk = ded_ms / 1e3 / nblk
    achieved = alg_blk / ded_s_blk / 1e9
    from mdsm_amd.profile import read_traffic
    prof = os.path.join(PROFILE_DIR, "served_pipeline_cfg5.md")
    tr = read_traffic(prof)
    lds_bytes_blk = units_blk * 2.0
    occ = {"occupancy_theoretical": tp.occupancy(0)}
    lds_rw = lds_roof(*tp.lds_bytes(0), ded_s_blk, probe_lds_rates(M))
    # latency at the real arrival rate (host pushes: PCIe included)
    print("[bench] cfg5: latency at the real arrival rate", file=sys.stderr, flush=True)

    def poisson_phase(seconds, speed, device):
        results.clear()
        npulse[0] = 0
        events = []
        er = np.random.default_rng(7 if device else 6)
        for i in range(B):
            t = er.exponential(w["mean_interval_s"]) / speed
            while t < seconds:
                events.append((t, i))
                t += er.exponential(w["mean_interval_s"]) / speed
        events.sort()
        t_start = now()
        for (ta, i) in events:
            while now() - t_start < ta:
                collect(1)
            arr = t_start + ta
            if device:
                while srv.push_device(beams[i], d_pool[pick(sched)], N, arrival=arr) is None:
                    collect(-1)
            else:
                push_host(beams[i], pick(sched), arr)
        srv.drain()
        while collect(0) is not None:
            pass
        lat = np.array([r[4] - r[2] for r in results])
        bt = np.array([r[5] for r in results])
        return {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                "max": float(lat.max()), "n": int(lat.size), "seconds_of_arrivals": seconds,
                "mean_batch": float(bt.mean()), "pulse_blocks": npulse[0]}

    lat_real = poisson_phase(args.serve_seconds, 1.0, device=False)
    lat_real["arrivals"] = "Poisson, mean one block per 4.194 s per beam, host pushes (PCIe incl.)"
    print(f"[bench] cfg5: real-rate p99 {lat_real['p99'] * 1e3:.1f} ms; compressed x{args.speedup}",
          file=sys.stderr, flush=True)
    lat_fast = poisson_phase(args.serve_seconds / 4, args.speedup, device=True)
    lat_fast["arrivals"] = f"Poisson compressed x{args.speedup}, device pushes"
    st = srv.stats()
    srv.close()
    del d_pool
    # single-beam float32 comparisons on the same block: this build's staged
    # kernel (the headline ratio) and the reference-shaped plain kernel
    x32 = pool[0].astype(np.float32)
    plan = M.Plan.brute(C, w["fch1"], w["foff"], w["tsamp"], N + Mo, "f32", w["ranges"])
    xin = np.concatenate([np.zeros((C, Mo), np.float32), x32], axis=1)
    din = M.DeviceBuffer.from_host(xin)
    dout = M.DeviceBuffer(plan.info.out_elems * 4)
    reps = 5

    def per_block(fn):
        fn()
        M.synchronize()
        t1 = now()
        for _ in range(reps):
            fn()
        M.synchronize()
        return (now() - t1) / reps

    plain_s = per_block(lambda: plan.dedisperse_plain(din, N + Mo, dout))
    staged_s = per_block(lambda: plan.dedisperse(din, N + Mo, dout))
    plain_rate = plan.info.dm_chan_samples / plain_s
    staged_rate = plan.info.dm_chan_samples / staged_s
    out = {
        "metric": BASELINE_METRIC, "value": value, "unit": "DM*chan*samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8->f32",
        "data": "synthetic (seeded uint8 noise N(128,16), pulses in 1% of blocks)",
        "config": {"workload": f"cfg5: {B} beams x {C} ch u8 blocks of {N} samples "
                   f"(tsamp 64 us, 1200-1500 MHz, overlap {Mo}), {D} DMs 0-1000, clip 5 sigma, "
                   "boxcar w=1..64 S/N>6; step = one block per beam, blocks HBM-resident",
                   "beams": B, "nchans": C, "block_samples": N, "overlap": Mo, "ndms": D,
                   "dm_chan_samples_per_block": units_blk},
        "rt_factor": rt_factor, "block_seconds": block_s,
        "latency_s": lat_real, "latency_compressed_s": lat_fast,
        "p99_below_block_duration": lat_real["p99"] < block_s,
        "vs_plain": value / staged_rate,
        "vs_plain_note": "served uint8 rate / this build's staged single-beam float32 kernel on "
                         "the same block",
        "staged_f32_single_beam": {"value": staged_rate, "ms_per_block": staged_s * 1e3},
        "vs_reference_shaped_plain": value / plain_rate,
        "plain_f32_single_beam": {"value": plain_rate, "ms_per_block": plain_s * 1e3,
                                  "kernel": "reference-shaped: thread per (DM, sample), "
                                            "global gathers"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": tr["traffic_bytes_per_block"] * nblk / max(ded_n, 1)
                     if tr and "traffic_bytes_per_block" in tr else None,
                     "traffic_source": os.path.relpath(prof, os.path.dirname(PROFILE_DIR))
                     if tr else None,
                     "kernel": "dedisperse_kernel", "launches": int(ded_n),
                     "blocks_per_launch": nblk / max(ded_n, 1),
                     "avg_launch_ms": ded_ms / max(ded_n, 1),
                     "alg_bytes_per_block": alg_blk,
                     "alg_bytes_per_launch": alg_blk * nblk / max(ded_n, 1),
                     "note": "brute-force dedispersion is LDS/VALU bound (about 220 adds per "
                             "algorithmic byte here): the 60% HBM target is not reachable under "
                             "algorithmic bytes; the LDS figure is the binding one"},
        "lds": {"bytes_per_block": lds_bytes_blk,
                "achieved_tbs": lds_bytes_blk / ded_s_blk / 1e12, "peak_tbs": LDS_PEAK_TBS,
                "frac": lds_bytes_blk / ded_s_blk / 1e12 / LDS_PEAK_TBS,
                "peak_source": LDS_PEAK_SOURCE, **occ, **lds_rw},
        "kernels_ms_per_block": {k: v[0] / nblk for k, v in kern.items()},
        "mean_batch": mean_batch_thr, "pulse_blocks": pulses_thr, "tuning": tuning,
        "server_stats": {k: int(v) for k, v in st.items()},
        "throughput_phase_stats": {k: int(v) for k, v in st_thr.items()},
    }
    if not args.no_cpu:
        print("[bench] cfg5: CPU baseline", file=sys.stderr, flush=True)
        ndm0 = plan.ranges[0].ndms
        ndm_mt = min(ndm0, 4 * args.cpu_dms)
        rows = dout.download(np.float32, (max(min(args.cpu_dms, ndm0), ndm_mt),
                                          plan.ranges[0].nout))
        out["cpu_baseline"] = cpu_baseline(plan, xin, w, rows, budget_dms=min(args.cpu_dms, ndm0))
        out["cpu_baseline_all_cores"] = cpu_baseline_all_cores(plan, xin, w, rows, ndm_mt)
        print("[bench] cfg5: parity leg", file=sys.stderr, flush=True)
        out.update(serve_parity(M, w, synth, Mo, widths))
    return out


def serve_parity(M, w, synth, Mo, widths, nblk=3):
    """One beam, three consecutive blocks of one stream with pulses inside a
    block and across both block edges, served with retained outputs; checked
    against the oracle: (1) every output row's sum over all its samples
    equals the exact integer sum from channel prefix sums of the clipped
    input (a checksum of every sample); (2) sampled rows equal the FP64
    oracle sample by sample; (3) the candidate set equals the oracle's
    streaming search of the concatenated series."""
    import oracle
    from mdsm_amd.server import Server
    C, N = w["nchans"], w["nsamp"]
    # output sample t of block k is stream sample k N - M + t: a pulse at
    # stream sample k N - M - 8 lies across the edge of blocks k-1 and k
    x = synth(nblk * N, 4242, [(300.0, N - Mo - 8, 16, 1.5), (700.0, 2 * N - Mo - 3, 4, 3.0),
                               (120.0, 30000, 8, 1.5)])
    srv = Server()
    b = srv.open_beam(C, w["fch1"], w["foff"], w["tsamp"], N, "u8", w["ranges"],
                      clip_nsigma=5.0, snr_thr=6.0, widths=widths, cand_cap=1 << 18,
                      retain_output=True)
    outs, cands = [], []
    for k in range(nblk):
        srv.push(b, np.ascontiguousarray(x[:, k * N:(k + 1) * N]))
        r, c = srv.poll(wait_ms=-1)
        assert r.status == 0 and r.seq == k
        outs.append(srv.output(b, k))
        srv.release(b, k)
        cands.append(c)
    R = b.info
    srv.close()
    (lo, step, D, _) = w["ranges"][0]
    sh = oracle.shift_table_ref(w["fch1"], w["foff"], C, lo, step, D, np.float32(w["tsamp"]))
    nout = outs[0].size // D
    # clipped stream with the zero history in front
    clipped = np.zeros((C, Mo + nblk * N), np.uint8)
    for k in range(nblk):
        o, *_ = oracle.clip_u8(np.ascontiguousarray(x[:, k * N:(k + 1) * N]), 0, N, 5.0)
        clipped[:, Mo + k * N:Mo + (k + 1) * N] = o
    rows_ok, worst_abs, worst_rel = True, 0.0, 0.0
    sample_rows = np.linspace(0, D - 1, 12).astype(int)
    for k in range(nblk):
        seg = clipped[:, k * N:k * N + Mo + N]
        P = np.zeros((C, seg.shape[1] + 1), np.int64)
        np.cumsum(seg, axis=1, out=P[:, 1:])
        ci = np.arange(C)[None, :]
        exp = (P[ci, sh + nout] - P[ci, sh]).sum(axis=1)
        got = outs[k].reshape(D, nout).astype(np.float64).sum(axis=1)
        rows_ok = rows_ok and bool(np.array_equal(got, exp.astype(np.float64)))
        ref = oracle.range_f64(np.ascontiguousarray(seg), 1, sh[sample_rows], nout)
        g = outs[k].reshape(D, nout)[sample_rows].astype(np.float64)
        ab = float(np.max(np.abs(g - ref)))
        worst_abs = max(worst_abs, ab)
        worst_rel = max(worst_rel, ab / max(float(np.max(np.abs(ref))), 1e-300))
    series = np.concatenate([o.reshape(D, nout)[:, :N] for o in outs], axis=1)
    ref_c = oracle.boxcar_stream(series, N, list(widths), 6.0)
    key_ref = sorted(zip(ref_c["blk"].tolist(), ref_c["dm"].tolist(), ref_c["t"].tolist(),
                         ref_c["width"].tolist()))
    key_got = sorted((k, int(d), int(t), int(wd)) for k, c in enumerate(cands)
                     for d, t, wd in zip(c["dm"], c["t"], c["width"]))
    edge = [kk for kk in key_got if kk[2] < 0]
    snr_ref = sorted(zip(ref_c["blk"].tolist(), ref_c["dm"].tolist(), ref_c["t"].tolist(),
                         ref_c["width"].tolist(), ref_c["snr"].tolist()))
    snr_got = sorted((k, int(d), int(t), int(wd), float(sn)) for k, c in enumerate(cands)
                     for d, t, wd, sn in zip(c["dm"], c["t"], c["width"], c["snr"]))
    return {
        "error_vs_fp64": {"max_abs": worst_abs, "rel": worst_rel, "tolerance_rel": 1e-5,
                          "rel_kind": REL_KIND,
                          "rows": int(sample_rows.size), "samples_per_row": int(nout),
                          "blocks": nblk, "note": "uint8 input: exact (integer sums)"},
        "row_checksums_exact": rows_ok,
        "candidates": {"match_oracle": key_got == key_ref, "snr_match_oracle": snr_got == snr_ref,
                       "n": len(key_got), "across_block_edges": len(edge),
                       "oracle": "streaming boxcar search of the concatenated series "
                                 "(per-block statistics), FP64"},
    }


def stages_bench(M, args):
    """Every stage beside dedispersion, timed with the library's launch events
    at the size it has in the configurations: corner turn and 2-bit unpack of a
    config-2 block (1024 ch x 2^20), the fused uint8 decimation of config 2
    (k = 2 and 4), and, on a config-5 block, the clipper (1024 ch x 65536 u8)
    and the detectors and detrend on its DM x time output (2048 x 65536 f32).
    Bytes are algorithmic (each input read once, each output written once)."""
    reps = max(1, args.steps // 2)
    res = {}

    def timed(name, key, fn, nbytes, note):
        fn()
        M.synchronize()
        M.timing_reset()
        M.timing_enable(True)
        for _ in range(reps):
            fn()
        M.synchronize()
        ms, n = M.timing_get(key)
        M.timing_enable(False)
        per = ms / max(n, 1) * (n / reps)  # ms per call of fn
        res[name] = {"ms": per, "bytes": nbytes, "gbs": nbytes / per / 1e6,
                     "hbm_frac": nbytes / per / 1e6 / HBM_PEAK_GBS, "note": note}
        print(f"[bench] stage {name}: {per:.3f} ms", file=sys.stderr, flush=True)

    C, S = 1024, 1 << 20
    x = M.synth_filterbank(C, S, "u8", seed=5, mean=128, sigma=16, nthreads=16)
    tm = np.ascontiguousarray(x.T)
    d_tm = M.DeviceBuffer.from_host(tm)
    d_cm = M.DeviceBuffer(C * S)
    timed("corner_turn_u8", "corner_turn",
          lambda: M.corner_turn(d_tm, "u8", S, C, C, d_cm, S),
          2 * C * S, "time-major [2^20][1024] u8 -> channel-major [1024][2^20] u8")
    raw2 = np.random.default_rng(6).integers(0, 256, S * C // 4, dtype=np.uint8)
    d_raw = M.DeviceBuffer.from_host(raw2)
    timed("unpack_2bit", "unpack",
          lambda: M.unpack_spectra(d_raw, 2, S, C, C // 4, d_cm, S, dtype_out="u8"),
          C * S // 4 + C * S, "2-bit spectra [2^20][1024] -> channel-major u8 [1024][2^20]")
    del d_tm, d_raw, tm, raw2
    w2 = WORKLOADS["cfg2"]
    plan = make_plan(M, w2)
    d_in = M.DeviceBuffer.from_host(x)
    d_out = M.DeviceBuffer(plan.info.out_elems * 4)
    d_w = M.DeviceBuffer(max(plan.info.work_bytes, 256))
    timed("decimate_u8_k2_k4", "decimate",
          lambda: plan.dedisperse(d_in, S, d_out, d_w),
          C * S + C * (S // 2) * 2 + C * (S // 4) * 2,
          "config 2: uint8 -> uint16 2- and 4-sums in one pass (timed inside a dedispersion step)")
    del d_in, d_out, d_w, plan
    N5, D5 = 1 << 16, 2048
    blk = np.ascontiguousarray(x[:, :N5])
    d_blk = M.DeviceBuffer.from_host(blk)
    d_clip = M.DeviceBuffer(C * N5)
    timed("clip_u8", "clip",
          lambda: M.clip_u8(d_blk, N5, C, 0, N5, 5.0, d_clip, N5),
          2 * C * N5, "config-5 block: per-channel median/MAD clip, 1024 x 65536 u8")
    y = np.random.default_rng(7).standard_normal((D5, N5), dtype=np.float32)
    d_y = M.DeviceBuffer.from_host(y)
    cap = 1 << 20
    d_c, d_n = M.DeviceBuffer(cap * 16), M.DeviceBuffer(4)
    d_mean, d_sig = M.DeviceBuffer(D5 * 8), M.DeviceBuffer(D5 * 8)
    timed("boxcar_w1_64", "detect",
          lambda: M.detect_boxcar(d_y, D5, N5, N5, [1, 2, 4, 8, 16, 32, 64], 6.0, d_c, cap, d_n,
                                  d_mean, d_sig),
          D5 * N5 * 4, "config-5 output 2048 x 65536 f32: row statistics + boxcar w = 1..64, S/N > 6")
    Hc = M.boxcar_carry_len([1, 2, 4, 8, 16, 32, 64])
    d_carry = M.DeviceBuffer(D5 * Hc * 4)
    d_carry.fill_u32(0)
    timed("boxcar_stream_w1_64", "detect",
          lambda: M.detect_boxcar_stream(d_y, D5, N5, N5, d_carry, Hc, True,
                                         [1, 2, 4, 8, 16, 32, 64], 6.0, d_c, cap, d_n, d_mean,
                                         d_sig),
          D5 * N5 * 4, "config-5 output: streaming search (68 carried samples per row, then the "
          "block), row statistics + boxcar w = 1..64, S/N > 6")
    dmv = M.DeviceBuffer.from_host(np.linspace(0.0, 1000.0, D5, dtype=np.float32))
    d_st = M.DeviceBuffer(M.median5_stats_len(D5, N5) * 8)
    timed("median5", "median5",
          lambda: M.detect_median5(d_y, D5, N5, N5, dmv, 6.0, d_c, cap, d_n, d_st, per_row=True),
          D5 * N5 * 4, "config-5 output: median-of-five detector, per-row statistics")
    timed("detrend_seg1024", "detrend",
          lambda: M.detrend(d_y, D5, N5, N5, 1024),
          2 * D5 * N5 * 4, "config-5 output in place, 1024-sample segments")
    return {"metric": "stage throughput (algorithmic bytes / kernel time)", "unit": "GB/s",
            "n_gpus": 1, "data": "synthetic", "reps": reps,
            "config": {"workload": "stages: corner turn, unpack, decimation at config 2; "
                                   "clip, boxcar, median-of-five, detrend at config 5"},
            "stages": res}


def batch_bench(M, w, args):
    """Config 4: 64 beams of uneven shapes -- channels from {256, 1024, 4096},
    samples uniform in [2^16, 2^19], uint8 or float16, 1-3 DM ranges with
    256-4096 trials in total and decimation 1-8 -- run as one batch: the
    autotuner picks the launch shape for the batch, then steady-state steps
    replay a captured HIP graph.  Every beam has its own input block in HBM,
    made on the device by the seeded generator (uint8: N(128, 16); float16:
    N(0, 1) with per-channel gains in [0.5, 2]) with 0-4 seeded pulses
    inside its plan's DM span."""
    rng = np.random.default_rng(w["seed"])
    fch1, bw, tsamp = 1500.0, 300.0, 64e-6
    plans, beams = [], []
    while len(plans) < w["beams"]:
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
            step = 0.1 * float(k)
            ranges.append((lo, step, int(nd), int(k)))
            lo += nd * step
        try:
            p = M.Plan.brute(C, fch1, -bw / C, tsamp, S, dt, ranges)
        except M.MdsmError:
            continue                      # block too short for that plan's delay: redraw
        plans.append(p)
        beams.append(dict(C=C, S=S, dtype=dt, ranges=ranges))
    units = sum(p.info.dm_chan_samples for p in plans)
    out_bytes = sum(p.info.out_elems * 4 for p in plans)
    print(f"[bench] cfg4: {len(plans)} beams, {units:.3e} DM*chan*samples per step, "
          f"outputs {out_bytes / 1e9:.1f} GB; generating the inputs on the device",
          file=sys.stderr, flush=True)
    prng = np.random.default_rng(w["seed"] + 1)
    ins, npulses = [], 0
    for i, (b, p) in enumerate(zip(beams, plans)):
        C, S, dt = b["C"], b["S"], b["dtype"]
        hi = b["ranges"][-1][0] + b["ranges"][-1][1] * b["ranges"][-1][2]
        pulses = [(float(prng.uniform(0.1 * hi, 0.9 * hi)),
                   int(prng.integers(0, max(1, S - p.info.maxshift))),
                   int(prng.integers(1, 33)), 1.0 if dt == "u8" else 0.5)
                  for _ in range(int(prng.integers(0, 5)))]
        npulses += len(pulses)
        d = M.DeviceBuffer(C * S * (1 if dt == "u8" else 2))
        extra = dict(mean=128.0, sigma=16.0) if dt == "u8" else dict(gain=(0.5, 2.0))
        M.synth_filterbank_device(d, C, S, dt, seed=4000 + i, fch1=fch1, foff=-bw / C,
                                  tsamp=tsamp, pulses=pulses, **extra)
        ins.append(d)
    M.synchronize()
    in_bytes = sum(d.nbytes for d in ins)
    strides = [b["S"] for b in beams]
    outs = [M.DeviceBuffer(p.info.out_elems * 4) for p in plans]
    batch = M.Batch(plans)
    s = M.Stream()
    t0 = time.perf_counter()
    tuning = batch.autotune(ins, strides, outs, stream=s, reps=1)
    tune_s = time.perf_counter() - t0
    print(f"[bench] cfg4: autotuned {tuning} in {tune_s:.1f} s", file=sys.stderr, flush=True)
    batch.run(ins, strides, outs, stream=s)          # eager step, timed
    s.synchronize()
    t0 = time.perf_counter()
    batch.run(ins, strides, outs, stream=s)
    s.synchronize()
    eager_s = time.perf_counter() - t0
    batch.capture(ins, strides, outs, stream=s)
    for _ in range(args.warmup):
        batch.replay(s)
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.replay(s)
    s.synchronize()
    wall = time.perf_counter() - t0
    n_u8 = sum(1 for b in beams if b["dtype"] == "u8")
    return {
        "metric": "DM*chan*samples/s (batched, graph replay)",
        "value": units * args.steps / wall, "unit": "DM*chan*samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8/f16->f32", "data": "synthetic (seeded, generated on the device, per beam)",
        "config": {"workload": f"cfg4: {len(plans)} beams (uint8 {n_u8}, float16 "
                   f"{len(plans) - n_u8}), C in {{256,1024,4096}}, S in [2^16,2^19], "
                   "1-3 ranges, 256-4096 trials, k 1-8; a distinct input block per beam "
                   f"with 0-4 pulses ({npulses} in all)", "seed": w["seed"],
                   "dm_chan_samples_per_step": units, "input_bytes": in_bytes,
                   "output_bytes": out_bytes},
        "tuning": tuning, "autotune_seconds": tune_s,
        "eager_step_ms": eager_s * 1e3, "graph_step_ms": wall * 1e3 / args.steps,
    }


def make_plan(M, w):
    if "passes" in w:
        return M.Plan.subband(w["nchans"], w["fch1"], w["foff"], w["tsamp"], w["nsamp"],
                              w["nsubs"], w["passes"], dtype=w["dtype"])
    return M.Plan.brute(w["nchans"], w["fch1"], w["foff"], w["tsamp"], w["nsamp"], w["dtype"],
                        w["ranges"])


def make_input(M, w, seed=20240601):
    rng = np.random.default_rng(seed)
    if w["dtype"] == "f16":
        pulses = [(float(rng.uniform(100, 1400)), int(rng.integers(10_000, w["nsamp"] - 60_000)),
                   int(rng.integers(1, 33)), 0.5) for _ in range(3)]
        x = M.synth_filterbank(w["nchans"], w["nsamp"], "f16", seed=seed, fch1=w["fch1"],
                               foff=w["foff"], tsamp=w["tsamp"], gain=(0.5, 2.0),
                               pulses=pulses, nthreads=16)
        return x, pulses
    if w["dtype"] == "u8":
        pulses = [(float(rng.uniform(10, 900)), int(rng.integers(50_000, w["nsamp"] - 500_000)),
                   int(rng.integers(1, 33)), 1.0) for _ in range(5)]
        x = M.synth_filterbank(w["nchans"], w["nsamp"], "u8", seed=seed, fch1=w["fch1"],
                               foff=w["foff"], tsamp=w["tsamp"], mean=128.0, sigma=16.0,
                               pulses=pulses, nthreads=16)
    else:
        pulses = [(56.8, 30000, 4, 0.5)]
        x = M.synth_filterbank(w["nchans"], w["nsamp"], "f32", seed=seed, fch1=w["fch1"],
                               foff=w["foff"], tsamp=w["tsamp"], pulses=pulses, nthreads=16)
    return x, pulses


def dedisperse_bytes(plan, w):
    """Algorithmic HBM bytes of the dedispersion launches of one step: every
    input sample read once (the raw block for k=1 ranges, each decimated copy
    for k>1), every output sample written once; for the two-stage subband
    plan the plan's own count (input, stage-1 series written and read back,
    output, tables)."""
    if "passes" in w:
        return float(plan.info.bytes)
    esize = {"u8": 1, "f16": 2, "f32": 4}[w["dtype"]]
    C, S = w["nchans"], w["nsamp"]
    b = 0.0
    ks = sorted({R.downsample for R in plan.ranges})
    for k in ks:
        if k == 1:
            b += C * S * esize
        else:
            b += C * (S // k) * (2 if w["dtype"] == "u8" else 4)
    for R in plan.ranges:
        b += R.ndms * R.nout * 4.0
    return b


def cpu_baseline(plan, x, w, gpu_out, budget_dms=16, nout_sample=1 << 17):
    """Time the reference CPU dedisperser's loop (src/prototypes/cpu_dedisp.c:
    89-96) on a bounded sample of range 0: `budget_dms` DM trials x all
    channels x `nout_sample` output samples.  Prefers the reference's own code
    (oracle/_ref, its main() run in-process on this sample); falls back to the
    oracle's restatement of the same loop.  Returns the JSON object."""
    import oracle
    R = plan.ranges[0]
    C = w["nchans"]
    sh = plan.shift_table(0)[:budget_dms]
    nout = min(nout_sample, R.nout)
    M0 = int(sh[-1, -1])
    x_tm = np.ascontiguousarray(x[:, :nout + M0].T.astype(np.float32))
    adds = float(budget_dms) * C * nout
    sample = (f"range 0 (DM {R.lowdm}..): {budget_dms} DM trials x {C} channels x {nout} "
              f"output samples, time-major float32 input as cpu_dedisp.c")
    res = None
    if oracle.ref() is not None and (C & (C - 1)) == 0:
        import tempfile
        fd_out = os.dup(1)
        with tempfile.TemporaryFile(mode="w+") as tf:
            sys.stdout.flush()
            os.dup2(tf.fileno(), 1)
            try:
                t0 = time.perf_counter()
                shr, out_tm, _ = oracle.ref_cpu_dedisp(w["fch1"], w["foff"], w["tsamp"], C, nout,
                                                       R.lowdm, R.dmstep, budget_dms, x_tm)
                wall = time.perf_counter() - t0
            finally:
                sys.stdout.flush()
                os.dup2(fd_out, 1)
                os.close(fd_out)
            tf.seek(0)
            txt = tf.read()
        ms = None
        for line in txt.splitlines():
            if line.startswith("Time:"):
                ms = float(line.split()[1])
        secs = ms / 1e3 if ms else wall
        ok = bool(np.array_equal(shr, sh)) and bool(
            np.array_equal(out_tm.T, gpu_out[:budget_dms, :nout]))
        res = {"value": adds / secs, "unit": "DM*chan*samples/s", "cores": 1,
               "kind": "reference", "sample": sample + "; reference cpu_dedisp.c main() built "
               "from its own source (-O2), loop time from its own timer",
               "seconds": secs, "matches_gpu": ok}
    else:
        t0 = time.perf_counter()
        out_tm = oracle.dedisperse_ref_tm(x_tm, sh, nout)
        secs = time.perf_counter() - t0
        ok = bool(np.array_equal(out_tm.T, gpu_out[:budget_dms, :nout]))
        res = {"value": adds / secs, "unit": "DM*chan*samples/s", "cores": 1, "kind": "port",
               "sample": sample + "; oracle restatement of the loop", "seconds": secs,
               "matches_gpu": ok}
    return res


def cpu_baseline_all_cores(plan, x, w, gpu_out, ndms, nout_sample=1 << 17):
    """The same loop (oracle restatement of cpu_dedisp.c:89-96) over ndms
    trials of range 0, shared by OpenMP threads over (trial, sample-run)
    pairs: the multi-core CPU figure next to the reference's single thread."""
    import oracle
    R = plan.ranges[0]
    C = w["nchans"]
    sh = plan.shift_table(0)[:ndms]
    nout = min(nout_sample, R.nout)
    M0 = int(sh[-1, -1])
    x_tm = np.ascontiguousarray(x[:, :nout + M0].T.astype(np.float32))
    try:
        nthreads = len(os.sched_getaffinity(0))
    except AttributeError:
        nthreads = os.cpu_count() or 1
    t0 = time.perf_counter()
    out_tm = oracle.dedisperse_ref_tm_mt(x_tm, sh, nout, nthreads)
    secs = time.perf_counter() - t0
    return {"value": float(ndms) * C * nout / secs, "unit": "DM*chan*samples/s",
            "cores": nthreads, "host_cpus": os.cpu_count(), "kind": "port",
            "sample": f"range 0: {ndms} DM trials x {C} channels x {nout} output samples, "
                      f"oracle restatement of cpu_dedisp.c:89-96 on {nthreads} OpenMP threads",
            "seconds": secs, "matches_gpu": bool(np.array_equal(out_tm.T, gpu_out[:ndms, :nout]))}


REL_KIND = ("normwise: max|gpu - fp64| / max|fp64| over the compared rows (per-value relative "
            "error is unbounded where a zero-mean sum crosses zero)")


def error_vs_fp64(plan, x, dout, rows=8, nsamp=1 << 15):
    """Max-abs and normwise relative error (max|gpu - fp64| / max|fp64|) of
    the first `rows` DM rows of every range over their first `nsamp` output
    samples, against the FP64 oracle on the same input."""
    import oracle
    worst_abs, worst_rel, per = 0.0, 0.0, []
    for r, R in enumerate(plan.ranges):
        nr, n = min(rows, R.ndms), min(nsamp, R.nout)
        sh = plan.shift_table(r)[:nr]
        need = min(x.shape[1], (n + int(sh.max()) + 1) * R.downsample)
        ref = oracle.range_f64(np.ascontiguousarray(x[:, :need]), R.downsample, sh, n)
        got = dout.download_rows(np.float32, nr, n, R.nout, offset=R.out_offset * 4)
        ab = float(np.max(np.abs(got.astype(np.float64) - ref)))
        rel = ab / max(float(np.max(np.abs(ref))), 1e-300)
        worst_abs, worst_rel = max(worst_abs, ab), max(worst_rel, rel)
        per.append({"range": r, "rows": nr, "samples": n, "max_abs": ab, "rel": rel})
    return {"max_abs": worst_abs, "rel": worst_rel, "tolerance_rel": 1e-5, "rel_kind": REL_KIND,
            "ranges": per}


def subband_error_vs_fp64(plan, x, w, dout, nsamp=2048):
    """Max-abs and normwise relative error of every DM row of a one-pass
    subband plan over its first `nsamp` output samples, against the FP64
    oracle of the two-stage rule on the same input window."""
    import oracle
    R = plan.ranges[0]
    q = w["passes"][0]
    n = min(nsamp, R.nout)
    co = plan.shift_table(0)
    fi = plan.fine_table(0, w["nsubs"])
    win = np.ascontiguousarray(x[:, :R.maxshift + n])
    ref = oracle.subband_f64(win, 1, w["nsubs"], q["ncalls"], q["calldms"], co, fi,
                             n + int(fi.max()), n)
    got = dout.download_rows(np.float32, R.ndms, n, R.nout)
    ab = float(np.max(np.abs(got.astype(np.float64) - ref)))
    return {"max_abs": ab, "rel": ab / max(float(np.max(np.abs(ref))), 1e-300),
            "tolerance_rel": 1e-5, "rel_kind": REL_KIND, "rows": int(R.ndms), "samples": n}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cfg5", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-dms", type=int, default=64,
                    help="DM trials of the single-thread CPU sample (the all-cores sample takes 4x)")
    ap.add_argument("--serve-seconds", type=float, default=30.0,
                    help="cfg5: seconds of Poisson arrivals at the real rate (the compressed "
                         "phase runs a quarter of that)")
    ap.add_argument("--max-batch", type=int, default=8,
                    help="cfg5: blocks per dedispersion launch of the server")
    ap.add_argument("--max-inflight", type=int, default=16,
                    help="cfg5: blocks on the device at once")
    ap.add_argument("--persistent", type=int, default=0,
                    help="cfg5: server launches on a persistent grid (1: 4 per CU, n > 1: n "
                         "workgroups; 0: one workgroup per item)")
    ap.add_argument("--speedup", type=float, default=20.0,
                    help="cfg5: arrival-time compression of the latency phase")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(beams and DM ranges shard without exchange; see the sharding notes)",
              file=sys.stderr)
        sys.exit(2)
    import mdsm_amd as M
    w = WORKLOADS[args.workload]
    if args.workload == "stages":
        print("[bench] stages", file=sys.stderr, flush=True)
        print(json.dumps(stages_bench(M, args)), flush=True)
        return
    if args.workload == "cfg4":
        print("[bench] cfg4 batched", file=sys.stderr, flush=True)
        print(json.dumps(batch_bench(M, w, args)), flush=True)
        return
    if args.workload == "cfg5":
        print("[bench] cfg5 serving", file=sys.stderr, flush=True)
        print(json.dumps(serve(M, w, args)), flush=True)
        return
    print(f"[bench] workload {args.workload}: generating input", file=sys.stderr, flush=True)
    x, pulses = make_input(M, w)
    plan = make_plan(M, w)
    info = plan.info
    stream = M.Stream()
    din = M.DeviceBuffer.from_host(x)
    dout = M.DeviceBuffer(info.out_elems * 4)
    dwork = M.DeviceBuffer(max(info.work_bytes, 256))
    print(f"[bench] input {x.nbytes/1e9:.2f} GB, output {info.out_elems*4/1e9:.2f} GB, "
          f"work {info.work_bytes/1e9:.2f} GB; warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        plan.dedisperse(din, w["nsamp"], dout, dwork, stream)
    stream.synchronize()
    M.timing_reset()
    M.timing_enable(True)
    M.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.dedisperse(din, w["nsamp"], dout, dwork, stream)
    stream.synchronize()
    M.synchronize()
    wall = time.perf_counter() - t0
    M.timing_enable(False)
    ded_ms, ded_n = M.timing_get("dedisperse")
    dec_ms, dec_n = M.timing_get("decimate")
    print(f"[bench] {args.steps} steps in {wall:.3f} s", file=sys.stderr, flush=True)

    ms_per_step = wall * 1e3 / args.steps
    value = info.dm_chan_samples * args.steps / wall
    # one dedispersion launch per step (two for the subband plan): the
    # roofline is per step's launches
    launches_per_step = max(ded_n, 1) / args.steps
    avg_launch_s = ded_ms / max(ded_n, 1) / 1e3 * launches_per_step
    alg_bytes = dedisperse_bytes(plan, w)
    achieved_gbs = alg_bytes / avg_launch_s / 1e9
    # HBM traffic of the dedispersion kernel from the committed rocprofv3
    # FETCH_SIZE/WRITE_SIZE summary of this workload (copy-calibrated)
    from mdsm_amd.profile import read_traffic
    prof = os.path.join(PROFILE_DIR, f"dedisperse_kernel_{args.workload}.md")
    tr = read_traffic(prof)
    # LDS: the integer path reads one uint16 per (DM, channel, sample), the
    # float path one float32
    dec_adds = sum(w["nchans"] * (w["nsamp"] // k) * (k - 1)
                   for k in sorted({R.downsample for R in plan.ranges}) if k > 1)
    adds_per_launch = info.adds - dec_adds       # one LDS read per dedispersion add
    lds_bytes = adds_per_launch * (2.0 if w["dtype"] == "u8" else 4.0)
    out = {
        "metric": "DM*chan*samples/s",
        "value": value,
        "unit": "DM*chan*samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": w["dtype"] + "->f32",
        "data": "synthetic (seeded Gaussian filterbank with injected pulses)",
        "config": {"workload": f"{args.workload}: {w['nchans']} ch x {w['nsamp']} samples "
                   f"{w['dtype']}, " + (f"ranges {w['ranges']}" if "ranges" in w else
                                        f"{w['nsubs']} subbands, passes {w['passes']}"),
                   "nchans": w["nchans"],
                   "nsamp": w["nsamp"], "ndms": int(info.ndms_total),
                   "dm_chan_samples_per_step": info.dm_chan_samples},
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved_gbs / HBM_PEAK_GBS,
                     "traffic": tr["traffic_bytes_per_launch"] * launches_per_step if tr else None,
                     "traffic_source": os.path.relpath(prof, os.path.dirname(PROFILE_DIR))
                     if tr else None,
                     "kernel": "dedisperse_kernel", "avg_launch_ms": avg_launch_s * 1e3,
                     "alg_bytes_per_launch": alg_bytes},
        "lds": {"bytes_per_launch": lds_bytes,
                "achieved_tbs": lds_bytes / avg_launch_s / 1e12, "peak_tbs": LDS_PEAK_TBS,
                "frac": lds_bytes / avg_launch_s / 1e12 / LDS_PEAK_TBS,
                "note": "binding resource: LDS gathers, one per DM*chan*sample",
                # the read+write roof for integer rows (ds_read_b64 trial reads);
                # float32 rows read through ds_read2st64_b64 at another rate
                **(lds_roof(*[sum(v) for v in zip(*[plan.lds_bytes(i) for i in
                                                     range(2 if info.is_subband else 1)])],
                            avg_launch_s, probe_lds_rates(M)) if w["dtype"] == "u8" else {})},
        "valu": {"adds_per_launch": adds_per_launch,
                 "achieved_tadds": adds_per_launch / avg_launch_s / 1e12,
                 "peak_tadds": VALU_ADD_PEAK_T,
                 "frac": adds_per_launch / avg_launch_s / 1e12 / VALU_ADD_PEAK_T,
                 "note": "one add per DM*chan*sample against one 32-bit add per lane per cycle; "
                         "uint8 rows add two channels of two samples per v_add3_u32 (packed "
                         "16-bit lanes), so the VALU issue rate is below this fraction"},
        "kernels_ms_per_step": {"dedisperse": ded_ms / args.steps,
                                "decimate": dec_ms / args.steps},
    }
    if not args.no_cpu and "ranges" in w:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        ndm0 = plan.ranges[0].ndms
        ndm_mt = min(ndm0, 4 * args.cpu_dms)
        gpu_rows = dout.download(np.float32, (max(min(args.cpu_dms, ndm0), ndm_mt),
                                              plan.ranges[0].nout))
        out["cpu_baseline"] = cpu_baseline(plan, x, w, gpu_rows, budget_dms=min(args.cpu_dms, ndm0))
        out["cpu_baseline_all_cores"] = cpu_baseline_all_cores(plan, x, w, gpu_rows, ndm_mt)
        # the CPU leg also checks sampled rows of every range against FP64
        out["error_vs_fp64"] = error_vs_fp64(plan, x, dout)
    elif not args.no_cpu and "passes" in w and len(w["passes"]) == 1:
        print("[bench] FP64 check", file=sys.stderr, flush=True)
        out["error_vs_fp64"] = subband_error_vs_fp64(plan, x, w, dout)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
