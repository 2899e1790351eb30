"""DM plan generation (the reference's python/ddplan.py restated in C++):
equal to the oracle restatement, and the properties the algorithm fixes."""
import math

import numpy as np
import pytest

import oracle

ALLOW = oracle._ALLOW_DDM
CASES = [  # lodm, hidm, f_ctr, bw, nchans, dt, nsubs
    (0.0, 1000.0, 1350.0, 300.0, 1024, 64e-6, 0),
    (0.0, 1500.0, 1300.0, 400.0, 4096, 54.6e-6, 0),
    (0.0, 1500.0, 1300.0, 400.0, 4096, 54.6e-6, 64),
    (10.0, 500.0, 150.0, 6.0, 256, 5.12e-6, 32),
    (0.0, 100.0, 350.0, 100.0, 2048, 81.92e-6, 0),
]


def _dm_smear(dm, bw, f):
    return 1000.0 * dm * bw / (0.0001205 * f ** 3.0)


@pytest.mark.parametrize("case", CASES)
def test_ddplan_equals_oracle(case):
    import mdsm_amd as M
    lo, hi, f, bw, nch, dt, ns = case
    got = M.ddplan(lo, hi, f, bw, nch, dt, ns)
    exp = oracle.ddplan(lo, hi, f, bw, nch, dt, ns)
    assert len(got) == len(exp)
    for g, e in zip(got, exp):
        for k in ("binsize", "ndms", "ncalls", "calldms"):
            assert g[k] == e[k], (k, g, e)
        for k in ("lowdm", "highdm", "dmstep", "sub_dmstep"):
            assert g[k] == np.float32(e[k]), (k, g, e)


@pytest.mark.parametrize("case", CASES)
def test_ddplan_properties(case):
    lo, hi, f, bw, nch, dt, ns = case
    ms = oracle.ddplan(lo, hi, f, bw, nch, dt, ns)
    dtms = 1000.0 * dt
    assert ms[0]["lowdm"] == lo and ms[-1]["highdm"] >= hi
    for a, b in zip(ms, ms[1:]):
        assert b["lowdm"] == a["highdm"]                     # contiguous
        assert b["binsize"] == 2 * a["binsize"]              # one more factor of two per pass
        assert b["dmstep"] >= a["dmstep"]
        # the step is the last allowed one whose full-band smearing stays below
        # 1.2x the effective sampling time (unless kept from the previous pass)
        i = ALLOW.index(b["dmstep"])
        nxt = ALLOW[i + 1]
        assert _dm_smear(0.5 * nxt, bw, f) >= 1.2 * dtms * b["binsize"] or nxt > 1000
    for m in ms:
        assert m["dmstep"] in ALLOW and m["binsize"] & (m["binsize"] - 1) == 0
        assert math.isclose(m["highdm"] - m["lowdm"], m["ndms"] * m["dmstep"], rel_tol=1e-12)
        if ns:
            assert m["ndms"] == m["ncalls"] * m["calldms"] and m["calldms"] % 2 == 0
            assert math.isclose(m["sub_dmstep"], m["calldms"] * m["dmstep"], rel_tol=1e-12)
            # the next even number of DMs per call would make the subband smearing
            # exceed 0.8 of the smaller of the step smearing and the sampling time
            nxt = _dm_smear(0.5 * (m["calldms"] + 2) * m["dmstep"], bw / ns, f)
            lim = 0.8 * min(_dm_smear(0.5 * m["dmstep"], bw, f), dtms * m["binsize"])
            assert nxt > lim
        else:
            assert m["ncalls"] == 1 and m["calldms"] == m["ndms"]


def test_first_pass_downsampling_closed_form():
    # coarse channels at low frequency: channel smearing dominates, the first pass
    # downsamples to the smallest power of two reaching 1.2x the channel smearing
    lo, hi, f, bw, nch, dt = 50.0, 200.0, 150.0, 6.0, 32, 5.12e-6
    ms = oracle.ddplan(lo, hi, f, bw, nch, dt)
    chan = _dm_smear(lo + 0.5 * ALLOW[0], bw / nch, f)
    okval = max(chan, _dm_smear(0.5 * ALLOW[0], bw, f), 1000 * dt)
    okval = okval if okval > 1.2 * chan else 1.2 * chan
    ds = 1
    while 1000 * dt * ds * 2 < okval:
        ds *= 2
    assert ms[0]["binsize"] == ds and ds > 1


def test_generated_plan_builds():
    import mdsm_amd as M
    ms = M.ddplan(0.0, 1000.0, 1350.0, 300.0, 1024, 64e-6)
    ranges = [(m["lowdm"], m["dmstep"], m["ndms"], m["binsize"]) for m in ms]
    p = M.Plan.brute(1024, 1500.0, -300.0 / 1024, 64e-6, 1 << 20, "u8", ranges)
    assert p.info.ndms_total == sum(m["ndms"] for m in ms)


# ---------------------------------------------------------------- hand-worked plans
# Two observation setups worked by hand from the smearing formulas of the
# reference's python/ddplan.py (dm_smear :126-132, BW_smear :134-141,
# guess_DMstep :143-149, subband_smear :151-161, dedisp_method :22-63 and
# dm_steps :166-253), with dt = 100 us, f_ctr = 1400 MHz, BW = 100 MHz, 256
# channels (channel width 0.390625 MHz); K = 0.0001205 * 1400^3 = 330652.0.
#
# Setup A, DM 0-2000, no subbands.  dm_smear(DM, bw) = 1000 DM bw / K ms.
#   Minimum smearings 5.9e-8 (channel), 1.5e-5 (band) < dt = 0.1 ms: no
#   downsampling at first.  guess_DMstep = 1e-4 K / 50 = 0.6613; the allowed
#   steps below 1.2 x 0.6613 = 0.794 end at 0.5.
#   Pass 1 (k 1, dDM 0.5): BW smearing dm_smear(0.25, 100) = 0.07561 ms; other
#     smearing sqrt(0.1^2 + 0.1^2 + 0.07561^2) = 0.16036; cross DM
#     2 x 0.001 x 0.16036 / 0.390625 x K = 271.49 -> ceil(271.49 / 0.5) = 543
#     DMs, up to 271.5.
#   Pass 2 (k 2, eff dt 0.2): BW_smear(1.0) = 0.1512 < 0.24 <= BW_smear(2.0) =
#     0.3024: dDM 1.0; other sqrt(0.01 + 0.04 + 0.1512^2) = 0.26994; cross
#     456.99 -> ceil(185.49) = 186 DMs, up to 457.5.
#   Pass 3 (k 4, 0.4): BW_smear(3.0) = 0.4536 < 0.48 <= BW_smear(5.0): dDM 3;
#     other 0.61302; cross 1037.8 -> ceil(193.44) = 194 DMs, up to 1039.5.
#   Pass 4 (k 8, 0.8): BW_smear(5.0) = 0.7561 < 0.96 <= BW_smear(10): dDM 5;
#     other 1.10529; cross 1871.2 -> ceil(166.34) = 167 DMs, up to 1874.5.
#   Pass 5 (k 16, 1.6): BW_smear(10) = 1.5122 < 1.92 <= BW_smear(20): dDM 10;
#     cross 3730.9 > 2000 -> 2000: ceil(12.55) = 13 DMs, up to 2004.5.
# Setup B, DM 0-100, 16 subbands.  dDM 0.5 as above; subband smearing of
#   (n + 2) x 0.5 over 100/16 = 6.25 MHz: n = 2..10 stay below 0.8 x
#   min(0.07561, 0.1) = 0.06049 (n = 10: dm_smear(3.0, 6.25) = 0.05671), n = 12
#   reaches dm_smear(3.5, 6.25) = 0.06616: 12 DMs per call, subband step 6.0,
#   subband smearing 0.05671; cross DM 287.96 > 100 -> 100: ceil(100 / 0.5) =
#   200 DMs -> ceil(200 x 0.5 / 6.0) = 17 calls of 12 = 204 DMs, up to 102.
HAND_A = [  # lowdm, highdm, dmstep, binsize, ndms
    (0.0, 271.5, 0.5, 1, 543), (271.5, 457.5, 1.0, 2, 186), (457.5, 1039.5, 3.0, 4, 194),
    (1039.5, 1874.5, 5.0, 8, 167), (1874.5, 2004.5, 10.0, 16, 13)]


def test_ddplan_hand_worked_brute_force():
    import mdsm_amd as M
    ms = M.ddplan(0.0, 2000.0, 1400.0, 100.0, 256, 1e-4)
    got = [(m["lowdm"], m["highdm"], m["dmstep"], m["binsize"], m["ndms"]) for m in ms]
    assert got == HAND_A
    assert all(m["ncalls"] == 1 and m["calldms"] == m["ndms"] for m in ms)
    # the same setup up to DM 100 stops after the first pass, at exactly 100
    one = M.ddplan(0.0, 100.0, 1400.0, 100.0, 256, 1e-4)
    assert [(m["lowdm"], m["highdm"], m["ndms"]) for m in one] == [(0.0, 100.0, 200)]


def test_ddplan_hand_worked_subband():
    import mdsm_amd as M
    (m,) = M.ddplan(0.0, 100.0, 1400.0, 100.0, 256, 1e-4, 16)
    assert (m["lowdm"], m["highdm"], m["dmstep"], m["sub_dmstep"], m["binsize"]) == \
        (0.0, 102.0, 0.5, 6.0, 1)
    assert (m["ndms"], m["ncalls"], m["calldms"]) == (204, 17, 12)


@pytest.mark.gpu
def test_generated_plans_run_on_the_gpu(gpu):
    """A generated multi-pass brute-force plan and a generated subband plan,
    dedispersed on the GPU: every row's total exact (uint8) and sampled rows
    bit-exact against the oracle; the subband plan within 1e-5 of the FP64
    two-stage oracle."""
    from test_gpu_configs import _check_totals, _gpu_row_totals, _row_totals
    M = gpu
    print("case generated plans on the GPU", flush=True)
    C, f_ctr, bw, dt = 256, 1350.0, 300.0, 64e-6
    fch1, foff = f_ctr + bw / 2, -bw / C
    ms = M.ddplan(0.0, 500.0, f_ctr, bw, C, dt)
    assert len(ms) >= 3
    ranges = [(m["lowdm"], m["dmstep"], m["ndms"], m["binsize"]) for m in ms]
    S = 1 << 15
    x = M.synth_filterbank(C, S, "u8", seed=21, mean=128, sigma=16, fch1=fch1, foff=foff,
                           pulses=[(120.0, 4000, 8, 3.0), (400.0, 9000, 16, 2.0)])
    p = M.Plan.brute(C, fch1, foff, dt, S, "u8", ranges)
    din = M.DeviceBuffer.from_host(x)
    dout = M.DeviceBuffer(p.info.out_elems * 4)
    dw = M.DeviceBuffer(max(p.info.work_bytes, 256))
    p.dedisperse(din, S, dout, dw)
    M.synchronize()
    for r, R in enumerate(p.ranges):
        sh = p.shift_table(r)
        tot, _ = _row_totals(x, sh, R.nout, R.downsample)
        got = _gpu_row_totals(dout, R.ndms, R.nout, R.out_offset, R.downsample, True)
        _check_totals(f"generated pass {r} (k {R.downsample})", got, tot, tot, True)
        rows = [0, R.ndms - 1]
        g = dout.download_rows(np.float32, R.ndms, R.nout, R.nout, offset=R.out_offset * 4)
        ref = oracle.range_f32(x, R.downsample, sh[rows], R.nout)
        assert np.array_equal(g[rows], ref)
    sbs = M.ddplan(0.0, 100.0, f_ctr, bw, C, dt, 16)
    assert len(sbs) >= 2 and sbs[1]["binsize"] == 2
    xs = M.synth_filterbank(C, S, "f16", seed=22, fch1=fch1, foff=foff,
                            pulses=[(50.0, 6000, 4, 1.0)])
    ps = M.Plan.subband(C, fch1, foff, dt, S, 16, sbs, dtype="f16")
    din2 = M.DeviceBuffer.from_host(xs)
    dout2 = M.DeviceBuffer(ps.info.out_elems * 4)
    ps.dedisperse(din2, S, dout2)
    M.synchronize()
    for r, sb in enumerate(sbs):
        R = ps.ranges[r]
        k = R.downsample
        co, fi = ps.shift_table(r), ps.fine_table(r, 16)
        got = dout2.download(np.float32, (R.ndms, R.nout), offset=R.out_offset * 4)
        # stage-1 series run nout + max(fine) samples: zeros past the block
        # keep the oracle's reads in bounds (no valid output reaches them)
        pad = np.zeros((C, k * (int(fi.max()) + int(co.max()) + 1)), xs.dtype)
        ref = oracle.subband_f64(np.concatenate([xs, pad], axis=1), k, 16, sb["ncalls"],
                                 sb["calldms"], co, fi, R.nout + int(fi.max()), R.nout)
        e = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        print(f"generated subband pass {r} (k {k}): {R.ndms} DMs, rel {e:.3e}")
        assert e <= 1e-5
