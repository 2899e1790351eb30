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
