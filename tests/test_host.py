"""Host-side logic of the product library on the CPU (no GPU calls):
plan shift tables and layouts against the oracle, validation, the synthetic
generator, and the C ABI export list."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import mdsm_amd as M
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_abi_exports_every_declared_symbol():
    decl = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            decl |= set(re.findall(r"\b(mdsm_\w+|initialiseMDSM|processSurveyParameters|"
                                   r"next_chunk|start_processing|tearDownMDSM|call_\w+)\s*\(", txt))
    assert len(decl) > 20
    out = subprocess.run(["nm", "-D", "--defined-only", M.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (\w+)", out))
    missing = sorted(decl - exported)
    assert not missing, f"declared but not exported: {missing}"
    lib = M.lib()
    for s in decl:
        assert getattr(lib, s) is not None


@pytest.mark.parametrize("cfg", [
    dict(C=256, fch1=1500.0, foff=-0.390625, ts=64e-6, S=65536,
         ranges=[(0.0, 0.2, 512, 1)]),
    dict(C=1024, fch1=1500.0, foff=-300.0 / 1024, ts=64e-6, S=1 << 20,
         ranges=[(0.0, 0.2, 1000, 1), (200.0, 0.5, 600, 2), (500.0, 1.0, 500, 4)]),
    dict(C=1024, fch1=1500.0, foff=-300.0 / 1024, ts=64e-6, S=65536 + 16176,
         ranges=[(0.0, 1000.0 / 2048, 2048, 1)]),
])
@pytest.mark.parametrize("rule", [0, 1])
def test_plan_tables_and_layout_match_oracle(cfg, rule):
    p = M.Plan.brute(cfg["C"], cfg["fch1"], cfg["foff"], cfg["ts"], cfg["S"], "u8",
                     cfg["ranges"], shift_rule=rule)
    off = 0
    row = 0
    for r, (lo, step, nd, k) in enumerate(cfg["ranges"]):
        R = p.ranges[r]
        ts = np.float32(cfg["ts"]) * np.float32(k)
        if rule == 0:
            exp = oracle.shift_table_ref(cfg["fch1"], cfg["foff"], cfg["C"], lo, step, nd, ts)
        else:
            exp = oracle.shift_table_ns(np.float32(cfg["fch1"]), np.float32(cfg["foff"]), cfg["C"],
                                        np.float32(lo), np.float32(step), nd,
                                        float(np.float32(cfg["ts"])) * k)
        got = p.shift_table(r)
        assert np.array_equal(got, exp)
        assert R.maxshift == int(exp.max())
        assert R.nout == cfg["S"] // k - R.maxshift
        assert R.out_offset == off and R.first_dm == row
        off += nd * R.nout
        row += nd
    assert p.info.out_elems == off and p.info.ndms_total == row
    dmv = p.dm_values()
    lo, step = cfg["ranges"][0][0], cfg["ranges"][0][1]
    assert dmv[1] == np.float32(np.float32(lo) + np.float32(step))


def test_config_maxshifts_match_reference():
    """maxshift of each benchmark configuration (configs 1-3) under the reference rule."""
    p1 = M.Plan.brute(256, 1500.0, -0.390625, 64e-6, 65536, "f32", [(0.0, 0.2, 512, 1)])
    assert p1.ranges[0].maxshift == 433
    p2 = M.Plan.brute(1024, 1500.0, -300.0 / 1024, 64e-6, 1 << 20, "u8",
                      [(0.0, 0.2, 1000, 1), (200.0, 0.5, 600, 2), (500.0, 1.0, 500, 4)])
    assert [R.maxshift for R in p2.ranges] == [3233, 4041, 4041]
    assert [R.nout for R in p2.ranges] == [1045343, 520247, 258103]


def test_subband_tables_match_oracle():
    C, nsubs, ncalls, calldms = 256, 16, 8, 16
    dmstep = 0.5
    p = M.Plan.subband(C, 1500.0, -0.390625, 64e-6, 40000, nsubs,
                       [dict(lowdm=4.0, dmstep=dmstep, sub_dmstep=dmstep * calldms, ncalls=ncalls,
                             calldms=calldms, binsize=2)], dtype="f16")
    co, fi = oracle.subband_tables(1500.0, -0.390625, C, nsubs, 4.0, dmstep * calldms, dmstep,
                                   ncalls, calldms, np.float32(64e-6) * np.float32(2))
    assert np.array_equal(p.shift_table(0), co)
    assert np.array_equal(p.fine_table(0, nsubs), fi)
    cps = C // nsubs
    M1 = int((co - co[:, (np.arange(C) // cps) * cps]).max())
    M2 = int(fi.max())
    assert p.ranges[0].nout == 40000 // 2 - M1 - M2
    assert p.ranges[0].maxshift == M1 + M2


@pytest.mark.parametrize("bad", [
    dict(nchans=0), dict(tsamp=0.0), dict(foff=0.5), dict(nsamp_in=10), dict(dtype=7),
    dict(ranges=[(0.0, 1.0, 0, 1)]), dict(ranges=[(0.0, 1.0, 10, 0)]),
    dict(ranges=[(-1.0, 1.0, 10, 1)]), dict(ranges=[(0.0, 1.0, 10, 1000)]),
])
def test_plan_validation_errors(bad):
    kw = dict(nchans=64, fch1=1500.0, foff=-1.0, tsamp=64e-6, nsamp_in=4096, dtype="u8",
              ranges=[(0.0, 1.0, 10, 1)])
    kw.update(bad)
    with pytest.raises(M.MdsmError) as ei:
        M.Plan.brute(**kw)
    assert ei.value.code == -1
    assert M.lib().mdsm_last_error()


def test_subband_validation():
    with pytest.raises(M.MdsmError):
        M.Plan.subband(100, 1500.0, -1.0, 64e-6, 4096, 7,
                       [dict(dmstep=1.0, sub_dmstep=4.0, ncalls=2, calldms=4)])


@pytest.mark.parametrize("bad", [dict(chunk_rows=1), dict(chunk_rows=9), dict(lds_budget=200000),
                                 dict(lds_budget=100), dict(persistent=2), dict(grid=1 << 21),
                                 dict(tile_group=5000)])
def test_tuning_validation(bad):
    p = M.Plan.brute(64, 1500.0, -1.0, 64e-6, 4096, "u8", [(0.0, 1.0, 10, 1)])
    before = p.tuning()
    t = dict(before)
    t.update(bad)
    with pytest.raises(M.MdsmError) as ei:
        p.set_tuning(**t)
    assert ei.value.code == -1
    assert p.tuning() == before


def test_tuning_round_trip_keeps_plan_info():
    p = M.Plan.brute(256, 1500.0, -0.390625, 64e-6, 20000, "u8",
                     [(0.0, 0.5, 100, 1), (50.0, 1.0, 40, 2)])
    info = (p.info.out_elems, p.info.ndms_total, p.info.maxshift, p.info.nblocks)
    for t in (dict(chunk_rows=2), dict(chunk_rows=6, lds_budget=40000, persistent=0),
              dict(chunk_rows=8, lds_budget=162816, tile_group=128, grid=1000)):
        full = dict(chunk_rows=0, lds_budget=0, persistent=1, grid=0, tile_group=64)
        full.update(t)
        p.set_tuning(**full)
        assert p.tuning() == full
        assert (p.info.out_elems, p.info.ndms_total, p.info.maxshift, p.info.nblocks) == info


def test_synth_is_seeded_and_thread_independent():
    a = M.synth_filterbank(37, 1001, "f32", seed=9, nthreads=1)
    b = M.synth_filterbank(37, 1001, "f32", seed=9, nthreads=5)
    c = M.synth_filterbank(37, 1001, "f32", seed=10, nthreads=5)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    z = M.synth_filterbank(64, 200000, "f32", seed=4)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    u = M.synth_filterbank(16, 50000, "u8", seed=4, mean=128, sigma=16)
    assert abs(u.mean() - 128) < 0.2 and abs(u.std() - 16) < 0.2
    h = M.synth_filterbank(8, 1000, "f16", seed=4)
    assert np.array_equal(h.astype(np.float32),
                          M.synth_filterbank(8, 1000, "f32", seed=4).astype(np.float16)
                          .astype(np.float32))


def test_synth_pulse_on_ns_track():
    C, S = 32, 4000
    base = M.synth_filterbank(C, S, "f32", seed=1, sigma=0.0)
    x = M.synth_filterbank(C, S, "f32", seed=1, sigma=1.0, mean=0.0,
                           pulses=[(300.0, 1000, 3, 5.0)])
    z = M.synth_filterbank(C, S, "f32", seed=1, sigma=1.0, mean=0.0)
    d = x - z
    ns = oracle.shift_table_ns(1500.0, -0.390625, C, 300.0, 0.0, 1, float(np.float32(64e-6)))[0]
    for c in range(C):
        on = np.nonzero(np.abs(d[c]) > 1e-3)[0]
        assert on.tolist() == list(range(1000 + ns[c], 1003 + ns[c]))
    assert np.all(base == 0)


def test_library_reports_no_device_or_a_device():
    # loading never fails; device_count is 0 without a GPU and >= 1 with one
    assert M.device_count() >= 0
    assert "gfx950" in M.version()


def test_plan_lds_bytes_from_shift_table():
    """mdsm_plan_lds_bytes against a count made here from the plan's own
    shift table: blocks of up to 64 trials, each read by 64 trial slots of
    8 bytes per lane (256 samples, 16-bit entries) per row and tile, and each
    row staged once per tile at 8 bytes per window sample, the row's window
    being the 256-sample tile plus the row's own shift spread over the
    block's trials above its 16-aligned base, rounded up to 64 (uint8,
    windows below 1024)."""
    C, S = 256, 40000
    ranges = [(0.0, 0.5, 150, 1), (75.0, 1.0, 70, 2)]
    p = M.Plan.brute(C, 1500.0, -300.0 / C, 64e-6, S, "u8", ranges)
    reads = writes = 0
    for r in range(len(ranges)):
        sh = p.shift_table(r)  # [ndms][C]
        ntiles = -(-p.ranges[r].nout // 256)
        for j0 in range(0, sh.shape[0], 64):
            blk = sh[j0:j0 + 64]
            base = blk.min(axis=0) & ~15
            Wrow = -(-(256 + (blk.max(axis=0) - base)) // 64) * 64
            assert Wrow.max() <= 1024 and Wrow.min() >= 256
            assert Wrow.min() < Wrow.max(), "rows of a block should differ in width here"
            reads += C * ntiles * 64 * 256 * 2
            writes += ntiles * 8 * int(Wrow.sum())
    assert p.lds_bytes(0) == (reads, writes)


def test_plan_lds_bytes_float16_two_planes():
    """The same count for float16 rows, which are staged in two planes of
    16-bit pairs: 4 bytes of LDS per window sample (half the four-phase
    layout of integer rows), 8 bytes read per lane and trial slot; a float
    block of at most 32 trials runs 4 trials per wave (32 slots)."""
    C, S = 256, 40000
    p = M.Plan.brute(C, 1500.0, -300.0 / C, 64e-6, S, "f16", [(0.0, 0.5, 150, 1)])
    sh = p.shift_table(0)
    ntiles = -(-p.ranges[0].nout // 256)
    reads = writes = 0
    for j0 in range(0, sh.shape[0], 64):
        blk = sh[j0:j0 + 64]
        base = blk.min(axis=0) & ~15
        Wrow = -(-(256 + (blk.max(axis=0) - base)) // 64) * 64
        assert Wrow.max() <= 1024
        slots = 32 if blk.shape[0] <= 32 else 64
        reads += C * ntiles * slots * 256 * 2
        writes += ntiles * 4 * int(Wrow.sum())
    assert sh.shape[0] % 64 <= 32, "the last block should run 4 trials per wave here"
    assert p.lds_bytes(0) == (reads, writes)


def test_lds_probes_reject_bad_arguments():
    """The LDS probes validate their arguments before touching a device."""
    for f in (lambda: M.probe_lds_read(3, 4), lambda: M.probe_lds_write(32, 4),
              lambda: M.probe_lds_read(8, 0), lambda: M.probe_lds_read(8, 9),
              lambda: M.probe_lds_read_form(0, 4), lambda: M.probe_lds_read_form(5, 4),
              lambda: M.probe_lds_read_form(1, 9)):
        with pytest.raises(M.MdsmError):
            f()

