"""mdsm_dedisperse_noalloc (workspace and counters in the library's static
per-GPU array) against mdsm_dedisperse on the same plans and inputs: two
calls in a row, the second on a smaller plan and input, then a subband plan
whose stage-1 series live in the array; outputs NaN-poisoned before each
call.  uint8 input: bit-exact; float16: identical (the same kernels run in
the same order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _both(M, plan, x, stride):
    din = M.DeviceBuffer.from_host(x)
    outs = []
    for mode in ("alloc", "noalloc"):
        dout = M.DeviceBuffer(plan.info.out_elems * 4)
        dout.fill_u32(0xFFFFFFFF)
        if mode == "alloc":
            plan.dedisperse(din, stride, dout)
        else:
            plan.dedisperse_noalloc(din, stride, dout)
        M.synchronize()
        outs.append(dout.download(np.float32, (plan.info.out_elems,)))
    return outs


def test_noalloc_matches_dedisperse_two_calls(gpu):
    M = gpu
    print("case noalloc: large then small uint8 plan", flush=True)
    assert M.lib().mdsm_noalloc_max_work() == (256 << 20) - 256
    cases = [
        (256, 1 << 17, [(0.0, 0.5, 128, 1), (64.0, 1.0, 96, 2), (160.0, 2.0, 64, 4)], 11),
        (64, 30000, [(10.0, 1.0, 40, 2)], 12),
    ]
    for C, S, ranges, seed in cases:
        plan = M.Plan.brute(C, 1500.0, -300.0 / C, 64e-6, S, "u8", ranges)
        assert 0 < plan.info.work_bytes <= M.lib().mdsm_noalloc_max_work()
        x = M.synth_filterbank(C, S, "u8", seed=seed, mean=128, sigma=16, fch1=1500.0,
                               foff=-300.0 / C, pulses=[(50.0, S // 3, 4, 3.0)])
        plan.upload()
        a, b = _both(M, plan, x, S)
        assert not np.isnan(a).any()
        assert np.array_equal(a, b), f"C={C} S={S}: noalloc output differs"


def test_noalloc_subband_matches(gpu):
    M = gpu
    print("case noalloc: float16 subband plan", flush=True)
    C, S = 256, 1 << 16
    passes = [dict(lowdm=16.0, dmstep=1.0, sub_dmstep=32.0, ncalls=8, calldms=32)]
    plan = M.Plan.subband(C, 1500.0, -300.0 / C, 64e-6, S, 16, passes, dtype="f16")
    x = M.synth_filterbank(C, S, "f16", seed=13, fch1=1500.0, foff=-300.0 / C,
                           pulses=[(100.0, 20000, 8, 1.0)])
    plan.upload()
    a, b = _both(M, plan, x, S)
    assert not np.isnan(a).any() and np.array_equal(a, b)


def test_noalloc_refuses_oversize_and_missing_upload(gpu):
    M = gpu
    print("case noalloc: refusals", flush=True)
    big = M.Plan.brute(4096, 1500.0, -300.0 / 4096, 64e-6, 1 << 19, "u8", [(0.0, 1.0, 64, 2)])
    assert big.info.work_bytes > M.lib().mdsm_noalloc_max_work()
    d = M.DeviceBuffer(256)
    with pytest.raises(M.MdsmError) as ei:
        big.dedisperse_noalloc(d, 1 << 19, d)
    assert ei.value.code == -1 and "workspace" in str(ei.value)
    small = M.Plan.brute(64, 1500.0, -300.0 / 64, 64e-6, 20000, "u8", [(0.0, 1.0, 16, 1)])
    with pytest.raises(M.MdsmError) as ei:
        small.dedisperse_noalloc(d, 20000, d)
    assert ei.value.code == -6
