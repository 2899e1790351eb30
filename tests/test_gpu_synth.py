"""The device generator (mdsm_synth_filterbank_device) against the host one:
the same seeded noise, gains and pulses, equal to the rounding of the
device's log and cos."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", ["u8", "f16", "f32"])
def test_device_synth_matches_host(gpu, dtype):
    M = gpu
    print(f"case device synth {dtype}", flush=True)
    C, S, stride = 96, 20011, 20032
    kw = dict(seed=123, fch1=1500.0, foff=-300.0 / C, tsamp=64e-6,
              pulses=[(80.0, 3000, 8, 3.0), (400.0, 9000, 2, 5.0), (10.0, 19990, 40, 2.0)])
    kw.update(dict(mean=128.0, sigma=16.0) if dtype == "u8" else dict(gain=(0.5, 2.0)))
    host = M.synth_filterbank(C, S, dtype, stride=stride, **kw)
    d = M.DeviceBuffer(C * stride * host.itemsize)
    M.synth_filterbank_device(d, C, S, dtype, stride=stride, **kw)
    M.synchronize()
    dev = d.download(host.dtype, (C, stride))[:, :S]
    h = host[:, :S]
    if dtype == "u8":
        diff = np.abs(dev.astype(np.int32) - h.astype(np.int32))
        assert diff.max() <= 1 and np.count_nonzero(diff) <= h.size * 1e-5
    else:
        a, b = dev.astype(np.float64), h.astype(np.float64)
        tol = 1e-6 if dtype == "f32" else 1e-3
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)) <= tol
    # the pulses sit on the same samples: the brightest column of the top channel
    assert int(np.argmax(dev[0].astype(np.float64))) == int(np.argmax(h[0].astype(np.float64)))
