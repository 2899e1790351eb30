"""GPU: the LDS probes behind the roofs the dedispersion kernel is measured
against, and the ordering of the read forms its row layouts were chosen by
(an aligned ds_read_b64 ahead of ds_read2st64_b32, which the float16 rows
use, ahead of either way to read 8 bytes at a 4-byte-aligned address)."""
import math

import pytest

pytestmark = pytest.mark.gpu


def test_lds_read_forms_order(gpu):
    M = gpu
    print("case lds_read_forms", flush=True)
    best = lambda f: max(f() for _ in range(2))
    b64 = best(lambda: M.probe_lds_read(8, 4))
    st32 = best(lambda: M.probe_lds_read_form(1, 4))
    st64 = best(lambda: M.probe_lds_read_form(2, 4))
    mis = best(lambda: M.probe_lds_read_form(3, 4))
    adj = best(lambda: M.probe_lds_read_form(4, 4))
    wr = best(lambda: M.probe_lds_write(8, 4))
    print(f"b64 {b64:.1f} read2st64_b32 {st32:.1f} read2st64_b64 {st64:.1f} "
          f"misaligned b64 {mis:.1f} adjacent read2_b32 {adj:.1f} write b64 {wr:.1f} TB/s")
    for v in (b64, st32, st64, mis, adj, wr):
        assert math.isfinite(v) and 0.5 < v < 400.0
    assert b64 > st32 > adj > mis
    assert b64 > wr
