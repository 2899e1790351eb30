#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ from the reference's own code.

Runs in the build container, where oracle/_ref/libmdsm_ref.so has been built
from the reference sources (make -C oracle ref).  The fixtures are data only
(inputs and the reference's outputs):

  ref_shift_digests.json   shift tables computed by the reference's
                           cpu_dedisp.c main() for every DM range of the five
                           benchmark configurations: maxshift, checksum,
                           SHA-256 of the int32 table, sampled entries.
  ref_cpu_dedisp_small.npz a seeded time-major input, the reference's shift
                           table and its dedispersed time-major output.
  ref_median5.npz          medianOfFive (src/mdsm/dedispersion_output.cpp:14-53)
                           on every tuple of {0..4}^5 and on random tuples.
  ref_process_brute.json   the candidate lines of the reference's process_brute
                           (dedispersion_output.cpp:135-213, one thread, no noise
                           arrays) on seeded [tdms][nsamp] blocks: fewer values
                           than one chunk, whole chunks, a partial last chunk.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

# (name, fch1, foff, tsamp, nchans, startdm, dmstep, ndms)
SHIFT_CASES = [
    ("cfg1", 1500.0, -0.390625, 64e-6, 256, 0.0, 0.2, 512),
    ("cfg2_r0_k1", 1500.0, -300.0 / 1024, 64e-6, 1024, 0.0, 0.2, 1000),
    ("cfg2_r1_k2", 1500.0, -300.0 / 1024, 128e-6, 1024, 200.0, 0.5, 600),
    ("cfg2_r2_k4", 1500.0, -300.0 / 1024, 256e-6, 1024, 500.0, 1.0, 500),
    ("cfg3_brute_equiv", 1500.0, -400.0 / 4096, 54.6e-6, 4096, 0.0, 1500.0 / 4096, 4096),
    ("cfg5", 1500.0, -300.0 / 1024, 64e-6, 1024, 0.0, 1000.0 / 2048, 2048),
]


# (seed, tdms, nsamp, threshold): 3000 values (< one chunk of 32768), 65536
# (two whole chunks: the second is never read), 200000 (6 of 7 chunks read)
PROCESS_BRUTE_CASES = [(31, 3, 1000, 3.0), (32, 16, 4096, 2.5), (33, 40, 5000, 3.0)]


def process_brute_input(seed, tdms, nsamp):
    rng = np.random.default_rng(seed)
    x = (100.0 + 10.0 * rng.standard_normal((tdms, nsamp))).astype(np.float32)
    x[tdms // 2, nsamp // 3:nsamp // 3 + 12] += 150.0
    return x


def digest(tab):
    tab = np.ascontiguousarray(tab, np.int32)
    idx = np.random.default_rng(0).integers(0, tab.size, 64)
    return {"sha256": hashlib.sha256(tab.tobytes()).hexdigest(), "sum": int(tab.sum(dtype=np.int64)),
            "maxshift": int(tab[-1, -1]), "max": int(tab.max()), "shape": list(tab.shape),
            "sample_idx": idx.tolist(), "sample_val": tab.reshape(-1)[idx].tolist()}


def main():
    R = oracle.ref()
    if R is None:
        sys.exit("oracle/_ref not built: run make -C oracle ref")
    out = {}
    for name, fch1, foff, tsamp, C, lo, step, nd in SHIFT_CASES:
        sh, _, _ = oracle.ref_cpu_dedisp(fch1, foff, tsamp, C, 1, lo, step, nd)
        d = digest(sh)
        d.update({"fch1": fch1, "foff": foff, "tsamp": tsamp, "nchans": C, "startdm": lo,
                  "dmstep": step, "ndms": nd})
        out[name] = d
        print(name, d["maxshift"], d["sha256"][:16])
    with open(os.path.join(HERE, "ref_shift_digests.json"), "w") as f:
        json.dump(out, f, indent=1)

    # small end-to-end case through the reference loop
    C, N, nd = 64, 600, 24
    fch1, foff, tsamp, lo, step = 1500.0, -1.5625, 64e-6, 0.0, 2.5
    sh, _, _ = oracle.ref_cpu_dedisp(fch1, foff, tsamp, C, N, lo, step, nd)
    M = int(sh[-1, -1])
    rng = np.random.default_rng(1234)
    x_tm = rng.standard_normal((N + M, C)).astype(np.float32)
    sh2, out_tm, inp = oracle.ref_cpu_dedisp(fch1, foff, tsamp, C, N, lo, step, nd, x_tm)
    assert np.array_equal(sh, sh2) and np.array_equal(inp, x_tm)
    np.savez_compressed(os.path.join(HERE, "ref_cpu_dedisp_small.npz"), x_tm=x_tm, shift=sh2,
                        out_tm=out_tm, params=np.array([fch1, foff, tsamp, lo, step]),
                        dims=np.array([C, N, nd]))

    # medianOfFive
    grid = np.stack(np.meshgrid(*([np.arange(5, dtype=np.float32)] * 5), indexing="ij"),
                    -1).reshape(-1, 5)
    rnd = np.random.default_rng(99).standard_normal((2000, 5)).astype(np.float32)
    tup = np.concatenate([grid, rnd])
    med = np.array([R.median_of_five(*map(float, t)) for t in tup], np.float32)
    np.savez_compressed(os.path.join(HERE, "ref_median5.npz"), tuples=tup, median=med)

    # process_brute: inputs regenerated from the seed, 100 + 10 N(0,1) float32
    import tempfile
    cases = []
    for seed, tdms, nsamp, thr in PROCESS_BRUTE_CASES:
        x = process_brute_input(seed, tdms, nsamp)
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "cands.dat")
            oracle.ref_process_brute(x, nsamp, 1.5, 0.25, thr, 1234.5, 64e-6, path)
            with open(path) as f:
                lines = f.read().strip().splitlines()
        cases.append({"seed": seed, "tdms": tdms, "nsamp": nsamp, "thr": thr, "lowdm": 1.5,
                      "dmstep": 0.25, "timestamp": 1234.5, "block_rate": 64e-6, "lines": lines})
        print("process_brute", tdms, nsamp, len(lines), "lines")
    with open(os.path.join(HERE, "ref_process_brute.json"), "w") as f:
        json.dump(cases, f, indent=0)
    print("fixtures written")


if __name__ == "__main__":
    main()
