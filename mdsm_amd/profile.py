"""Profiling support for the dedispersion path (rocprofv3 under the hood).

Programs to run under the profiler (the program itself follows ``--``):

    rocprofv3 --kernel-trace --stats -d OUT/kt -o run -- python3 -m mdsm_amd.profile steps --workload cfg2 --steps 3
    rocprofv3 --pmc FETCH_SIZE       -d OUT/fetch -o run -- python3 -m mdsm_amd.profile steps --workload cfg2 --steps 1
    rocprofv3 --pmc WRITE_SIZE       -d OUT/write -o run -- python3 -m mdsm_amd.profile steps --workload cfg2 --steps 1
    rocprofv3 --pmc FETCH_SIZE       -d OUT/cfetch -o run -- python3 -m mdsm_amd.profile copy
    rocprofv3 --pmc WRITE_SIZE       -d OUT/cwrite -o run -- python3 -m mdsm_amd.profile copy
    rocprofv3 --pmc SQ_INSTS_LDS SQ_INSTS_LDS_LOAD_BANDWIDTH SQ_LDS_BANK_CONFLICT SQ_LDS_IDX_ACTIVE \
        SQ_WAVES SQ_BUSY_CYCLES SQ_WAVE_CYCLES GRBM_GUI_ACTIVE -d OUT/lds -o run -- \
        python3 -m mdsm_amd.profile steps --workload cfg2 --steps 1

and the summary of their SQLite results:

    python3 -m mdsm_amd.profile summarize --dir OUT --workload cfg2 --kernel dedisperse_kernel > summary.md

FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are scaled by what the same
counter reads for a plain copy kernel of known size: the 1 GiB float32
copy of ``copy`` (mdsm_decimate with k = 1 reads and writes 1 GiB each).
"""
from __future__ import annotations

import argparse
import collections
import glob
import os
import shlex
import sqlite3
import sys

COPY_C, COPY_S = 1024, 1 << 18          # float32 [1024][262144] = 1 GiB


def _steps(workload: str, steps: int, tile_group=None) -> None:
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import mdsm_amd as M
    import bench
    w = bench.WORKLOADS[workload]
    x, _ = bench.make_input(M, w)
    plan = bench.make_plan(M, w)
    if tile_group is not None:
        t = plan.tuning()
        t["tile_group"] = tile_group
        plan.set_tuning(**t)
    din = M.DeviceBuffer.from_host(x)
    dout = M.DeviceBuffer(plan.info.out_elems * 4)
    dw = M.DeviceBuffer(max(plan.info.work_bytes, 256))
    for i in range(steps):
        plan.dedisperse(din, w["nsamp"], dout, dw)
        M.synchronize()
        print(f"step {i + 1}/{steps} done", flush=True)


def _serve(steps: int, warmup: int) -> None:
    """The timed phase of the served config-5 bench (32 beams, HBM-resident
    blocks pushed back to back): warmup + steps blocks per beam."""
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import mdsm_amd as M
    import bench
    ns = argparse.Namespace(steps=steps, warmup=warmup, throughput_only=True)
    r = bench.serve(M, bench.WORKLOADS["cfg5"], ns)
    print(f"served blocks {r['blocks']} in the profiled run", flush=True)


def _lds_probe() -> None:
    """LDS read and write bandwidth of ds_read/ds_write _b32/_b64/_b128 at 1,
    2, 4 and 8 waves per SIMD, then the other read forms at 4 and 8."""
    import mdsm_amd as M
    print("| bytes per lane | waves per SIMD | read TB/s | write TB/s |")
    print("|---|---|---|---|")
    for w in (4, 8, 16):
        for occ in (1, 2, 4, 8):
            r = max(M.probe_lds_read(w, occ) for _ in range(3))
            wr = max(M.probe_lds_write(w, occ) for _ in range(3))
            print(f"| {w} | {occ} | {r:.1f} | {wr:.1f} |", flush=True)
    print()
    print("| read form | bytes per lane | waves per SIMD | read TB/s |")
    print("|---|---|---|---|")
    forms = {1: ("ds_read2st64_b32, two dwords 256 B apart", 8),
             2: ("ds_read2st64_b64, two 8-byte entries 512 B apart", 16),
             3: ("ds_read_b64 at a 4-byte-aligned address", 8),
             4: ("ds_read2_b32, adjacent dwords at a 4-byte-aligned address", 8)}
    for f, (name, b) in forms.items():
        for occ in (4, 8):
            r = max(M.probe_lds_read_form(f, occ) for _ in range(3))
            print(f"| {name} | {b} | {occ} | {r:.1f} |", flush=True)


def _copy() -> None:
    import mdsm_amd as M
    a = M.DeviceBuffer(COPY_C * COPY_S * 4)
    b = M.DeviceBuffer(COPY_C * COPY_S * 4)
    a.fill_u32(0)
    M.synchronize()
    for _ in range(3):
        M.check(M.lib().mdsm_decimate(a.ptr, 2, COPY_C, COPY_S, COPY_S, 1, b.ptr, COPY_S, None))
    M.synchronize()
    print("copy done", COPY_C * COPY_S * 4, "bytes each way", flush=True)


def _db(d: str) -> str:
    dbs = sorted(glob.glob(os.path.join(d, "**", "*_results.db"), recursive=True))
    if not dbs:
        raise SystemExit(f"no *_results.db under {d}")
    return dbs[-1]


def kernel_stats(db: str) -> list:
    con = sqlite3.connect(db)
    return con.execute("select name, total_calls, total_duration, average, percentage "
                       "from top_kernels order by total_duration desc").fetchall()


def kernel_dispatches(db: str, pattern: str) -> list:
    con = sqlite3.connect(db)
    return con.execute("select duration, grid_x, workgroup_x, lds_size, vgpr_count, sgpr_count "
                       "from kernels where name like ?", (f"%{pattern}%",)).fetchall()


def counter_per_dispatch(db: str, pattern: str, counter: str) -> list:
    """Sum of `counter` over its instances, per dispatch of kernels matching `pattern`."""
    con = sqlite3.connect(db)
    rows = con.execute("select dispatch_id, value from counters_collection "
                       "where kernel_name like ? and counter_name = ?",
                       (f"%{pattern}%", counter)).fetchall()
    per = collections.defaultdict(float)
    for did, v in rows:
        per[did] += v
    return [per[k] for k in sorted(per)]


def summarize(args) -> None:
    kt, fe, wr = (_db(os.path.join(args.dir, s)) for s in ("kt", "fetch", "write"))
    cf, cw = (_db(os.path.join(args.dir, s)) for s in ("cfetch", "cwrite"))
    copy_bytes = COPY_C * COPY_S * 4
    cfk = counter_per_dispatch(cf, "decimate", "FETCH_SIZE")
    cwk = counter_per_dispatch(cw, "decimate", "WRITE_SIZE")
    f_scale = copy_bytes / (sum(cfk) / len(cfk) * 1024.0)
    w_scale = copy_bytes / (sum(cwk) / len(cwk) * 1024.0)
    fk = counter_per_dispatch(fe, args.kernel, "FETCH_SIZE")
    wk = counter_per_dispatch(wr, args.kernel, "WRITE_SIZE")
    fetch_b = sum(fk) / len(fk) * 1024.0 * f_scale
    write_b = sum(wk) / len(wk) * 1024.0 * w_scale
    d = kernel_dispatches(kt, args.kernel)
    print(f"# rocprofv3 summary: {args.kernel}, workload {args.workload}\n")
    print("Commands (repository root, one GPU):\n")
    print("```")
    if args.workload == "cfg5":
        prog = "python3 -m mdsm_amd.profile serve"
        kt_args, pmc_args = "--steps 4 --warmup 1", "--steps 1 --warmup 1"
    else:
        prog = f"python3 -m mdsm_amd.profile steps --workload {args.workload}"
        kt_args, pmc_args = "--steps 3", "--steps 1"
    print(f"rocprofv3 --kernel-trace --stats -d {args.dir}/kt -o run -- {prog} {kt_args}")
    print(f"rocprofv3 --pmc FETCH_SIZE -d {args.dir}/fetch -o run -- {prog} {pmc_args}")
    print(f"rocprofv3 --pmc WRITE_SIZE -d {args.dir}/write -o run -- {prog} {pmc_args}")
    print(f"rocprofv3 --pmc FETCH_SIZE -d {args.dir}/cfetch -o run -- python3 -m mdsm_amd.profile copy")
    print(f"rocprofv3 --pmc WRITE_SIZE -d {args.dir}/cwrite -o run -- python3 -m mdsm_amd.profile copy")
    print(f"python3 -m mdsm_amd.profile summarize --dir {args.dir} --workload {args.workload} "
          f"--kernel {shlex.quote(args.kernel)}" + (f" --blocks {args.blocks}" if args.blocks else "")
          + (f" --kt-blocks {args.kt_blocks}" if args.kt_blocks else ""))
    print("```\n")
    print("## Kernel statistics (--kernel-trace --stats)\n")
    print("| kernel | calls | total ms | average ms | % |")
    print("|---|---|---|---|---|")
    for name, calls, tot, avg, pct in kernel_stats(kt):
        print(f"| `{name[:70]}` | {calls} | {tot / 1e3:.3f} | {avg / 1e3:.3f} | {pct:.1f} |")
    if d:
        durs = [r[0] / 1e6 for r in d]
        print(f"\n`{args.kernel}` dispatches: {len(d)}, duration ms min {min(durs):.3f} "
              f"avg {sum(durs) / len(durs):.3f} max {max(durs):.3f}; grid {d[0][1]} threads, "
              f"workgroup {d[0][2]}, LDS {d[0][3]} B, VGPR {d[0][4]}, SGPR {d[0][5]}")
    print("\n## HBM traffic per dispatch (--pmc, one counter group per run)\n")
    print(f"* copy calibration: 1 GiB float32 copy reads FETCH_SIZE "
          f"{sum(cfk) / len(cfk):.0f} KiB and WRITE_SIZE {sum(cwk) / len(cwk):.0f} KiB per dispatch; "
          f"scale factors fetch x{f_scale:.3f}, write x{w_scale:.3f}")
    print(f"* `{args.kernel}` FETCH_SIZE {sum(fk) / len(fk):.0f} KiB -> {fetch_b / 1e9:.3f} GB read")
    print(f"* `{args.kernel}` WRITE_SIZE {sum(wk) / len(wk):.0f} KiB -> {write_b / 1e9:.3f} GB written")
    ld = os.path.join(args.dir, "lds")
    if os.path.isdir(ld):
        ldb = _db(ld)
        names = ("SQ_INSTS_LDS", "SQ_INSTS_LDS_LOAD_BANDWIDTH", "SQ_LDS_BANK_CONFLICT",
                 "SQ_LDS_IDX_ACTIVE", "SQ_WAVES", "SQ_BUSY_CYCLES", "SQ_WAVE_CYCLES",
                 "GRBM_GUI_ACTIVE")
        vals = {}
        for n in names:
            v = counter_per_dispatch(ldb, args.kernel, n)
            if v:
                vals[n] = sum(v) / len(v)
        print("\n## LDS and occupancy counters per dispatch (--pmc, one run)\n")
        print(f"`rocprofv3 --pmc {' '.join(names)} -d {args.dir}/lds -o run -- {prog} {pmc_args}`\n")
        for n in names:
            if n in vals:
                print(f"* {n} = {vals[n]:.4e}")
        if "SQ_INSTS_LDS_LOAD_BANDWIDTH" in vals and d:
            lb = vals["SQ_INSTS_LDS_LOAD_BANDWIDTH"] * 64.0
            avg_s = sum(r[0] for r in d) / len(d) / 1e9
            if args.blocks:   # dispatches of the served run carry different numbers of blocks
                ltot = sum(counter_per_dispatch(ldb, args.kernel, "SQ_INSTS_LDS_LOAD_BANDWIDTH"))
                kt_tot = sum(r[0] for r in d) / 1e9
                kt_blocks = args.kt_blocks or args.blocks
                print(f"* LDS load bytes per block = {ltot * 64.0 / args.blocks:.4e}; at the "
                      f"traced kernel time per block ({kt_tot / kt_blocks * 1e3:.3f} ms) that is "
                      f"{ltot * 64.0 / args.blocks / (kt_tot / kt_blocks) / 1e12:.1f} TB/s")
            else:
                print(f"* LDS load bytes (64-byte units x 64) = {lb:.4e}; at the traced average "
                      f"duration that is {lb / avg_s / 1e12:.1f} TB/s")
        if "SQ_LDS_BANK_CONFLICT" in vals and vals.get("SQ_LDS_IDX_ACTIVE"):
            print(f"* bank-conflict cycles / LDS-active cycles = "
                  f"{vals['SQ_LDS_BANK_CONFLICT'] / vals['SQ_LDS_IDX_ACTIVE']:.4f}")
        if vals.get("GRBM_GUI_ACTIVE") and "SQ_WAVE_CYCLES" in vals:
            # GRBM_GUI_ACTIVE is summed over the 8 XCDs; SQ_WAVE_CYCLES counts
            # resident-wave time in units of 4 cycles over all 256 CUs
            cyc = vals["GRBM_GUI_ACTIVE"] / 8.0
            w = vals["SQ_WAVE_CYCLES"] * 4.0 / cyc / 256.0
            print(f"* mean resident waves per CU = SQ_WAVE_CYCLES x 4 / (GRBM_GUI_ACTIVE / 8) / 256 "
                  f"= {w:.1f} ({w / 4:.2f} per SIMD)")
    if args.blocks:
        # the served run's dispatches carry different numbers of blocks: per block
        fb = sum(fk) * 1024.0 * f_scale / args.blocks
        wb = sum(wk) * 1024.0 * w_scale / args.blocks
        print(f"* counted: the dispatches of `{args.kernel}` only (the served launches; the "
              f"autotuner's trial launches before them take the instance without row statistics)")
        print(f"* over the {args.blocks} blocks of the counter runs: {fb / 1e9:.4f} GB read and "
              f"{wb / 1e9:.4f} GB written per block")
        if d:
            kt_blocks = args.kt_blocks or args.blocks
            print(f"* traced `{args.kernel}` time per block: "
                  f"{sum(r[0] for r in d) / 1e6 / kt_blocks:.3f} ms ({kt_blocks} blocks)")
        print(f"\ntraffic_bytes_per_block = {fb + wb:.6e}")
        print(f"fetch_bytes_per_block = {fb:.6e}")
        print(f"write_bytes_per_block = {wb:.6e}")
        return
    print(f"\ntraffic_bytes_per_launch = {fetch_b + write_b:.6e}")
    print(f"fetch_bytes_per_launch = {fetch_b:.6e}")
    print(f"write_bytes_per_launch = {write_b:.6e}")


def read_traffic(path: str) -> dict | None:
    """Parse the `*_bytes_per_launch = v` lines of a summary written by `summarize`."""
    try:
        with open(path) as f:
            out = {}
            for line in f:
                if "_bytes_per_launch = " in line or "_bytes_per_block = " in line:
                    k, v = line.split(" = ")
                    out[k.strip()] = float(v)
            return out or None
    except OSError:
        return None


def kernel_table(args) -> None:
    """Per-kernel statistics of one --kernel-trace --stats run (no counters)."""
    rows = kernel_stats(_db(args.dir))
    print(f"# rocprofv3 kernel statistics: {args.title}\n")
    print("Command (repository root, one GPU):\n")
    print("```")
    print(args.command)
    print("```\n")
    print("| kernel | calls | total ms | average ms | % |")
    print("|---|---|---|---|---|")
    for name, calls, total, avg, pct in rows:
        print(f"| `{name[:70]}` | {calls} | {total / 1e3:.3f} | {avg / 1e3:.4f} | {pct:.1f} |")  # us -> ms


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("steps")
    s.add_argument("--workload", default="cfg2")
    s.add_argument("--steps", type=int, default=3)
    s.add_argument("--tile-group", type=int, default=None,
                   help="override the plan's tile group (items sharing one tile window)")
    sub.add_parser("copy")
    z = sub.add_parser("summarize")
    z.add_argument("--dir", required=True)
    z.add_argument("--workload", default="cfg2")
    z.add_argument("--kernel", default="dedisperse_kernel")
    z.add_argument("--blocks", type=int, default=0,
                   help="served run: blocks of each counter run (traffic per block)")
    z.add_argument("--kt-blocks", type=int, default=0,
                   help="served run: blocks of the kernel-trace run")
    sv = sub.add_parser("serve")
    sv.add_argument("--steps", type=int, default=4)
    sv.add_argument("--warmup", type=int, default=1)
    sub.add_parser("lds-probe")
    kt = sub.add_parser("kernels")
    kt.add_argument("--dir", required=True)
    kt.add_argument("--title", default="")
    kt.add_argument("--command", default="")
    args = ap.parse_args()
    if args.cmd == "steps":
        _steps(args.workload, args.steps, args.tile_group)
    elif args.cmd == "copy":
        _copy()
    elif args.cmd == "serve":
        _serve(args.steps, args.warmup)
    elif args.cmd == "lds-probe":
        _lds_probe()
    elif args.cmd == "kernels":
        kernel_table(args)
    else:
        summarize(args)


if __name__ == "__main__":
    main()
