"""Summaries of rocprofv3 output for the lattice kernels.

  python -m dmftpack_amd.rocprof_summary trace <results.db> "<command>"
      per-kernel table (calls, total, average) and kernel resources from a
      --kernel-trace --stats SQLite result.

  python -m dmftpack_amd.rocprof_summary pmc --fetch <dir> --write <dir> (--copy-mib M | --factors F W)
          --method NAME [--json profiles/pmc_traffic.json] [--command "..."]
      HBM bytes per launch from two --pmc runs (FETCH_SIZE, then WRITE_SIZE;
      SQLite or csv output), calibrated by a plain copy of known size:
      the copy kernel (__amd_rocclr_copyBuffer, its largest dispatch) moves M MiB
      each way, so factor = M MiB / (counter KiB * 1024) for each counter, and a
      kernel's traffic per launch is the mean over its dispatches of
      FETCH * factor_fetch + WRITE * factor_write.

  python -m dmftpack_amd.rocprof_summary sq --run <dir> --method NAME [--json profiles/pmc_flops.json]
      executed FP64/FP32 VALU FLOPs per launch (SQ_INSTS_VALU_FLOPS_* x 64 lanes)
      and the instruction mix, from one --pmc pass of SQ counters.

  python -m dmftpack_amd.rocprof_summary mfma --run <dir> [--command "..."]
      FP64 MFMA and VALU FLOPs per launch and MFMA busy cycles, from one --pmc
      pass of MFMA_COUNTERS.
"""
import argparse
import csv
import glob
import json
import os
import re
import sqlite3
import sys
from collections import defaultdict


# ---------------------------------------------------------------------------
def trace(db_path, command):
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, total_calls, total_duration, average, percentage from top_kernels").fetchall()
    print(f"# rocprofv3 kernel-trace summary\n# command: {command}\n# durations in microseconds")
    print(f"{'calls':>6} {'total_us':>12} {'avg_us':>12} {'pct':>6}  kernel")
    for name, calls, tot, avg, pct in rows:
        print(f"{calls:>6} {tot:>12.1f} {avg:>12.1f} {pct:>6.2f}  {name[:150]}")
    res = db.execute("select name, vgpr_count, accum_vgpr_count, scratch_size, lds_size, grid_x, grid_y, workgroup_x "
                     "from kernels group by name").fetchall()
    print("\n# resources (vgpr, agpr, scratch B/lane, lds B, grid x, grid y, wg)")
    for r in res:
        if is_ours(r[0]):
            print(f"{r[1]:>4} {r[2]:>4} {r[3]:>6} {r[4]:>7} {r[5]:>8} {r[6]:>5} {r[7]:>4}  {r[0][:110]}")


# ---------------------------------------------------------------------------
def read_counter(run_dir, counter):
    acc = defaultdict(float)
    dbs = glob.glob(os.path.join(run_dir, "**", "*results.db"), recursive=True)
    for f in dbs:
        db = sqlite3.connect(f)
        for name, did, v in db.execute("select kernel_name, dispatch_id, value from counters_collection "
                                       "where counter_name = ?", (counter,)):
            acc[(name, f, did)] += float(v)
    files = [] if dbs else glob.glob(os.path.join(run_dir, "**", "*counter_collection.csv"), recursive=True)
    if not dbs and not files:
        sys.exit(f"no rocprofv3 results under {run_dir}")
    for f in files:
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter:
                    continue
                # rows of one dispatch (one per instance) are summed
                acc[(row["Kernel_Name"], f, row.get("Dispatch_Id", ""))] += float(row["Counter_Value"])
    out = defaultdict(list)
    for (name, _, _), v in acc.items():
        out[name].append(v)
    return out


def is_ours(name):
    """Kernels of this library: dg::k_..., including those in anonymous namespaces."""
    return re.search(r"(^|::|\s)k_\w+[(<]", name) is not None


def lattice_mode(name):
    m = re.search(r"k_lattice_(schur|plain)\w*<([^>]*)>", name)
    if not m:
        return None
    args = [x.strip() for x in m.group(2).split(",")]
    idx = 2 if m.group(1) == "plain" else 1
    return {"0": "FWD", "1": "DENS", "2": "VJP", "3": "FWD_HERM", "4": "DENS_HERM", "5": "FWD_REAL"}.get(
        args[idx] if len(args) > idx else "")


def add_calls(d, keys):
    """A forward or adjoint call of the one-matrix-per-lane Schur path can be more
    than one launch: FWD below N_freq plus FWD_HERM or DENS_HERM (fused with the
    adjoint's tr G^2) above; VJP below N_freq plus DENS above when not fused (the
    bench step makes no other launch of those modes).  FWD_CALL and VJP_CALL sum them."""
    for call, parts in (("FWD_CALL", ("FWD", "FWD_HERM", "DENS_HERM")), ("VJP_CALL", ("VJP", "DENS"))):
        have = [m for m in parts if m in d]
        if have:
            d[call] = {k: sum(d[m][k] for m in have) for k in keys}
            d[call]["kernel"] = " + ".join(d[m]["kernel"] for m in have)


def pmc(argv):
    ap = argparse.ArgumentParser(prog="rocprof_summary pmc")
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--copy-mib", type=float, default=0.0)
    ap.add_argument("--factors", type=float, nargs=2, default=None, metavar=("FETCH", "WRITE"),
                    help="calibration factors of an earlier run on the same GPU type, for runs without a copy")
    ap.add_argument("--method", required=True)
    ap.add_argument("--json", default=None)
    ap.add_argument("--command", default="")
    a = ap.parse_args(argv)
    fetch = read_counter(a.fetch, "FETCH_SIZE")
    write = read_counter(a.write, "WRITE_SIZE")
    copy_bytes = a.copy_mib * (1 << 20)

    def copy_kib(d):
        vals = [max(v) for k, v in d.items() if "copyBuffer" in k]
        return max(vals) if vals else None

    print("# HBM traffic per launch from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (separate runs)")
    if a.command:
        print(f"# command: {a.command}")
    if a.factors:
        ff, fw = a.factors
        print(f"# calibration: factors {ff:.4f} (FETCH_SIZE) and {fw:.4f} (WRITE_SIZE) of a calibration copy in an "
              f"earlier run on the same GPU type")
    else:
        cf, cw = copy_kib(fetch), copy_kib(write)
        if not cf or not cw or a.copy_mib <= 0:
            sys.exit("calibration copy not found in the runs (or --copy-mib missing); give --factors")
        ff, fw = copy_bytes / (cf * 1024.0), copy_bytes / (cw * 1024.0)
        print(f"# calibration: a {a.copy_mib:.0f} MiB copy reads FETCH_SIZE {cf:.0f} KiB (factor {ff:.4f}) and "
              f"WRITE_SIZE {cw:.0f} KiB (factor {fw:.4f})")
    print(f"{'dispatches':>10} {'fetch_MB':>10} {'write_MB':>10} {'total_MB':>10}  kernel")
    result = {}
    for name in sorted(set(fetch) | set(write)):
        fv, wv = fetch.get(name, []), write.get(name, [])
        if not fv or not wv:
            continue
        mf = sum(fv) / len(fv) * 1024.0 * ff
        mw = sum(wv) / len(wv) * 1024.0 * fw
        print(f"{len(fv):>10} {mf / 1e6:>10.2f} {mw / 1e6:>10.2f} {(mf + mw) / 1e6:>10.2f}  {name[:120]}")
        mode = lattice_mode(name)
        if mode:
            result[mode] = dict(bytes_per_launch=mf + mw, fetch_bytes=mf, write_bytes=mw, kernel=name)
    add_calls(result, ("bytes_per_launch", "fetch_bytes", "write_bytes"))
    if a.json:
        db = json.load(open(a.json)) if os.path.exists(a.json) else {}
        db[a.method] = {k: v["bytes_per_launch"] for k, v in result.items()}
        db.setdefault("_detail", {})[a.method] = result
        db["_calibration"] = dict(copy_mib=a.copy_mib, fetch_factor=ff, write_factor=fw)
        with open(a.json, "w") as fh:
            json.dump(db, fh, indent=1)


def sq(argv):
    """Executed FP64/FP32 VALU FLOPs and instruction mix per launch from one
    --pmc pass of SQ counters.  SQ_INSTS_VALU_FLOPS_* count FLOPs per wave
    instruction, so they are scaled by the 64 lanes of a wavefront."""
    ap = argparse.ArgumentParser(prog="rocprof_summary sq")
    ap.add_argument("--run", required=True)
    ap.add_argument("--command", default="")
    ap.add_argument("--json", default=None)
    ap.add_argument("--method", default="")
    a = ap.parse_args(argv)
    rows = defaultdict(lambda: defaultdict(float))
    dur = defaultdict(dict)
    for f in glob.glob(os.path.join(a.run, "**", "*results.db"), recursive=True):
        db = sqlite3.connect(f)
        for name, did, cn, v, d in db.execute("select kernel_name, dispatch_id, counter_name, value, duration "
                                              "from counters_collection"):
            rows[name][cn] += v
            dur[name][(f, did)] = d
    print("# executed VALU work per launch from rocprofv3 --pmc SQ_INSTS_VALU_FLOPS_FP64 SQ_INSTS_VALU_FLOPS_FP32 "
          "SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_BUSY_CYCLES SQ_WAVE_CYCLES SQ_INSTS_LDS SQ_INSTS_SMEM")
    if a.command:
        print(f"# command: {a.command}")
    print("# FLOPs = SQ_INSTS_VALU_FLOPS_* x 64 lanes; TFLOP/s over the launch duration under the profiler")
    print(f"{'launch_ms':>9} {'fp64_GF':>9} {'fp32_GF':>9} {'TFLOP/s':>8} {'valu_Minst':>10} {'lds_Minst':>9} "
          f"{'smem_Minst':>10}  kernel")
    out = {}
    for name in sorted(rows):
        if not is_ours(name):
            continue
        nd = len(dur[name])
        t = sum(dur[name].values()) / nd * 1e-9
        c = {k: v / nd for k, v in rows[name].items()}
        f64 = c.get("SQ_INSTS_VALU_FLOPS_FP64", 0.0) * 64
        f32 = c.get("SQ_INSTS_VALU_FLOPS_FP32", 0.0) * 64
        rate = (f64 + f32) / t / 1e12 if t > 0 else 0.0
        print(f"{t * 1e3:>9.2f} {f64 / 1e9:>9.1f} {f32 / 1e9:>9.1f} {rate:>8.1f} {c.get('SQ_INSTS_VALU', 0) / 1e6:>10.1f} "
              f"{c.get('SQ_INSTS_LDS', 0) / 1e6:>9.1f} {c.get('SQ_INSTS_SMEM', 0) / 1e6:>10.1f}  {name[:100]}")
        mode = lattice_mode(name)
        if mode:
            out[mode] = dict(fp64_flops=f64, fp32_flops=f32, launch_s=t, kernel=name)
    add_calls(out, ("fp64_flops", "fp32_flops", "launch_s"))
    if a.json:
        db = json.load(open(a.json)) if os.path.exists(a.json) else {}
        db[a.method] = out
        with open(a.json, "w") as fh:
            json.dump(db, fh, indent=1)


MFMA_COUNTERS = ("SQ_INSTS_VALU_MFMA_MOPS_F64 SQ_INSTS_VALU_MFMA_F64 SQ_VALU_MFMA_BUSY_CYCLES SQ_INSTS_VALU_FLOPS_FP64 "
                 "SQ_INSTS_VALU SQ_BUSY_CYCLES GRBM_GUI_ACTIVE")


def mfma(argv):
    """FP64 work of kernels that use the matrix cores, from one --pmc pass of
    MFMA_COUNTERS: MFMA FLOPs = SQ_INSTS_VALU_MFMA_MOPS_F64 x 512 (one MOP is
    512 FLOPs), VALU FLOPs = SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes, both over the
    launch duration under the profiler against the 78.6 TFLOP/s FP64 peak
    (matrix and vector alike on MI355X); MFMA busy = SQ_VALU_MFMA_BUSY_CYCLES
    over the launch's GRBM_GUI_ACTIVE cycles (that counter is summed over the
    8 XCDs, so divided by 8) x 1024 SIMDs, which matches MFMA FLOPs / (time x
    FP64 matrix peak) on the kernels measured here."""
    ap = argparse.ArgumentParser(prog="rocprof_summary mfma")
    ap.add_argument("--run", required=True)
    ap.add_argument("--command", default="")
    a = ap.parse_args(argv)
    rows = defaultdict(lambda: defaultdict(float))
    dur = defaultdict(dict)
    for f in glob.glob(os.path.join(a.run, "**", "*results.db"), recursive=True):
        db = sqlite3.connect(f)
        for name, did, cn, v, d in db.execute("select kernel_name, dispatch_id, counter_name, value, duration "
                                              "from counters_collection"):
            rows[name][cn] += v
            dur[name][(f, did)] = d
    print(f"# FP64 matrix-core and vector work per launch from rocprofv3 --pmc {MFMA_COUNTERS}")
    if a.command:
        print(f"# command: {a.command}")
    print("# mfma_GF = SQ_INSTS_VALU_MFMA_MOPS_F64 x 512, valu_GF = SQ_INSTS_VALU_FLOPS_FP64 x 64; frac of 78.6 "
          "TFLOP/s over the launch time under the profiler; mfma_busy = SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE/8 "
          "x 1024 SIMDs)")
    print(f"{'launch_ms':>9} {'mfma_GF':>9} {'valu_GF':>9} {'TFLOP/s':>8} {'frac':>6} {'mfma_Minst':>10} "
          f"{'mfma_busy':>9}  kernel")
    for name in sorted(rows):
        if not is_ours(name):
            continue
        nd = len(dur[name])
        t = sum(dur[name].values()) / nd * 1e-9
        c = {k: v / nd for k, v in rows[name].items()}
        fm = c.get("SQ_INSTS_VALU_MFMA_MOPS_F64", 0.0) * 512
        fv = c.get("SQ_INSTS_VALU_FLOPS_FP64", 0.0) * 64
        rate = (fm + fv) / t / 1e12 if t > 0 else 0.0
        gui = c.get("GRBM_GUI_ACTIVE", 0.0) / 8.0
        busy = c.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0) / (gui * 1024) if gui > 0 else 0.0
        print(f"{t * 1e3:>9.3f} {fm / 1e9:>9.1f} {fv / 1e9:>9.1f} {rate:>8.1f} {rate / 78.6:>6.3f} "
              f"{c.get('SQ_INSTS_VALU_MFMA_F64', 0) / 1e6:>10.2f} {busy:>9.3f}  {name[:100]}")


def main():
    if len(sys.argv) < 2 or sys.argv[1] not in ("trace", "pmc", "sq", "mfma"):
        sys.exit(__doc__)
    if sys.argv[1] == "trace":
        trace(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else "")
    elif sys.argv[1] == "sq":
        sq(sys.argv[2:])
    elif sys.argv[1] == "mfma":
        mfma(sys.argv[2:])
    else:
        pmc(sys.argv[2:])


if __name__ == "__main__":
    main()
