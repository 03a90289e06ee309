"""Summarise a rocprofv3 SQLite result (kernel-trace) into a per-kernel table.

usage: python profiles/summarize_rocpd.py <results.db> "<command that produced it>" > profiles/<name>.txt
"""
import sqlite3
import sys


def main(db_path, command):
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
        if r[0].startswith("void dg::k_lattice") or "herm_eig" in r[0] or "static" in r[0]:
            print(f"{r[1]:>4} {r[2]:>4} {r[3]:>6} {r[4]:>7} {r[5]:>8} {r[6]:>5} {r[7]:>4}  {r[0][:110]}")


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2] if len(sys.argv) > 2 else "")
