"""Summarise a rocprofv3 SQLite (rocpd) kernel trace: per-kernel calls,
total/average duration, registers, scratch and grid size.

Usage: python profiles/summarize_rocpd.py <run_results.db> [title]
"""
import sqlite3
import sys

db = sqlite3.connect(sys.argv[1])
title = sys.argv[2] if len(sys.argv) > 2 else sys.argv[1]
rows = db.execute(
    "select name, count(*), sum(duration), avg(duration), max(vgpr_count), max(accum_vgpr_count), "
    "max(scratch_size), max(lds_size), max(grid_x)/max(workgroup_x) from kernels group by name "
    "order by sum(duration) desc").fetchall()
total = sum(r[2] for r in rows)
print(f"# {title}\n")
print("| kernel | calls | total ms | avg us | % | vgpr | agpr | scratch B/lane | lds B | workgroups |")
print("|---|---|---|---|---|---|---|---|---|---|")
for r in rows:
    name = r[0].replace("void ", "").replace("(pm::JKArgs)", "")
    print(f"| `{name[:70]}` | {r[1]} | {r[2] / 1e6:.3f} | {r[3] / 1e3:.1f} | {100 * r[2] / total:.1f} | {r[4]} | {r[5]} | {r[6]} | {r[7]} | {r[8]} |")
