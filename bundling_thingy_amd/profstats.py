"""Per-kernel duration summary of a rocprofv3 kernel trace.

Usage: python -m bundling_thingy_amd.profstats <run_results.db | kernel_trace.csv>
Reads the rocpd SQLite database (rocprofv3's default output) or a
kernel_trace CSV and prints calls, average, total, min and max per kernel.
"""
import csv
import sqlite3
import sys


def rows_from_db(path):
    db = sqlite3.connect(path)
    q = "select name, duration from kernels"
    return [(n, float(d)) for n, d in db.execute(q)]


def rows_from_csv(path):
    out = []
    with open(path) as f:
        for r in csv.DictReader(f):
            out.append((r["Kernel_Name"], float(r["End_Timestamp"]) - float(r["Start_Timestamp"])))
    return out


def summarize(rows):
    agg = {}
    for name, d in rows:
        a = agg.setdefault(name, [0, 0.0, float("inf"), 0.0])
        a[0] += 1
        a[1] += d
        a[2] = min(a[2], d)
        a[3] = max(a[3], d)
    lines = ["| kernel | calls | avg us | total ms | min us | max us |", "|---|---|---|---|---|---|"]
    for name, (c, t, mn, mx) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"| {name[:48]} | {c} | {t / c / 1e3:.2f} | {t / 1e6:.3f} | "
                     f"{mn / 1e3:.2f} | {mx / 1e3:.2f} |")
    return "\n".join(lines)


if __name__ == "__main__":
    p = sys.argv[1]
    print(summarize(rows_from_db(p) if p.endswith(".db") else rows_from_csv(p)))
