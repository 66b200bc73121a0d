"""Summaries of rocprofv3 output for the profiles/ notes (measurement
tooling beside bench.py; nothing in the library imports it).

Kernel names are shown without "void " and the parameter list, so template
instances keep their arguments ("bt_neighbor_mean<16, 4>"); the <kernels>
arguments below match the base name, every template instance of it.

  python profstats.py stats <kernel_trace.csv | results.db>
      per-kernel calls / average / total / min / max durations
  python profstats.py last <kernel_trace.csv> <kernels> <n>
      average duration of the last n dispatches; <kernels> is one name or a
      comma-separated list whose averages are added (one stage made of
      several launches, each dispatched once per stage)
  python profstats.py series <kernel_trace.csv> <kernel> <n>
      durations (ms) of the last n dispatches of one kernel, in order
  python profstats.py window <kernel_trace.csv> <t0> <t1>
      the dispatches that start between t0 and t1 (seconds of the monotonic
      clock, as bench.py prints them): span from the first start to the
      last end, busy time (union of the dispatch intervals), idle gaps, and
      per-kernel totals, also summed over each kernel's template instances
  python profstats.py queues <kernel_trace.csv> <t0> <t1> <marker>
      per hardware queue: busy time and the spans of the ticks that begin at
      each dispatch of <marker> (a kernel run once per iteration), with the
      gaps between them; then the time two or more queues run kernels
  python profstats.py traffic_series <fetch.csv> <write.csv> <kernels> <n> <copy_bytes>
      the same counters per dispatch (MB fetched and written per kernel, the
      last n dispatches in order), as a table
  python profstats.py valu <counters.csv> <main> <heavy> <n>
      per dispatch (the last n of each of two kernels, in order): duration
      under counter collection (the same rows' timestamps), SQ_INSTS_VALU and
      VALU issue utilization = instructions x 4 cycles / (duration x 2.4 GHz
      x 1,024 SIMDs), with the main kernel's template arguments
  python profstats.py traffic <fetch.csv> <write.csv> <kernels> <n> <copy_bytes>
      HBM bytes per dispatch from FETCH_SIZE / WRITE_SIZE (KiB per dispatch),
      each scaled by what the counter reads for HIP's device copy of
      copy_bytes bytes (read once, written once) in the same run; a kernel
      list is summed as in 'last'; prints JSON
"""
import csv
import json
import sqlite3
import sys


def _short(name):
    """'void bt_neighbor_mean<16, 4>(BtJob const*, ...)' -> 'bt_neighbor_mean<16, 4>'."""
    name = name[5:] if name.startswith("void ") else name
    name = name.replace("(anonymous namespace)::", "")
    depth = 0
    for i, c in enumerate(name):
        if c == "<":
            depth += 1
        elif c == ">":
            depth -= 1
        elif c == "(" and depth == 0:
            return name[:i]
    return name


def _base(name):
    """The kernel's base name, without template arguments."""
    return _short(name).split("<")[0]


def _trace_rows(path):
    if path.endswith(".db"):
        db = sqlite3.connect(path)
        return [(_short(n), float(d)) for n, d in db.execute("select name, duration from kernels")]
    out = []
    with open(path) as f:
        for r in csv.DictReader(f):
            out.append((_short(r["Kernel_Name"]),
                        float(r["End_Timestamp"]) - float(r["Start_Timestamp"])))
    return out


def stats(path):
    agg = {}
    for name, d in _trace_rows(path):
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


def last(path, kernels, n):
    rows = _trace_rows(path)
    total = 0.0
    for k in kernels.split(","):
        d = [t for name, t in rows if _base(name) == k][-n:]
        total += sum(d) / len(d) / 1e6
    return total


def window(path, t0, t1):
    iv = []
    with open(path) as f:
        for r in csv.DictReader(f):
            a, b = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
            if t0 * 1e9 <= a <= t1 * 1e9:
                iv.append((a, b, _short(r["Kernel_Name"])))
    if not iv:
        return "no dispatches in the window"
    iv.sort()
    busy, gaps = 0, []
    cs, ce = iv[0][0], iv[0][1]
    for a, b, _ in iv[1:]:
        if a > ce:
            busy += ce - cs
            gaps.append(a - ce)
            cs, ce = a, b
        else:
            ce = max(ce, b)
    busy += ce - cs
    span = max(b for _, b, _ in iv) - iv[0][0]
    gaps.sort()
    big = [g for g in gaps if g > 500_000]
    med = gaps[len(gaps) // 2] if gaps else 0
    agg = {}
    for a, b, name in iv:
        agg[name] = agg.get(name, 0) + (b - a)
    base = {}
    for a, b, name in iv:
        k = _base(name)
        c = base.setdefault(k, [0, 0])
        c[0] += 1
        c[1] += b - a
    lines = [f"dispatches {len(iv)}, span {span / 1e6:.1f} ms, busy {busy / 1e6:.1f} ms "
             f"({100.0 * busy / span:.0f} %)",
             f"idle gaps: {len(gaps)}, {sum(gaps) / 1e6:.1f} ms in all, median {med / 1e3:.1f} us; "
             f"{len(big)} over 0.5 ms, {sum(big) / 1e6:.1f} ms",
             "| kernel | total ms | share of busy |", "|---|---|---|"]
    for name, t in sorted(agg.items(), key=lambda kv: -kv[1]):
        lines.append(f"| {name[:48]} | {t / 1e6:.2f} | {100.0 * t / busy:.1f} % |")
    lines += ["", "By base name (template instances together):", "",
              "| kernel | dispatches | total ms | avg us | share of busy |", "|---|---|---|---|---|"]
    for name, (c, t) in sorted(base.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"| {name[:48]} | {c} | {t / 1e6:.2f} | {t / c / 1e3:.1f} | "
                     f"{100.0 * t / busy:.1f} % |")
    return "\n".join(lines)


def queues(path, t0, t1, marker):
    """Per hardware queue, over the dispatches starting in [t0, t1]: busy
    time, and the spans of the stretches that begin at each dispatch of
    `marker` (one per tick when the marker runs once per iteration): from
    the marker's start to the last end before the next marker, and the gap
    to the next marker. Last, the time at least two queues run kernels."""
    by = {}
    allv = []
    with open(path) as f:
        for r in csv.DictReader(f):
            a, b = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
            if t0 * 1e9 <= a <= t1 * 1e9:
                v = (a, b, _base(r["Kernel_Name"]), r["Queue_Id"])
                by.setdefault(r["Queue_Id"], []).append(v)
                allv.append(v)

    def pct(xs, q):
        xs = sorted(xs)
        return xs[min(len(xs) - 1, int(q / 100.0 * len(xs)))] if xs else 0.0

    lines = ["| queue | dispatches | busy ms | ticks | span p50 / p90 / p99 ms | gap p50 / p99 ms |",
             "|---|---|---|---|---|---|"]
    for qid, iv in sorted(by.items()):
        iv.sort()
        busy = sum(b - a for a, b, _, _ in iv)
        marks = [k for k, v in enumerate(iv) if v[2] == marker]
        spans, gaps = [], []
        for j in range(len(marks) - 1):
            seg = iv[marks[j]:marks[j + 1]]
            end = max(b for _, b, _, _ in seg)
            spans.append((end - seg[0][0]) / 1e6)
            gaps.append((iv[marks[j + 1]][0] - end) / 1e6)
        if not spans:
            continue
        lines.append(f"| {qid} | {len(iv)} | {busy / 1e6:.1f} | {len(spans)} | "
                     f"{pct(spans, 50):.2f} / {pct(spans, 90):.2f} / {pct(spans, 99):.2f} | "
                     f"{pct(gaps, 50):.3f} / {pct(gaps, 99):.3f} |")
    ev = sorted([(a, 1, q) for a, _, _, q in allv] + [(b, -1, q) for _, b, _, q in allv])
    active, last, two = {}, None, 0
    for t, d, q in ev:
        if last is not None and sum(1 for c in active.values() if c > 0) >= 2:
            two += t - last
        active[q] = active.get(q, 0) + d
        last = t
    lines.append(f"two or more queues running kernels: {two / 1e6:.1f} ms")
    return "\n".join(lines)


def traffic_series(fetch_csv, write_csv, kernels, n, copy_bytes):
    """Per dispatch (the last n of each kernel, in order): MB fetched and
    written, each counter scaled as in `traffic`."""
    cols, heads = [], []
    for path, counter in ((fetch_csv, "FETCH_SIZE"), (write_csv, "WRITE_SIZE")):
        rows = _counter(path, counter)
        calib_kib = max(float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"])
        factor = (copy_bytes / 1024.0) / calib_kib
        for k in kernels.split(","):
            v = [float(r["Counter_Value"]) for r in rows if _base(r["Kernel_Name"]) == k][-n:]
            cols.append([x * 1024.0 * factor / 1e6 for x in v])
            heads.append(f"{k} {'fetch' if counter == 'FETCH_SIZE' else 'write'} MB")
    lines = ["| dispatch | " + " | ".join(heads) + " | total MB |",
             "|---|" + "---|" * (len(heads) + 1)]
    for i in range(min(len(c) for c in cols)):
        vals = [c[i] for c in cols]
        lines.append(f"| {i} | " + " | ".join(f"{x:.1f}" for x in vals) + f" | {sum(vals):.1f} |")
    return "\n".join(lines)


def valu(counter_csv, main, heavy, n):
    rows = _counter(counter_csv, "SQ_INSTS_VALU")
    cols = {}
    for k in (main, heavy):
        mine = [r for r in rows if _base(r["Kernel_Name"]) == k][-n:]
        inst = [float(r["Counter_Value"]) for r in mine]
        dur = [(_short(r["Kernel_Name"]),
                float(r["End_Timestamp"]) - float(r["Start_Timestamp"])) for r in mine]
        cols[k] = (inst, dur)
    peak = 2.4e9 * 1024 / 4.0  # wave64 VALU instructions per second
    lines = ["| iteration | main shape | main us | main VALU instr | main util | heavy us | "
             "heavy VALU instr | heavy util |", "|---|---|---|---|---|---|---|---|"]
    tot = {main: [0.0, 0.0], heavy: [0.0, 0.0]}
    for i in range(n):
        cells = []
        for k in (main, heavy):
            inst, dur = cols[k]
            name, t = dur[i]
            if k == main:
                cells.append("<" + name.split("<", 1)[1] if "<" in name else "")
            cells += [f"{t / 1e3:.0f}", f"{inst[i]:.3e}", f"{inst[i] / (t * 1e-9) / peak:.2f}"]
            tot[k][0] += t
            tot[k][1] += inst[i]
        lines.append(f"| {i} | " + " | ".join(cells) + " |")
    lines.append(f"| all | | {tot[main][0] / 1e3:.0f} | {tot[main][1]:.3e} | "
                 f"{tot[main][1] / (tot[main][0] * 1e-9) / peak:.2f} | {tot[heavy][0] / 1e3:.0f} | "
                 f"{tot[heavy][1]:.3e} | {tot[heavy][1] / (tot[heavy][0] * 1e-9) / peak:.2f} |")
    t_all = tot[main][0] + tot[heavy][0]
    lines.append(f"stage: {t_all / 1e6:.2f} ms for {n} iterations, VALU issue "
                 f"{(tot[main][1] + tot[heavy][1]) / (t_all * 1e-9) / peak:.2f} of peak")
    return "\n".join(lines)


def _counter(path, name):
    with open(path) as f:
        return [r for r in csv.DictReader(f) if r["Counter_Name"] == name]


def traffic(fetch_csv, write_csv, kernel, n, copy_bytes):
    out = {"kernel": kernel, "dispatches": n, "copy_bytes": copy_bytes}
    total = 0.0
    for key, path, counter in (("fetch", fetch_csv, "FETCH_SIZE"), ("write", write_csv, "WRITE_SIZE")):
        rows = _counter(path, counter)
        copies = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]]
        calib_kib = max(copies)  # the large calibration copy
        factor = (copy_bytes / 1024.0) / calib_kib
        kib = 0.0
        for k in kernel.split(","):
            vals = [float(r["Counter_Value"]) for r in rows if _base(r["Kernel_Name"]) == k][-n:]
            kib += sum(vals) / len(vals)
        out[key] = {"kib_per_dispatch": kib, "calibration_kib": calib_kib, "factor": factor,
                    "bytes_per_dispatch": kib * 1024.0 * factor}
        total += kib * 1024.0 * factor
    out["bytes_per_dispatch"] = total
    return out


if __name__ == "__main__":
    cmd = sys.argv[1]
    if cmd == "stats":
        print(stats(sys.argv[2]))
    elif cmd == "last":
        print(f"{last(sys.argv[2], sys.argv[3], int(sys.argv[4])):.4f} ms")
    elif cmd == "series":
        d = [t for name, t in _trace_rows(sys.argv[2]) if _base(name) == sys.argv[3]]
        d = d[-int(sys.argv[4]):]
        print(" ".join(f"{t / 1e6:.2f}" for t in d))
    elif cmd == "window":
        print(window(sys.argv[2], float(sys.argv[3]), float(sys.argv[4])))
    elif cmd == "queues":
        print(queues(sys.argv[2], float(sys.argv[3]), float(sys.argv[4]), sys.argv[5]))
    elif cmd == "traffic_series":
        print(traffic_series(sys.argv[2], sys.argv[3], sys.argv[4], int(sys.argv[5]),
                             int(sys.argv[6])))
    elif cmd == "valu":
        print(valu(sys.argv[2], sys.argv[3], sys.argv[4], int(sys.argv[5])))
    elif cmd == "traffic":
        print(json.dumps(traffic(sys.argv[2], sys.argv[3], sys.argv[4], int(sys.argv[5]),
                                 int(sys.argv[6])), indent=1))
    else:
        raise SystemExit(__doc__)
