"""Summarise rocprofv3 output databases (rocpd SQLite) into the short text
tables kept under profiles/.

    python -m simradar_amd.rocprof_summary kernels DB [DB ...]
    python -m simradar_amd.rocprof_summary counters DB [DB ...]

`kernels` prints per-kernel call counts and average / total durations (the
kernel-trace statistics); `counters` prints the per-dispatch average of every
PMC counter collected, per kernel.
"""
import re
import sqlite3
import sys


def short(name):
    """Readable name of a kernel: library template instantiations are reduced
    to the library's kernel function (plus the lambda index that tells two
    dispatches of it apart)."""
    if len(name) <= 100:
        return name
    m = re.search(r"detail::(radix_sort_onesweep_(?:iteration|global_offsets))", name) or \
        re.search(r"detail::(radix_sort_\w+|device_\w+|\w+_kernel)", name)
    lam = re.findall(r"lambda\(auto:1\)#(\d)", name)
    base = name.split("::")[0] + "::" + (m.group(1) if m else name[:60])
    return base + (" #" + lam[-1] if lam else "")


def kernels(dbs):
    rows = {}
    for db in dbs:
        c = sqlite3.connect(db)
        for name, calls, total, avg in c.execute("select name, total_calls, total_duration, average from top_kernels"):
            r = rows.setdefault(short(name), [0, 0.0])
            r[0] += calls
            r[1] += total
    total_all = sum(v[1] for v in rows.values()) or 1.0
    out = ["| kernel | calls | avg (us) | total (ms) | share |", "|---|---:|---:|---:|---:|"]
    for name, (calls, total) in sorted(rows.items(), key=lambda kv: -kv[1][1]):
        out.append("| %s | %d | %.3f | %.3f | %.1f%% |" % (name, calls, total / calls, total / 1000.0,
                                                        100.0 * total / total_all))
    return "\n".join(out)


def counters(dbs):
    rows = {}
    for db in dbs:
        c = sqlite3.connect(db)
        q = "select kernel_name, counter_name, value, duration from counters_collection"
        for name, counter, value, dur in c.execute(q):
            r = rows.setdefault((short(name), counter), [0, 0.0, 0.0])
            r[0] += 1
            r[1] += value
            r[2] += dur
    out = ["| kernel | counter | dispatches | avg value per dispatch | avg duration (us) |", "|---|---|---:|---:|---:|"]
    for (name, counter), (n, v, d) in sorted(rows.items()):
        out.append("| %s | %s | %d | %.1f | %.3f |" % (name, counter, n, v / n, d / n / 1000.0))
    return "\n".join(out)


if __name__ == "__main__":
    mode, dbs = sys.argv[1], sys.argv[2:]
    print(kernels(dbs) if mode == "kernels" else counters(dbs))
