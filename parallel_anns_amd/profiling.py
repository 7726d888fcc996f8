"""Summaries of rocprofv3 output databases (rocpd SQLite, ROCm 7.x).

  python -m parallel_anns_amd.profiling trace   <results.db> [kernel-substring] [--last N]
  python -m parallel_anns_amd.profiling counter <results.db> <COUNTER> [kernel-substring] [--last N]
  python -m parallel_anns_amd.profiling scale   <copy_fetch.db> <copy_write.db> <bytes>

`trace` prints per-kernel call counts and mean/min/max durations (all
dispatches, and the last N dispatches of the selected kernel, e.g. the timed
steps of a benchmark); `counter` prints a hardware counter per dispatch;
`scale` turns the FETCH_SIZE / WRITE_SIZE (KiB) readings of a copy kernel of
known size into bytes-per-counted-KiB factors.
"""
import argparse
import sqlite3


def kernel_durations(db, match=None):
    c = sqlite3.connect(db)
    rows = list(c.execute("select name, duration from kernels order by start"))
    if match:
        rows = [r for r in rows if match in r[0]]
    return rows


def counter_values(db, counter, match=None):
    c = sqlite3.connect(db)
    rows = list(c.execute("select dispatch_id, kernel_name, value, duration from counters_collection "
                          "where counter_name=? order by dispatch_id", (counter,)))
    if match:
        rows = [r for r in rows if match in r[1]]
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["trace", "counter", "scale"])
    ap.add_argument("args", nargs="+")
    ap.add_argument("--last", type=int, default=0)
    a = ap.parse_args()
    if a.mode == "trace":
        db = a.args[0]
        match = a.args[1] if len(a.args) > 1 else None
        rows = kernel_durations(db)
        names = sorted({r[0] for r in rows})
        print(f"{'kernel':70s} {'calls':>6s} {'mean_us':>12s} {'min_us':>12s} {'max_us':>12s}")
        for n in names:
            d = [r[1] / 1e3 for r in rows if r[0] == n]
            print(f"{n[:70]:70s} {len(d):6d} {sum(d) / len(d):12.3f} {min(d):12.3f} {max(d):12.3f}")
        if match and a.last:
            d = [r[1] / 1e3 for r in kernel_durations(db, match)][-a.last:]
            print(f"last {len(d)} dispatches of *{match}*: mean {sum(d) / len(d):.3f} us; "
                  + " ".join(f"{x:.1f}" for x in d))
    elif a.mode == "counter":
        db, counter = a.args[0], a.args[1]
        match = a.args[2] if len(a.args) > 2 else None
        rows = counter_values(db, counter, match)
        if a.last:
            rows = rows[-a.last:]
        for r in rows:
            print(f"dispatch {r[0]:5d} {r[1][:50]:50s} {counter} {r[2]:14.1f} KiB  duration {r[3] / 1e3:10.3f} us")
        if rows:
            print(f"mean {counter}: {sum(r[2] for r in rows) / len(rows):.1f} KiB per dispatch")
    else:
        fdb, wdb, nbytes = a.args[0], a.args[1], int(a.args[2])
        f = [r[2] for r in counter_values(fdb, "FETCH_SIZE", "copy_kernel")]
        w = [r[2] for r in counter_values(wdb, "WRITE_SIZE", "copy_kernel")]
        ff = nbytes / (sum(f) / len(f) * 1024)
        wf = nbytes / (sum(w) / len(w) * 1024)
        print(f"copy of {nbytes} bytes: FETCH_SIZE {sum(f) / len(f):.1f} KiB -> factor {ff:.4f} bytes per "
              f"counted byte; WRITE_SIZE {sum(w) / len(w):.1f} KiB -> factor {wf:.4f}")


if __name__ == "__main__":
    main()
