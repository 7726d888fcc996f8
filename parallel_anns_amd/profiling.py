"""Summaries of rocprofv3 output databases (rocpd SQLite, ROCm 7.x).

  python -m parallel_anns_amd.profiling trace   <results.db> [kernel-substring] [--last N]
  python -m parallel_anns_amd.profiling counter <results.db> <COUNTER> [kernel-substring] [--last N]
  python -m parallel_anns_amd.profiling scale   <copy_fetch.db> <copy_write.db> <bytes>
  python -m parallel_anns_amd.profiling report [uneven] --trace T.db --fetch F.db --write W.db --cal-fetch CF.db
                                                --cal-write CW.db --copy-bytes B --run run.json
                                                --command "..." --md out.md --records out.json
                                                [--counters P1.db,P2.db]

`trace` prints per-kernel call counts and mean/min/max durations (all
dispatches, and the last N dispatches of the selected kernel, e.g. the timed
steps of a benchmark); `counter` prints a hardware counter per dispatch;
`scale` turns the FETCH_SIZE / WRITE_SIZE (KiB) readings of a copy kernel of
known size into bytes-per-counted-KiB factors.  `report` combines a kernel
trace and the two counter passes of a profile_run command (the last `steps`
dispatches of each search-kernel variant) with the copy-kernel factors into a
markdown summary and the HBM-traffic records bench.py reports.  `report
uneven` does the same for a profile_run --uneven command: the last
`steps` runs of all its buckets' launches, each dispatch matched to its
bucket by its LDS size and grid (in launch order where those do not tell
buckets apart).
"""
import argparse
import sqlite3
import sys


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


def _columns(c, view):
    return [d[0] for d in c.execute(f"select * from {view} limit 1").description]


def _dispatches(db, view, match, extra=""):
    """(duration ns, key, value) per dispatch of kernels matching `match`, in
    dispatch order; key = (LDS bytes, grid size) where the view has them."""
    c = sqlite3.connect(db)
    cols = _columns(c, view)
    pick = lambda names: next((n for n in names if n in cols), None)
    lds = pick(("lds_size", "lds_block_size", "group_segment_size"))
    grid = pick(("grid_size", "grid_size_x"))
    name = "name" if view == "kernels" else "kernel_name"
    order = "start" if view == "kernels" else "dispatch_id"
    key = ", ".join(x for x in (lds, grid) if x) or "0"
    val = "value" if view != "kernels" else "0"
    q = f"select {name}, duration, {val}, {key} from {view} {extra} order by {order}"
    return [(r[1], tuple(r[3:]), r[2]) for r in c.execute(q) if match in r[0]]


def _by_run(rows, nb, steps, buckets, lds_col=True):
    """The last `steps` runs of nb dispatches each, as lists indexed by bucket."""
    rows = rows[-steps * nb:]
    assert len(rows) == steps * nb, (len(rows), steps, nb)
    runs = []
    for s in range(steps):
        grp = rows[s * nb:(s + 1) * nb]
        keys = [g[1] for g in grp]
        want = [(b["lds_bytes"],) for b in buckets]
        out = [None] * nb
        if len(set(keys)) == nb and keys[0] != (0,):
            # LDS bytes first; the grid (in work-items or workgroups) splits ties
            free = list(range(nb))
            for g in grp:
                cand = [i for i in free if want[i][0] == g[1][0]] or free
                if len(cand) > 1 and len(g[1]) > 1:
                    gs = g[1][1]
                    cand = [i for i in cand if gs in (buckets[i]["workgroups"], 64 * buckets[i]["workgroups"])] or cand
                out[cand[0]] = g
                free.remove(cand[0])
        else:
            out = grp
        runs.append(out)
    return runs


def report_uneven(a):
    import json
    run = json.loads(open(a.run).read().strip().splitlines()[-1])
    steps, bk = run["steps"], run["buckets"]
    nb = len(bk)
    cf = [r[2] for r in counter_values(a.cal_fetch, "FETCH_SIZE", "copy_kernel")]
    cw = [r[2] for r in counter_values(a.cal_write, "WRITE_SIZE", "copy_kernel")]
    ff = a.copy_bytes / (sum(cf) / len(cf) * 1024)
    wf = a.copy_bytes / (sum(cw) / len(cw) * 1024)
    # the fixed-L squared-L2 instantiations, with and without the deferred
    # merge: search_kernel<D, learned = false, norm = false, trace = false, defer>
    match = f"search_kernel<{a.dim}, false, false, false, "
    tr = _by_run(_dispatches(a.trace, "kernels", match), nb, steps, bk)
    fe = _by_run(_dispatches(a.fetch, "counters_collection", match, "where counter_name='FETCH_SIZE'"), nb, steps, bk)
    wr = _by_run(_dispatches(a.write, "counters_collection", match, "where counter_name='WRITE_SIZE'"), nb, steps, bk)
    o = run["options"]
    lines = ["# " + a.title, "", "Command (each pass):", "", "```", a.command, "```", "",
             f"Launch: {nb} queue-size buckets (largest L {o['bucket_max_L']}), "
             f"{'side by side' if o['concurrent'] else 'one after another'}, one hipGraph per run; "
             f"visited tables {'by each bucket' + chr(39) + 's largest L' if o['table_eighths'] == 0 else str(o['table_eighths']) + '/8'}, "
             f"{o['rows_per_wave'] or 96} rows per distance step, {o['overflow_slots']} overflow entries; "
             f"M = {run['M']}, mean L {run['mean_L']}.", "",
             f"Copy-kernel calibration ({a.copy_bytes} bytes copied): FETCH_SIZE factor {ff:.4f}, "
             f"WRITE_SIZE factor {wf:.4f} bytes per counted byte.", "",
             f"Per bucket, mean of the last {steps} runs (algorithmic bytes: distance rows and results exact, "
             "adjacency rows at the batch's mean read slots per expansion)"
             + (": the buckets run side by side, so their kernel durations overlap and each bucket's rate is "
                "the one it gets while sharing the GPU with the others" if o["concurrent"] else "") + ":", "",
             "| bucket | largest L | queries | workgroups | LDS B | visited slots | trace mean us | FETCH KiB | "
             "WRITE KiB | traffic GB | algorithmic GB | traffic / alg. | achieved GB/s |",
             "|---|---|---|---|---|---|---|---|---|---|---|---|---|"]
    mean = lambda xs: sum(xs) / len(xs)
    tot_us = tot_t = 0.0
    records = []
    for i, b in enumerate(bk):
        us = mean([r[i][0] for r in tr]) / 1e3
        fk = mean([r[i][2] for r in fe])
        wk = mean([r[i][2] for r in wr])
        traffic = fk * 1024 * ff + wk * 1024 * wf
        alg = b["algorithmic_bytes"]
        tot_us += us
        tot_t += traffic
        lines.append(f"| {i} | {b['max_L']} | {b['nq']:,} | {b['workgroups']:,} | {b['lds_bytes']:,} | "
                     f"{b['hash_slots']:,} | {us:,.1f} | {fk:,.0f} | {wk:,.0f} | {traffic / 1e9:.3f} | "
                     f"{alg / 1e9:.3f} | {traffic / alg:.3f} | {alg / (us * 1e-6) / 1e9:,.1f} |")
    alg = run["algorithmic_bytes_per_run"]
    lines += ["", f"Whole run: {run['run_ms']} ms device time per run (HIP events around the graph; "
              f"sum of bucket kernel durations {tot_us / 1e3:,.3f} ms), algorithmic {alg / 1e9:.3f} GB, "
              f"HBM traffic {tot_t / 1e9:.3f} GB (traffic / alg. {tot_t / alg:.3f}); achieved "
              f"{alg / (run['run_ms'] * 1e-3) / 1e9:,.1f} GB/s algorithmic = "
              f"{alg / (run['run_ms'] * 1e-3) / 8e12:.3f} of 8 TB/s.", ""]
    records.append({"config": run["config"], "kernel": f"search_kernel<{a.dim}, false>", "M": run["M"],
                    "uneven": True, "buckets": [b["max_L"] for b in bk], "traffic_bytes": int(tot_t),
                    "algorithmic_bytes": int(alg), "run_ms": run["run_ms"], "fetch_factor": round(ff, 4),
                    "write_factor": round(wf, 4), "command": a.command})
    if a.counters and all("iterations" in b for b in bk):
        per = {}
        for db in a.counters.split(","):
            names = [r[0] for r in sqlite3.connect(db).execute("select distinct counter_name from counters_collection")]
            for cn in names:
                rr = _by_run(_dispatches(db, "counters_collection", match, f"where counter_name='{cn}'"), nb, steps,
                             bk)
                per[cn] = [mean([r[i][2] for r in rr]) for i in range(nb)]
        lines += ["Hardware counters per search iteration (sum over the launch / the bucket's iterations; a "
                  "--pmc pass runs the dispatches one at a time, so each bucket runs alone here):", "",
                  "| counter | " + " | ".join(f"L ≤ {b['max_L']}" for b in bk) + " |",
                  "|---|" + "---|" * nb,
                  "| iterations | " + " | ".join(f"{b['iterations']:,}" for b in bk) + " |"]
        for cn in sorted(per):
            lines.append(f"| {cn} | " + " | ".join(f"{per[cn][i] / max(1, bk[i]['iterations']):,.1f}"
                                                   for i in range(nb)) + " |")
        lines.append("")
    lines += ["Kernel trace, all kernels of the traced run:", "", "```"]
    rows = kernel_durations(a.trace)
    for n in sorted({x[0] for x in rows}):
        dd = [x[1] / 1e3 for x in rows if x[0] == n]
        lines.append(f"{n[:60]:60s} calls {len(dd):4d} mean {sum(dd) / len(dd):10.1f} us "
                     f"min {min(dd):10.1f} max {max(dd):10.1f}")
    lines += ["```", ""]
    open(a.md, "w").write("\n".join(lines))
    json.dump({"records": records}, open(a.records, "w"), indent=1)
    print("\n".join(lines))


def report(a):
    import json
    run = json.loads(open(a.run).read().strip().splitlines()[-1])
    steps = run["steps"]
    cf = [r[2] for r in counter_values(a.cal_fetch, "FETCH_SIZE", "copy_kernel")]
    cw = [r[2] for r in counter_values(a.cal_write, "WRITE_SIZE", "copy_kernel")]
    ff = a.copy_bytes / (sum(cf) / len(cf) * 1024)
    wf = a.copy_bytes / (sum(cw) / len(cw) * 1024)
    lines = ["# " + a.title, "", "Command (each pass):", "", "```", a.command, "```", "",
             f"Copy-kernel calibration ({a.copy_bytes} bytes copied): FETCH_SIZE factor {ff:.4f}, "
             f"WRITE_SIZE factor {wf:.4f} bytes per counted byte.", "",
             "| variant | M | L | table 8ths | trace mean us (last %d) | run ms | FETCH KiB | WRITE KiB | "
             "traffic GB | algorithmic GB | traffic / alg. | achieved GB/s |" % steps,
             "|---|---|---|---|---|---|---|---|---|---|---|---|"]
    records = []
    dim = a.dim
    for name, variant in (("fixed", "false"), ("learned", "true")):
        if name not in run:
            continue
        r = run[name]
        # squared-L2 search instantiation: search_kernel<D, learned, norm = false, trace = false, defer>
        match = f"search_kernel<{dim}, {variant}, false, false, "
        kd = kernel_durations(a.trace, match)
        shown = kd[-1][0].split("(")[0].replace("void panns::", "") if kd else match
        d = [x[1] / 1e3 for x in kd][-steps:]
        fv = [x[2] for x in counter_values(a.fetch, "FETCH_SIZE", match)][-steps:]
        wv = [x[2] for x in counter_values(a.write, "WRITE_SIZE", match)][-steps:]
        mean_us = sum(d) / len(d)
        fk, wk = sum(fv) / len(fv), sum(wv) / len(wv)
        traffic = fk * 1024 * ff + wk * 1024 * wf
        alg = r["algorithmic_bytes_per_launch"]
        lines.append(f"| {shown} | {r['M']} | {r['L']} | {r['table_eighths']} | {mean_us:.1f} | {r['kernel_ms']} | "
                     f"{fk:,.0f} | {wk:,.0f} | {traffic / 1e9:.3f} | {alg / 1e9:.3f} | {traffic / alg:.3f} | "
                     f"{alg / (mean_us * 1e-6) / 1e9:,.1f} |")
        rec = {"config": run["config"], "kernel": f"search_kernel<{dim}, {variant}>", "M": r["M"], "L": r["L"],
               "traffic_bytes": int(traffic), "algorithmic_bytes": int(alg), "fetch_kib": round(fk, 1),
               "write_kib": round(wk, 1),
               "fetch_factor": round(ff, 4), "write_factor": round(wf, 4), "trace_mean_us": round(mean_us, 1),
               "command": a.command}
        if "threshold" in r:
            rec["threshold"] = r["threshold"]
        records.append(rec)
    lines += ["", "Kernel trace, all kernels of the traced run:", "", "```"]
    rows = kernel_durations(a.trace)
    for n in sorted({x[0] for x in rows}):
        dd = [x[1] / 1e3 for x in rows if x[0] == n]
        lines.append(f"{n[:60]:60s} calls {len(dd):4d} mean {sum(dd) / len(dd):10.1f} us "
                     f"min {min(dd):10.1f} max {max(dd):10.1f}")
    lines += ["```", ""]
    open(a.md, "w").write("\n".join(lines))
    json.dump({"records": records}, open(a.records, "w"), indent=1)
    print("\n".join(lines))


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "report":
        ap = argparse.ArgumentParser()
        ap.add_argument("mode")
        ap.add_argument("kind", nargs="?", default="fixed", choices=["fixed", "uneven"])
        for k in ("--trace", "--fetch", "--write", "--cal-fetch", "--cal-write", "--run", "--command", "--md",
                  "--records"):
            ap.add_argument(k, required=True)
        ap.add_argument("--copy-bytes", type=int, required=True)
        ap.add_argument("--counters", default="", help="comma-separated --pmc databases of the same command "
                        "(uneven reports: a per-iteration table per bucket)")
        ap.add_argument("--dim", type=int, default=96)
        ap.add_argument("--title", default="search_kernel, 1M x 96, 10k queries")
        a = ap.parse_args()
        return report_uneven(a) if a.kind == "uneven" else report(a)
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
