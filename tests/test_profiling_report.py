"""The reports of parallel_anns_amd.profiling on synthetic rocprofv3
databases.  Uneven launches: each search-kernel dispatch is matched to its
bucket by LDS size (launch order differs from the bucket list), FETCH_SIZE
and WRITE_SIZE (KiB) are scaled by the copy kernel's bytes-per-counted-byte
factors, and extra counter passes are divided by each bucket's iterations.
Single launches (the records the bench's roofline traffic comes from): the
last `steps` dispatches of each variant, warm-up ones ignored.  Expected
values are computed by hand from the numbers planted here."""
import argparse
import json
import sqlite3

import pytest

from parallel_anns_amd import profiling

KERNEL = "void panns::search_kernel<96, false, false, false, {}>(panns::KArgs)"
COPY = "copy_kernel(float const*, float*, unsigned long)"
STEPS = 2
COPY_BYTES = 1 << 20

# two buckets: (largest L, LDS bytes, workgroups, iterations, algorithmic bytes)
BUCKETS = [(1999, 28736, 1280, 1000, 4_000_000_000), (417, 14080, 2048, 4000, 9_000_000_000)]
# per dispatch (launch order: the L <= 417 bucket first), per counter
TRACE_NS = {417: 20_000_000, 1999: 27_000_000}
FETCH_KIB = {417: 3_000_000.0, 1999: 1_500_000.0}
WRITE_KIB = {417: 400_000.0, 1999: 900_000.0}
SQ_VALU = {417: 4_800_000.0, 1999: 1_300_000.0}


def _counters_db(path, rows):
    """rows: (dispatch_id, kernel_name, counter, value, lds, grid)."""
    c = sqlite3.connect(path)
    c.execute("create table counters_collection (dispatch_id integer, kernel_name text, counter_name text, "
              "value real, duration integer, lds_block_size integer, grid_size integer)")
    c.executemany("insert into counters_collection values (?, ?, ?, ?, 1000, ?, ?)", rows)
    c.commit()
    c.close()


def _launches(counter, values):
    rows, did = [], 0
    for step in range(STEPS):
        for L in (417, 1999):  # launch order differs from the bucket list
            lds, wg = next((b[1], b[2]) for b in BUCKETS if b[0] == L)
            did += 1
            rows.append((did, KERNEL.format("true" if L > 1000 else "false"), counter, values[L], lds, 64 * wg))
    return rows


@pytest.fixture
def dbs(tmp_path):
    trace = str(tmp_path / "trace.db")
    c = sqlite3.connect(trace)
    c.execute("create table kernels (name text, duration integer, start integer, lds_block_size integer, "
              "grid_size integer)")
    t = 0
    for step in range(STEPS):
        for L in (417, 1999):
            lds, wg = next((b[1], b[2]) for b in BUCKETS if b[0] == L)
            t += 1
            c.execute("insert into kernels values (?, ?, ?, ?, ?)",
                      (KERNEL.format("true" if L > 1000 else "false"), TRACE_NS[L], t, lds, 64 * wg))
    c.commit()
    c.close()
    out = {"trace": trace}
    for name, counter, vals in (("fetch", "FETCH_SIZE", FETCH_KIB), ("write", "WRITE_SIZE", WRITE_KIB),
                                ("sq", "SQ_INSTS_VALU", SQ_VALU)):
        out[name] = str(tmp_path / f"{name}.db")
        _counters_db(out[name], _launches(counter, vals))
    # copy kernel: 2 x 512 KiB counted for FETCH_SIZE, 1 x 1024 KiB for WRITE_SIZE
    out["cal_fetch"] = str(tmp_path / "cf.db")
    _counters_db(out["cal_fetch"], [(1, COPY, "FETCH_SIZE", 512.0, 0, 1024)])
    out["cal_write"] = str(tmp_path / "cw.db")
    _counters_db(out["cal_write"], [(1, COPY, "WRITE_SIZE", 1024.0, 0, 1024)])
    run = {"config": "c4", "steps": STEPS, "M": 4, "uneven": True, "mean_L": 634.1, "run_ms": 30.0,
           "options": {"bucket_max_L": [417, 1999], "concurrent": 1, "table_eighths": 3, "rows_per_wave": 64,
                       "overflow_slots": 0},
           "algorithmic_bytes_per_run": sum(b[4] for b in BUCKETS),
           "buckets": [{"max_L": L, "nq": 100, "workgroups": wg, "lds_bytes": lds, "hash_slots": 4112,
                        "algorithmic_bytes": alg, "iterations": it} for L, lds, wg, it, alg in BUCKETS]}
    out["run"] = str(tmp_path / "run.json")
    with open(out["run"], "w") as f:
        f.write(json.dumps(run) + "\n")
    out["md"] = str(tmp_path / "r.md")
    out["records"] = str(tmp_path / "r.json")
    return out


def test_uneven_report_scales_matches_and_divides(dbs):
    a = argparse.Namespace(trace=dbs["trace"], fetch=dbs["fetch"], write=dbs["write"], cal_fetch=dbs["cal_fetch"],
                           cal_write=dbs["cal_write"], copy_bytes=COPY_BYTES, run=dbs["run"], command="cmd",
                           md=dbs["md"], records=dbs["records"], dim=96, title="t",
                           counters=dbs["sq"])
    profiling.report_uneven(a)
    md = open(dbs["md"]).read()
    rows = {int(l.split("|")[2]): [x.strip() for x in l.split("|")[1:-1]]
            for l in md.splitlines() if l.startswith("| 0 ") or l.startswith("| 1 ")}
    ff = COPY_BYTES / (512.0 * 1024)  # 2.0 bytes per counted byte
    wf = COPY_BYTES / (1024.0 * 1024)  # 1.0
    total = 0.0
    for L, lds, wg, it, alg in BUCKETS:
        r = rows[L]
        assert int(r[3].replace(",", "")) == wg and int(r[4].replace(",", "")) == lds
        assert float(r[6].replace(",", "")) == pytest.approx(TRACE_NS[L] / 1e3)
        traffic = FETCH_KIB[L] * 1024 * ff + WRITE_KIB[L] * 1024 * wf
        total += traffic
        assert float(r[9]) == pytest.approx(traffic / 1e9, abs=5e-4)
        assert float(r[11]) == pytest.approx(traffic / alg, abs=5e-4)
    rec = json.load(open(dbs["records"]))["records"][0]
    assert rec["traffic_bytes"] == pytest.approx(total, rel=1e-9)
    assert rec["fetch_factor"] == 2.0 and rec["write_factor"] == 1.0
    # the per-iteration table: columns in bucket-list order (L <= 1999, then L <= 417)
    line = next(l for l in md.splitlines() if l.startswith("| SQ_INSTS_VALU |"))
    vals = [float(x.replace(",", "")) for x in line.split("|")[2:-1]]
    assert vals == pytest.approx([SQ_VALU[1999] / 1000, SQ_VALU[417] / 4000], abs=0.05)


def test_fixed_report_takes_the_last_steps_of_each_variant(tmp_path):
    """report() (one launch per step, fixed-L and learned-stop variants):
    the mean of the last `steps` dispatches of each variant, earlier
    (warm-up) dispatches ignored, traffic from the scaled counters."""
    steps = 3
    fixed = KERNEL.format("false")
    learned = "void panns::search_kernel<96, true, false, false, false>(panns::KArgs)"
    # warm-up dispatch first (different values), then `steps` timed ones
    spec = {fixed: ([9e9, 3.0e6, 3.2e6, 3.4e6], [1.0, 5.0e6, 5.0e6, 5.0e6], [7.0, 30e3, 30e3, 30e3]),
            learned: ([9e9, 2.9e6, 3.0e6, 3.1e6], [1.0, 4.5e6, 4.5e6, 4.5e6], [7.0, 10e3, 10e3, 10e3])}
    trace = str(tmp_path / "trace.db")
    c = sqlite3.connect(trace)
    c.execute("create table kernels (name text, duration integer, start integer)")
    t = 0
    for name, (dur, _, _) in spec.items():
        for d in dur:
            t += 1
            c.execute("insert into kernels values (?, ?, ?)", (name, int(d), t))
    c.commit()
    c.close()
    paths = {}
    for key, counter, idx in (("fetch", "FETCH_SIZE", 1), ("write", "WRITE_SIZE", 2)):
        rows, did = [], 0
        for name, vals in spec.items():
            for v in vals[idx]:
                did += 1
                rows.append((did, name, counter, v, 0, 0))
        paths[key] = str(tmp_path / f"{key}.db")
        _counters_db(paths[key], rows)
    paths["cf"] = str(tmp_path / "cf.db")
    _counters_db(paths["cf"], [(1, COPY, "FETCH_SIZE", 512.0, 0, 1024)])
    paths["cw"] = str(tmp_path / "cw.db")
    _counters_db(paths["cw"], [(1, COPY, "WRITE_SIZE", 1024.0, 0, 1024)])
    run = {"config": "c1", "steps": steps,
           "fixed": {"M": 4, "L": 289, "table_eighths": 5, "kernel_ms": 3.2,
                     "algorithmic_bytes_per_launch": 22_000_000_000},
           "learned": {"M": 4, "L": 289, "table_eighths": 4, "kernel_ms": 3.0, "threshold": 0.44,
                       "algorithmic_bytes_per_launch": 20_000_000_000}}
    rp = str(tmp_path / "run.json")
    with open(rp, "w") as f:
        f.write(json.dumps(run) + "\n")
    a = argparse.Namespace(trace=trace, fetch=paths["fetch"], write=paths["write"], cal_fetch=paths["cf"],
                           cal_write=paths["cw"], copy_bytes=COPY_BYTES, run=rp, command="cmd",
                           md=str(tmp_path / "r.md"), records=str(tmp_path / "r.json"), dim=96, title="t")
    profiling.report(a)
    recs = {r["kernel"]: r for r in json.load(open(a.records))["records"]}
    for name, variant, alg in ((fixed, "false", 22e9), (learned, "true", 20e9)):
        dur, fk, wk = spec[name]
        r = recs[f"search_kernel<96, {variant}>"]
        assert r["trace_mean_us"] == pytest.approx(sum(dur[1:]) / 3 / 1e3, abs=0.05)
        traffic = fk[-1] * 1024 * 2.0 + wk[-1] * 1024 * 1.0
        assert r["traffic_bytes"] == pytest.approx(traffic, rel=1e-6)
        assert r["algorithmic_bytes"] == alg
    assert recs["search_kernel<96, true>"]["threshold"] == 0.44
