"""Summaries of rocprofv3 result databases (rocpd SQLite).

    python -m warpsph_amd.profsum kernels <results.db> [--steps N]
    python -m warpsph_amd.profsum dispatches <kernel_trace.csv> [--steps N]
    python -m warpsph_amd.profsum traffic <fetch.db> <write.db> <copy_fetch.db> <copy_write.db> [--copy-bytes B]
    python -m warpsph_amd.profsum pipeline <results.db> [...] [--json out.json]

``kernels`` prints per-kernel call counts and durations (the --stats view);
``traffic`` prints HBM bytes per dispatch of the Jacobi kernels from
FETCH_SIZE and WRITE_SIZE runs, scaled by what the same counters read for a
plain device copy of known size; ``pipeline`` the vector-memory pipeline
figures of the Jacobi kernels (VALU issue, TA/TD busy, cache hit rates).
"""
from __future__ import annotations

import argparse
import collections
import json
import sqlite3


def kernels(db: str, steps: int | None = None, limit: int = 25) -> str:
    cur = sqlite3.connect(db).cursor()
    rows = list(cur.execute(
        "select name, total_calls, total_duration, average, percentage from top_kernels"))
    out = [f"{'kernel':64s} {'calls':>6s} {'total_us':>11s} {'avg_us':>9s} {'%':>6s}"
           + ("  us/step" if steps else "")]
    for name, calls, tot, avg, pct in rows[:limit]:
        line = f"{name[:64]:64s} {calls:6d} {tot:11.1f} {avg:9.2f} {pct:6.2f}"
        if steps:
            line += f"  {tot / steps:8.1f}"
        out.append(line)
    return "\n".join(out)


def dispatches(csv_path: str, steps: int | None = None, limit: int = 25) -> str:
    """Per-kernel durations from a rocprofv3 kernel-trace CSV, with the mean
    over working dispatches beside the plain mean: the device-driven Jacobi
    loop also issues launches that return at once after its stop test, and a
    dispatch counts as working when it takes more than a quarter of the
    kernel's 90th-percentile duration."""
    import csv
    per = collections.defaultdict(list)
    with open(csv_path, newline="") as fh:
        rows = csv.DictReader(fh)
        for r in rows:
            keys = {k.lower(): k for k in r}
            name = r[keys.get("kernel_name", "kernel_name")]
            t0 = int(r[keys["start_timestamp"]])
            t1 = int(r[keys["end_timestamp"]])
            per[name].append((t1 - t0) / 1000.0)
    total = sum(sum(v) for v in per.values())
    out = [f"{'kernel':60s} {'calls':>6s} {'total_us':>11s} {'avg_us':>9s} {'work':>6s} "
           f"{'work_avg_us':>11s} {'%':>6s}" + ("  us/step" if steps else "")]
    for name, ds in sorted(per.items(), key=lambda kv: -sum(kv[1]))[:limit]:
        srt = sorted(ds)
        p90 = srt[int(0.9 * (len(srt) - 1))]
        work = [d for d in ds if d > 0.25 * p90]
        line = (f"{name.split('(')[0][:60]:60s} {len(ds):6d} {sum(ds):11.1f} {sum(ds) / len(ds):9.2f} "
                f"{len(work):6d} {sum(work) / max(1, len(work)):11.2f} {100 * sum(ds) / total:6.2f}")
        if steps:
            line += f"  {sum(ds) / steps:8.1f}"
        out.append(line)
    return "\n".join(out)


def _counter(db: str, counter: str):
    """{kernel name: [value per dispatch]} of one counter (summed over instances)."""
    cur = sqlite3.connect(db).cursor()
    per = collections.defaultdict(float)
    names = {}
    for d, k, c, v in cur.execute(
            "select dispatch_id, kernel_name, counter_name, value from counters_collection"):
        if c == counter:
            per[d] += v
            names[d] = k.split("(")[0]
    res = collections.defaultdict(list)
    for d, v in per.items():
        res[names[d]].append(v)
    return res


def traffic(fetch_db, write_db, copy_fetch_db, copy_write_db, copy_bytes):
    f, w = _counter(fetch_db, "FETCH_SIZE"), _counter(write_db, "WRITE_SIZE")
    cf, cw = _counter(copy_fetch_db, "FETCH_SIZE"), _counter(copy_write_db, "WRITE_SIZE")

    def big(d):  # the copy kernel's dispatches of the calibration run
        vals = [v for vs in d.values() for v in vs]
        return max(vals) if vals else 0.0
    kib_f, kib_w = big(cf), big(cw)
    scale_f = copy_bytes / (kib_f * 1024) if kib_f else float("nan")
    scale_w = copy_bytes / (kib_w * 1024) if kib_w else float("nan")
    out = {"copy_bytes": copy_bytes, "copy_fetch_kib": kib_f, "copy_write_kib": kib_w,
           "fetch_scale": scale_f, "write_scale": scale_w, "kernels": {}}
    for k in sorted(set(f) | set(w)):
        fv = [v for v in f.get(k, []) if v > 0]
        wv = [v for v in w.get(k, []) if v > 0]
        if not fv and not wv:
            continue
        # dispatches doing real work (the device-driven Jacobi loop also
        # issues launches that return at once after the stop test)
        mf = max(fv) if fv else 0.0
        mw = max(wv) if wv else 0.0
        rf = [v for v in fv if v > 0.5 * mf]
        rw = [v for v in wv if v > 0.5 * mw]
        out["kernels"][k] = {"dispatches": max(len(fv), len(wv)), "working_dispatches": len(rf),
                             "fetch_bytes_mean": sum(rf) / max(1, len(rf)) * 1024 * scale_f,
                             "write_bytes_mean": sum(rw) / max(1, len(rw)) * 1024 * scale_w}
    return out


SQ_COLUMNS = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD",
              "SQ_WAVE_CYCLES", "SQ_WAIT_INST_ANY", "SQ_BUSY_CYCLES")


def sq(db: str) -> str:
    """Per kernel, from its dispatch with the most VALU instructions (a working
    one): waves, then VALU, SALU, LDS and vector-read instructions, wave cycles
    and wait cycles per wave, and the dispatch's busy cycles."""
    cur = sqlite3.connect(db).cursor()
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    names = {}
    for d, k, c, v in cur.execute(
            "select dispatch_id, kernel_name, counter_name, value from counters_collection"):
        per[d][c] += v
        names[d] = k.split("(")[0]
    best = {}
    for d, cs in per.items():
        k = names[d]
        if cs.get("SQ_WAVES", 0) <= 0:
            continue
        if k not in best or cs.get("SQ_INSTS_VALU", 0) > best[k].get("SQ_INSTS_VALU", 0):
            best[k] = cs
    out = [f"{'kernel':36s} {'waves':>8s} {'valu/w':>8s} {'salu/w':>8s} {'lds/w':>7s} "
           f"{'vmem/w':>7s} {'wcyc/w':>8s} {'wait/w':>8s} {'busy':>9s}"]
    for k, cs in sorted(best.items(), key=lambda kv: -kv[1].get("SQ_BUSY_CYCLES", 0)):
        w = cs["SQ_WAVES"]
        if w < 1000:
            continue
        out.append(f"{k[:36]:36s} {w:8.0f} {cs['SQ_INSTS_VALU'] / w:8.0f} "
                   f"{cs['SQ_INSTS_SALU'] / w:8.0f} {cs['SQ_INSTS_LDS'] / w:7.0f} "
                   f"{cs['SQ_INSTS_VMEM_RD'] / w:7.0f} {cs['SQ_WAVE_CYCLES'] / w:8.0f} "
                   f"{cs['SQ_WAIT_INST_ANY'] / w:8.0f} {cs['SQ_BUSY_CYCLES']:9.0f}")
    return "\n".join(out)


PIPELINE = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_VMEM_RD", "GRBM_GUI_ACTIVE", "TA_TA_BUSY",
            "TA_ADDR_STALLED_BY_TC_CYCLES", "TD_TD_BUSY", "TD_TC_STALL",
            "TCP_TOTAL_CACHE_ACCESSES", "TCP_TCC_READ_REQ", "TCC_HIT", "TCC_MISS")


def pipeline(dbs, kernels=("k_j1_dev", "k_j2_dev"), cus: int = 256, xcds: int = 8) -> dict:
    """Per kernel, medians over its working dispatches (those with at least
    half the largest VALU count) of the counters in ``PIPELINE``, gathered
    from several runs of the same command (one counter group per run), and
    the derived fractions: VALU issue = 4 cycles per wave instruction over
    the 4 SIMDs of every CU; TA and TD busy per CU; L1 (TCP) and L2 (TCC)
    hit rates; L1 cache accesses per vector read instruction. Cycles per
    dispatch = GRBM_GUI_ACTIVE / number of XCDs."""
    import statistics
    runs = []
    for db in dbs:
        cur = sqlite3.connect(db).cursor()
        per = collections.defaultdict(lambda: collections.defaultdict(float))
        names = {}
        for d, k, c, v in cur.execute(
                "select dispatch_id, kernel_name, counter_name, value from counters_collection"):
            per[d][c] += v
            names[d] = k.split("(")[0]
        runs.append((per, names))
    out = {}
    for kern in kernels:
        vals = collections.defaultdict(list)
        for per, names in runs:
            mine = {d: cs for d, cs in per.items() if kern in names[d]}
            if not mine:
                continue
            # working dispatches: a group without SQ_INSTS_VALU keeps those with
            # the run's largest GRBM_GUI_ACTIVE half (the early-return launches
            # of the device-driven loop last a few microseconds)
            key = "SQ_INSTS_VALU" if any("SQ_INSTS_VALU" in cs for cs in mine.values()) \
                else "GRBM_GUI_ACTIVE"
            top = max(cs.get(key, 0.0) for cs in mine.values())
            for cs in mine.values():
                if key in cs and cs[key] >= 0.5 * top:
                    for c, v in cs.items():
                        vals[c].append(v)
        med = {c: statistics.median(v) for c, v in vals.items()}
        cyc = med.get("GRBM_GUI_ACTIVE", 0.0) / xcds
        d = dict(counters=med, cycles=cyc)
        if cyc:
            if "SQ_INSTS_VALU" in med:
                d["valu_issue"] = med["SQ_INSTS_VALU"] * 4 / (4 * cus) / cyc
            for c, name in (("TA_TA_BUSY", "ta_busy"), ("TD_TD_BUSY", "td_busy"),
                            ("TA_ADDR_STALLED_BY_TC_CYCLES", "ta_stalled_by_tc"),
                            ("TD_TC_STALL", "td_stalled_by_tc")):
                if c in med:
                    d[name] = med[c] / cus / cyc
        if med.get("TCP_TOTAL_CACHE_ACCESSES"):
            d["l1_hit"] = 1 - med.get("TCP_TCC_READ_REQ", 0.0) / med["TCP_TOTAL_CACHE_ACCESSES"]
            if med.get("SQ_INSTS_VMEM_RD"):
                d["l1_accesses_per_vmem_read"] = med["TCP_TOTAL_CACHE_ACCESSES"] / med["SQ_INSTS_VMEM_RD"]
        if med.get("TCC_HIT", 0) + med.get("TCC_MISS", 0) > 0:
            d["l2_hit"] = med["TCC_HIT"] / (med["TCC_HIT"] + med["TCC_MISS"])
        out[kern] = d
    return out


def pipeline_table(res: dict) -> str:
    cols = (("valu_issue", "VALU issue"), ("ta_busy", "TA busy"), ("td_busy", "TD busy"),
            ("ta_stalled_by_tc", "TA<-TC stall"), ("td_stalled_by_tc", "TD<-TC stall"),
            ("l1_hit", "L1 hit"), ("l2_hit", "L2 hit"),
            ("l1_accesses_per_vmem_read", "L1 acc/vmem"))
    out = [f"{'kernel':12s} " + " ".join(f"{h:>12s}" for _, h in cols)]
    for k, d in res.items():
        out.append(f"{k:12s} " + " ".join(
            f"{d[c]:12.3f}" if c in d and c != "l1_accesses_per_vmem_read" else
            (f"{d[c]:12.1f}" if c in d else f"{'-':>12s}") for c, _ in cols))
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("db")
    k.add_argument("--steps", type=int)
    dsp = sub.add_parser("dispatches")
    dsp.add_argument("csv")
    dsp.add_argument("--steps", type=int)
    q = sub.add_parser("sq")
    q.add_argument("db")
    pl = sub.add_parser("pipeline")
    pl.add_argument("dbs", nargs="+")
    pl.add_argument("--json", help="also write the figures to this file")
    pl.add_argument("--kernels", nargs="+", default=["k_j1_dev", "k_j2_dev"],
                    help="kernel name substrings to summarise")
    t = sub.add_parser("traffic")
    t.add_argument("fetch")
    t.add_argument("write")
    t.add_argument("copy_fetch")
    t.add_argument("copy_write")
    t.add_argument("--copy-bytes", type=int, default=1024 * 1024 * 1024)
    a = ap.parse_args()
    if a.cmd == "kernels":
        print(kernels(a.db, a.steps))
    elif a.cmd == "dispatches":
        print(dispatches(a.csv, a.steps))
    elif a.cmd == "sq":
        print(sq(a.db))
    elif a.cmd == "pipeline":
        res = pipeline(a.dbs, tuple(a.kernels))
        print(pipeline_table(res))
        if a.json:
            with open(a.json, "w") as fh:
                json.dump(res, fh, indent=1)
    else:
        print(json.dumps(traffic(a.fetch, a.write, a.copy_fetch, a.copy_write, a.copy_bytes),
                         indent=1))


if __name__ == "__main__":
    main()
