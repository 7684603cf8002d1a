"""Short text summaries of rocprofv3 CSV output (``--output-format csv``) for the ERI kernels.

    python -m unomol_amd.profsum trace <dir>/run_kernel_stats.csv <dir>/run_kernel_trace.csv \
        --kernel "<demangled name>" --last L
        per-kernel table of the whole run, then the named kernel's average duration over
        its last L dispatches (bench.py runs its timed steps last, so with ``--steps 1``
        those are the timed step's launches) and over all of its dispatches;

    python -m unomol_amd.profsum pmc <dir>/run_counter_collection.csv --counter FETCH_SIZE \
        --kernel "<demangled name>" --last L --copy-bytes 1073741824
        the counter (KiB per dispatch) averaged over the kernel's last L dispatches, the
        same counter read on the streaming-copy kernel ``k_copy_probe`` (which moves
        ``--copy-bytes`` each way), and the kernel's bytes per dispatch scaled by
        (copy bytes / copy reading).

    python -m unomol_amd.profsum flops <dir>/run_counter_collection.csv <bench stderr>
        per ERI kernel of the bench's timed step (exact kernel names): executed VALU FLOPs
        from SQ_INSTS_VALU_FLOPS_FP64 / _FP32 (plus the _TRANS counters) over the kernel's
        last L dispatches (L = the step's launches of that kernel) against the bench's
        algorithmic FLOPs for the same launches.

    python -m unomol_amd.profsum bytes <fetch counters csv> <write counters csv> \
        <dir>/run_kernel_trace.csv --match k_gemm --match k_sp2_step
        per kernel whose name contains one of the --match strings: dispatches, FETCH_SIZE
        and WRITE_SIZE bytes per dispatch (each scaled by its copy-kernel reading), the
        average duration from a kernel trace of the same command, and the HBM rate.

All print Markdown; ``pmc`` ends with one JSON line holding the numbers.
"""
import argparse
import csv
import json
import sys

COPY_KERNEL = "k_copy_probe"


def _dispatches(rows, name_key, kernel):
    out = [r for r in rows if r[name_key] == kernel]
    out.sort(key=lambda r: int(r.get("Dispatch_Id") or r.get("Correlation_Id") or 0))
    return out


def trace(stats_csv, trace_csv, kernel, last, top=25):
    stats = list(csv.DictReader(open(stats_csv)))
    stats.sort(key=lambda r: -float(r["TotalDurationNs"]))
    print("| kernel | calls | total ms | avg us | % |")
    print("|---|---|---|---|---|")
    for r in stats[:top]:
        print("| `%s` | %s | %.2f | %.1f | %.2f |" % (r["Name"][:100], r["Calls"], float(r["TotalDurationNs"]) / 1e6,
                                                   float(r["AverageNs"]) / 1e3, float(r["Percentage"])))
    if not kernel:
        return
    rows = _dispatches(list(csv.DictReader(open(trace_csv))), "Kernel_Name", kernel)
    dur = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6 for r in rows]
    print()
    print("kernel `%s`: %d dispatches in the run" % (kernel, len(dur)))
    if dur:
        n = min(last or len(dur), len(dur))
        print("- last %d dispatches: average %.4f ms, total %.3f ms" % (n, sum(dur[-n:]) / n, sum(dur[-n:])))
        print("- all dispatches: average %.4f ms, total %.3f ms" % (sum(dur) / len(dur), sum(dur)))


def pmc(counter_csv, counter, kernel, last, copy_bytes):
    rows = [r for r in csv.DictReader(open(counter_csv)) if r["Counter_Name"] == counter]
    kr = _dispatches(rows, "Kernel_Name", kernel)
    cr = [r for r in rows if COPY_KERNEL in r["Kernel_Name"]]
    if not kr or not cr:
        sys.exit("no %s dispatches of %s or of the copy kernel in %s" % (counter, kernel, counter_csv))
    n = min(last or len(kr), len(kr))
    kib = sum(float(r["Counter_Value"]) for r in kr[-n:]) / n
    kib_all = sum(float(r["Counter_Value"]) for r in kr) / len(kr)
    copy_kib = sum(float(r["Counter_Value"]) for r in cr) / len(cr)
    scale = copy_bytes / (copy_kib * 1024.0)
    print("counter %s, kernel `%s`" % (counter, kernel))
    print("- last %d dispatches: %.1f KiB per dispatch (all %d: %.1f)" % (n, kib, len(kr), kib_all))
    print("- copy kernel (%d bytes each way): %.1f KiB per dispatch -> scale %.4f" % (copy_bytes, copy_kib, scale))
    print("- kernel bytes per dispatch, scaled: %.4e" % (kib * 1024.0 * scale))
    print(json.dumps({"counter": counter, "kernel": kernel, "dispatches": n, "kib_per_dispatch": kib,
                      "copy_kib": copy_kib, "scale": scale, "bytes_per_dispatch": kib * 1024.0 * scale}))


def _bench_classes(path):
    """The per-class table bench.py writes to stderr ('[bench] eri_classes [...]'), or the
    eri_classes field of a JSON line (bench.py --fock-only)."""
    for line in open(path):
        if line.startswith("[bench] eri_classes "):
            return json.loads(line[len("[bench] eri_classes "):])
    for line in open(path):
        line = line.strip()
        if line.startswith("{"):
            rec = json.loads(line)
            if "eri_classes" in rec:
                return rec["eri_classes"]
    sys.exit("no eri_classes in %s" % path)


def flops(counter_csv, bench_out, lanes=64):
    """Per kernel (exact demangled name; every (class, precision, launch kind) entry of the
    bench table names one kernel instantiation): the bench's algorithmic FLOPs of the timed
    step's launches against the executed FP64/FP32 VALU FLOP counters over the kernel's last
    L dispatches (L = the step's launches of that kernel), counter x lanes = lane-slot FLOPs
    (inactive lanes of a wave included)."""
    classes = _bench_classes(bench_out)
    by_kernel = {}
    for r in csv.DictReader(open(counter_csv)):
        by_kernel.setdefault(r["Kernel_Name"], []).append(r)
    model = {}
    for c in classes:
        m = model.setdefault(c["kernel"], {"gflop": 0.0, "launches": 0, "label": [], "f32": c["precision"] == "f32"})
        m["gflop"] += c["gflop"]
        m["launches"] += int(round(c["launches"]))
        m["label"].append("%s %s k%d" % (c["class"], c["precision"], c.get("kind", 0)))
    print("| kernel | entries | launches | model GFLOP | counter GFLOP (x%d lanes) | counter/model |" % lanes)
    print("|---|---|---|---|---|---|")
    tot_m = tot_c = 0.0
    rows_out = []
    for name, m in sorted(model.items(), key=lambda kv: -kv[1]["gflop"]):
        rows = by_kernel.get(name, [])
        if not rows:
            continue
        names = ("SQ_INSTS_VALU_FLOPS_FP32", "SQ_INSTS_VALU_FLOPS_FP32_TRANS") if m["f32"] else \
                ("SQ_INSTS_VALU_FLOPS_FP64", "SQ_INSTS_VALU_FLOPS_FP64_TRANS")
        disp = sorted({int(r["Dispatch_Id"]) for r in rows})[-m["launches"]:]
        keep = set(disp)
        cnt = lanes * sum(float(r["Counter_Value"]) for r in rows
                          if int(r["Dispatch_Id"]) in keep and r["Counter_Name"] in names)
        tot_m += m["gflop"]
        tot_c += cnt / 1e9
        rows_out.append({"kernel": name, "model_gflop": m["gflop"], "counter_gflop": cnt / 1e9,
                         "launches": len(disp)})
        print("| `%s` | %s | %d | %.3f | %.3f | %.3f |" % (name, ", ".join(m["label"]), len(disp), m["gflop"],
                                                         cnt / 1e9, cnt / 1e9 / max(m["gflop"], 1e-30)))
    print()
    print("all ERI kernels: model %.2f GFLOP, counters %.2f GFLOP, ratio %.3f" % (tot_m, tot_c, tot_c / max(tot_m, 1e-30)))
    print(json.dumps({"model_gflop": tot_m, "counter_gflop": tot_c, "kernels": rows_out[:12]}))


def _copy_scale(rows, copy_bytes):
    cr = [r for r in rows if COPY_KERNEL in r["Kernel_Name"]]
    if not cr:
        sys.exit("no copy-kernel dispatch in the counter file")
    return copy_bytes / (sum(float(r["Counter_Value"]) for r in cr) / len(cr) * 1024.0)


def hbm_bytes(fetch_csv, write_csv, trace_csv, match, copy_bytes, peak_gbps=8000.0):
    """Per-kernel HBM bytes (scaled counters) and rate against the average duration."""
    per = {}
    for path, counter in ((fetch_csv, "FETCH_SIZE"), (write_csv, "WRITE_SIZE")):
        rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == counter]
        scale = _copy_scale(rows, copy_bytes)
        for r in rows:
            name = r["Kernel_Name"]
            if not any(m in name for m in match):
                continue
            e = per.setdefault(name, {})
            e.setdefault(counter, []).append(float(r["Counter_Value"]) * 1024.0 * scale)
    dur = {}
    for r in csv.DictReader(open(trace_csv)):
        name = r["Kernel_Name"]
        if any(m in name for m in match):
            dur.setdefault(name, []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9)
    print("| kernel | dispatches | fetched B / dispatch | written B / dispatch | avg us | GB/s | of 8 TB/s |")
    print("|---|---|---|---|---|---|---|")
    out = []
    for name, e in sorted(per.items()):
        f = e.get("FETCH_SIZE", [0.0])
        w = e.get("WRITE_SIZE", [0.0])
        fb, wb = sum(f) / len(f), sum(w) / len(w)
        d = dur.get(name)
        avg = sum(d) / len(d) if d else float("nan")
        rate = (fb + wb) / avg / 1e9 if d else float("nan")
        print("| `%s` | %d | %.4e | %.4e | %.1f | %.1f | %.3f |" % (name[:90], len(f), fb, wb, avg * 1e6, rate,
                                                                rate / peak_gbps))
        out.append({"kernel": name, "dispatches": len(f), "fetch_bytes": fb, "write_bytes": wb, "avg_s": avg,
                    "GBps": rate})
    print(json.dumps(out))


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("stats_csv")
    t.add_argument("trace_csv")
    t.add_argument("--kernel", default="")
    t.add_argument("--last", type=int, default=0)
    p = sub.add_parser("pmc")
    p.add_argument("counter_csv")
    p.add_argument("--counter", required=True)
    p.add_argument("--kernel", required=True)
    p.add_argument("--last", type=int, default=0)
    p.add_argument("--copy-bytes", type=float, default=float(1 << 30))
    f = sub.add_parser("flops")
    f.add_argument("counter_csv")
    f.add_argument("bench_out")
    f.add_argument("--lanes", type=int, default=64)
    hb = sub.add_parser("bytes")
    hb.add_argument("fetch_csv")
    hb.add_argument("write_csv")
    hb.add_argument("trace_csv")
    hb.add_argument("--match", action="append", required=True)
    hb.add_argument("--copy-bytes", type=float, default=float(1 << 30))
    a = ap.parse_args(argv)
    if a.cmd == "bytes":
        hbm_bytes(a.fetch_csv, a.write_csv, a.trace_csv, a.match, a.copy_bytes)
    elif a.cmd == "trace":
        trace(a.stats_csv, a.trace_csv, a.kernel, a.last)
    elif a.cmd == "flops":
        flops(a.counter_csv, a.bench_out, a.lanes)
    else:
        pmc(a.counter_csv, a.counter, a.kernel, a.last, a.copy_bytes)


if __name__ == "__main__":
    main()
