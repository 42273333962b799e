"""Summaries of rocprofv3 output for the residual / SSP-RK3 kernels.

    python -m optimisation2forme_amd.profiling stats  <kernel_stats.csv>
    python -m optimisation2forme_amd.profiling traffic <fetch_dir> <write_dir> \
        <copy_fetch_dir> <copy_write_dir> <copy_bytes> [--json out.json]

`stats` prints the per-kernel table of a `rocprofv3 --kernel-trace --stats
--output-format csv` run.  `traffic` turns two `--pmc FETCH_SIZE` /
`--pmc WRITE_SIZE` runs (KiB per dispatch) into HBM bytes per launch: each
counter is scaled by copy_bytes / (what the same counter reads for one plain
device copy of copy_bytes bytes), measured in the two calibration runs
(`python bench.py --copy-calibration`).
"""
import csv
import glob
import json
import os
import sys
from collections import defaultdict


def _csv(path_or_dir, suffix):
    if os.path.isdir(path_or_dir):
        hits = sorted(glob.glob(os.path.join(path_or_dir, "**", "*" + suffix), recursive=True))
        if not hits:
            raise FileNotFoundError("no *%s under %s" % (suffix, path_or_dir))
        return hits[0]
    return path_or_dir


def short_name(name):
    """'void o2f::k_cell_reconstruct(o2f::Topo, ...)' -> 'o2f::k_cell_reconstruct'."""
    n = name.replace("(anonymous namespace)::", "").split("(")[0]
    if n.startswith("void "):
        n = n[5:]
    return n.strip()


def counter_per_dispatch(path_or_dir, counter):
    """{kernel: (dispatches, mean counter value per dispatch)}; the value of a
    dispatch is summed over the rows rocprofv3 writes for it."""
    per = defaultdict(float)
    names = {}
    with open(_csv(path_or_dir, "counter_collection.csv")) as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] != counter:
                continue
            d = int(row["Dispatch_Id"])
            per[d] += float(row["Counter_Value"])
            names[d] = short_name(row["Kernel_Name"])
    acc = defaultdict(list)
    for d, v in per.items():
        acc[names[d]].append(v)
    return {k: (len(v), sum(v) / len(v)) for k, v in acc.items()}


def copy_kib(path_or_dir, counter, pattern="copyBuffer"):
    """Mean counter value (KiB) of one calibration kernel of the
    calibration run (`bench.py --copy-calibration`): the runtime's device
    copy by default; "elementwise_kernel<4, at::native::AUnaryFunctor" is
    the elementwise copy, "gather_kernel" the 128-byte block gather."""
    res = counter_per_dispatch(path_or_dir, counter)
    hits = [k for k in res if pattern in k]
    if not hits:
        raise KeyError("no calibration kernel matching %r in %s" % (pattern, path_or_dir))
    name = max(hits, key=lambda k: res[k][1])
    return name, res[name][1]


def traffic(fetch_dir, write_dir, cfetch_dir, cwrite_dir, copy_bytes):
    fetch = counter_per_dispatch(fetch_dir, "FETCH_SIZE")
    write = counter_per_dispatch(write_dir, "WRITE_SIZE")
    cf_name, cf = copy_kib(cfetch_dir, "FETCH_SIZE")
    cw_name, cw = copy_kib(cwrite_dir, "WRITE_SIZE")
    fscale = copy_bytes / (cf * 1024.0)
    wscale = copy_bytes / (cw * 1024.0)
    out = {"copy_kernel": cf_name, "copy_bytes": copy_bytes, "copy_fetch_kib": cf, "copy_write_kib": cw,
           "fetch_scale": fscale, "write_scale": wscale, "kernels": {}}
    # the other calibration kernels, where the run has them
    for label, pat in (("elementwise", "AUnaryFunctor"), ("block_gather", "gather_kernel")):
        try:
            _, f = copy_kib(cfetch_dir, "FETCH_SIZE", pat)
            _, w = copy_kib(cwrite_dir, "WRITE_SIZE", pat)
            out[label + "_fetch_scale"] = copy_bytes / (f * 1024.0)
            out[label + "_write_scale"] = copy_bytes / (w * 1024.0)
        except KeyError:
            pass
    for k in sorted(set(fetch) & set(write)):
        fb = fetch[k][1] * 1024.0 * fscale
        wb = write[k][1] * 1024.0 * wscale
        fb1 = fetch[k][1] * 1024.0 * out.get("block_gather_fetch_scale", fscale)
        out["kernels"][k] = {"dispatches": fetch[k][0], "fetch_kib": fetch[k][1], "write_kib": write[k][1],
                             "traffic_bytes": fb + wb, "traffic_bytes_gather_scaled": fb1 + wb}
    return out


def stats_table(path):
    rows = []
    with open(_csv(path, "kernel_stats.csv")) as f:
        for row in csv.DictReader(f):
            rows.append((short_name(row["Name"]), int(row["Calls"]), float(row["TotalDurationNs"]),
                         float(row["AverageNs"]), float(row["Percentage"])))
    rows.sort(key=lambda r: -r[2])
    lines = ["| kernel | calls | total ms | avg us | % |", "|---|---|---|---|---|"]
    for n, c, t, a, p in rows:
        lines.append("| %s | %d | %.3f | %.3f | %.2f |" % (n, c, t * 1e-6, a * 1e-3, p))
    return "\n".join(lines)


def main(argv):
    if len(argv) >= 2 and argv[0] == "stats":
        print(stats_table(argv[1]))
        return 0
    if len(argv) >= 6 and argv[0] == "traffic":
        out = traffic(argv[1], argv[2], argv[3], argv[4], float(argv[5]))
        if "--json" in argv:
            with open(argv[argv.index("--json") + 1], "w") as f:
                json.dump(out, f, indent=1)
        print("copy calibration: %s, %.0f bytes -> FETCH_SIZE %.1f KiB (x%.3f), WRITE_SIZE %.1f KiB (x%.3f)"
              % (out["copy_kernel"], out["copy_bytes"], out["copy_fetch_kib"], out["fetch_scale"],
                 out["copy_write_kib"], out["write_scale"]))
        for label in ("elementwise", "block_gather"):
            if label + "_fetch_scale" in out:
                print("%s calibration: FETCH_SIZE x%.3f, WRITE_SIZE x%.3f"
                      % (label, out[label + "_fetch_scale"], out[label + "_write_scale"]))
        print("| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes / launch (copy scale) "
              "| (block-gather scale) |")
        print("|---|---|---|---|---|---|")
        for k, v in sorted(out["kernels"].items(), key=lambda kv: -kv[1]["traffic_bytes"]):
            print("| %s | %d | %.1f | %.1f | %.0f | %.0f |" % (k, v["dispatches"], v["fetch_kib"], v["write_kib"],
                                                             v["traffic_bytes"], v["traffic_bytes_gather_scaled"]))
        return 0
    print(__doc__, file=sys.stderr)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
