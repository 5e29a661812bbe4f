"""rocprofv3 helpers: per-kernel statistics and HBM traffic of the bench.

``python -m gauxc_amd.profiling stats OUT.md -- <bench args>`` runs
``rocprofv3 --kernel-trace --stats`` on bench.py and writes the per-kernel
summary (calls, total and average duration) as markdown.

``python -m gauxc_amd.profiling traffic OUT.json -- <bench args>`` runs two
``rocprofv3 --pmc`` passes (FETCH_SIZE, then WRITE_SIZE) on bench.py and
the same two passes on ``bench.py --copy-calibration`` (three 1 GiB device
copies), and writes per-kernel KiB per dispatch plus the factors that map
each counter onto the copy's known byte count.  bench.py reads that file
to report measured HBM bytes (``roofline.traffic``).  A markdown table of
the same numbers is written beside it (``OUT.md``); ``python -m
gauxc_amd.profiling table OUT.json`` rewrites that table from the file.

rocprofv3 runs as a child process with the program itself after ``--``;
this module never touches the GPU.  Raw rocprofv3 output goes under
``$XCQ_PROFILE_DIR`` (default: a directory in the system's temporary
directory).
"""
import csv
import json
import os
import re
import sqlite3
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def _short(name):
    name = re.sub(r"\(.*", "", name).replace("void ", "").replace("xcq::", "")
    return name.strip()


def _run(cmd, timeout):
    print("[profiling] " + " ".join(cmd), flush=True)
    env = dict(os.environ, TMPDIR=tempfile.gettempdir())
    r = subprocess.run(cmd, timeout=timeout, env=env, cwd=tempfile.gettempdir())
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}")


def _pmc(counter, out_dir, bench_args, timeout=240):
    d = os.path.join(out_dir, counter)
    _run(["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--",
          sys.executable, BENCH] + bench_args, timeout)
    per = {}
    for root, _, files in os.walk(d):
        for f in files:
            if f.endswith("counter_collection.csv"):
                with open(os.path.join(root, f)) as fh:
                    for row in csv.DictReader(fh):
                        if row["Counter_Name"] != counter:
                            continue
                        k = _short(row["Kernel_Name"])
                        e = per.setdefault(k, [0.0, set()])
                        e[0] += float(row["Counter_Value"])
                        e[1].add(row["Dispatch_Id"])
    return {k: (v[0], len(v[1])) for k, v in per.items()}


def traffic(out_json, bench_args, work_dir):
    fetch = _pmc("FETCH_SIZE", work_dir, bench_args)
    write = _pmc("WRITE_SIZE", work_dir, bench_args)
    cf = _pmc("FETCH_SIZE", os.path.join(work_dir, "copy"), ["--copy-calibration"])
    cw = _pmc("WRITE_SIZE", os.path.join(work_dir, "copy"), ["--copy-calibration"])
    copy_bytes = 3 * (1 << 30)  # three copies of 1 GiB
    ck = [k for k in cf if "copyBuffer" in k]
    if not ck:
        raise RuntimeError("copy calibration kernel not found in the profile")
    fetch_scale = copy_bytes / (cf[ck[0]][0] * 1024)
    write_scale = copy_bytes / (cw[ck[0]][0] * 1024)
    kernels = {}
    for k in sorted(set(fetch) & set(write)):
        kernels[k] = {"fetch_kib": fetch[k][0] / fetch[k][1], "write_kib": write[k][0] / write[k][1],
                      "dispatches": fetch[k][1]}
    doc = {"command_fetch": "rocprofv3 --pmc FETCH_SIZE --output-format csv -- python3 bench.py "
                            + " ".join(bench_args),
           "command_write": "rocprofv3 --pmc WRITE_SIZE --output-format csv -- python3 bench.py "
                            + " ".join(bench_args),
           "command_copy": "rocprofv3 --pmc FETCH_SIZE|WRITE_SIZE -- python3 bench.py "
                           "--copy-calibration",
           "fetch_scale": fetch_scale, "write_scale": write_scale,
           "copy_fetch_kib": cf[ck[0]][0], "copy_write_kib": cw[ck[0]][0],
           "kernels": kernels}
    with open(out_json, "w") as f:
        json.dump(doc, f, indent=1)
    table(out_json)
    print(json.dumps({"fetch_scale": fetch_scale, "write_scale": write_scale}), flush=True)


def table(in_json):
    """Markdown rendering of a traffic summary: per-kernel counters per
    dispatch and the scaled HBM bytes, written to the same path with .md."""
    with open(in_json) as f:
        doc = json.load(f)
    fs, ws = doc["fetch_scale"], doc["write_scale"]
    rows = []
    for k, e in doc["kernels"].items():
        gb = (e["fetch_kib"] * fs + e["write_kib"] * ws) * 1024 / 1e9
        rows.append((gb, k, e))
    rows.sort(key=lambda r: -r[0])
    out_md = os.path.splitext(in_json)[0] + ".md"
    name = os.path.basename(in_json)
    with open(out_md, "w") as f:
        f.write("# HBM traffic per kernel launch (rocprofv3 PMC)\n\n")
        f.write("Commands (one pass per counter, then the same two counters on a plain copy):\n\n")
        for c in ("command_fetch", "command_write", "command_copy"):
            f.write("    " + doc[c] + "\n")
        f.write(f"\n(driven by `python -m gauxc_amd.profiling traffic {name} -- <bench args>`; "
                f"bench.py reads `{name}` for `roofline.traffic`).\n\n")
        f.write(f"Copy calibration: three 1 GiB copies read {doc['copy_fetch_kib']:.0f} KiB by "
                f"FETCH_SIZE and wrote {doc['copy_write_kib']:.0f} KiB by WRITE_SIZE, so "
                f"FETCH_SIZE is scaled by {fs:.4f} and WRITE_SIZE by {ws:.4f}.\n\n")
        f.write("| kernel | dispatches | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch "
                "| HBM GB/dispatch (scaled) |\n|---|---|---|---|---|\n")
        for gb, k, e in rows:
            f.write(f"| `{k}` | {e['dispatches']} | {e['fetch_kib']:.0f} | {e['write_kib']:.0f} "
                    f"| {gb:.2f} |\n")
    print(f"[profiling] wrote {out_md}", flush=True)


def stats(out_md, bench_args, work_dir, timeout=300):
    d = os.path.join(work_dir, "stats")
    cmd = ["rocprofv3", "--kernel-trace", "--stats", "-d", d, "-o", "run", "--",
           sys.executable, BENCH] + bench_args
    _run(cmd, timeout)
    rows = []
    for root, _, files in os.walk(d):
        for f in files:
            p = os.path.join(root, f)
            if f.endswith("kernel_stats.csv"):
                with open(p) as fh:
                    for r in csv.DictReader(fh):
                        rows.append((_short(r["Name"]), int(r["Calls"]),
                                     float(r["TotalDurationNs"]) / 1e3,
                                     float(r["AverageNs"]) / 1e3))
            elif f.endswith(".db") and not rows:
                con = sqlite3.connect(p)
                for r in con.execute("select name, total_calls, total_duration, average "
                                     "from top_kernels"):
                    rows.append((_short(r[0]), int(r[1]), float(r[2]), float(r[3])))
    rows.sort(key=lambda r: -r[2])
    with open(out_md, "w") as f:
        f.write("# rocprofv3 kernel statistics\n\n")
        f.write("Command: `rocprofv3 --kernel-trace --stats -- python3 bench.py "
                + " ".join(bench_args) + "`\n\n")
        f.write("| kernel | calls | total (us) | average (us) |\n|---|---|---|---|\n")
        for r in rows[:25]:
            f.write(f"| `{r[0]}` | {r[1]} | {r[2]:.1f} | {r[3]:.1f} |\n")
    print(f"[profiling] wrote {out_md} ({len(rows)} kernels)", flush=True)


def main(argv):
    if len(argv) == 2 and argv[0] == "table":
        table(os.path.abspath(argv[1]))
        return 0
    if len(argv) < 3 or argv[0] not in ("stats", "traffic") or "--" not in argv:
        print(__doc__)
        return 2
    mode, out = argv[0], os.path.abspath(argv[1])
    bench_args = argv[argv.index("--") + 1:]
    work = os.path.abspath(os.environ.get("XCQ_PROFILE_DIR",
                                          os.path.join(tempfile.gettempdir(), "xcq_profiling")))
    os.makedirs(work, exist_ok=True)
    if mode == "stats":
        stats(out, bench_args, work)
    else:
        traffic(out, bench_args, work)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
