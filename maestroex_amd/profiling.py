"""HBM traffic of a kernel from rocprofv3 counter runs.

Two `rocprofv3 --pmc` runs of the same command, one collecting FETCH_SIZE and
one WRITE_SIZE (KiB per dispatch), each also covering a plain device-to-device
copy of a known size (`bench.py --calib-copy-mib N`).  The copy's counter
values calibrate the counters: scale = known bytes / counted bytes.  The
kernel's traffic per launch is then the mean over its dispatches of
fetch * fetch_scale + write * write_scale.

    python -m maestroex_amd.profiling --fetch F.csv --write W.csv \\
        --copy-mib 1024 --kernel k_es_fused --title "..." --command "..."

prints a markdown summary whose line `traffic_bytes_per_launch: <int>` is
what bench.py reports as roofline.traffic.
"""
import argparse
import csv
import re
import statistics


def short_name(name):
    n = re.sub(r"\(anonymous namespace\)::", "", name)
    n = re.sub(r"^void ", "", n)
    return n.split("(")[0].split("<")[0].strip()


def per_dispatch(csv_path, counter):
    """{kernel short name: [value per dispatch, in dispatch order]} for one
    counter."""
    out = {}
    acc = {}
    for r in csv.DictReader(open(csv_path)):
        if r["Counter_Name"] != counter:
            continue
        key = (int(r["Dispatch_Id"]), r["Kernel_Name"])
        acc[key] = acc.get(key, 0.0) + float(r["Counter_Value"])
    for (_, name), v in sorted(acc.items()):
        out.setdefault(short_name(name), []).append(v)
    return out


def calibration_copy(counts, copy_kib):
    """The counter value of the calibration copy: the first copy dispatch
    that moves at least an eighth of the calibration size (the calibration
    copy runs before any other large copy; bench.py's later copy-bandwidth
    copies are larger and must not be taken for it)."""
    for name, vs in counts.items():
        if "copy" not in name.lower():
            continue
        for v in vs:
            if v >= copy_kib / 8:
                return v
    raise SystemExit("no calibration copy in the counter file")


def summarize(fetch_csv, write_csv, copy_mib, kernel):
    f = per_dispatch(fetch_csv, "FETCH_SIZE")
    w = per_dispatch(write_csv, "WRITE_SIZE")
    copy_kib = copy_mib * 1024.0
    cf = calibration_copy(f, copy_kib)
    cw = calibration_copy(w, copy_kib)
    fs = copy_kib / cf
    ws = copy_kib / cw
    kf = statistics.mean(f[kernel])
    kw = statistics.mean(w[kernel])
    traffic = (kf * fs + kw * ws) * 1024.0
    return dict(copy_fetch_kib=cf, copy_write_kib=cw, fetch_scale=fs, write_scale=ws,
                fetch_kib=kf, write_kib=kw, dispatches=len(f[kernel]), traffic=traffic)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--copy-mib", type=float, required=True)
    ap.add_argument("--kernel", required=True)
    ap.add_argument("--title", default="HBM traffic")
    ap.add_argument("--command", default="")
    ap.add_argument("--algorithmic-bytes", type=float, default=None,
                    help="algorithmic bytes per launch, for the ratio")
    a = ap.parse_args()
    s = summarize(a.fetch, a.write, a.copy_mib, a.kernel)
    print(f"# {a.title}\n")
    if a.command:
        print(f"Commands (from the repository root, on one MI355X):\n\n```\n{a.command}\n```\n")
    print("| quantity | value |\n|---|---|")
    print(f"| copy kernel FETCH_SIZE (KiB) for {a.copy_mib:g} MiB | {s['copy_fetch_kib']:.1f} |")
    print(f"| copy kernel WRITE_SIZE (KiB) for {a.copy_mib:g} MiB | {s['copy_write_kib']:.1f} |")
    print(f"| FETCH scale (known / counted) | {s['fetch_scale']:.4f} |")
    print(f"| WRITE scale (known / counted) | {s['write_scale']:.4f} |")
    print(f"| `{a.kernel}` FETCH_SIZE per dispatch (KiB, mean of {s['dispatches']}) | "
          f"{s['fetch_kib']:.1f} |")
    print(f"| `{a.kernel}` WRITE_SIZE per dispatch (KiB) | {s['write_kib']:.1f} |")
    print(f"| HBM bytes per launch (scaled) | {s['traffic']:.4g} |")
    if a.algorithmic_bytes:
        print(f"| algorithmic bytes per launch | {a.algorithmic_bytes:.4g} |")
        print(f"| traffic / algorithmic | {s['traffic'] / a.algorithmic_bytes:.3f} |")
    print(f"\ntraffic_bytes_per_launch: {int(round(s['traffic']))}")


if __name__ == "__main__":
    main()
