"""Summarise rocprofv3 CSV output into the short text tables kept under
profiles/.

  python -m mif_amd.rocprof_summary stats  <kernel_stats.csv>
  python -m mif_amd.rocprof_summary traffic <fetch counter_collection.csv> \
        <write counter_collection.csv> --copy-mib 1024 --unknowns N

`traffic` scales FETCH_SIZE and WRITE_SIZE (KiB per dispatch) by what the
same counters read for the plain device copy of --copy-mib MiB that
`bench.py --copy-calibration` performs, and prints per kernel the measured
bytes per launch next to the algorithmic bytes (words per pressure unknown
x unknowns x element size; the element size is read from the kernel's
template type, float or double, unless --esize is given).  The last line is a JSON object with the same
figures, which bench.py reads to fill the roofline's `traffic` field.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re

# algorithmic words per pressure unknown per launch (read once + written once)
ALG_WORDS = [
    # stages 2 and 3 with the previous stage's projection fused in (the last
    # template argument): phi read and the incremented pressure written too
    (r"k_stage<[^,]+, 2, \w+, \w+, \w+, true>", 15.0, "stage 2 + projection 1: Y2*(3),R1(3),p,phi -> Y3(3), S(3), p"),
    (r"k_stage<[^,]+, 3, \w+, \w+, \w+, true>", 12.0, "stage 3 + projection 2: Y3*(3),S(3),p,phi -> u(3), p"),
    (r"k_stage<[^,]+, 1, \w+, \w+, \w+, true>", 12.0,
     "stage 1 + the previous step's last projection: u*(3),p,phi -> Y2(3), R1(3), p"),
    (r"k_stage<[^,]+, 1(, \w+)*>", 10.0, "stage 1: u,v,w,p -> Y2(3), R1(3)"),
    (r"k_stage<[^,]+, 2(, \w+)*>", 13.0, "stage 2: Y2(3),R1(3),p -> Y3(3), S(3)"),
    (r"k_stage<[^,]+, 3(, \w+)*>", 10.0, "stage 3: Y3(3),S(3),p -> u(3)"),
    # the fused defect-correction x passes of the mixed form (FP32 kernels;
    # words of 4 bytes, an FP64 array counting two)
    (r"k_fft_p2<float, \d+, \d+, 0, 0, true, float, true(, false)?>", 6.0, "x forward + divergence: u,v,w -> B, b64"),
    (r"k_fft_p2<float, \d+, \d+, 0, 1, true, float, true(, false)?>", 3.0, "x inverse + unpack: B, B32 -> phi"),
    (r"k_fft_p2<float, \d+, \d+, 0, 0, false, float, false, true>", 4.0,
     "x forward of the defect: phi32 (7-point), b64 -> B"),
    (r"k_fft_p2<[^>]*, 0, 0, true(, \w+)*>", 4.0, "x forward + divergence: u,v,w -> B"),
    (r"k_fft_p2<[^>]*, 0, 1, true(, \w+)*>", 2.0, "x inverse + unpack: B -> phi"),
    (r"k_fft_p2<[^>]*, 0, [01], false(, \w+)*>", 2.0, "x pass: B -> B"),
    (r"k_fft_p2<[^>]*, 1, [01], false(, \w+)*>", 2.0, "y pass: B -> B"),
    (r"k_fft_p2<[^>]*, 2, 2, false(, \w+)*>", 2.0, "z forward/scale/inverse: B -> B"),
    (r"k_project<", 9.0, "u,v,w,phi,p -> u,v,w,p"),
]


def short(name: str) -> str:
    name = name.replace("void ", "")
    return name.split("(")[0]


def alg_words(name: str):
    for pat, words, what in ALG_WORDS:
        if re.search(pat, name):
            return words, what
    return None, ""


def stats(path: str):
    rows = list(csv.DictReader(open(path)))
    print("| kernel | calls | avg us | total ms | % |")
    print("|---|---:|---:|---:|---:|")
    for r in rows:
        print(f"| `{short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
              f"{float(r['TotalDurationNs']) / 1e6:.3f} | {float(r['Percentage']):.1f} |")


def _counter(path: str, ctr: str):
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == ctr:
            d[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return d


def element_size(name: str) -> int:
    m = re.search(r"<(float|double)", name)
    return 4 if m and m.group(1) == "float" else 8


def traffic(fetch_csv: str, write_csv: str, copy_mib: int, unknowns: int, esize: int):
    f = _counter(fetch_csv, "FETCH_SIZE")
    w = _counter(write_csv, "WRITE_SIZE")
    copy_kib = copy_mib * 1024.0
    cf = [v for v in f.get("__amd_rocclr_copyBuffer", []) if v > 0.25 * copy_kib]
    cw = [v for v in w.get("__amd_rocclr_copyBuffer", []) if v > 0.5 * copy_kib]
    if not cf or not cw:
        raise SystemExit("no calibration copy found in the counter files")
    sf = copy_kib / (sum(cf) / len(cf))
    sw = copy_kib / (sum(cw) / len(cw))
    print(f"calibration: {copy_mib} MiB copy reads FETCH_SIZE {sum(cf) / len(cf):.0f} KiB and WRITE_SIZE "
          f"{sum(cw) / len(cw):.0f} KiB per launch -> scale fetch x{sf:.3f}, write x{sw:.3f}")
    print()
    print("| kernel | launches | measured read MB | measured write MB | measured total MB | algorithmic MB | ratio |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    out = {}
    for name in f:
        words, what = alg_words(name)
        if words is None:
            continue
        rd = sum(f[name]) / len(f[name]) * sf * 1024
        wr = sum(w.get(name, [0])) / max(1, len(w.get(name, []))) * sw * 1024
        alg = words * unknowns * (esize or element_size(name))
        key = short(name)
        out[key] = {"read_bytes": rd, "write_bytes": wr, "total_bytes": rd + wr, "alg_bytes": alg}
        print(f"| `{key}` | {len(f[name])} | {rd / 1e6:.1f} | {wr / 1e6:.1f} | {(rd + wr) / 1e6:.1f} | "
              f"{alg / 1e6:.1f} | {(rd + wr) / alg:.2f} |")
    print()
    print(json.dumps({"fetch_scale": sf, "write_scale": sw, "kernels": out}))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_csv")
    t.add_argument("write_csv")
    t.add_argument("--copy-mib", type=int, required=True)
    t.add_argument("--unknowns", type=int, required=True)
    t.add_argument("--esize", type=int, default=0, help="element size in bytes (default: from the kernel type)")
    a = ap.parse_args()
    if a.cmd == "stats":
        stats(a.csv)
    else:
        traffic(a.fetch_csv, a.write_csv, a.copy_mib, a.unknowns, a.esize)


if __name__ == "__main__":
    main()
