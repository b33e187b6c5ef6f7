"""Summaries of rocprofv3 CSV output for the propagator kernels (text for profiles/).

    python wormlike-chain-triblock_amd/profile_summary.py kernels <kernel_stats.csv>
        markdown table of per-kernel calls / average / min / max / share
    python wormlike-chain-triblock_amd/profile_summary.py last <kernel_trace.csv> <pairs>
        per bench kernel, the average duration of its launches in the last of <pairs> sweep
        pairs (bench.py's per-kernel timing pass), to compare with the bench line
    python wormlike-chain-triblock_amd/profile_summary.py traffic <fetch.csv> <write.csv> \
            <copy_fetch.csv> <copy_write.csv>
        HBM bytes per launch from FETCH_SIZE and WRITE_SIZE counter runs of bench.py (KiB
        per dispatch), each scaled by what the same counter reads for the 256 MiB device
        copy of `bench.py --copy-only` (profiled the same way, in a process of its own)

The traffic output carries `traffic <kernel> <bytes>` lines, the format bench.py reads.
"""
import collections
import csv
import re
import sys

# bench.py kernel keys -> substrings of the HIP kernel names
KEYS = [
    ("inv_rhs_fft_xy", ("k_inv_rhs64", "k_inv_rhs128", "k_inv_rhs<")),
    ("rhs_fft_xy", ("k_rhs64", "k_rhs128", "k_rhs_fft_xy")),
    ("ifft_xy", ("k_inv64", "k_inv128", "k_ifft_xy")),
    ("zsolve", ("k_zsolve",)),
    ("density_acc", ("k_density_acc",)),
]
COPY_BYTES = 256 * 1024 * 1024


def key_of(name):
    for key, pats in KEYS:
        if any(p in name for p in pats):
            return key
    return None


def last(path, pairs):
    """Average launch duration per bench kernel over its last len / pairs launches."""
    per = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        k = key_of(r["Kernel_Name"])
        if k:
            per[k].append((int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    print("# kernel            launches_in_last_pair  average_us")
    for k in sorted(per):
        v = sorted(per[k])
        n = len(v) // pairs
        tail = v[len(v) - n:]
        avg = sum(b - a for a, b in tail) / max(n, 1) / 1e3
        print(f"# {k:<18} {n:>10}  {avg:10.2f}")


def kernels(path):
    rows = list(csv.DictReader(open(path)))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    print("| kernel | calls | average us | min us | max us | share % |")
    print("|---|---|---|---|---|---|")
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        name = re.sub(r"\s+", " ", r["Name"])[:90]
        print(f"| `{name}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
              f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | "
              f"{100 * float(r['TotalDurationNs']) / tot:.2f} |")


def per_dispatch(path, counter):
    """{dispatch id: (kernel name, value in bytes)}; values summed over the rows of a
    dispatch (one per agent dimension when the CSV splits them)."""
    out = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        d = r["Dispatch_Id"]
        name, v = out.get(d, (r["Kernel_Name"], 0.0))
        out[d] = (name, v + float(r["Counter_Value"]) * 1024.0)
    return out


def calibration(disp):
    vals = [v for n, v in disp.values() if "copyBuffer" in n and v > 0.25 * COPY_BYTES]
    if not vals:
        raise SystemExit("no 256 MiB copy dispatch found: profile bench.py --copy-only")
    return COPY_BYTES / (sum(vals) / len(vals)), sum(vals) / len(vals)


def traffic(fetch_csv, write_csv, cal_fetch_csv, cal_write_csv):
    f, w = per_dispatch(fetch_csv, "FETCH_SIZE"), per_dispatch(write_csv, "WRITE_SIZE")
    fs, fraw = calibration(per_dispatch(cal_fetch_csv, "FETCH_SIZE"))
    ws, wraw = calibration(per_dispatch(cal_write_csv, "WRITE_SIZE"))
    acc = collections.defaultdict(lambda: [0.0, 0.0, 0, 0])
    for d, (n, v) in f.items():
        k = key_of(n)
        if k:
            acc[k][0] += v * fs
            acc[k][2] += 1
    for d, (n, v) in w.items():
        k = key_of(n)
        if k:
            acc[k][1] += v * ws
            acc[k][3] += 1
    print(f"# calibration: the 256 MiB device copy reads FETCH_SIZE {fraw / 1024:.0f} KiB, "
          f"WRITE_SIZE {wraw / 1024:.0f} KiB -> scale FETCH x{fs:.3f}, WRITE x{ws:.3f}")
    print("# kernel            fetch_bytes  write_bytes  total_bytes  launches")
    for k, (fb, wb, nf, nw) in sorted(acc.items()):
        fa, wa = fb / max(nf, 1), wb / max(nw, 1)
        print(f"# {k:16s} {fa:12.4e} {wa:12.4e} {fa + wa:12.4e} {nf:9d}")
    for k, (fb, wb, nf, nw) in sorted(acc.items()):
        print(f"traffic {k} {fb / max(nf, 1) + wb / max(nw, 1):.0f}")


if __name__ == "__main__":
    if len(sys.argv) >= 3 and sys.argv[1] == "kernels":
        kernels(sys.argv[2])
    elif len(sys.argv) >= 4 and sys.argv[1] == "last":
        last(sys.argv[2], int(sys.argv[3]))
    elif len(sys.argv) >= 6 and sys.argv[1] == "traffic":
        traffic(*sys.argv[2:6])
    else:
        print(__doc__)
        sys.exit(2)
