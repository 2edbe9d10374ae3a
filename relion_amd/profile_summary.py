"""Summarise rocprofv3 databases (rocpd SQLite) into short per-kernel tables.

Kernel trace:  python relion_amd/profile_summary.py <kt_results.db> [--command "..."] > summary.md
HBM traffic:   python relion_amd/profile_summary.py --fetch <fetch.db> --write <write.db> [--fetch-scale 2.0]
               [--write-scale 1.0] [--kernel-fetch-scale NAME=SCALE ...] [--command "..."] > traffic.md
The traffic mode averages FETCH_SIZE and WRITE_SIZE (KiB per dispatch) per kernel, multiplies them by the
calibration scales (what the counters read for a device-to-device copy of known size) and reports GB per
launch and the resulting bandwidth over the dispatch durations of the counter runs, then one line
``traffic_bytes_per_launch <kernel> <bytes>`` per kernel. FETCH_SIZE counts a 128-byte read request as 64 bytes
and a 64-byte one as 64 bytes, so the copy calibration (x2) holds for streaming kernels; --kernel-fetch-scale sets
the scale of a kernel whose reads are 64-byte gathers to 1.
"""
import argparse
import sqlite3


def summarise(db):
    con = sqlite3.connect(db)
    rows = con.execute(
        "select name, count(*), sum(duration), avg(duration), min(duration), max(duration), max(vgpr_count), "
        "max(lds_size), max(scratch_size) from kernels group by name order by sum(duration) desc").fetchall()
    total = sum(r[2] for r in rows) or 1
    return rows, total


def counter(db, name):
    """{kernel: (dispatches, mean counter value, mean duration ns)} of one PMC counter."""
    con = sqlite3.connect(db)
    rows = con.execute(
        "select name, count(*), avg(counter_value), avg(duration) from pmc_events where counter_name = ? "
        "group by name", (name,)).fetchall()
    return {r[0]: (r[1], r[2], r[3]) for r in rows}


def short(name):
    name = name.split("(")[0]
    return name[5:] if name.startswith("void ") else name


def kernel_table(a):
    rows, total = summarise(a.db)
    if a.command:
        print(f"command: `{a.command}`\n")
    print("| kernel | calls | total ms | avg us | min us | max us | % | VGPR | LDS B | scratch B |")
    print("|---|---|---|---|---|---|---|---|---|---|")
    for name, n, tot, avg, mn, mx, vgpr, lds, scr in rows[:a.top]:
        print(f"| {short(name)} | {n} | {tot / 1e6:.2f} | {avg / 1e3:.1f} | {mn / 1e3:.1f} | {mx / 1e3:.1f} | "
              f"{100.0 * tot / total:.1f} | {vgpr} | {lds} | {scr} |")
    print(f"\ntotal kernel time: {total / 1e6:.2f} ms")


def traffic_table(a):
    f = counter(a.fetch, "FETCH_SIZE")
    w = counter(a.write, "WRITE_SIZE")
    if a.command:
        print(f"command: `{a.command}`\n")
    print(f"FETCH_SIZE x {a.fetch_scale:.3f}, WRITE_SIZE x {a.write_scale:.3f} (copy calibration)")
    for k, v in a.kernel_scale.items():
        print(f"FETCH_SIZE x {v:.3f} for {k} (64-byte gathers)")
    print()
    print("| kernel | launches | avg us | fetch GB/launch | write GB/launch | GB/s |")
    print("|---|---|---|---|---|---|")
    rows = []
    for name, (n, fk, dur) in f.items():
        wk = w.get(name, (0, 0.0, dur))[1]
        fb = fk * 1024.0 * a.kernel_scale.get(short(name), a.fetch_scale)
        wb = wk * 1024.0 * a.write_scale
        rows.append((n * dur, name, n, dur, fb, wb))
    top = sorted(rows, reverse=True)[:a.top]
    for _, name, n, dur, fb, wb in top:
        bw = (fb + wb) / (dur * 1e-9) / 1e9 if dur else 0.0
        print(f"| {short(name)} | {n} | {dur / 1e3:.1f} | {fb / 1e9:.3f} | {wb / 1e9:.3f} | {bw:.0f} |")
    # machine-readable lines (bench.py reads the one of its dominant kernel)
    print()
    for _, name, n, dur, fb, wb in top:
        print(f"traffic_bytes_per_launch {short(name)} {fb + wb:.0f}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db", nargs="?")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--fetch-scale", type=float, default=2.0)
    ap.add_argument("--write-scale", type=float, default=1.0)
    ap.add_argument("--kernel-fetch-scale", action="append", default=[], metavar="NAME=SCALE")
    ap.add_argument("--command", default="")
    ap.add_argument("--top", type=int, default=12)
    a = ap.parse_args()
    a.kernel_scale = {k: float(v) for k, v in (x.rsplit("=", 1) for x in a.kernel_fetch_scale)}
    if a.fetch and a.write:
        traffic_table(a)
    else:
        kernel_table(a)


if __name__ == "__main__":
    main()
