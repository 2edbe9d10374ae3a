"""Summarise a rocprofv3 kernel-trace database (rocpd SQLite) into a short per-kernel table.

Usage: python -m relion_amd.profile_summary <results.db> [--command "..."] > summary.md
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--command", default="")
    ap.add_argument("--top", type=int, default=12)
    a = ap.parse_args()
    rows, total = summarise(a.db)
    if a.command:
        print(f"command: `{a.command}`\n")
    print("| kernel | calls | total ms | avg us | min us | max us | % | VGPR | LDS B | scratch B |")
    print("|---|---|---|---|---|---|---|---|---|---|")
    for name, n, tot, avg, mn, mx, vgpr, lds, scr in rows[:a.top]:
        short = name.split("(")[0]
        print(f"| {short} | {n} | {tot / 1e6:.2f} | {avg / 1e3:.1f} | {mn / 1e3:.1f} | {mx / 1e3:.1f} | "
              f"{100.0 * tot / total:.1f} | {vgpr} | {lds} | {scr} |")
    print(f"\ntotal kernel time: {total / 1e6:.2f} ms")


if __name__ == "__main__":
    main()
