"""Summaries of rocprofv3 result databases (SQLite, rocpd schema) as markdown.

Kernel trace: per-kernel dispatch count and mean/min/max duration.  HBM traffic: FETCH_SIZE
and WRITE_SIZE (KiB per dispatch) from two separate --pmc runs, each scaled by the reading of
the same counter on the calibration copy kernel of profile_step (dns::k_copy moves a known
number of bytes each way).  The machine-readable lines `traffic <step-kernel> <bytes>` are
what bench.py reads for its roofline entry.

    python -m f90_dns_amd.rocprof_summary --trace T.db [--fetch F.db --write W.db] \\
        [--mfma M.db] --title "..." --command "..." > profiles/NAME.md

MFMA utilisation: SQ_VALU_MFMA_BUSY_CYCLES per dispatch (summed over the SIMDs) from a --pmc run,
divided by SIMDs x engine clock x the dispatch's mean duration from the kernel trace.
"""
import argparse
import collections
import re
import sqlite3

from f90_dns_amd.profile_step import COPY_DOUBLES

# device kernel -> name of the step kernel (dns_profile_kernel_name)
STEP_NAMES = [
    (r"k_zderiv(_rag)?<\d+, 1", "zderiv_vort"),
    (r"k_c2r(_rag)?<\d+, 1,", "c2r_nl"),
    (r"k_c2r_nl_pers<", "c2r_nl"),
    (r"k_r2c(_rag)?<\d+, 1,", "r2c_cross"),
    (r"k_flow_control", "flow_control"),
    (r"k_helmholtz", "helmholtz"),
    (r"k_poisson(_rag|_w2)?<", "poisson_fused"),
]


def step_name(kernel):
    for pat, name in STEP_NAMES:
        if re.search(pat, kernel):
            return name
    return None


def short(kernel):
    return kernel.split("(")[0].replace("void ", "")


def trace_table(db):
    c = sqlite3.connect(db)
    d = collections.defaultdict(list)
    for name, dur in c.execute("select name, duration from kernels"):
        d[short(name)].append(dur)
    return d


def counter_table(db, counter):
    c = sqlite3.connect(db)
    d = collections.defaultdict(list)
    for name, cn, v in c.execute("select kernel_name, counter_name, value from counters_collection"):
        if cn == counter:
            d[short(name)].append(v)
    return d


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--mfma", help="--pmc SQ_VALU_MFMA_BUSY_CYCLES SQ_BUSY_CYCLES result database")
    ap.add_argument("--simds", type=int, default=1024, help="SIMDs of the device (256 CUs x 4)")
    ap.add_argument("--clock-ghz", type=float, default=2.4, help="peak engine clock")
    ap.add_argument("--title", default="kernel statistics")
    ap.add_argument("--command", default="")
    args = ap.parse_args()
    tr = trace_table(args.trace)
    print(f"# {args.title}\n")
    if args.command:
        print("Made by:\n\n```\n" + args.command + "\n```\n")
    print("| kernel | step kernel | dispatches | mean us | min us | max us |")
    print("|---|---|---|---|---|---|")
    rows = sorted(tr.items(), key=lambda kv: -sum(kv[1]))
    for name, v in rows:
        if "rocclr" in name:
            continue
        print(f"| `{name}` | {step_name(name) or ''} | {len(v)} | {sum(v) / len(v) / 1e3:.1f} | "
              f"{min(v) / 1e3:.1f} | {max(v) / 1e3:.1f} |")
    if args.fetch and args.write:
        fe = counter_table(args.fetch, "FETCH_SIZE")
        wr = counter_table(args.write, "WRITE_SIZE")
        copy = COPY_DOUBLES * 8
        cf = [v for k, v in fe.items() if "k_copy" in k]
        cw = [v for k, v in wr.items() if "k_copy" in k]
        if not cf or not cw:
            raise SystemExit("no calibration copy kernel in the counter runs")
        f_fetch = copy / (sum(cf[0]) / len(cf[0]) * 1024.0)
        f_write = copy / (sum(cw[0]) / len(cw[0]) * 1024.0)
        print("\n## HBM traffic per dispatch\n")
        print(f"Calibration (dns::k_copy, {copy / 2**20:.0f} MiB each way): FETCH_SIZE x {f_fetch:.3f}, "
              f"WRITE_SIZE x {f_write:.3f} give bytes.\n")
        print("| kernel | fetch MB | write MB | total MB |")
        print("|---|---|---|---|")
        lines = []
        for name, v in rows:
            if name not in fe or name not in wr or "rocclr" in name or "k_copy" in name:
                continue
            fb = sum(fe[name]) / len(fe[name]) * 1024.0 * f_fetch
            wb = sum(wr[name]) / len(wr[name]) * 1024.0 * f_write
            print(f"| `{name}` | {fb / 1e6:.1f} | {wb / 1e6:.1f} | {(fb + wb) / 1e6:.1f} |")
            sn = step_name(name)
            if sn:
                lines.append(f"traffic {sn} {fb + wb:.0f}")
        print("\nMachine-readable (bytes per dispatch):\n\n```")
        for l in lines:
            print(l)
        print("```")

    if args.mfma:
        mb = counter_table(args.mfma, "SQ_VALU_MFMA_BUSY_CYCLES")
        print("\n## MFMA utilisation per dispatch\n")
        print(f"SQ_VALU_MFMA_BUSY_CYCLES / ({args.simds} SIMDs x {args.clock_ghz} GHz x mean duration):\n")
        print("| kernel | MFMA busy cycles (sum over SIMDs) | mean us | MFMA busy fraction |")
        print("|---|---|---|---|")
        for name, v in rows:
            if name not in mb or "rocclr" in name or "k_copy" in name:
                continue
            busy = sum(mb[name]) / len(mb[name])
            dur_us = sum(v) / len(v) / 1e3
            frac = busy / (args.simds * args.clock_ghz * 1e3 * dur_us)
            print(f"| `{name}` | {busy:.3e} | {dur_us:.1f} | {frac:.2f} |")


if __name__ == "__main__":
    main()
