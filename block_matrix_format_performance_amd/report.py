"""Speed-up and roofline report for the driver's result files (the role of the
reference's plot/plot.py:19-54, which loads results/result.json, divides the
"CPU CSR" time by every format's time and plots the speed-up against the
block size).

    python -m block_matrix_format_performance_amd.report result.json \\
        [--gbs result_gbs.json] [--peak 8000] [--png-prefix report]

result.json is {"<bs>": {"<format>": seconds}} as written by bmfp_main (and by
the reference's main.cpp); result_gbs.json (bmfp_main --out-gbs) holds each
format's effective bandwidth in the same layout, from which the fraction of
the HBM roofline is reported.  Prints the tables; with --png-prefix also
writes <prefix>_speedup.png and <prefix>_roofline.png (matplotlib, Agg).
"""
import argparse
import json
import math


def load(path):
    """{bs (int): {format: value}} sorted by block size."""
    with open(path) as f:
        raw = json.load(f)
    return {int(bs): rows for bs, rows in sorted(raw.items(), key=lambda kv: int(kv[0]))}


def speedup(results, base="CPU CSR"):
    """{bs: {format: base time / format time}} (plot.py:calculate_speedup)."""
    out = {}
    for bs, rows in results.items():
        if base not in rows:
            raise KeyError(f"no '{base}' row for bs {bs}")
        ref = rows[base]
        out[bs] = {name: (ref / t if t and t > 0 and ref is not None else math.nan) for name, t in rows.items()}
    return out


def roofline_fraction(gbs, peak_gbs):
    """{bs: {format: GB/s / peak}}."""
    return {bs: {name: (v / peak_gbs if v is not None else math.nan) for name, v in rows.items()}
            for bs, rows in gbs.items()}


def formats(table):
    names = []
    for rows in table.values():
        for n in rows:
            if n not in names:
                names.append(n)
    return names


def render(table, title, fmt="{:.2f}"):
    """Plain-text table: one row per format, one column per block size."""
    sizes = list(table)
    names = formats(table)
    width = max([len(n) for n in names] + [len(title)])
    lines = [title.ljust(width) + "".join(f"{bs:>10}" for bs in sizes)]
    for n in names:
        cells = []
        for bs in sizes:
            v = table[bs].get(n)
            cells.append(f"{'-' if v is None or (isinstance(v, float) and math.isnan(v)) else fmt.format(v):>10}")
        lines.append(n.ljust(width) + "".join(cells))
    return "\n".join(lines)


def plot(table, ylabel, path, log=False):
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    fig, ax = plt.subplots(figsize=(10, 6))
    sizes = list(table)
    for n in formats(table):
        ys = [table[bs].get(n, math.nan) for bs in sizes]
        ax.plot(sizes, ys, marker="o", label=n)
    ax.set(xlabel="Block size", ylabel=ylabel)
    if log:
        ax.set_yscale("log")
    ax.legend(prop={"size": 7})
    fig.savefig(path, dpi=150, bbox_inches="tight")
    plt.close(fig)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("result", help="result.json ({bs: {format: seconds}})")
    ap.add_argument("--gbs", help="bandwidth file from bmfp_main --out-gbs")
    ap.add_argument("--peak", type=float, default=8000.0, help="HBM peak in GB/s (MI355X: 8000)")
    ap.add_argument("--base", default="CPU CSR", help="format the speed-up is relative to")
    ap.add_argument("--png-prefix", help="write <prefix>_speedup.png (and _roofline.png with --gbs)")
    args = ap.parse_args(argv)
    results = load(args.result)
    sp = speedup(results, args.base)
    print(render(sp, f"speed-up vs {args.base}"))
    roof = None
    if args.gbs:
        roof = roofline_fraction(load(args.gbs), args.peak)
        print()
        print(render(roof, f"fraction of {args.peak:.0f} GB/s", "{:.3f}"))
    if args.png_prefix:
        plot(sp, f"Speed-up vs {args.base}", f"{args.png_prefix}_speedup.png", log=True)
        if roof is not None:
            plot(roof, f"Fraction of the {args.peak:.0f} GB/s HBM roofline", f"{args.png_prefix}_roofline.png")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
