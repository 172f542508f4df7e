"""Regenerate the library's built-in launch shapes for runtime-size grids.

Runs the autotuner (bri17_autotune: middle-axis split, pencil tile width and
radix orders of the runtime-size kernels, timed on the device) on the cubic
fp32 grids of the serving workload that have no fixed-size kernels, and
writes bri17_amd/csrc/tuned_defaults.inc, the table that operators created
without an explicit tuning fall back to (after tunings registered with
bri17_tuning_cache_set).  Run on the target GPU, then rebuild the library:

    python -m bri17_amd.tune_defaults [--out PATH] [--reps N]
"""
import argparse
import json
import os
import sys
import time

SIDES = (32, 36, 48, 54, 72, 81, 96, 108, 144, 162, 192, 216, 243)
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "csrc", "tuned_defaults.inc"))
    ap.add_argument("--json", default=None, help="also write the tunings and timings as JSON")
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    from bri17_amd import autotune

    rows, report = [], {}
    for n in SIDES:
        shape = (n, n, n)
        autotune.clear()
        t0 = time.time()
        best, ms, tried = autotune.tune(shape, "float32", reps=args.reps, register_result=False)
        print("%d^3 f32: %.4f ms, %d candidates, %.1f s: %s" % (n, ms, tried, time.time() - t0, best),
              flush=True)
        report["3:%dx%dx%d:f32" % (n, n, n)] = dict(best, ms=ms, candidates=tried)
        if not best["force_generic"]:
            continue  # the planner's own choice won: nothing to record
        rad = [(r + [0] * 8)[:8] for r in best["radices"]]
        rows.append("  {3, {%d, %d, %d}, 0, %d, %d, {{%s}, {%s}, {%s}}},  // %.4f ms" % (
            n, n, n, best["plane_split"], best["pencil_width"], ", ".join(map(str, rad[0])),
            ", ".join(map(str, rad[1])), ", ".join(map(str, rad[2])), ms))
    with open(args.out, "w") as f:
        f.write("// Written by bri17_amd/tune_defaults.py (autotuner on an MI355X, fp32, cubic\n"
                "// grids of the serving workload): {dim, {N0, N1, N2}, dtype, plane_split,\n"
                "// pencil_width, radix orders of the row / middle-axis / axis-0 transforms}.\n")
        f.write("\n".join(rows) + "\n")
    if args.json:
        with open(args.json, "w") as f:
            json.dump(report, f, indent=1)
    print("wrote", args.out, flush=True)


if __name__ == "__main__":
    main()
