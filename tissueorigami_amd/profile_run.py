"""Profiling driver: one or more device-resident NR steps of a benchmark
workload, preceded by a plain device copy of known size.

The copy (the library's tf_copy_kernel over --copy-mib MiB: reads and writes
exactly that many bytes) is the calibration of the HBM byte counters: under
    rocprofv3 --pmc FETCH_SIZE -- python3 tissueorigami_amd/profile_run.py ...
(and again with WRITE_SIZE) its dispatches give the counter reading of a known
number of bytes, and the counter readings of the solver's kernels, scaled by
that factor, are their HBM traffic per launch.

    python3 tissueorigami_amd/profile_run.py --config c3 --steps 1 --precision mixed
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from tissueorigami_amd import nr  # noqa: E402
from tissueorigami_amd.workloads import make_workload  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="c3")
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--rtol", type=float, default=None)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--no-graphs", action="store_true",
                    help="direct kernel launches instead of captured hipGraph chunks")
    args = ap.parse_args()
    rtol = args.rtol if args.rtol else (1e-3 if args.precision == "mixed" else 1e-10)
    n = args.copy_mib * (1 << 20)
    gbs = nr.copy_bandwidth(args.copy_mib, 3)
    print(f"[profile_run] copy calibration: 4 x tf_copy_kernel over {n} bytes each way "
          f"({gbs:.0f} GB/s)", file=sys.stderr, flush=True)
    w = make_workload(args.config)
    m = w["mesh"]
    nn, ne = len(m["nodes"]), len(m["conn"])
    s = nr.NewtonRaphsonSolver(nn, m["conn"], m["refpos"])
    s.set_material(w["E"], w["nu"], w["eta"], w["ext"])
    s.set_growth_rates(w["rates"])
    s.set_surface_flags(w["place"], w["etype"], w["spans"])
    if args.no_graphs:
        s.set_graphs(False)
    n_blocks, n_slots = s.pattern_info()
    print(f"[profile_run] {args.config}: {ne} prisms, {nn} nodes, {n_blocks} blocks",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    _, stats = s.advance(args.steps, w["dt"], u0=m["nodes"].reshape(-1), precision=args.precision,
                         krylov_rtol=rtol, want_positions=False)
    wall = time.perf_counter() - t0
    for st in stats:
        print(f"[profile_run] step: NR {st['nr_iterations']} krylov {st['krylov_iterations']}",
              file=sys.stderr, flush=True)
    s.close()
    print(json.dumps(dict(config=args.config, precision=args.precision, krylov_rtol=rtol,
                          prisms=ne, nodes=nn, n_blocks=n_blocks, n_slots=n_slots,
                          copy_bytes_each_way=n, copy_gbs=gbs, wall_s=wall,
                          krylov_iterations=int(np.sum([st["krylov_iterations_total"]
                                                        for st in stats])))))


if __name__ == "__main__":
    main()
