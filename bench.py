"""Benchmark of the implicit Newton-Raphson step on prism meshes (one MI355X).

Workload (default): large ellipse, semi-axes 200 x 140 um, triangle side 1 um,
8 prism layers over 30 um (1.62 M prisms, 0.91 M nodes, 2.74 M DOF);
E = 1000 Pa in the apical layer and 100 Pa below, nu = 0.3, internal viscosity
10 Pa s; external drag 16000 (apical surface) / 10 (basal surface); in-plane
growth from a seeded Gaussian random field (mean 0.02/h, sd 0.01/h,
correlation length 15 um); dt = 60 s.

A bench "step" is one implicit time step (a full NR solve to |du|_2 <= 1e-8,
as Simulation::updateStepNR runs it); the metric counts NR iterations
("Newton steps") per second over the timed time steps.  The mixed-precision
step is timed first, then the plain FP64 step from the same state with the
same growth; the line reports the mixed rate as `value`, the FP64 rate, their
ratio and the relative position error of mixed against FP64.

    python bench.py [--steps K] [--warmup W] [--config c3|c2|c1] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from tissueorigami_amd import nr, tissue  # noqa: E402

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    # (kind, geometry, layers, thickness, dt, E_apical, E_other, nu, eta, growth)
    "c1": dict(kind="rectangle", lx=10.0, ly=5.0, side=1.0, layers=2, thickness=2.0, dt=10.0,
               E_ap=1000.0, E_other=1000.0, nu=0.3, eta=0.0, grf=False,
               desc="rectangle 10x5 um, side 1 um, 2 layers"),
    "c2": dict(kind="ellipse", a=100.0, b=70.0, side=2.0, layers=4, thickness=20.0, dt=60.0,
               E_ap=1000.0, E_other=100.0, nu=0.3, eta=10.0, grf=True,
               desc="ellipse 100x70 um, side 2 um, 4 layers over 20 um"),
    "c3": dict(kind="ellipse", a=200.0, b=140.0, side=1.0, layers=8, thickness=30.0, dt=60.0,
               E_ap=1000.0, E_other=100.0, nu=0.3, eta=10.0, grf=True,
               desc="ellipse 200x140 um, side 1 um, 8 layers over 30 um"),
}


def make_workload(name, seed=2024):
    c = CONFIGS[name]
    if c["kind"] == "rectangle":
        m = tissue.rectangle(c["lx"], c["ly"], c["side"], c["thickness"], c["layers"])
    else:
        m = tissue.ellipse(c["a"], c["b"], c["side"], c["thickness"], c["layers"])
    ne, nn = len(m["conn"]), len(m["nodes"])
    top = m["layer"] == m["n_layers"] - 1
    E = np.where(top, c["E_ap"], c["E_other"])
    nu = np.full(ne, c["nu"])
    eta = np.full(ne, c["eta"])
    ext = np.zeros((nn, 3))
    ext[m["node_placement"] == 1] = 16000.0
    ext[m["node_placement"] == 0] = 10.0
    if c["grf"]:
        cent = m["nodes"][m["conn"]].mean(1)[:, :2]
        rate = tissue.gaussian_random_field(cent, 0.02, 0.01, 15.0, seed)
    else:
        rate = np.where(m["layer"] == 0, 0.01, 0.0)
    place, spans = tissue.element_placement(m["conn"], m["node_placement"])
    etype = tissue.element_type(m["conn"], m["node_type"])
    inc = tissue.growth_increment(rate, c["dt"])
    rates = np.zeros((ne, 3))
    rates[:, 0] = rate / 3600.0
    rates[:, 1] = rate / 3600.0
    return dict(mesh=m, E=E, nu=nu, eta=eta, ext=ext, inc=inc, rates=rates, dt=c["dt"],
                place=place, spans=spans, etype=etype, cfg=c, seed=seed)


def invert_batch(F):
    return np.linalg.inv(F).reshape(-1, 9)


def run_steps(solver, w, u0, Fg0, n_steps, precision, profile=False, krylov_rtol=1e-10):
    """n_steps device-resident implicit steps (growth, viscosity surfaces and
    the NR solve all on the device) from positions u0 and growth tensor Fg0;
    returns the final state and accumulated statistics."""
    solver.growth_tensor(Fg0)
    tot = dict(nr_iterations=0, krylov_iterations=0, ms_element=0.0, ms_assembly=0.0,
               ms_krylov=0.0, converged=True, per_step=[])
    if profile:
        solver.profile(True)
    # positions and growth state are uploaded before the timed region
    solver.advance(0, w["dt"], u0=u0, want_positions=False)
    t0 = time.perf_counter()
    _, stats = solver.advance(n_steps, w["dt"], None, 1e-8, 20, precision, krylov_rtol,
                              want_positions=False)
    wall = time.perf_counter() - t0
    for st in stats:
        tot["nr_iterations"] += st["nr_iterations"]
        tot["krylov_iterations"] += st["krylov_iterations_total"]
        for k in ("ms_element", "ms_assembly", "ms_krylov"):
            tot[k] += st[k]
        tot["converged"] &= st["converged"]
        tot["per_step"].append(dict(nr=st["nr_iterations"], krylov=st["krylov_iterations"],
                                    du=[float("%.3e" % d) for d in st["du_norms"]]))
        print(f"[bench]   {precision} step: NR {st['nr_iterations']} krylov {st['krylov_iterations']} "
              f"du {['%.2e' % d for d in st['du_norms']]}", file=sys.stderr, flush=True)
    kt = solver.kernel_times() if profile else None
    if profile:
        solver.profile(False)
    u, _ = solver.advance(0, w["dt"])
    return u, solver.growth_tensor(), tot, wall, kt


def spmv_bytes(n_blocks, n_nodes, real_bytes):
    # algorithmic bytes of one SpMV: block values + block column index per
    # stored block, one row pointer per node row, x read and y written once
    return n_blocks * (9 * real_bytes + 4) + n_nodes * 4 + 2 * 3 * n_nodes * real_bytes


def cpu_baseline(w, u, Fg, krylov_its_per_nr, budget_s=25.0):
    """The CPU oracle's sparse path (OpenMP BSR assembly + block-Jacobi
    BiCGStab, the reference's algorithm with Pardiso substituted) on the host
    cores, timed on a bounded sample: one assembly of the full workload and a
    capped number of BiCGStab iterations, extrapolated to the measured
    iterations per NR iteration."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    m = w["mesh"]
    lam, mu = oracle.lame(w["E"], w["nu"])
    A = oracle.visc_surfaces(len(m["nodes"]), m["conn"], u, w["place"], w["etype"], w["spans"])
    Fg1 = w["inc"] @ Fg.reshape(-1, 3, 3)
    P = oracle.Problem(m["conn"], m["refpos"], invert_batch(Fg1), lam, mu, w["eta"], w["ext"], A)
    # probe one iteration-capped solve to size the sample
    t0 = time.perf_counter()
    du, k, ta, ts = P.nr_iteration_sparse(u, u, w["dt"], 1e-30, 2)
    probe = time.perf_counter() - t0
    per_it = max(ts / max(k, 1), 1e-6)
    cap = int(max(2, min(400, (budget_s - probe) / per_it)))
    du, k, ta, ts = P.nr_iteration_sparse(u, u, w["dt"], 1e-30, cap)
    t_nr = ta + ts / max(k, 1) * krylov_its_per_nr
    return dict(value=1.0 / t_nr, unit="NR iterations/s", cores=oracle.num_threads(),
                kind="port",
                sample=(f"one NR iteration of the same workload on the host: BSR assembly "
                        f"{ta:.2f} s + {k} BiCGStab iterations {ts:.2f} s, extrapolated to "
                        f"{krylov_its_per_nr:.0f} BiCGStab iterations per NR iteration (measured on "
                        f"the GPU run); Pardiso substituted by the same block-Jacobi BiCGStab"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=list(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--mixed-rtol", type=float, default=1e-5,
                    help="relative residual of each FP32 linear solve (the NR loop refines)")
    ap.add_argument("--fp64-rtol", type=float, default=1e-10)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(a tissue would shard by node partition with a halo exchange; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)
    w = make_workload(args.config)
    m = w["mesh"]
    nn, ne = len(m["nodes"]), len(m["conn"])
    print(f"[bench] {args.config}: {w['cfg']['desc']}: {ne} prisms, {nn} nodes, {3*nn} DOF",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    solver = nr.NewtonRaphsonSolver(nn, m["conn"], m["refpos"])
    solver.set_material(w["E"], w["nu"], w["eta"], w["ext"])
    solver.set_growth_rates(w["rates"])
    solver.set_surface_flags(w["place"], w["etype"], w["spans"])
    setup_s = time.perf_counter() - t0
    n_blocks, n_slots = solver.pattern_info()
    print(f"[bench] setup {setup_s:.2f} s, {n_blocks} blocks, {n_slots} SELL slots",
          file=sys.stderr, flush=True)
    u0 = m["nodes"].reshape(-1).copy()
    Fg0 = np.tile(np.eye(3).reshape(9), (ne, 1))
    # warm-up (mixed) from the initial state, then the timed steps continue from it
    u_w, Fg_w, _, wall_w, _ = run_steps(solver, w, u0, Fg0, args.warmup, "mixed",
                                        krylov_rtol=args.mixed_rtol)
    print(f"[bench] warmup {args.warmup} steps {wall_w:.2f} s", file=sys.stderr, flush=True)
    u_m, _, tm, wall_m, kt_m = run_steps(solver, w, u_w, Fg_w, args.steps, "mixed", profile=True,
                                         krylov_rtol=args.mixed_rtol)
    print(f"[bench] mixed: {tm['nr_iterations']} NR its, {tm['krylov_iterations']} Krylov its, "
          f"{wall_m:.2f} s", file=sys.stderr, flush=True)
    res = dict(metric="Newton steps/s at 1.6M prisms, mixed vs FP64 @1e-8 rel err; adjoint cost/fwd",
               unit="NR iterations/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
               higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32/f64 mixed",
               data="synthetic")
    res["value"] = tm["nr_iterations"] / wall_m
    res["ms_per_step"] = 1000.0 * wall_m / args.steps
    fp64 = None
    if not args.no_fp64:
        u_d, _, td, wall_d, kt_d = run_steps(solver, w, u_w, Fg_w, args.steps, "fp64", profile=True,
                                             krylov_rtol=args.fp64_rtol)
        print(f"[bench] fp64: {td['nr_iterations']} NR its, {td['krylov_iterations']} Krylov its, "
              f"{wall_d:.2f} s", file=sys.stderr, flush=True)
        err = float(np.abs(u_m - u_d).max() / np.abs(u_d).max())
        fp64 = dict(value=td["nr_iterations"] / wall_d, nr_iterations=td["nr_iterations"],
                    krylov_iterations=td["krylov_iterations"], wall_s=wall_d,
                    ms_element=td["ms_element"], ms_assembly=td["ms_assembly"],
                    ms_krylov=td["ms_krylov"], converged=td["converged"])
        res["fp64"] = fp64
        res["mixed_vs_fp64_speedup"] = res["value"] / fp64["value"]
        res["mixed_vs_fp64_time_steps_speedup"] = wall_d / wall_m
        res["mixed_rel_pos_error_vs_fp64"] = err
    res["mixed"] = dict(nr_iterations=tm["nr_iterations"], krylov_iterations=tm["krylov_iterations"],
                        wall_s=wall_m, ms_element=tm["ms_element"], ms_assembly=tm["ms_assembly"],
                        ms_krylov=tm["ms_krylov"], converged=tm["converged"],
                        time_steps_per_s=args.steps / wall_m, per_step=tm["per_step"])
    res["adjoint_cost_per_forward"] = None
    res["config"] = dict(workload=f"{args.config}: {w['cfg']['desc']}", prisms=ne, nodes=nn,
                         dof=3 * nn, dt_s=w["dt"], nr_tol=1e-8, krylov_rtol_fp64=args.fp64_rtol,
                         krylov_rtol_mixed=args.mixed_rtol, precond="node-block Jacobi",
                         krylov="BiCGStab (reduced synchronisation)",
                         growth_seed=w["seed"], setup_s=setup_s)
    # roofline of the dominant kernel: the FP32 SpMV of the mixed solve
    if kt_m and kt_m["spmv_launches_fp32"] > 0:
        avg_s = kt_m["spmv_ms_fp32"] / kt_m["spmv_launches_fp32"] / 1000.0
        b = spmv_bytes(n_blocks, nn, 4)
        ach = b / avg_s / 1e9
        res["roofline"] = dict(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                               frac=ach / HBM_PEAK_GBS, traffic=None, kernel="spmv_kernel<float>",
                               algorithmic_bytes=b, avg_launch_ms=avg_s * 1e3,
                               launches=kt_m["spmv_launches_fp32"])
    if fp64 and kt_d and kt_d["spmv_launches_fp64"] > 0:
        avg_s = kt_d["spmv_ms_fp64"] / kt_d["spmv_launches_fp64"] / 1000.0
        b = spmv_bytes(n_blocks, nn, 8)
        res["roofline_fp64_spmv"] = dict(achieved=b / avg_s / 1e9, frac=b / avg_s / 1e9 / HBM_PEAK_GBS,
                                         avg_launch_ms=avg_s * 1e3, algorithmic_bytes=b)
    if not args.no_cpu_baseline:
        k_per_nr = (fp64 or {}).get("krylov_iterations", tm["krylov_iterations"]) / max(
            1, (fp64 or {}).get("nr_iterations", tm["nr_iterations"]))
        res["cpu_baseline"] = cpu_baseline(w, u_w, Fg_w, k_per_nr)
    solver.close()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
