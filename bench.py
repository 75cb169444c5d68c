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

Batched sweep (reported beside the headline line, --no-batch skips it): 64
independent elliptical tissues, prism counts log-uniform in 2 k - 40 k,
2-6 layers, E ~ U[100, 2000] Pa per tissue, per-tissue seeded growth fields;
all tissues in one handle (one launch per solver stage for the whole batch,
tile size chosen by the autotuner), mixed precision; the same tissues stepped
one after another in their own handles give the sequential rate.

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
from tissueorigami_amd.workloads import CONFIGS, make_workload  # noqa: E402

HBM_PEAK_GBS = 8000.0

def invert_batch(F):
    return np.linalg.inv(F).reshape(-1, 9)


def run_steps(solver, w, u0, Fg0, n_steps, precision, krylov_rtol=1e-10):
    """n_steps device-resident implicit steps (growth, viscosity surfaces and
    the NR solve all on the device) from positions u0 and growth tensor Fg0;
    returns the final state and accumulated statistics."""
    solver.growth_tensor(Fg0)
    tot = dict(nr_iterations=0, krylov_iterations=0, ms_element=0.0, ms_assembly=0.0,
               ms_krylov=0.0, converged=True, per_step=[], fp64_fallbacks=0, cg_fallbacks=0)
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
        tot["fp64_fallbacks"] += st["fp64_fallbacks"]
        tot["cg_fallbacks"] += st["cg_fallbacks"]
        tot["per_step"].append(dict(nr=st["nr_iterations"], krylov=st["krylov_iterations"],
                                    du=[float("%.3e" % d) for d in st["du_norms"]]))
        print(f"[bench]   {precision} step: NR {st['nr_iterations']} krylov {st['krylov_iterations']} "
              f"du {['%.2e' % d for d in st['du_norms']]}", file=sys.stderr, flush=True)
    u, _ = solver.advance(0, w["dt"])
    return u, solver.growth_tensor(), tot, wall


def spmv_v_bytes(n_blocks, n_nodes, s):
    """Algorithmic bytes of one tz_spmv_v launch (v = J ph fused with
    s = r - alpha v, sh = M^-1 s): per stored 3x3 block its 9 values and its
    column index; per node row ph (gathered, counted once), r and the 3x3
    block-Jacobi inverse read, v, s and sh written."""
    return n_blocks * (9 * s + 4) + n_nodes * s * (3 + 3 + 9 + 9)


def spmv_t_bytes(n_blocks, n_nodes, s):
    """tz_spmv_t (t = J sh fused with the dots (t, s), (t, t)): block values and
    columns; per node sh (gathered once) and s read, t written."""
    return n_blocks * (9 * s + 4) + n_nodes * s * (3 + 3 + 3)


def bicgstab_iteration_bytes(n_blocks, n_nodes, s):
    """Algorithmic bytes of one BiCGStab iteration of the tiled solver: the
    two fused SpMVs, tz_update_x (reads x, ph, sh, s, t, rh; writes x, r) and
    tz_update_p (reads r, p, v, w and the block-Jacobi inverse; writes p, ph)."""
    return (spmv_v_bytes(n_blocks, n_nodes, s) + spmv_t_bytes(n_blocks, n_nodes, s)
            + n_nodes * s * (8 * 3) + n_nodes * s * (6 * 3 + 9))


def assembly_bytes(n_prisms, n_blocks, n_nodes, s, visc, combined=False):
    """Algorithmic bytes of the element kernel and the slot-map assembly of one
    NR iteration (tangent precision s bytes): element inputs (connectivity,
    reference shape, InvFg, Lame/viscosity, positions at two time levels),
    18x18 tangents and 18-forces written, then each block gathered once with
    its source index, the operator (9 values per block) and the block-Jacobi
    inverse written, node forces read and the residual written.  With viscosity
    the FP64 path writes and gathers an elastic and a viscous tangent and an
    elastic and a viscous force; the mixed path (combined) one FP32 tangent
    and one force holding both."""
    nk = 2 if visc and not combined else 1
    nf = 1 if combined else 2
    elem = n_prisms * (24 + 144 + 72 + 24 + 2 * 144 + nk * 324 * s + nf * 144)
    asm = n_prisms * 36 * (4 + nk * 9 * s) + n_blocks * 9 * s + n_nodes * 9 * (s + 8)
    forces = n_prisms * nf * 144 + n_nodes * (3 * 8 * 4)
    return elem + asm + forces


def cg_spmv_bytes(n_blocks, n_nodes, s, fused=True):
    """tzc_spmv: block values and columns, and per node: with the search
    direction formed on the fly (fused: p = z + beta p_old, no coarse space) z
    and p_old gathered once each, p and q written; with the coarse space p
    gathered once and q written."""
    return n_blocks * (9 * s + 4) + n_nodes * s * ((3 + 3 + 3 + 3) if fused else (3 + 3))


def cg_iteration_bytes(n_blocks, n_nodes, s, n_coarse=0):
    """One CG iteration: tzc_spmv and tzc_update (reads x, p, r, q and the
    block-Jacobi inverse; writes x, r, z), and with a coarse space of n_coarse
    DOF: restriction (list index, offsets, mask, r gathered), the dense inverse
    (FP32, or FP64 for the FP64 CG) read once, and the search-direction update
    tzc_pupdate (aggregate, offsets, flags; z read, p read and written)."""
    b = (cg_spmv_bytes(n_blocks, n_nodes, s, fused=not n_coarse)
         + n_nodes * s * (6 + 3 + 6 + 3 + 9 + 3))
    if n_coarse:
        b += (n_nodes * (4 + 12 + 1 + 3 * s) + s * n_coarse * n_coarse
              + n_nodes * (4 + 12 + 3) + n_nodes * s * 9)
    return b


def coarse_setup_bytes(n_blocks, n_coarse, gemms, s=4):
    """Per assembly: A_c from the operator (values of s bytes, slot and row
    lists) and `gemms` dense FP64 products of the Newton-Schulz refinement."""
    return n_blocks * (9 * s + 4 + 4) + gemms * 3 * 8 * n_coarse * n_coarse


def profile_step(solver, w, u, Fg, precision, krylov_rtol):
    """One extra, untimed step with per-launch event timing of the SpMVs
    (direct launches, no graphs); returns the kernel times and the step's
    BiCGStab iterations (the launches that did work: a chunk can end with
    launches that return at once after convergence)."""
    solver.growth_tensor(Fg)
    solver.profile(True)
    _, st = solver.advance(1, w["dt"], u0=u, precision=precision, krylov_rtol=krylov_rtol,
                           want_positions=False)
    kt = solver.kernel_times()
    solver.profile(False)
    kt["krylov_iterations"] = st[0]["krylov_iterations_total"]
    return kt


def stream_copy_ceiling(mib=1024, reps=10):
    """Device copy bandwidth on this GPU (GB/s, bytes read + written): the
    best of the library's plain 16-byte-vector copy kernels over `mib` MiB."""
    return nr.copy_bandwidth(mib, reps)


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the rocprofv3
    FETCH_SIZE / WRITE_SIZE passes summarised in profiles/ (None if absent)."""
    path = os.path.join(ROOT, "profiles", "spmv_traffic_c3.json")
    try:
        d = json.load(open(path))
    except (OSError, ValueError):
        return None
    k = d.get("kernels", {}).get(kernel)
    if not k or d.get("workload") != workload:
        return None
    return k.get("hbm_bytes_per_launch")


def run_adjoint(krylov_rtol=1e-10):
    """C5: gradient of the apical shape loss after 20 steps with respect to
    the 64 growth-bump coefficients; cost of the backward sweep against the
    forward solve (device time and wall time)."""
    from tissueorigami_amd.workloads import c5_config, growth_steps
    c = c5_config()
    m = c["mesh"]
    s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
    s.set_material(c["E"], c["nu"], c["eta"], c["ext_visc"])
    inv, dinv = growth_steps(c["phi"] @ c["coef"], c["dt"], c["n_steps"])
    u0 = m["nodes"].reshape(-1)
    # warm-up pair (first launches, graph capture)
    s.adj_forward(u0, inv, c["dt"], c["place"], c["etype"], c["spans"], tol=1e-8, maxit=20,
                  krylov_rtol=krylov_rtol)
    s.adj_backward(dinv, np.ones(3 * len(m["nodes"])), krylov_rtol=krylov_rtol)
    t0 = time.perf_counter()
    u, stf = s.adj_forward(u0, inv, c["dt"], c["place"], c["etype"], c["spans"], tol=1e-8,
                           maxit=20, krylov_rtol=krylov_rtol)
    t1 = time.perf_counter()
    x = u.reshape(-1, 3)
    dL = np.zeros_like(x)
    dL[c["apical"]] = 2 * (x[c["apical"]] - c["target"])
    dL_dr, stb = s.adj_backward(dinv, dL.reshape(-1), krylov_rtol=krylov_rtol)
    t2 = time.perf_counter()
    s.close()
    return dict(workload="c5: c1 mesh, 64 Gaussian growth bumps, 20 steps of 10 s",
                forward_wall_s=t1 - t0, backward_wall_s=t2 - t1,
                forward_device_ms=stf["ms_total"], backward_device_ms=stb["ms_total"],
                cost_per_forward_wall=(t2 - t1) / (t1 - t0),
                cost_per_forward_device=stb["ms_total"] / stf["ms_total"],
                forward_nr_iterations=stf["nr_iterations"],
                backward_krylov_iterations=stb["krylov_iterations_total"])


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
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return dict(value=1.0 / t_nr, unit="NR iterations/s", cores=oracle.num_threads(),
                kind="port", cpu_model=cpu_model, nproc=os.cpu_count(),
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"),
                sample=(f"one NR iteration of the same workload on the host: BSR assembly "
                        f"{ta:.2f} s + {k} BiCGStab iterations {ts:.2f} s, extrapolated to "
                        f"{krylov_its_per_nr:.0f} BiCGStab iterations per NR iteration (measured on "
                        f"the GPU run); Pardiso substituted by the same block-Jacobi BiCGStab"))


def run_batch(steps, warmup, mixed_rtol, mixed_krylov="cg", seed=4, n=64, sequential_sample=8):
    """The batched sweep: returns a dict for the JSON line."""
    ts = tissue.batch_ellipses(n, seed=seed)
    dt = 60.0
    meshes = [(len(c["mesh"]["nodes"]), c["mesh"]["conn"], c["mesh"]["refpos"]) for c in ts]
    cat = lambda k: np.concatenate([c[k] for c in ts])
    t0 = time.perf_counter()
    b = nr.TissueBatch(meshes)
    b.set_material(cat("E"), cat("nu"), cat("eta"), cat("ext_visc"))
    b.set_growth_rates(np.concatenate([tissue.rates_xyz(c["rate"]) for c in ts]))
    b.set_surface_flags(cat("place"), cat("etype"), cat("spans"))
    b.set_krylov("mixed", mixed_krylov)
    setup_s = time.perf_counter() - t0
    u0 = np.concatenate([c["mesh"]["nodes"].reshape(-1) for c in ts])
    b.advance(max(warmup, 1), dt, u0=u0, precision="mixed", krylov_rtol=mixed_rtol,
              want_positions=False)
    spt, ms_it = b.autotune("mixed")
    print(f"[bench] batch: {n} tissues, {b.n_prisms} prisms, autotuned {spt} slices per tile "
          f"({ms_it:.3f} ms per solver iteration)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    _, stats = b.advance(steps, dt, None, 1e-8, 20, "mixed", mixed_rtol, want_positions=False)
    wall = time.perf_counter() - t0
    nr_its = sum(st["nr_iterations"] for step in stats for st in step)
    conv = all(st["converged"] for step in stats for st in step)
    fb = sum(st["fp64_fallbacks"] for step in stats for st in step)
    cgfb = sum(st["cg_fallbacks"] for step in stats for st in step)
    # launches are shared: the batch's wall time per step is that of its
    # slowest tissue; the rate counts every tissue's NR iterations
    # kernel evidence: one more step with every CG SpMV launch bracketed by HIP
    # events on the handle's stream (direct launches); the achieved rate counts
    # each tissue's algorithmic SpMV bytes once per CG iteration it ran
    b.profile(True)
    _, pst = b.advance(1, dt, None, 1e-8, 20, "mixed", mixed_rtol, want_positions=False)
    kt = b.kernel_times()
    b.profile(False)
    its = [st["krylov_iterations_total"] for st in pst[0]]
    blocks = [tissue.block_count(c["mesh"]["conn"], len(c["mesh"]["nodes"])) for c in ts]
    algo = sum(k * cg_spmv_bytes(nb, len(c["mesh"]["nodes"]), 4)
               for k, nb, c in zip(its, blocks, ts))
    spmv_s = kt["spmv_ms_fp32"] / 1e3
    kname = "tzc_spmv<float>" if mixed_krylov == "cg" else "tz_spmv_v<float>"
    profiled = dict(kernel=kname, launches=kt["spmv_launches_fp32"],
                    spmv_ms=kt["spmv_ms_fp32"], avg_launch_ms=kt["spmv_ms_fp32"] /
                    max(kt["spmv_launches_fp32"], 1), tissue_iterations=int(sum(its)),
                    max_tissue_iterations=int(max(its)), algorithmic_bytes=int(algo),
                    achieved=algo / spmv_s / 1e9 if spmv_s > 0 else None,
                    frac=(algo / spmv_s / 1e9 / HBM_PEAK_GBS) if spmv_s > 0 else None,
                    element_ms=kt["element_ms"], assembly_ms=kt["assembly_ms"],
                    assemblies=kt["assemblies"],
                    fallbacks=sum(st["cg_fallbacks"] + st["fp64_fallbacks"] for st in pst[0]))
    b.close()
    # sequential reference: a sample of the tissues, each in its own handle,
    # same steps, extrapolated by prism count
    idx = np.linspace(0, n - 1, sequential_sample).astype(int)
    seq_wall, seq_prisms = 0.0, 0
    for i in idx:
        c = ts[i]
        m = c["mesh"]
        s1 = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
        s1.set_material(c["E"], c["nu"], c["eta"], c["ext_visc"])
        s1.set_growth_rates(tissue.rates_xyz(c["rate"]))
        s1.set_surface_flags(c["place"], c["etype"], c["spans"])
        s1.set_krylov("mixed", mixed_krylov)
        s1.advance(max(warmup, 1), dt, u0=m["nodes"].reshape(-1), precision="mixed",
                   krylov_rtol=mixed_rtol, want_positions=False)
        t1 = time.perf_counter()
        s1.advance(steps, dt, None, 1e-8, 20, "mixed", mixed_rtol, want_positions=False)
        seq_wall += time.perf_counter() - t1
        seq_prisms += len(m["conn"])
        s1.close()
    seq_est = seq_wall * sum(len(c["mesh"]["conn"]) for c in ts) / seq_prisms
    out = dict(workload=f"{n} ellipses, prisms log-uniform 2k-40k, 2-6 layers, E~U[100,2000] Pa",
               tissues=n, prisms=int(sum(len(c["mesh"]["conn"]) for c in ts)),
               nodes=int(sum(len(c["mesh"]["nodes"]) for c in ts)), seed=seed, dt_s=dt,
               steps=steps, slices_per_tile=spt, nr_iterations=nr_its,
               nr_iterations_per_s=nr_its / wall, time_steps_per_s=steps / wall, wall_s=wall,
               converged=conv, fp64_fallbacks=fb, cg_fallbacks=cgfb, krylov_mixed=mixed_krylov,
               setup_s=setup_s, profiled_step=profiled,
               sequential_estimate_s=seq_est, batch_speedup_vs_sequential=seq_est / wall,
               sequential_sample=f"{sequential_sample} of the {n} tissues, scaled by prism count")
    print(f"[bench] batch: {steps} steps {wall:.2f} s, {nr_its} NR its; sequential estimate "
          f"{seq_est:.2f} s", file=sys.stderr, flush=True)
    return out


def run_bound(w, solver_steps, mixed_rtol, mixed_krylov):
    """C3 with the reference's circumference binding (x and y of every
    circumference node bound to the basal node of its column): mixed steps
    on the explicit bound operator N^T J N + I_slave."""
    m = w["mesh"]
    nn, ne, n2 = len(m["nodes"]), len(m["conn"]), m["n2d"]
    circ = np.where(m["at_circumference"][:n2] == 1)[0]
    pairs = np.array([[3 * (L * n2 + k) + d, 3 * k + d] for k in circ
                      for L in range(1, m["n_layers"] + 1) for d in (0, 1)], np.int32)
    t0 = time.perf_counter()
    s = nr.NewtonRaphsonSolver(nn, m["conn"], m["refpos"])
    s.set_material(w["E"], w["nu"], w["eta"], w["ext"], None, pairs)
    s.set_growth_rates(w["rates"])
    s.set_surface_flags(w["place"], w["etype"], w["spans"])
    s.set_krylov("mixed", mixed_krylov)
    setup_s = time.perf_counter() - t0
    _, n_slots_bound = s.pattern_info()
    u0 = m["nodes"].reshape(-1).copy()
    s.advance(1, w["dt"], u0=u0, precision="mixed", krylov_rtol=mixed_rtol, want_positions=False)
    t0 = time.perf_counter()
    u, stats = s.advance(solver_steps, w["dt"], None, 1e-8, 20, "mixed", mixed_rtol)
    wall = time.perf_counter() - t0
    d = u - u0
    s.close()
    nr_its = sum(st["nr_iterations"] for st in stats)
    return dict(workload="c3 with circumference binding (x, y of circumference nodes to the "
                         "basal node of their column)", pairs=int(len(pairs)),
                steps=solver_steps, nr_iterations=nr_its, nr_iterations_per_s=nr_its / wall,
                time_steps_per_s=solver_steps / wall,
                krylov_iterations=int(sum(st["krylov_iterations_total"] for st in stats)),
                converged=all(st["converged"] for st in stats), setup_s=setup_s,
                max_slave_master_gap=float(np.abs(d[pairs[:, 0]] - d[pairs[:, 1]]).max()))


def run_packing(steps, warmup, mixed_rtol, mixed_krylov, name="c2", gap=0.6):
    """Node-node packing at scale: two copies of a workload's ellipse facing
    apical to apical (workloads.contact_workload); mixed steps with the couples
    detected on the device at the start of every step, the FP64 steps from the
    same state beside them, and the cost of one detection."""
    from tissueorigami_amd.workloads import contact_workload
    w = contact_workload(name, gap)
    m = w["mesh"]
    nn, ne = len(m["nodes"]), len(m["conn"])
    t0 = time.perf_counter()
    s = nr.NewtonRaphsonSolver(nn, m["conn"], m["refpos"])
    s.set_material(w["E"], w["nu"], w["eta"], w["ext"], w["fixed"])
    s.set_growth_rates(w["rates"])
    s.set_surface_flags(w["place"], w["etype"], w["spans"])
    s.set_krylov("mixed", mixed_krylov)
    s.set_packing(m["node_placement"], capacity=2)
    setup_s = time.perf_counter() - t0
    blocks, slots = s.pattern_info()
    u0 = m["nodes"].reshape(-1).copy()
    Fg0 = np.tile(np.eye(3).reshape(9), (ne, 1))
    out = dict(workload=w["desc"], prisms=ne, nodes=nn, steps=steps, setup_s=setup_s,
               sell_slots=int(slots), blocks=int(blocks))
    for prec, rtol in (("mixed", mixed_rtol), ("fp64", 1e-10)):
        for _ in range(warmup):
            s.growth_tensor(Fg0)
            s.advance(1, w["dt"], u0=u0, precision=prec, krylov_rtol=rtol, want_positions=False)
        s.growth_tensor(Fg0)
        s.advance(0, w["dt"], u0=u0, want_positions=False)
        couples, nr_its, kr_its, conv, fb = [], 0, 0, True, [0, 0, 0]
        t0 = time.perf_counter()
        for k in range(steps):
            _, st = s.advance(1, w["dt"], None, 1e-8, 20, prec, rtol, want_positions=False)
            couples.append(s.packing_count())
            nr_its += st[0]["nr_iterations"]
            kr_its += st[0]["krylov_iterations_total"]
            conv &= st[0]["converged"]
            fb[0] += st[0]["cg_fallbacks"]
            fb[1] += st[0]["fp64_fallbacks"]
            fb[2] += st[0].get("gmres_fallbacks", 0)
        wall = time.perf_counter() - t0
        u, _ = s.advance(0, w["dt"])
        out[prec] = dict(nr_iterations=nr_its, krylov_iterations=kr_its, wall_s=wall,
                         nr_iterations_per_s=nr_its / wall, time_steps_per_s=steps / wall,
                         couples_per_step=couples, converged=bool(conv), cg_fallbacks=fb[0],
                         fp64_fallbacks=fb[1], gmres_fallbacks=fb[2])
        out[prec + "_u"] = u
        print(f"[bench] packing {prec}: couples {couples}, {nr_its} NR its, {wall:.2f} s",
              file=sys.stderr, flush=True)
    um, uf = out.pop("mixed_u"), out.pop("fp64_u")
    out["mixed_rel_pos_error_vs_fp64"] = float(np.abs(um - uf).max() / np.abs(uf).max())
    out["mixed_vs_fp64_speedup"] = (out["mixed"]["nr_iterations_per_s"] /
                                    out["fp64"]["nr_iterations_per_s"])
    # one detection at the step-start positions (device resident), wall time
    s.detect_packing(u0)
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        n = s.detect_packing()
    out["detect_ms"] = (time.perf_counter() - t0) / reps * 1e3
    out["couples_at_start"] = s.detect_packing(u0)
    s.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=list(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-binding", action="store_true")
    ap.add_argument("--no-packing", action="store_true")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every solver kernel directly (no hipGraph replay)")
    ap.add_argument("--mixed-rtol", type=float, default=1e-3,
                    help="relative residual of each FP32 linear solve (inexact Newton forcing "
                         "term; the NR loop with the FP64 residual refines)")
    ap.add_argument("--fp64-rtol", type=float, default=1e-10)
    ap.add_argument("--mixed-krylov", default="cg", choices=["cg", "bicgstab"],
                    help="Krylov method of the FP32 inner solves (FP64 uses BiCGStab)")
    ap.add_argument("--coarse-cell", type=float, default=-1.0,
                    help="aggregate cell size of the two-level CG preconditioner (um; negative: "
                         "automatic, about sqrt(blocks) coarse DOF; 0: block Jacobi only)")
    ap.add_argument("--coarse-bands", type=int, default=3,
                    help="bands through the thickness per aggregate cell of the coarse space")
    ap.add_argument("--coarse-every-iteration", action="store_true",
                    help="form the coarse inverse at every NR iteration (default: once per step)")
    ap.add_argument("--no-warm-start", action="store_true",
                    help="start every Krylov solve from zero (default: a step's first solve "
                         "starts from the displacement of the last steps)")
    ap.add_argument("--no-forcing-safeguard", action="store_true",
                    help="run every inexact-Newton solve to its forcing term (default: a solve "
                         "predicted to end the step is not oversolved)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(a tissue would shard by a 2D partition of its triangle mesh, every part keeping "
              "all its layers, with one halo exchange per SpMV; batched tissues whole, by "
              "bin packing)",
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
    if args.no_graphs:
        solver.set_graphs(False)
    solver.set_krylov("mixed", args.mixed_krylov)
    solver.set_coarse(args.coarse_cell)
    solver.set_coarse_bands(args.coarse_bands)
    solver.set_coarse_lag(not args.coarse_every_iteration)
    solver.set_warm_start(not args.no_warm_start)
    solver.set_forcing_safeguard(not args.no_forcing_safeguard)
    n_blocks, n_slots = solver.pattern_info()
    print(f"[bench] setup {setup_s:.2f} s, {n_blocks} blocks, {n_slots} SELL slots",
          file=sys.stderr, flush=True)
    u0 = m["nodes"].reshape(-1).copy()
    Fg0 = np.tile(np.eye(3).reshape(9), (ne, 1))
    # warm-up (mixed) from the initial state, then the timed steps continue from it
    u_w, Fg_w, _, wall_w = run_steps(solver, w, u0, Fg0, args.warmup, "mixed",
                                     krylov_rtol=args.mixed_rtol)
    print(f"[bench] warmup {args.warmup} steps {wall_w:.2f} s", file=sys.stderr, flush=True)
    spt, ms_it = solver.autotune("mixed")
    print(f"[bench] autotuned tile: {spt} slices ({ms_it:.3f} ms per FP32 solver iteration)",
          file=sys.stderr, flush=True)
    u_m, _, tm, wall_m = run_steps(solver, w, u_w, Fg_w, args.steps, "mixed",
                                   krylov_rtol=args.mixed_rtol)
    print(f"[bench] mixed: {tm['nr_iterations']} NR its, {tm['krylov_iterations']} Krylov its, "
          f"{wall_m:.2f} s", file=sys.stderr, flush=True)
    n_agg, coarse_on = solver.coarse_info()  # of the mixed steps (an FP64 step turns it off)
    res = dict(metric="Newton steps/s at 1.6M prisms, mixed vs FP64 @1e-8 rel err; adjoint cost/fwd",
               unit="NR iterations/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
               higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32/f64 mixed",
               data="synthetic")
    res["value"] = tm["nr_iterations"] / wall_m
    res["ms_per_step"] = 1000.0 * wall_m / args.steps
    fp64 = None
    td = tc = None
    if not args.no_fp64:
        u_d, _, td, wall_d = run_steps(solver, w, u_w, Fg_w, args.steps, "fp64",
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
        # the same two-level CG and forcing term as the mixed step, in FP64
        # (FP64 operator, residual, vectors and coarse inverse): what precision
        # alone buys
        solver.set_krylov("fp64", "cg")
        u_c, _, tc, wall_c = run_steps(solver, w, u_w, Fg_w, args.steps, "fp64",
                                       krylov_rtol=args.mixed_rtol)
        solver.set_krylov("fp64", "bicgstab")
        print(f"[bench] fp64 two-level CG: {tc['nr_iterations']} NR its, "
              f"{tc['krylov_iterations']} Krylov its, {wall_c:.2f} s", file=sys.stderr, flush=True)
        res["fp64_cg"] = dict(value=tc["nr_iterations"] / wall_c, nr_iterations=tc["nr_iterations"],
                              krylov_iterations=tc["krylov_iterations"], wall_s=wall_c,
                              ms_element=tc["ms_element"], ms_assembly=tc["ms_assembly"],
                              ms_krylov=tc["ms_krylov"], converged=tc["converged"],
                              krylov_rtol=args.mixed_rtol,
                              rel_pos_error_vs_fp64=float(np.abs(u_c - u_d).max()
                                                          / np.abs(u_d).max()))
        res["mixed_vs_fp64_cg_time_steps_speedup"] = wall_c / wall_m
        res["mixed_vs_best_fp64_time_steps_speedup"] = min(wall_c, wall_d) / wall_m
    res["mixed"] = dict(nr_iterations=tm["nr_iterations"], krylov_iterations=tm["krylov_iterations"],
                        wall_s=wall_m, ms_element=tm["ms_element"], ms_assembly=tm["ms_assembly"],
                        ms_krylov=tm["ms_krylov"], converged=tm["converged"],
                        fp64_fallbacks=tm["fp64_fallbacks"], cg_fallbacks=tm["cg_fallbacks"],
                        time_steps_per_s=args.steps / wall_m, per_step=tm["per_step"],
                        coarse=dict(aggregates=n_agg, coarse_dof=6 * n_agg, active=coarse_on))
    res["adjoint_cost_per_forward"] = None
    res["config"] = dict(workload=f"{args.config}: {w['cfg']['desc']}", prisms=ne, nodes=nn,
                         dof=3 * nn, dt_s=w["dt"], nr_tol=1e-8, krylov_rtol_fp64=args.fp64_rtol,
                         krylov_rtol_mixed=args.mixed_rtol,
                         precond_fp64="node-block Jacobi",
                         precond_mixed=("node-block Jacobi + coarse correction (rigid-body modes "
                                        f"of {n_agg} aggregates: (x, y) cells x "
                                        f"{args.coarse_bands} thickness bands; inverse formed "
                                        + ("every NR iteration)" if args.coarse_every_iteration
                                           else "once per time step)")
                                        if coarse_on else "node-block Jacobi"),
                         krylov_mixed=args.mixed_krylov, krylov_fp64="bicgstab",
                         solve_policies=dict(
                             warm_start=("off" if args.no_warm_start else
                                         "every leg: a step's first solve starts from the "
                                         "displacement the last two steps extrapolate to"),
                             forcing_safeguard=("off" if args.no_forcing_safeguard else
                                                "legs with rtol >= 1e-6 (mixed, fp64_cg): a solve "
                                                "predicted to end the step runs to min(0.1, "
                                                "max(rtol, 0.5 tol / e))")),
                         growth_seed=w["seed"], setup_s=setup_s, slices_per_tile=spt)
    # roofline of the dominant kernel: the FP32 SpMV v = J ph of the mixed
    # solve (tz_spmv_v<float>), timed with HIP events on the solver's stream
    # in one extra profiled step from the warm state
    kt_m = profile_step(solver, w, u_w, Fg_w, "mixed", args.mixed_rtol)
    res["kernel_times_mixed_step"] = kt_m
    workload = res["config"]["workload"]
    if kt_m["spmv_launches_fp32"] > 0:
        avg_s = kt_m["spmv_ms_fp32"] / kt_m["krylov_iterations"] / 1000.0
        if args.mixed_krylov == "cg":
            kname, b = "tzc_spmv<float>", cg_spmv_bytes(n_blocks, nn, 4, fused=not coarse_on)
        else:
            kname, b = "tz_spmv_v<float>", spmv_v_bytes(n_blocks, nn, 4)
        ach = b / avg_s / 1e9
        res["roofline"] = dict(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                               frac=ach / HBM_PEAK_GBS, traffic=measured_traffic(kname, workload),
                               kernel=kname, algorithmic_bytes=b, avg_launch_ms=avg_s * 1e3,
                               launches_with_work=kt_m["krylov_iterations"],
                               launches=kt_m["spmv_launches_fp32"])
    if kt_m["spmv_t_launches_fp32"] > 0:
        avg_t = kt_m["spmv_t_ms_fp32"] / kt_m["krylov_iterations"] / 1000.0
        bt = spmv_t_bytes(n_blocks, nn, 4)
        res["roofline_spmv_t_fp32"] = dict(kernel="tz_spmv_t<float>", achieved=bt / avg_t / 1e9,
                                           frac=bt / avg_t / 1e9 / HBM_PEAK_GBS,
                                           avg_launch_ms=avg_t * 1e3, algorithmic_bytes=bt,
                                           traffic=measured_traffic("tz_spmv_t<float>", workload))
    if fp64:
        kt_d = profile_step(solver, w, u_w, Fg_w, "fp64", args.fp64_rtol)
        if kt_d["spmv_launches_fp64"] > 0:
            avg_s = kt_d["spmv_ms_fp64"] / kt_d["krylov_iterations"] / 1000.0
            b = spmv_v_bytes(n_blocks, nn, 8)
            res["roofline_fp64_spmv"] = dict(kernel="tz_spmv_v<double>", achieved=b / avg_s / 1e9,
                                             frac=b / avg_s / 1e9 / HBM_PEAK_GBS,
                                             avg_launch_ms=avg_s * 1e3, algorithmic_bytes=b)
    # whole NR iteration: element kernel + assembly (+ coarse setup) + k Krylov iterations
    visc = bool(np.any(w["eta"] != 0))
    for key, t, sz, meth in (("mixed", tm, 4, args.mixed_krylov),
                             ("fp64", td if fp64 else None, 8, "bicgstab"),
                             ("fp64_cg", tc if fp64 else None, 8, "cg")):
        if t is None:
            continue
        k_per = t["krylov_iterations"] / t["nr_iterations"]
        nco = 6 * n_agg if (key != "fp64" and coarse_on and meth == "cg") else 0
        if meth == "cg":
            it_bytes = cg_iteration_bytes(n_blocks, nn, sz, nco)
        else:
            it_bytes = bicgstab_iteration_bytes(n_blocks, nn, sz)
        bnr = (assembly_bytes(ne, n_blocks, nn, sz, visc, combined=key == "mixed")
               + k_per * it_bytes)
        if nco:
            # two Newton-Schulz steps (4 GEMMs) and a residual product per coarse setup:
            # one per time step (lagged) or one per assembly
            per_nr = 1.0 if args.coarse_every_iteration else args.steps / t["nr_iterations"]
            bnr += coarse_setup_bytes(n_blocks, nco, 5, sz) * per_nr
        secs = (t["ms_element"] + t["ms_assembly"] + t["ms_krylov"]) / 1e3 / t["nr_iterations"]
        res.setdefault("roofline_nr_iteration", {})[key] = dict(
            algorithmic_bytes=bnr, krylov_iterations=k_per, device_s=secs,
            achieved=bnr / secs / 1e9, frac=bnr / secs / 1e9 / HBM_PEAK_GBS)
    ceiling = stream_copy_ceiling()
    res["stream_copy_gbs"] = ceiling
    if "roofline" in res:
        res["roofline"]["frac_of_stream_copy"] = res["roofline"]["achieved"] / ceiling
    print(f"[bench] STREAM copy ceiling {ceiling:.0f} GB/s", file=sys.stderr, flush=True)
    if not args.no_adjoint:
        adj = run_adjoint()
        res["adjoint"] = adj
        res["adjoint_cost_per_forward"] = adj["cost_per_forward_device"]
        print(f"[bench] adjoint: backward/forward {adj['cost_per_forward_device']:.2f} (device), "
              f"{adj['cost_per_forward_wall']:.2f} (wall)", file=sys.stderr, flush=True)
    if not args.no_binding and args.config == "c3":
        res["bound"] = run_bound(w, args.steps, args.mixed_rtol, args.mixed_krylov)
        print(f"[bench] bound: {res['bound']['nr_iterations_per_s']:.2f} NR its/s, "
              f"{res['bound']['time_steps_per_s']:.2f} steps/s", file=sys.stderr, flush=True)
    if not args.no_packing and args.config == "c3":
        res["packing"] = run_packing(args.steps, args.warmup, args.mixed_rtol, args.mixed_krylov)
        print(f"[bench] packing: {res['packing']['mixed']['nr_iterations_per_s']:.2f} NR its/s "
              f"mixed, detection {res['packing']['detect_ms']:.2f} ms",
              file=sys.stderr, flush=True)
    if not args.no_batch and args.config == "c3":
        res["batch"] = run_batch(args.steps, args.warmup, args.mixed_rtol, args.mixed_krylov)
    if not args.no_cpu_baseline:
        k_per_nr = (fp64 or {}).get("krylov_iterations", tm["krylov_iterations"]) / max(
            1, (fp64 or {}).get("nr_iterations", tm["nr_iterations"]))
        res["cpu_baseline"] = cpu_baseline(w, u_w, Fg_w, k_per_nr)
    solver.close()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
