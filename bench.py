#!/usr/bin/env python
"""Benchmark of the MI355X residual-apply path on one GPU.

Workload (the headline configuration): FP64 residual apply of 3D isotropic linear elasticity
on P2 tetrahedra, a 64^3 box split 6 tets per cube (1,572,864 cells, 2,146,689 P2 nodes,
6,440,067 dofs), 14-point degree-5 quadrature, Vs ~ lognormal(ln 3000, 0.2) per cell,
Vp = 1.732 Vs, density 2500, displacement = sum of 8 seeded sinusoids of amplitude 1e-3 m.
One step = one residual apply (INSERT mode: zeroed local residual + assembly) with the
solution already resident in HBM.  value = dofs / time per apply (GDoF/s).

The timed kernel is the library's default path: the quadrature sum of the 14-point rule
evaluated through the linear structure of the pointwise kernel (vertex gradients and the rule's
P1 moment matrix; identical to the point-by-point sum up to rounding).  ``extra`` also reports
the point-by-point evaluation of the 14 quadrature points, the FP32 apply, the FP64 Jacobian
action, and the Q1 hex elasticity apply of the 32^3 cube (6 km, u ~ U[-1e-3, 1e-3] m with the
shear Dirichlet values inserted) and of a 192^3 cube (working set beyond the 256 MB Infinity
Cache).

The CPU baseline is the oracle (the C restatement of the reference's cell loop with the
pointwise kernel called through a function pointer) timed on this host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEED = 20261202


def p2_bytes(nn, nv, nc, s):
    """Algorithmic bytes of one P2 apply: u read + r write (s bytes per dof), vertex coordinates
    (FP64), closure (10 x int32 per cell), lambda/mu per cell (s bytes each)."""
    return 2 * s * 3 * nn + 24 * nv + 40 * nc + 2 * s * nc


def q1_bytes(nn, nc, s):
    """Algorithmic bytes of one Q1 apply: u read + r write + coordinates (FP64), closure (8 x
    int32 per cell), lambda/mu per cell."""
    return 2 * s * 3 * nn + 24 * nn + 32 * nc + 2 * s * nc




def time_applies(fn, steps, warmup, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def apply_percentiles(fn, steps, torch):
    """Per-apply wall times (torch events on the handle's stream, each apply bracketed): median
    and p99 in ms."""
    import numpy as np
    ts = []
    for _ in range(steps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        ts.append(a.elapsed_time(b))
    return float(np.median(ts)), float(np.percentile(ts, 99))


def graph_apply_ms(dm, fn, torch, per_graph=20, replays=10):
    """Per-apply time of `fn` captured `per_graph` times into one HIP graph (torch.cuda.CUDAGraph on
    a side stream the handle launches on), replayed `replays` times; the handle returns to the
    current stream afterwards."""
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    dm.set_stream(s)
    try:
        with torch.cuda.stream(s):
            fn()
        s.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            for _ in range(per_graph):
                fn()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(s):  # replay() launches on the current stream
            g.replay()
            s.synchronize()
            a.record(s)
            for _ in range(replays):
                g.replay()
            b.record(s)
        b.synchronize()
        return a.elapsed_time(b) / (per_graph * replays)
    finally:
        torch.cuda.synchronize()
        dm.set_stream(torch.cuda.current_stream())


def cpu_description():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def kernel_avg_ms(dm, fn, steps, torch, name="plx:apply_tiles"):
    dm.reset_event_stats()
    dm.set_event_logging(True)
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    n, ms = dm.event_stats(name)
    dm.set_event_logging(False)
    return ms / max(n, 1), n


def setup(mesh, aux, sizes):
    from pylith_amd.dm import DM
    dm = DM(mesh)
    dm.add_residual(0, 0, 0, 1)
    dm.add_jacobian(0, 0, 0, 0, 0, 2)
    dm.set_auxiliary(0, 0, aux, sizes)
    return dm


def cpu_baseline(problem, u, colors, budget_s=10.0, nthreads=None):
    """Oracle residual on the host: colored OpenMP over all usable cores (or nthreads) for
    ~budget_s (at least one apply)."""
    import numpy as np
    nthreads = nthreads or min(16, os.cpu_count() or 1)
    out = None
    n = 0
    t0 = time.perf_counter()
    while True:
        out = problem.residual_colored(u, 1, colors, nthreads=nthreads, out=np.zeros_like(u))
        n += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = (time.perf_counter() - t0) / n
    return out, dt, n, nthreads


def gm_bytes_update(nn, nc):
    """Generalized Maxwell state update per launch: state read + write (24 doubles per point x 8
    points), next-step state stress (6 x 8 doubles), per-cell material (8 doubles), closure (8
    int32), u and coordinates at the nodes (FP64)."""
    return nc * (2 * 24 * 8 * 8 + 6 * 8 * 8 + 8 * 8 + 32) + 48 * nn


def bench_c2_solve(torch, mg, L):
    """C2 numerics: P2 64^3 x 6 with the shear Dirichlet values, solved to 1e-8 relative residual
    by FGMRES with an FP32 PCG preconditioner (FP64 true residual) and by plain FP64 PCG, each
    with point-block Jacobi and with the multigrid V-cycle of the assembled P2 operator (set-up:
    Jacobian assembly on the device, Morton renumbering and aggregation on the host, timed
    apart)."""
    import numpy as np
    from pylith_amd.dm import Solver
    mesh = mg.tet_p2_box(64, (0.0, 0.0, 0.0), (6000.0, 6000.0, 6000.0))
    aux, sizes = mg.elastic_aux(mesh.num_cells, seed=SEED, vs_sigma=0.2, vp_vs=1.732)
    dm = setup(mesh, aux, sizes)
    dofs, vals = mg.shear_bcs(mesh.node_coords, 0.0, 6000.0)
    dm.set_constraints(dofs)
    solver = Solver(dm)
    solver.set_inner(1e-3, 20000, 30)  # each FP32 PCG solve gains 3 digits
    out = {}
    sols = {}
    for name, pc, method in (("fp64_cg", "pbjacobi", L.KSP_CG), ("fp32_fgmres_ir", "pbjacobi", L.KSP_FGMRES_IR),
                             ("fp64_cg_multigrid", "amg", L.KSP_CG),
                             ("fp32_fgmres_ir_multigrid", "amg", L.KSP_FGMRES_IR)):
        print(f"[bench] c2 {name}: preconditioner {pc}", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solver.set_preconditioner(pc)
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t0
        x = torch.zeros(3 * mesh.num_nodes, dtype=torch.float64, device="cuda")
        dm.insert_boundary_values(torch.from_numpy(vals).cuda(), x)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = solver.solve(x, method=method, rtol=1e-8, maxit=20000)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        sols[name] = x.cpu().numpy()
        out[name] = {"seconds": dt, "preconditioner_setup_s": t_setup, "converged": bool(st["converged"]),
                     "outer": st["outer_iterations"], "inner_fp32": st["inner_iterations"],
                     "rel_residual": st["final_residual"] / max(st["initial_residual"], 1e-300)}
        print(f"[bench] c2 {name}: {out[name]}", file=sys.stderr, flush=True)
    b = sols["fp64_cg"]
    for name in ("fp32_fgmres_ir", "fp64_cg_multigrid", "fp32_fgmres_ir_multigrid"):
        out[name]["rel_diff_vs_fp64_cg"] = float(np.abs(sols[name] - b).max() / np.abs(b).max())
        out[name]["speedup_vs_fp64_cg"] = out["fp64_cg"]["seconds"] / out[name]["seconds"]
    out["rel_diff_fp32ir_vs_fp64"] = out["fp32_fgmres_ir"]["rel_diff_vs_fp64_cg"]
    out["speedup_fp32ir_vs_fp64"] = out["fp32_fgmres_ir"]["speedup_vs_fp64_cg"]
    out["workload"] = "P2 64^3 x 6 tets, -x fixed, +x u_y = 1 m, rtol 1e-8"
    solver.close()
    dm.close()
    return out


def bench_c1_solve(torch, mg, L):
    """C1 full solve: Q1 32^3 (6 km), -x fixed, +x u_y = 1 m, uniform moduli (rho 2500, Vs 3000,
    Vp 5292), FP64 CG to 1e-8 relative residual with point-block Jacobi and with the multigrid
    V-cycle (set-up timed separately), and FGMRES with the FP32 multigrid PCG inside."""
    import numpy as np
    from pylith_amd.dm import Solver
    m = mg.hex_box(32, (0.0, 0.0, 0.0), (6000.0, 6000.0, 6000.0))
    aux, sizes = mg.elastic_aux(m.num_cells)
    dm = setup(m, aux, sizes)
    dofs, vals = mg.shear_bcs(m.node_coords, 0.0, 6000.0)
    dm.set_constraints(dofs)
    solver = Solver(dm)
    out, sols = {}, {}
    for name, pc, method in (("cg_pbjacobi", "pbjacobi", L.KSP_CG), ("cg_amg", "amg", L.KSP_CG),
                             ("fgmres_fp32_amg", "amg", L.KSP_FGMRES_IR)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solver.set_preconditioner(pc)
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t0
        x = torch.zeros(3 * m.num_nodes, dtype=torch.float64, device="cuda")
        dm.insert_boundary_values(torch.from_numpy(vals).cuda(), x)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = solver.solve(x, method=method, rtol=1e-8, maxit=20000)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        sols[name] = x.cpu().numpy()
        out[name] = {"seconds": dt, "preconditioner_setup_s": t_setup, "converged": bool(st["converged"]),
                     "iterations": st["outer_iterations"], "inner_fp32": st["inner_iterations"],
                     "rel_residual": st["final_residual"] / max(st["initial_residual"], 1e-300)}
    ref = sols["cg_pbjacobi"]
    for name in ("cg_amg", "fgmres_fp32_amg"):
        out[name]["rel_diff_vs_pbjacobi"] = float(np.abs(sols[name] - ref).max() / np.abs(ref).max())
        out[name]["speedup_vs_pbjacobi"] = out["cg_pbjacobi"]["seconds"] / out[name]["seconds"]
    out["workload"] = "Q1 32^3 (107,811 dofs), -x fixed, +x u_y = 1 m, rtol 1e-8"
    solver.close()
    dm.close()
    return out


def bench_c3(torch, mg, L, precond="pbjacobi"):
    """C3: generalized Maxwell (3 elements, eta log-uniform [1e17, 1e19] per cell) + vertical
    fault (2 m left-lateral on a 4 x 4 km patch: 2 x 2 km at full slip tapering to zero at the
    patch edge), 48^3 Q1 hexes in a 6 km cube, 10 steps of 0.1 yr: per step the state term,
    an FP64 solve (FGMRES + FP32 PCG), the state update."""
    import numpy as np
    from pylith_amd.dm import DM, Solver
    n = 48
    m, f = mg.hex_box_with_fault(n, (0.0, 0.0, -6000.0), (6000.0, 6000.0, 0.0), n // 2, brick=8)
    aux, sizes = mg.genmax_aux(m.num_cells, 8, seed=SEED, log_eta_range=(1e17, 1e19), ratios=(0.3, 0.3, 0.3))
    dt = 0.1 * mg.YEAR
    dm = DM(m)
    dm.add_residual(0, 0, 0, 3)
    dm.add_jacobian(0, 0, 0, 0, 0, 4)
    dm.set_auxiliary(0, 0, aux, sizes, per_qp=True)
    dm.set_constants(0, 0, [dt])
    dm.add_fault(f)
    slip = mg.tapered_patch_slip(f, (3000.0, -3000.0), 1000.0, 2.0, component=1, taper=1000.0)
    dm.set_fault_slip(slip)
    fixed = np.nonzero(np.isclose(m.node_coords[:, 0], 0.0) | np.isclose(m.node_coords[:, 0], 6000.0))[0]
    dofs = np.sort((3 * fixed[:, None] + np.arange(3)).ravel()).astype(np.int32)
    dm.set_constraints(dofs)
    solver = Solver(dm)
    t0 = time.perf_counter()
    solver.set_preconditioner(precond)
    t_pc = time.perf_counter() - t0
    nloc = dm.sizes()[0]
    nu = 3 * m.num_nodes
    x = torch.zeros(nloc, dtype=torch.float64, device="cuda")
    zero = torch.zeros_like(x)
    b = torch.zeros_like(x)
    res = torch.zeros_like(x)
    steps = []
    dm.reset_event_stats()
    dm.set_event_logging(True)
    t_all = time.perf_counter()
    for step in range(10):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dm.compute_residual(0, 0, zero, b, mode=L.INSERT_VALUES)  # state term F(0)
        st = solver.solve(x, b=b, method=L.KSP_FGMRES_IR, rtol=1e-10, maxit=200)
        if step == 9:  # full residual (domain + cohesive rows) at this step's solution
            dm.compute_residual(0, 0, x, res, mode=L.INSERT_VALUES)
            dm.compute_residual_cohesive(x, res, mode=L.ADD_VALUES)
            rn = res.cpu().numpy()
            rn[dofs] = 0.0
            fres = float(np.abs(rn).max() / max(np.abs(b.cpu().numpy()).max(), 1e-300))
        dm.update_state_vars(0, 0, x)
        torch.cuda.synchronize()
        steps.append({"seconds": time.perf_counter() - t0, "converged": bool(st["converged"]),
                      "inner_fp32": st["inner_iterations"]})
    total = time.perf_counter() - t_all
    n_upd, ms_upd = dm.event_stats("plx:genmax_update")
    n_coh, ms_coh = dm.event_stats("plx:cohesive")
    dm.set_event_logging(False)
    # the C3 unit: one residual apply (elastic operator + state term, cohesive rows) + one update
    r_apply = lambda: (dm.compute_residual(0, 0, x, res, mode=L.INSERT_VALUES),  # noqa: E731
                       dm.compute_residual_cohesive(x, res, mode=L.ADD_VALUES))
    t_res = time_applies(r_apply, 20, 3, torch)
    k_res, _ = kernel_avg_ms(dm, r_apply, 20, torch)
    t_upd = time_applies(lambda: dm.update_state_vars(0, 0, x), 20, 3, torch)
    k_upd, _ = kernel_avg_ms(dm, lambda: dm.update_state_vars(0, 0, x), 20, torch, "plx:genmax_update")
    bu = gm_bytes_update(m.num_nodes, m.num_cells)
    br = q1_bytes(m.num_nodes, m.num_cells, 8) + 6 * 8 * 8 * m.num_cells
    out = {
        "workload": "48^3 Q1 GenMaxwell (3 Maxwell elements/point) + vertical fault, 10 x 0.1 yr, FP64",
        "preconditioner": precond, "preconditioner_setup_s": t_pc,
        "inner_fp32_per_step": [s["inner_fp32"] for s in steps],
        "dofs": nu, "lagrange_dofs": nloc - nu, "cohesive_cells": f.num_cohesive,
        "steps_seconds_total": total, "step_ms_mean": 1e3 * total / 10,
        "all_converged": all(s["converged"] for s in steps),
        "unit_ms_residual_plus_update": 1e3 * (t_res + t_upd),
        "residual_apply_ms": 1e3 * t_res, "residual_kernel_ms": k_res,
        "residual_hbm_frac": br / (k_res * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "state_update_ms": 1e3 * t_upd, "state_update_kernel_ms": k_upd,
        "state_update_algorithmic_bytes": bu,
        "state_update_hbm_frac": bu / (k_upd * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "cohesive_kernel_ms_per_call": ms_coh / max(n_coh, 1),
        "last_step_full_residual_rel": fres,
    }
    solver.close()
    dm.close()
    return out


def bench_c4(torch, mg, L, n_impulses=512):
    """C4: 64^3 Q1 (80 km cube, h = 1.25 km) sheared so the mid-plane fault dips 60 degrees,
    x faces fixed; one unit left-lateral impulse per fault vertex of a 32 (strike) x 16 (dip)
    region (the per-vertex impulse basis of FaultCohesiveImpulses): 512 right-hand sides."""
    import numpy as np
    from pylith_amd import workloads
    from pylith_amd.dm import Solver
    from pylith_amd.greensfns import GreensFunctions
    prob = workloads.faulted_box(64, length=80000.0, dip_deg=60.0)
    f = prob.fault
    slip, verts = mg.fault_vertex_impulses(f, (16, 48), (48, 64), component=1)
    slip = slip[:n_impulses]
    nimp = slip.shape[0]
    nloc = prob.local_size
    nu = 3 * prob.mesh.num_nodes
    gf = GreensFunctions(prob.dm)
    t0 = time.perf_counter()
    best, probe = gf.autotune()
    t_tune = time.perf_counter() - t0
    rtol = 1e-8  # both paths to the same relative residual; errors vs a 1e-14 FP64 solve reported
    s_dev = torch.from_numpy(slip).cuda()
    U = torch.zeros(nimp, nloc, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = gf.solve(s_dev, U, rtol=rtol, maxit=30)
    torch.cuda.synchronize()
    t_gf = time.perf_counter() - t0
    # plain FP64 GPU path (matrix-free PCG, one impulse at a time) on a sample, and the FP64
    # reference solutions (tighter tolerance) for the error
    solver = Solver(prob.dm)
    sample = [0, nimp // 3, (2 * nimp) // 3, nimp - 1]
    errs, errs64, t64, refs = [], [], [], []
    for j in sample:
        prob.dm.set_fault_slip(slip[j])
        x = torch.zeros(nloc, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        solver.solve(x, method=L.KSP_CG, rtol=rtol, maxit=20000)
        torch.cuda.synchronize()
        t64.append(time.perf_counter() - t1)
        xr = torch.zeros_like(x)
        solver.solve(xr, method=L.KSP_CG, rtol=1e-14, maxit=40000)
        ref = xr.cpu().numpy()[:nu]
        refs.append(ref)
        errs.append(float(np.abs(U[j].cpu().numpy()[:nu] - ref).max() / np.abs(ref).max()))
        errs64.append(float(np.abs(x.cpu().numpy()[:nu] - ref).max() / np.abs(ref).max()))
    fp64_rate = 1.0 / float(np.mean(t64))
    gf_rate = nimp / t_gf
    # the plain FP64 path with the same multigrid preconditioner (the like-for-like baseline):
    # matrix-free FP64 CG, one impulse at a time, FP32 V-cycle as its preconditioner
    solver_mg = Solver(prob.dm)
    solver_mg.set_preconditioner("amg")
    t64mg, errs64mg = [], []
    for j, refv in zip(sample, refs):
        prob.dm.set_fault_slip(slip[j])
        x = torch.zeros(nloc, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        solver_mg.solve(x, method=L.KSP_CG, rtol=rtol, maxit=20000)
        torch.cuda.synchronize()
        t64mg.append(time.perf_counter() - t1)
        errs64mg.append(float(np.abs(x.cpu().numpy()[:nu] - refv).max() / np.abs(refv).max()))
    solver_mg.close()
    fp64_mg_rate = 1.0 / float(np.mean(t64mg))
    # the same block solves preconditioned by the smoothed-aggregation multigrid V-cycle
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gf.set_preconditioner("amg")
    t_amg_setup = time.perf_counter() - t0
    best_a, probe_a = gf.autotune()
    Ua = torch.zeros_like(U)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sta = gf.solve(s_dev, Ua, rtol=rtol, maxit=30)
    torch.cuda.synchronize()
    t_amg = time.perf_counter() - t0
    errs_a = []
    for j, refv in zip(sample, refs):
        errs_a.append(float(np.abs(Ua[j].cpu().numpy()[:nu] - refv).max() / np.abs(refv).max()))
    amg = {"setup_s": t_amg_setup, "block": sta["block"], "probe_us_per_rhs_iteration": probe_a,
           "seconds": t_amg, "solves_per_s": nimp / t_amg,
           "outer_iterations": sta["outer_iterations"], "inner_iterations": sta["inner_iterations"],
           "max_rel_residual": sta["max_relative_residual"], "max_rel_error_vs_fp64": max(errs_a),
           "speedup_vs_fp64": (nimp / t_amg) / fp64_rate,
           "speedup_vs_fp64_incl_setup": (nimp / (t_amg + t_amg_setup)) / fp64_rate,
           "speedup_vs_fp64_multigrid": (nimp / t_amg) / fp64_mg_rate,
           "speedup_vs_fp64_multigrid_incl_setup": (nimp / (t_amg + t_amg_setup)) / fp64_mg_rate}
    out = {
        "workload": "64^3 Q1, 80 km cube, fault dipping 60 deg, 32 x 16 vertex impulses (512 RHS)",
        "impulses": nimp, "block": st["block"], "autotune_s": t_tune,
        "probe_us_per_rhs_iteration": probe,
        "seconds": t_gf, "solves_per_s": gf_rate,
        "outer_iterations": st["outer_iterations"], "inner_iterations": st["inner_iterations"],
        "max_rel_residual": st["max_relative_residual"],
        "max_rel_error_vs_fp64": max(errs), "error_sample": len(sample), "rtol": rtol,
        "fp64_path_max_rel_error": max(errs64),
        "fp64_single_solves_per_s": fp64_rate,
        "fp64_single_solves_per_s_note": "FP64 CG, point-block Jacobi, one impulse at a time",
        "fp64_multigrid_single_solves_per_s": fp64_mg_rate,
        "fp64_multigrid_path_max_rel_error": max(errs64mg),
        "speedup_vs_fp64": gf_rate / fp64_rate,
        "multigrid": amg,
    }
    solver.close()
    gf.close()
    prob.dm.close()
    return out


def bench_c5(torch, n_requests=2000, max_block=64, n_capacity=400, inner_rtol=1e-5, load=0.6):
    """C5: Poisson stream of fault-slip solves on 16^3/32^3/64^3/96^3 meshes (40/30/20/10%),
    64 patches with N(0, 0.5 m) slip, lognormal Vs per cell (per mesh size), 200 GB budget,
    continuous batching (requests join the running block of their size as slots free up), each
    request solved to an FP64 relative residual of 1e-9 (errors vs FP64 <= 1e-8).  Capacity:
    `n_capacity` requests queued at once; the stream: `n_requests` Poisson arrivals at `load` of
    it (60%: the stream's served rate then keeps up with the offered rate, so p50/p99 are those of
    a steady state and not of a growing queue).  64 slots per mesh size and an inner FP32 tolerance
    of 1e-5 (one refinement step fewer than 1e-4) were the fastest of the measured settings."""
    import numpy as np
    from pylith_amd import _lib as L
    from pylith_amd import serving
    from pylith_amd import workloads
    from pylith_amd.dm import Solver
    sizes, probs = (16, 32, 64, 96), (0.4, 0.3, 0.2, 0.1)
    rtol = 1e-9
    eng = serving.ServingEngine(sizes, block=max_block, rtol=rtol, inner_rtol=inner_rtol)
    t0 = time.perf_counter()
    eng.warm()
    t_warm = time.perf_counter() - t0
    resident = {str(sz): eng.states[sz].bytes for sz in sizes}
    cap_reqs = serving.make_requests(n_capacity, sizes, probs, rate=None, seed=11)
    done = eng.run(cap_reqs)
    cap = serving.latency_stats(done)["solves_per_s"]
    rate = load * cap
    reqs = serving.make_requests(n_requests, sizes, probs, rate=rate, seed=12)
    done = eng.run(reqs)
    stats = serving.latency_stats(done)
    offered = n_requests / max(r.arrival for r in reqs)
    per = {}
    for sz in sizes:
        d = [r for r in done if r.size == sz]
        if d:
            per[str(sz)] = serving.latency_stats(d)
    joined = sum(r.running_at_admission > 0 for r in done)
    # plain FP64 GPU path on the same mix: one request at a time, matrix-free FP64 CG, with
    # point-block Jacobi and with the same multigrid preconditioner (like for like)
    t64, t64mg = {}, {}
    rng = np.random.default_rng(13)
    for sz in sizes:
        st = eng.states[sz]
        slip = workloads.patch_slip(st.prob.fault, st.patch, rng.normal(0.0, 0.5, st.n_patches))
        st.prob.dm.set_fault_slip(slip)
        for pc, store in (("pbjacobi", t64), ("amg", t64mg)):
            with torch.cuda.stream(st.stream):  # the stream the mesh state's DM launches on
                solver = Solver(st.prob.dm)
                solver.set_preconditioner(pc)
                x = torch.zeros(st.prob.local_size, dtype=torch.float64, device="cuda")
                solver.solve(x, method=L.KSP_CG, rtol=rtol, maxit=50000)  # warm (graph capture)
                x.zero_()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                solver.solve(x, method=L.KSP_CG, rtol=rtol, maxit=50000)
                torch.cuda.synchronize()
                store[str(sz)] = time.perf_counter() - t0
                solver.close()
    fp64_rate = 1.0 / sum(p * t64[str(sz)] for sz, p in zip(sizes, probs))
    fp64_mg_rate = 1.0 / sum(p * t64mg[str(sz)] for sz, p in zip(sizes, probs))
    eng.close()
    return {"workload": "Poisson stream, meshes 16^3/32^3/64^3/96^3 at 40/30/20/10%, 64 fault patches, "
                        "continuous batching",
            "requests": n_requests, "capacity_requests": n_capacity, "capacity_solves_per_s": cap,
            "arrival_rate_per_s": rate, "offered_rate_per_s": offered,
            "p50_ms": stats["p50_ms"], "p99_ms": stats["p99_ms"], "solves_per_s": stats["solves_per_s"],
            "errors": stats["errors"], "joined_a_running_block": joined,
            "per_size": per, "warm_seconds": t_warm, "resident_bytes_measured": resident,
            "rtol": rtol, "max_block": max_block, "inner_rtol": inner_rtol, "load_fraction_of_capacity": load,
            "fp64_single_seconds": t64, "fp64_solves_per_s_mix": fp64_rate,
            "fp64_multigrid_single_seconds": t64mg, "fp64_multigrid_solves_per_s_mix": fp64_mg_rate,
            "capacity_speedup_vs_fp64": cap / fp64_rate,
            "capacity_speedup_vs_fp64_multigrid": cap / fp64_mg_rate}


def bench_c5_sweep(torch):
    """Serving capacity (requests queued at once) of the C5 mix for several slot-block widths
    and inner FP32 tolerances, and of each mesh size served alone."""
    from pylith_amd import serving
    sizes, probs = (16, 32, 64, 96), (0.4, 0.3, 0.2, 0.1)
    out = {}
    for block, inner in ((32, 1e-4), (32, 1e-5), (64, 1e-4), (16, 1e-4)):
        eng = serving.ServingEngine(sizes, block=block, rtol=1e-9, inner_rtol=inner)
        eng.warm()
        done = eng.run(serving.make_requests(400, sizes, probs, rate=None, seed=11))
        st = serving.latency_stats(done)
        st["outer_mean"] = float(sum(r.stats["outer_iterations"] for r in done) / len(done))
        st["inner_mean"] = float(sum(r.stats["inner_iterations"] for r in done) / len(done))
        out[f"block{block}_inner{inner:g}"] = st
        if block == 32 and inner == 1e-4:
            for sz in sizes:
                done = eng.run(serving.make_requests(100, (sz,), (1.0,), rate=None, seed=21))
                out[f"alone_{sz}"] = serving.latency_stats(done)
        eng.close()
        torch.cuda.empty_cache()
    return out


def bench_p2(torch, mg, n_edge, steps, warmup, renumber_seed=None, algorithm=None):
    """P2 FP64 residual apply on an n^3 x 6 box (same synthetic inputs as the headline); with
    `renumber_seed` the cells, nodes and vertices are renumbered at random first, and
    `algorithm` forces an element path."""
    mesh = mg.tet_p2_box(n_edge, (0.0, 0.0, 0.0), (6000.0, 6000.0, 6000.0))
    if renumber_seed is not None:
        mesh = mg.renumbered(mesh, renumber_seed)
    aux, sizes = mg.elastic_aux(mesh.num_cells, seed=SEED, vs_sigma=0.2, vp_vs=1.732)
    dm = setup(mesh, aux, sizes)
    if algorithm is not None:
        dm.set_kernel_options(algorithm)
    x = torch.from_numpy(mg.sinusoid_field(mesh.node_coords, SEED)).cuda()
    r = torch.zeros_like(x)
    f = lambda: dm.compute_residual(0, 0, x, r, mode=1)  # noqa: E731
    sec = time_applies(f, steps, warmup, torch)
    kms, _ = kernel_avg_ms(dm, f, steps, torch)
    B = p2_bytes(mesh.num_nodes, mesh.num_vertices, mesh.num_cells, 8)
    dm.close()
    return {"cells": mesh.num_cells, "dofs": 3 * mesh.num_nodes, "gdofs": 3 * mesh.num_nodes / sec / 1e9,
            "ms": sec * 1e3, "kernel_ms": kms or None, "algorithmic_bytes": B,
            "hbm_frac_kernel": B / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS if kms else None,
            "hbm_frac_apply": B / sec / 1e9 / HBM_PEAK_GBS}


def bench_q1(torch, mg, n_edge, steps, warmup, renumber_seed=None, jitter=None, algorithm=None):
    """Q1 hex FP64 residual apply on an n^3 cube (6 km, u ~ U[-1e-3, 1e-3] m with the shear
    Dirichlet values inserted): wall time per apply, tile-kernel time and HBM fractions; with
    `renumber_seed` the cells and nodes are renumbered at random first, with `jitter` the interior
    nodes move by up to that fraction of the spacing (non-affine hexes), and `algorithm` forces
    an element path (the quadrature-point one: the pointwise stress at every Gauss point)."""
    hm = mg.hex_box(n_edge, (0.0, 0.0, 0.0), (6000.0, 6000.0, 6000.0))
    if renumber_seed is not None:
        hm = mg.renumbered(hm, renumber_seed)
    if jitter is not None:
        hm = mg.jittered(hm, jitter * 6000.0 / n_edge, 7)
    haux, hs = mg.elastic_aux(hm.num_cells)
    hu = mg.random_field(hm.num_nodes, SEED)
    dofs_bc, vals = mg.shear_bcs(hm.node_coords, 0.0, 6000.0)
    hu[dofs_bc] = vals
    hdm = setup(hm, haux, hs)
    if algorithm is not None:
        hdm.set_kernel_options(algorithm)
    hx = torch.from_numpy(hu).cuda()
    hr = torch.zeros_like(hx)
    f = lambda: hdm.compute_residual(0, 0, hx, hr, mode=1)  # noqa: E731
    s = time_applies(f, steps, warmup, torch)
    km, _ = kernel_avg_ms(hdm, f, steps, torch)
    hm_ms, _ = kernel_avg_ms(hdm, f, steps, torch, "plx:reduce_halo")
    med, p99 = apply_percentiles(f, steps, torch)
    hb = q1_bytes(hm.num_nodes, hm.num_cells, 8)
    gms = graph_apply_ms(hdm, f, torch)
    jv = torch.zeros_like(hx)
    sj = time_applies(lambda: hdm.compute_jacobian_action(0, 0, hx, jv, mode=1), steps, warmup, torch)
    out = {"gdofs": 3 * hm.num_nodes / s / 1e9, "ms": s * 1e3, "ms_median": med, "ms_p99": p99,
           "ms_graph": gms, "gdofs_graph": 3 * hm.num_nodes / (gms * 1e-3) / 1e9,
           "kernel_ms": km or None, "halo_reduce_ms": hm_ms or None,
           "algorithmic_bytes": hb, "hbm_frac_kernel": hb / (km * 1e-3) / 1e9 / HBM_PEAK_GBS if km else None,
           "hbm_frac_apply": hb / s / 1e9 / HBM_PEAK_GBS,
           "jacobian_action_gdofs": 3 * hm.num_nodes / sj / 1e9}
    hdm.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the secondary measurements")
    ap.add_argument("--q1-large", type=int, default=192, help="edge of the large Q1 cube")
    ap.add_argument("--no-systems", action="store_true",
                    help="skip the solver / viscoelastic / Green's-function / serving measurements")
    ap.add_argument("--only", choices=["c1solve", "c2solve", "c3", "c3amg", "c4", "c5", "c5b64", "c5b32", "c5sweep", "q1",
                                       "p2large", "pointfn"],
                    default=None,
                    help="run one system measurement alone and print its JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library runs on a single MI355X and has no "
              "multi-GPU path (a mesh partition per GPU with a halo exchange of shared nodes "
              "would be the way to shard it).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from pylith_amd import meshgen as mg

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.only:
        from pylith_amd import _lib as L
        fn = {"c1solve": lambda: bench_c1_solve(torch, mg, L),
              "c2solve": lambda: bench_c2_solve(torch, mg, L), "c3": lambda: bench_c3(torch, mg, L),
              "c3amg": lambda: bench_c3(torch, mg, L, "amg"),
              "c4": lambda: bench_c4(torch, mg, L), "c5": lambda: bench_c5(torch),
              "c5b64": lambda: bench_c5(torch, max_block=64), "c5sweep": lambda: bench_c5_sweep(torch),
              "c5b32": lambda: bench_c5(torch, max_block=32),
              "q1": lambda: bench_q1(torch, mg, args.q1_large, args.steps, args.warmup),
              "p2large": lambda: bench_p2(torch, mg, 96, args.steps, args.warmup),
              "pointfn": lambda: {"p2_64cube": bench_p2(torch, mg, 64, args.steps, args.warmup,
                                                        algorithm=L.ALGO_POINTFN),
                                  "q1_192cube": bench_q1(torch, mg, args.q1_large, args.steps, args.warmup,
                                                         algorithm=L.ALGO_POINTFN)}}[args.only]
        print(f"[bench] {args.only}", file=sys.stderr, flush=True)
        print(json.dumps({args.only: fn()}))
        return
    print("[bench] building the P2 64^3 x 6 workload", file=sys.stderr, flush=True)
    mesh = mg.tet_p2_box(64, (0.0, 0.0, 0.0), (6000.0, 6000.0, 6000.0))
    aux, sizes = mg.elastic_aux(mesh.num_cells, seed=SEED, vs_sigma=0.2, vp_vs=1.732)
    u = mg.sinusoid_field(mesh.node_coords, SEED)
    dm = setup(mesh, aux, sizes)
    ndofs = 3 * mesh.num_nodes
    x = torch.from_numpy(u).cuda()
    r = torch.zeros_like(x)

    def apply64():
        dm.compute_residual(0, 0, x, r, mode=1)

    print("[bench] timing FP64 residual apply", file=sys.stderr, flush=True)
    sec = time_applies(apply64, args.steps, args.warmup, torch)
    kms, kcount = kernel_avg_ms(dm, apply64, args.steps, torch)
    hms, _ = kernel_avg_ms(dm, apply64, args.steps, torch, "plx:reduce_halo")
    med, p99 = apply_percentiles(apply64, args.steps, torch)
    gms = graph_apply_ms(dm, apply64, torch)
    B = p2_bytes(mesh.num_nodes, mesh.num_vertices, mesh.num_cells, 8)
    achieved = B / (kms * 1e-3) / 1e9
    traffic = traffic_build = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        traffic = tj.get("p2_fp64_residual_64cube")
        builds = tj.get("_build")
        traffic_build = builds.get("p2_fp64_residual_64cube") if isinstance(builds, dict) else builds
    result = {
        "metric": "residual-apply GDoF/s (% HBM roofline)",
        "value": ndofs / sec / 1e9,
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "ms_per_step_median": med,
        "ms_per_step_p99": p99,
        "ms_per_step_graph": gms,
        "value_graph": ndofs / (gms * 1e-3) / 1e9,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "P2 tet elasticity residual apply, 64^3 box x 6 tets (1,572,864 cells, "
                        "6,440,067 dofs), 14-point quadrature, lognormal Vs, FP64",
            "cells": mesh.num_cells, "nodes": mesh.num_nodes, "dofs": ndofs,
            "quadrature_points": 14, "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "apply_tiles<double, TET_P2, factored, lds-atomic>",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "frac_basis": "the apply_tiles kernel alone (HIP events on its stream)",
            "frac_apply_kernels": B / ((kms + hms) * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "frac_apply_wall": B / sec / 1e9 / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_basis": "rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 of this kernel (profiles/traffic.json, "
                             f"build {traffic_build})",
            "algorithmic_bytes": B,
            "kernel_avg_ms": kms,
            "kernel_launches_timed": kcount,
            "halo_reduce_avg_ms": hms,
        },
    }

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py  # noqa: E402  (checker / CPU baseline only)
        print("[bench] CPU baseline (oracle)", file=sys.stderr, flush=True)
        problem = oracle_py.Problem(mesh, aux, sizes)
        ref, dt, n, nthreads = cpu_baseline(problem, u, mesh.colors)
        apply64()
        torch.cuda.synchronize()
        err = float(np.abs(r.cpu().numpy() - ref).max() / np.abs(ref).max())
        print("[bench] CPU baseline, one core", file=sys.stderr, flush=True)
        _, dt1, n1, _ = cpu_baseline(problem, u, mesh.colors, budget_s=4.0, nthreads=1)
        result["cpu_baseline"] = {
            "value": ndofs / dt / 1e9, "unit": "GDoF/s", "cores": nthreads, "kind": "port",
            "sample": f"{n} full applies of the same workload, oracle cell loop (PetscPointFn "
                      f"call per quadrature point), colored OpenMP over {nthreads} threads",
            "single_core": {"value": ndofs / dt1 / 1e9, "unit": "GDoF/s", "cores": 1,
                            "sample": f"{n1} full applies on one thread"},
            "host": cpu_description(),
        }
        result["parity"] = {"rel_err_inf_vs_oracle": err, "tol": 1e-12, "ok": err <= 1e-12}

    if not args.no_extra:
        extra = {}
        from pylith_amd import _lib as L
        print("[bench] point-by-point 14-point evaluation", file=sys.stderr, flush=True)
        dm.set_kernel_options(L.ALGO_QUADRATURE_POINTS, L.ACCUMULATE_AUTO)
        secq = time_applies(apply64, args.steps, args.warmup, torch)
        kq, _ = kernel_avg_ms(dm, apply64, args.steps, torch)
        rq = r.clone()
        dm.set_kernel_options(L.ALGO_AUTO, L.ACCUMULATE_AUTO)
        apply64()
        torch.cuda.synchronize()
        extra["p2_fp64_residual_quadrature_points"] = {
            "gdofs": ndofs / secq / 1e9, "ms": secq * 1e3, "kernel_ms": kq,
            "hbm_frac": B / (kq * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "rel_diff_vs_default": float((rq - r).abs().max() / r.abs().max())}
        print("[bench] coloured (bitwise reproducible) accumulation", file=sys.stderr, flush=True)
        rdef = r.clone()  # the default (LDS-atomic) result
        dm.set_kernel_options(L.ALGO_AUTO, L.ACCUMULATE_COLORED)
        secc = time_applies(apply64, args.steps, args.warmup, torch)
        kc, _ = kernel_avg_ms(dm, apply64, args.steps, torch)
        rc = r.clone()
        apply64()
        torch.cuda.synchronize()
        extra["p2_fp64_residual_colored"] = {
            "gdofs": ndofs / secc / 1e9, "ms": secc * 1e3, "kernel_ms": kc,
            "hbm_frac": B / (kc * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "bitwise_repeatable": bool(torch.equal(rc, r)),
            "rel_diff_vs_default": float((rc - rdef).abs().max() / rdef.abs().max())}
        dm.set_kernel_options(L.ALGO_AUTO, L.ACCUMULATE_AUTO)
        apply64()
        torch.cuda.synchronize()
        print("[bench] FP32 and Jacobian action", file=sys.stderr, flush=True)
        x32 = x.float()
        r32 = torch.zeros_like(x32)
        sec32 = time_applies(lambda: dm.compute_residual(0, 0, x32, r32, mode=1), args.steps,
                             args.warmup, torch)
        k32, _ = kernel_avg_ms(dm, lambda: dm.compute_residual(0, 0, x32, r32, mode=1), args.steps, torch)
        B32 = p2_bytes(mesh.num_nodes, mesh.num_vertices, mesh.num_cells, 4)
        extra["p2_fp32_residual"] = {"gdofs": ndofs / sec32 / 1e9, "ms": sec32 * 1e3,
                                     "kernel_ms": k32, "hbm_frac": B32 / (k32 * 1e-3) / 1e9 / HBM_PEAK_GBS}
        jv = torch.zeros_like(x)
        secj = time_applies(lambda: dm.compute_jacobian_action(0, 0, x, jv, mode=1), args.steps,
                            args.warmup, torch)
        extra["p2_fp64_jacobian_action"] = {"gdofs": ndofs / secj / 1e9, "ms": secj * 1e3}
        dm.close()
        del x, r, x32, r32, jv
        torch.cuda.empty_cache()
        print("[bench] P2 96^3 x 6 (working set beyond the Infinity Cache)", file=sys.stderr, flush=True)
        extra["p2_fp64_residual_96cube"] = bench_p2(torch, mg, 96, max(10, args.steps // 2), args.warmup)
        torch.cuda.empty_cache()
        # 32^3 as specified; 192^3 and 256^3 (2.0 GB per apply) where the working set is far
        # beyond the 256 MB Infinity Cache
        for n_edge in sorted({32, args.q1_large, 256}):
            print(f"[bench] Q1 hex {n_edge}^3", file=sys.stderr, flush=True)
            steps = args.steps if n_edge <= 64 else max(10, args.steps // 2)
            extra[f"q1_fp64_residual_{n_edge}cube"] = bench_q1(torch, mg, n_edge, steps, args.warmup)
            torch.cuda.empty_cache()
        # meshes numbered without locality (random renumbering of cells and nodes): the planner
        # tiles them in Morton order of the cell centroids
        print("[bench] renumbered meshes", file=sys.stderr, flush=True)
        extra["q1_fp64_residual_192cube_renumbered"] = bench_q1(torch, mg, args.q1_large, max(10, args.steps // 2),
                                                                args.warmup, renumber_seed=5)
        torch.cuda.empty_cache()
        extra["p2_fp64_residual_64cube_renumbered"] = bench_p2(torch, mg, 64, max(10, args.steps // 2), args.warmup,
                                                               renumber_seed=6)
        torch.cuda.empty_cache()
        # the quadrature-point Q1 paths: the pointwise stress at every Gauss point on affine cells,
        # and on non-affine cells (interior nodes moved by up to 0.2 h: the Jacobian at each point)
        from pylith_amd import _lib as L
        print("[bench] Q1 quadrature-point paths", file=sys.stderr, flush=True)
        extra["q1_fp64_residual_192cube_quadrature_points"] = bench_q1(
            torch, mg, args.q1_large, max(10, args.steps // 2), args.warmup, algorithm=L.ALGO_QUADRATURE_POINTS)
        torch.cuda.empty_cache()
        extra["q1_fp64_residual_192cube_nonaffine"] = bench_q1(torch, mg, args.q1_large, max(10, args.steps // 2),
                                                               args.warmup, jitter=0.2)
        torch.cuda.empty_cache()
        # the same operators through the PetscPointFn-form device path (the plug-in route for
        # kernels written to PyLith's pointwise signature: a thread per cell, colour passes)
        print("[bench] PetscPointFn-form path", file=sys.stderr, flush=True)
        extra["p2_fp64_residual_pointfn"] = bench_p2(torch, mg, 64, 5, 2, algorithm=L.ALGO_POINTFN)
        torch.cuda.empty_cache()
        extra["q1_fp64_residual_192cube_pointfn"] = bench_q1(torch, mg, args.q1_large, 5, 2,
                                                             algorithm=L.ALGO_POINTFN)
        torch.cuda.empty_cache()
        result["extra"] = extra
    if not args.no_systems:
        from pylith_amd import _lib as L
        systems = {}
        for name, fn in (("c1_solve", lambda: bench_c1_solve(torch, mg, L)),
                         ("c2_solve_fp32_refinement", lambda: bench_c2_solve(torch, mg, L)),
                         ("c3_genmaxwell_fault", lambda: bench_c3(torch, mg, L)),
                         ("c3_genmaxwell_fault_multigrid", lambda: bench_c3(torch, mg, L, "amg")),
                         ("c4_greens_functions", lambda: bench_c4(torch, mg, L)),
                         ("c5_serving", lambda: bench_c5(torch))):
            print(f"[bench] {name}", file=sys.stderr, flush=True)
            systems[name] = fn()
            torch.cuda.empty_cache()
        result["systems"] = systems
    print(json.dumps(result))


if __name__ == "__main__":
    main()
