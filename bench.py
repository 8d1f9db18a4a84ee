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


def cpu_baseline(problem, u, colors, budget_s=10.0):
    """Oracle residual on the host: colored OpenMP over all usable cores for ~budget_s."""
    import numpy as np
    nthreads = min(16, os.cpu_count() or 1)
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


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the secondary measurements")
    ap.add_argument("--q1-large", type=int, default=192, help="edge of the large Q1 cube")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X and has no "
              "multi-GPU path (sharding is described in DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from pylith_amd import meshgen as mg

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    print("[bench] building configs[1] workload", file=sys.stderr, flush=True)
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
    B = p2_bytes(mesh.num_nodes, mesh.num_vertices, mesh.num_cells, 8)
    achieved = B / (kms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("p2_fp64_residual_64cube")
    result = {
        "metric": "residual-apply GDoF/s (% HBM roofline)",
        "value": ndofs / sec / 1e9,
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
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
            "traffic": traffic,
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
        result["cpu_baseline"] = {
            "value": ndofs / dt / 1e9, "unit": "GDoF/s", "cores": nthreads, "kind": "port",
            "sample": f"{n} full applies of the same workload, oracle cell loop (PetscPointFn "
                      f"call per quadrature point), colored OpenMP over {nthreads} threads",
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
        for n_edge in (32, args.q1_large):
            print(f"[bench] Q1 hex {n_edge}^3", file=sys.stderr, flush=True)
            hm = mg.hex_box(n_edge, (0.0, 0.0, 0.0), (6000.0, 6000.0, 6000.0))
            haux, hs = mg.elastic_aux(hm.num_cells)
            hu = mg.random_field(hm.num_nodes, SEED)
            dofs_bc, vals = mg.shear_bcs(hm.node_coords, 0.0, 6000.0)
            hu[dofs_bc] = vals
            hdm = setup(hm, haux, hs)
            hx = torch.from_numpy(hu).cuda()
            hr = torch.zeros_like(hx)
            f = lambda: hdm.compute_residual(0, 0, hx, hr, mode=1)  # noqa: E731
            steps = args.steps if n_edge <= 64 else max(10, args.steps // 2)
            s = time_applies(f, steps, args.warmup, torch)
            km, _ = kernel_avg_ms(hdm, f, steps, torch)
            hb = q1_bytes(hm.num_nodes, hm.num_cells, 8)
            extra[f"q1_fp64_residual_{n_edge}cube"] = {
                "gdofs": 3 * hm.num_nodes / s / 1e9, "ms": s * 1e3, "kernel_ms": km,
                "algorithmic_bytes": hb,
                "hbm_frac_kernel": hb / (km * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "hbm_frac_apply": hb / s / 1e9 / HBM_PEAK_GBS,
            }
            hdm.close()
            del hx, hr
            torch.cuda.empty_cache()
        result["extra"] = extra
    print(json.dumps(result))


if __name__ == "__main__":
    main()
