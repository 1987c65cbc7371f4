"""Benchmark of the MI355X matrix-free Stokes solver.

Workload: 3D unit box, 128^3 Q2xQ1 cells (2x2x2 coarse box, 7 levels), 16
seeded sinkers with a 1e6 viscosity contrast, free-slip walls, viscosity per
quadrature point on the active level.

A step is one FP64 application of the fine-level Stokes operator (the metric's
"Stokes matvec"), inputs resident in HBM.  ``value`` is GDoF/s = (N_u + N_p) /
average device time per step, timed with hipEvents on the library's stream.
After the timed steps the script reports, outside ``value``: the FP32 level
A-block apply rate, and time-to-solve to 1e-8 relative residual with the FP32
V-cycle and with the FP64 V-cycle (both inside FP64 GMRES/FGMRES), the time
to solve an ensemble of 48 independent 2D problems together, and on general
meshes (the spherical shell with one adaptive level, hanging nodes, free
slip) the FP64 Stokes apply and the global-coarsening multigrid solve.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 128] [--no-solve] [--no-box-solve]
                      [--no-cpu] [--no-ensemble] [--no-tria] [--shell-refinement G] [--fgmres-profile K]
                      [--contraction-probe]

``--fgmres-profile K`` runs only K cheap iterations of the FP32-GMG solve and
prints their time (the command the per-iteration kernel profile is taken on).
``--contraction-probe`` times the 3-point basis contraction as VALU FMAs and
on the matrix cores and exits.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(size):
    """HBM bytes per FP64 Stokes apply (the four dispatches of the split apply:
    plane kernel, two velocity gathers, pressure gather) from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload, scaled by the
    copy-kernel calibration recorded beside it; None when no summary exists."""
    path = os.path.join(ROOT, "profiles", f"stokes_fp64_{size}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    return rec.get("hbm_bytes_per_apply")


def host_cpu():
    """lscpu's model, sockets and cores per socket of this host."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    sockets = int(info.get("Socket(s)", "0") or 0)
    cps = int(info.get("Core(s) per socket", "0") or 0)
    return {"model": info.get("Model name", "?"), "sockets": sockets, "cores_per_socket": cps,
            "physical_cores": sockets * cps, "usable_cpus": len(os.sched_getaffinity(0))}


def make_config(size):
    from aspect_amd import problems as P
    cfg = P.config_sinker_128()
    if size != 128:
        cfg.n_levels = int(np.log2(size // 2)) + 1
        cfg.name = "sinker_128"  # same material law, smaller mesh
    return cfg


def fgmres_profile(cfg, eta, rho, s, its):
    """K cheap iterations of the FP32-GMG solve (no expensive steps): one
    Arnoldi step = block preconditioner + Stokes apply + DCGS reduce/update."""
    import torch
    from aspect_amd import StokesMatrixFreeHandler

    h = StokesMatrixFreeHandler(cfg.box(s), "fp32")
    h.assemble(eta, "harmonic")
    h.build_preconditioner()
    sz = h.sizes(h.top)
    n = sz["nu"] + sz["np"]
    dev = torch.device("cuda:0")
    rho_d = torch.from_numpy(np.ascontiguousarray(rho.reshape(-1))).to(dev)
    rhs = torch.empty(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    h.assemble_buoyancy_rhs_device(rho_d.data_ptr(), cfg.gravity(), rhs.data_ptr())
    sol = torch.zeros(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()  # the library runs on its own stream
    rc, r = h.solve_device(rhs.data_ptr(), 0, sol.data_ptr(), tolerance=cfg.tolerance, n_cheap_steps=its,
                           n_expensive_steps=0)
    k = max(1, r["cheap_iterations"])
    print(json.dumps({"fgmres_iterations": r["cheap_iterations"], "solve_seconds": r["solve_seconds"],
                      "ms_per_iteration": r["solve_seconds"] / k * 1e3}))
    return 0


def contraction_probe():
    """FP32 3-point basis contractions (the sum-factorization kernel's inner
    operation) as VALU FMAs and on the matrix cores (v_mfma_f32_16x16x4f32
    with the 3 x 3 table packed block-diagonally into 16 x 16 tiles): useful
    FMA rates, and the cross-check of the two results."""
    import ctypes as C
    from aspect_amd.mfs import _check, load_library
    lib = load_library()
    n, reps = 80 * 65536, 256
    rec = {}
    outs = []
    for v, name in ((0, "valu_fma"), (1, "mfma_16x16x4f32")):
        sec = C.c_double(0)
        out = np.empty(3 * n)
        _check(lib.mfs_probe_contraction(v, C.c_longlong(n), reps, C.byref(sec),
                                         out.ctypes.data_as(C.c_void_p)))
        contractions = n * reps * (1 if v == 0 else 4)
        rec[name] = {"seconds": sec.value, "useful_tflops": 2 * 9 * contractions / sec.value / 1e12}
        outs.append(out)
    i = np.arange(3 * n, dtype=np.int64)
    x = (((i * 2654435761) % 1000).astype(np.float32) * np.float32(1e-3) - np.float32(0.5)).astype(float)
    sx = outs[0] - x
    rec["max_rel_diff_mfma_vs_fma"] = float(np.max(np.abs(outs[1] - 4 * sx)) / np.max(np.abs(4 * sx)))
    rec["lines"], rec["reps"] = n, reps
    print(json.dumps({"contraction_probe": rec}), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--contraction-probe", action="store_true",
                    help="time the 3-point basis contraction as VALU FMAs and on the matrix cores, then exit")
    ap.add_argument("--no-box-solve", action="store_true",
                    help="skip the 128^3 time-to-solve runs (the general-mesh solve still runs)")
    ap.add_argument("--no-tria", action="store_true",
                    help="skip the general-mesh (spherical shell with hanging nodes) apply and solve")
    ap.add_argument("--shell-refinement", type=int, default=3,
                    help="global refinement of the 96-cell shell for the general-mesh apply")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the ensemble of 48 independent 2D problems")
    ap.add_argument("--cheap-steps", type=int, default=1000,
                    help="Number of cheap Stokes solver steps (ASPECT's default 1000)")
    ap.add_argument("--fgmres-profile", type=int, default=0, metavar="K",
                    help="only run K cheap FP32-GMG FGMRES iterations of the solve (per-iteration kernel profile)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three device copies of 1 GiB (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.empty(2 ** 27, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        x.fill_(1.0)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 8, "copy_bytes_written": y.numel() * 8, "copies": 3}))
        return 0
    if args.contraction_probe:
        return contraction_probe()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only",
                          "n_gpus": args.gpus}))
        return 2

    import torch
    from aspect_amd import StokesMatrixFreeHandler, problems as P

    cfg = make_config(args.size)
    t0 = time.perf_counter()
    eta, rho, s = P.fields(cfg)
    log(f"fields {cfg.n_cells} cells in {time.perf_counter() - t0:.1f}s, pressure scaling {s:.4g}")
    if args.fgmres_profile:
        return fgmres_profile(cfg, eta, rho, s, args.fgmres_profile)

    h32 = StokesMatrixFreeHandler(cfg.box(s), "fp32")
    h32.assemble(eta, "none")
    t0 = time.perf_counter()
    h32.build_preconditioner()
    log(f"FP32 GMG set-up {time.perf_counter() - t0:.2f}s")
    sz = h32.sizes(h32.top)
    nu, npd, cells = sz["nu"], sz["np"], sz["cells"]
    n = nu + npd
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    x = torch.randn(n, dtype=torch.float64, device=dev, generator=gen)
    y = torch.empty_like(x)
    torch.cuda.synchronize()

    # warm-up, then exactly K timed steps bracketed by synchronization
    if args.warmup:
        h32.time_operator(0, 0, x.data_ptr(), y.data_ptr(), args.warmup)
    h32.synchronize()
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    t_step = h32.time_operator(0, 0, x.data_ptr(), y.data_ptr(), args.steps)
    h32.synchronize()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - w0) / args.steps
    gdofs = n / t_step / 1e9
    nq = 27
    bytes_stokes = 8 * 2 * n + 8 * cells * nq
    achieved = bytes_stokes / t_step / 1e9
    log(f"Stokes FP64 apply: {t_step * 1e3:.3f} ms (wall {wall * 1e3:.3f} ms) {gdofs:.3f} GDoF/s "
        f"{achieved:.0f} GB/s algorithmic")

    # FP32 fine-level A-block apply (cellwise level viscosity)
    u32 = torch.randn(nu, dtype=torch.float32, device=dev, generator=gen)
    v32 = torch.empty_like(u32)
    torch.cuda.synchronize()
    h32.time_operator(1, h32.top, u32.data_ptr(), v32.data_ptr(), 2)
    t_a32 = h32.time_operator(1, h32.top, u32.data_ptr(), v32.data_ptr(), max(5, args.steps // 2))
    bytes_a32 = 4 * 2 * nu + 4 * cells
    log(f"A-block FP32 fine level: {t_a32 * 1e3:.3f} ms {nu / t_a32 / 1e9:.3f} GDoF/s "
        f"{bytes_a32 / t_a32 / 1e9:.0f} GB/s algorithmic")

    # one A-block and one Schur V-cycle (FP32, FP64 vectors at the boundary)
    ud = torch.randn(nu, dtype=torch.float64, device=dev, generator=gen)
    pd = torch.randn(npd, dtype=torch.float64, device=dev, generator=gen)
    vd, qd = torch.empty_like(ud), torch.empty_like(pd)
    torch.cuda.synchronize()
    t_vc = {}
    for which, name, a_, b_ in ((3, "a_vcycle_fp32_ms", ud, vd), (4, "s_vcycle_fp32_ms", pd, qd)):
        h32.time_operator(which, h32.top, a_.data_ptr(), b_.data_ptr(), 1)
        t_vc[name] = h32.time_operator(which, h32.top, a_.data_ptr(), b_.data_ptr(), 3) * 1e3
    log(f"V-cycles FP32: A {t_vc['a_vcycle_fp32_ms']:.2f} ms, S {t_vc['s_vcycle_fp32_ms']:.2f} ms")
    del ud, pd, vd, qd

    result = {
        "metric": "Stokes matvec GDoF/s; solve time to 1e-8 residual, FP32-GMG vs FP64",
        "value": gdofs,
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"sinker box {args.size}^3 Q2xQ1, 16 seeded sinkers, contrast 1e6, "
                               "per-qp viscosity, free slip; step = one FP64 Stokes operator apply",
                   "cells": cells, "velocity_dofs": nu, "pressure_dofs": npd, "levels": cfg.n_levels},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.size)},
        "ablock_fp32_gdofs": nu / t_a32 / 1e9,
        "ablock_fp32_ms": t_a32 * 1e3,
        "wall_ms_per_step": wall * 1e3,
        **t_vc,
    }

    if not args.no_solve and not args.no_box_solve:
        # Time to solve: ASPECT's GMG requires material averaging (with averaging
        # disabled its default solver selection falls back to AMG,
        # source/simulator/helper_functions.cc:2833); the solve uses "harmonic
        # average only viscosity" as ASPECT's nsinker benchmark does, with the
        # FP32 V-cycle and with the FP64 V-cycle, both inside FP64 (F)GMRES.
        g = cfg.gravity()
        rhs = torch.empty(n, dtype=torch.float64, device=dev)
        rho_d = torch.from_numpy(np.ascontiguousarray(rho.reshape(-1))).to(dev)
        torch.cuda.synchronize()
        h32.assemble_buoyancy_rhs_device(rho_d.data_ptr(), g, rhs.data_ptr())
        del rho_d
        del h32
        torch.cuda.empty_cache()
        solves = {}
        sols = {}
        for prec in ("fp32", "fp64"):
            hs = StokesMatrixFreeHandler(cfg.box(s), prec)
            hs.assemble(eta, "harmonic")
            t0 = time.perf_counter()
            hs.build_preconditioner()
            t_setup = time.perf_counter() - t0
            sol = torch.zeros(n, dtype=torch.float64, device=dev)
            torch.cuda.synchronize()  # the library runs on its own stream
            rc, r = hs.solve_device(rhs.data_ptr(), 0, sol.data_ptr(), tolerance=cfg.tolerance,
                                    n_cheap_steps=args.cheap_steps)
            hs.synchronize()
            log(f"{prec.upper()}-GMG solve: {r['cheap_iterations']}+{r['expensive_iterations']} its "
                f"{r['solve_seconds']:.3f}s (set-up {t_setup:.2f}s) converged={r['converged']}")
            solves[prec] = dict(seconds=r["solve_seconds"], setup_seconds=t_setup,
                                cheap_iterations=r["cheap_iterations"],
                                expensive_iterations=r["expensive_iterations"], converged=bool(r["converged"]),
                                final_residual=r["final_linear_residual"], r0=r["r0_norm"],
                                cheap_phase_seconds=r["cheap_seconds"],
                                cheap_phase_final_residual=r["cheap_final_residual"],
                                pressure_normalization_adjustment=r["pressure_normalization_adjustment"])
            sols[prec] = sol
            del hs
            torch.cuda.empty_cache()
        du = torch.linalg.norm(sols["fp32"][:nu] - sols["fp64"][:nu]) / torch.linalg.norm(sols["fp64"][:nu])
        dp = torch.linalg.norm(sols["fp32"][nu:] - sols["fp64"][nu:]) / torch.linalg.norm(sols["fp64"][nu:])
        result["solve"] = {"averaging": "harmonic average only viscosity", "tolerance": cfg.tolerance,
                           "fp32_gmg": solves["fp32"], "fp64_gmg": solves["fp64"],
                           "speedup_fp32_over_fp64": solves["fp64"]["seconds"] / solves["fp32"]["seconds"],
                           "rel_l2_velocity_fp32_vs_fp64": float(du), "rel_l2_pressure_fp32_vs_fp64": float(dp)}
        if not args.no_cpu:
            # both GPU solutions under the oracle's FP64 operator on the CPU:
            # ||b - K x|| / ||r0|| with the scaled pressure unknown p/s
            rhs_h = rhs.cpu().numpy()
            for prec in ("fp32", "fp64"):
                solves[prec]["_x"] = sols[prec].cpu().numpy()
        del sols
    else:
        del h32

    if not args.no_ensemble and not args.no_solve:
        # Ensemble of 48 independent 2D problems of uneven sizes (16^2 .. 256^2
        # cells, each with its own sinker field), FP32 V-cycle, solved together:
        # one handle and HIP stream per problem, 16 host worker threads.
        from aspect_amd import Ensemble
        cfgs = P.config_ensemble(48, seed=4)
        fields = [P.fields(c) for c in cfgs]
        ens = Ensemble([c.box(f[2]) for c, f in zip(cfgs, fields)], "fp32")
        t0 = time.perf_counter()
        ens.setup([f[0] for f in fields], "harmonic")
        t_set = time.perf_counter() - t0
        rhss = [m.assemble_buoyancy_rhs(f[1], c.gravity()) for m, c, f in zip(ens.members, cfgs, fields)]
        _, outs, secs = ens.solve(rhss, tolerance=1e-8)
        its = sum(o.cheap_iterations + o.expensive_iterations for o in outs)
        result["ensemble"] = {"problems": len(cfgs), "dofs": int(sum(r.size for r in rhss)), "seconds": secs,
                              "setup_seconds": t_set, "iterations": its,
                              "converged": all(o.converged for o in outs), "tolerance": 1e-8}
        log(f"ensemble of {len(cfgs)} 2D problems: {secs:.2f}s, {its} iterations, "
            f"converged={result['ensemble']['converged']}")
        del ens

    if not args.no_tria:
        # General level meshes: the spherical shell (96 coarse cells, global
        # refinement, one adaptive level on the 20% of cells with the largest
        # |dT/dr|), free slip on both surfaces, per-point geometry and
        # viscosity, hanging-node and no-normal-flux constraints.  FP64 Stokes
        # apply; then the global-coarsening multigrid and the solve on the next
        # smaller shell (free slip on both spheres: the discrete rigid
        # rotations are only nearly in the kernel on curved cells, so the
        # true residual can end a little above the GMRES estimate).
        from aspect_amd import TriaMultigrid, TriaOperators
        from aspect_amd import tria as TR
        t0 = time.perf_counter()
        tr = P.config_shell(3, args.shell_refinement)
        mesh = tr.build({0: "tangential", 1: "tangential"})
        t_mesh = time.perf_counter() - t0
        op = TriaOperators(mesh, pressure_scaling=1.0)
        op.set_viscosity(P.shell_viscosity(mesh.x_q))
        nt = mesh.n_dofs_u + mesh.n_dofs_p
        gen = torch.Generator(device="cuda").manual_seed(5)
        xs = torch.randn(nt, dtype=torch.float64, device="cuda", generator=gen)
        ys = torch.empty_like(xs)
        torch.cuda.synchronize()
        op.time("stokes", xs.data_ptr(), ys.data_ptr(), 3)
        t_app = op.time("stokes", xs.data_ptr(), ys.data_ptr(), 20)
        nq = 27
        alg = 8 * 2 * nt + 8 * mesh.n_cells * nq + 8 * mesh.n_cells * nq * 10
        rec = {"workload": f"shell 96 x 8^{args.shell_refinement} cells + 1 adaptive level (top 20% |dT/dr|), "
                           "free slip, Q2xQ1, per-point J^-1/JxW and viscosity",
               "cells": mesh.n_cells, "dofs": nt, "hanging_velocity_nodes": int(len(mesh.hanging_u)),
               "constraint_lines": int(len(mesh.con_u[0]) + len(mesh.con_p[0])),
               "stokes_apply_ms": t_app * 1e3, "gdofs_per_s": nt / t_app / 1e9,
               "algorithmic_bytes": alg, "achieved_GBs": alg / t_app / 1e9,
               "frac_hbm": alg / t_app / 1e9 / HBM_PEAK_GBS, "mesh_setup_seconds": t_mesh}
        log(f"general mesh: {mesh.n_cells} cells, {nt} DoFs, Stokes apply {t_app * 1e3:.3f} ms "
            f"({nt / t_app / 1e9:.2f} GDoF/s, {alg / t_app / 1e9:.0f} GB/s algorithmic)")
        del op, xs, ys
        if not args.no_solve:
            t0 = time.perf_counter()
            tr2 = P.config_shell(3, max(args.shell_refinement - 1, 0))
            trias, meshes, pu, pp = TR.hierarchy(tr2, {0: "tangential", 1: "tangential"})
            parents = [None] + [TR.parent_map(trias[l], trias[l - 1]) for l in range(1, len(trias))]
            levels = [TriaOperators(m, pressure_scaling=1.0) for m in meshes]
            t_h = time.perf_counter() - t0
            mg = TriaMultigrid(levels, pu, pp, parents)
            t0 = time.perf_counter()
            # material evaluation on the finest level, harmonic averaging, DGQ0 level tables
            mg.set_viscosity(P.shell_viscosity(meshes[-1].x_q), "harmonic")
            mg.build()
            t_b = time.perf_counter() - t0
            top = meshes[-1]
            rng = np.random.default_rng(6)
            b = rng.normal(size=top.n_dofs_u + top.n_dofs_p)
            b[top.con_u[0]] = 0.0
            b[top.n_dofs_u + top.con_p[0]] = 0.0
            fp = np.ones(top.n_dofs_p, bool)
            fp[top.con_p[0]] = False
            bp = b[top.n_dofs_u:]
            bp[fp] -= bp[fp].mean()   # enclosed flow: right-hand side orthogonal to the constant pressure
            xsol, out, rc = mg.solve(b, tolerance=1e-8, n_cheap_steps=1000, restart=100)
            resid = np.linalg.norm(b - levels[-1].vmult("stokes", xsol)) / np.linalg.norm(b)
            rec["solve"] = {"levels": [m.n_cells for m in meshes], "dofs": int(b.size), "seconds": out.solve_seconds,
                            "cheap_iterations": out.cheap_iterations,
                            "expensive_iterations": out.expensive_iterations,
                            "cheap_phase_seconds": out.cheap_seconds, "converged": bool(out.converged),
                            "rel_residual": resid, "gmg_build_seconds": t_b, "host_setup_seconds": t_h,
                            "boundary": "free slip on both spheres",
                            "averaging": "harmonic average only viscosity, DGQ0 level tables"}
            log(f"general mesh solve: levels {[m.n_cells for m in meshes]}, {out.cheap_iterations}+"
                f"{out.expensive_iterations} iterations, {out.solve_seconds:.2f} s, residual {resid:.2e}")
            del mg, levels
        result["general_mesh"] = rec

    if not args.no_cpu:
        # CPU baseline (oracle/cpu_port.cpp, a CPU restatement, not the
        # reference binary): the FP64 Stokes apply of the same workload with 8
        # cells per SIMD vector (AVX-512 where the host has it) and OpenMP over
        # cell batches, on the host cores this job may use, plus one thread;
        # one FP64 A-block V-cycle of the oracle's multigrid (set-up excluded).
        import oracle as O
        threads = min(16, len(os.sched_getaffinity(0)))
        p = O.make_problem(3, 2, cfg.coarse, cfg.n_levels, cfg.extent, cfg.bc, pressure_scaling=s)
        port = O.CpuPort(p)
        xc = x.cpu().numpy()
        yc = np.empty_like(xc)

        def timed(nthreads, budget, max_reps):
            port.stokes_vmult(eta, xc, out=yc, threads=nthreads)
            reps, t0 = 0, time.perf_counter()
            while reps < max_reps and (time.perf_counter() - t0 < budget or reps < 2):
                port.stokes_vmult(eta, xc, out=yc, threads=nthreads)
                reps += 1
            return (time.perf_counter() - t0) / reps, reps

        tc, reps = timed(threads, 10.0, 40)
        t1, reps1 = timed(1, 0.0, 2)
        O.set_threads(threads)
        t_vset, t_vc = O.time_vcycle(p, 1.0 / np.mean(1.0 / eta, axis=1), 0, reps=1)
        host = host_cpu()
        result["cpu_baseline"] = {
            "value": n / tc / 1e9, "unit": "GDoF/s", "cores": threads, "kind": "port",
            "sample": f"{reps} FP64 Stokes applies of the {args.size}^3 workload ({tc:.3f} s each) on {threads} "
                      f"threads, SIMD across cells (8 per vector, {O.simd_width_name()}); 1 thread: "
                      f"{n / t1 / 1e9:.4f} GDoF/s ({reps1} applies); one FP64 A-block V-cycle of the oracle "
                      f"(scalar cell loop, {threads} threads): {t_vc:.2f} s",
            "value_1_thread": n / t1 / 1e9, "host": host, "vcycle_a_fp64_seconds": t_vc,
            "vcycle_setup_seconds": t_vset}
        log(f"CPU baseline {n / tc / 1e9:.4f} GDoF/s on {threads} threads ({reps} x {tc:.3f}s), "
            f"1 thread {n / t1 / 1e9:.4f} GDoF/s, A V-cycle {t_vc:.2f} s; host {host}")
        if "solve" in result:
            eta_h = 1.0 / np.mean(1.0 / eta, axis=1)
            for prec in ("fp32", "fp64"):
                rec = result["solve"][f"{prec}_gmg"]
                xs = rec.pop("_x")
                xs[nu:] = (xs[nu:] - rec["pressure_normalization_adjustment"]) / s
                r = rhs_h - O.stokes_vmult(p, cfg.n_levels - 1, eta_h, xs)
                rec["oracle_relative_residual"] = float(np.linalg.norm(r) / rec["r0"])
                log(f"{prec.upper()}-GMG solution under the oracle operator: ||b - Kx|| / ||r0|| = "
                    f"{rec['oracle_relative_residual']:.3e}")
    elif "solve" in result:
        for prec in ("fp32", "fp64"):
            result["solve"][f"{prec}_gmg"].pop("_x", None)

    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
