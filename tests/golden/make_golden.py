"""Writes the oracle fixtures the GPU tests compare against at sizes whose
oracle solve takes minutes (run once on the CPU, outputs committed):

  config3_sinker_32.npz  the 128^3 workload's field (16 seeded sinkers, 1e6
                         contrast, 2x2x2 coarse box, harmonic averaging) at
                         32^3: FP64-GMG solve to 1e-8, r0, iteration count,
                         solution norms and 4000 sampled solution entries
  ensemble48.npz         the 48 problems of the 2D ensemble (16^2 .. 256^2
                         cells) at their real sizes: per problem the FP64-GMG
                         iteration count and 400 sampled solution entries

  hanging2_{2,3}d.npz    two cells, one of them refined once (the smallest mesh
                         with hanging nodes), zero velocity on the x- face and
                         no normal flux on the others: cell lists, geometry,
                         closed constraint lines, viscosity, a source vector
                         and the oracle's Stokes, A, B^T and mass applies

    python tests/golden/make_golden.py [config3] [ensemble] [hanging]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from aspect_amd import problems as P  # noqa: E402


def harmonic(eta):
    return 1.0 / np.mean(1.0 / eta, axis=1)


def config3_32():
    cfg = P.config_sinker_128()
    cfg.n_levels = 5
    eta, rho, s = P.fields(cfg)
    p = O.make_problem(3, 2, cfg.coarse, cfg.n_levels, cfg.extent, cfg.bc, pressure_scaling=s)
    rhs = O.rhs_buoyancy(p, rho, cfg.gravity())
    t = time.time()
    sol, res = O.solve(p, harmonic(eta), rhs, params=O.default_params(tolerance=1e-8))
    assert res["converged"]
    nu = O.sizes(p, cfg.n_levels - 1)["nu"]
    rng = np.random.default_rng(32)
    idx = np.sort(np.concatenate([rng.choice(nu, 3000, replace=False),
                                  nu + rng.choice(sol.size - nu, 1000, replace=False)]))
    np.savez_compressed(os.path.join(HERE, "config3_sinker_32.npz"), idx=idx, values=sol[idx],
                        r0=res["r0_norm"], cheap=res["cheap_iterations"], norm_u=np.linalg.norm(sol[:nu]),
                        norm_p=np.linalg.norm(sol[nu:]), pressure_scaling=s)
    print(f"config3 32^3: {res['cheap_iterations']} iterations, {time.time() - t:.0f} s", flush=True)


def hanging2():
    from aspect_amd import tria as T
    from oracle import tria as OT
    for dim in (2, 3):
        t = T.Tria(T.BoxCoarse(dim, (2, 1, 1)[:dim], (2.0, 1.0, 1.0)[:dim]), 0)
        t.refine(np.array([True, False]))
        bc = {0: "zero"}
        bc.update({f: "tangential" for f in range(1, 2 * dim)})
        m = t.build(bc)
        rng = np.random.default_rng(200 + dim)
        eta = np.exp(rng.normal(size=(m.n_cells, 3 ** dim)))
        ops = OT.Operators(m, eta, pressure_scaling=1.25)
        x = rng.normal(size=ops.nu + ops.np)
        np.savez_compressed(os.path.join(HERE, f"hanging2_{dim}d.npz"), cell_u=m.cell_u, cell_p=m.cell_p,
                            n_nodes_u=m.n_nodes_u, n_nodes_p=m.n_nodes_p, geometry=m.geometry,
                            con_u_dofs=m.con_u[0], con_u_offsets=m.con_u[1], con_u_masters=m.con_u[2],
                            con_u_weights=m.con_u[3], con_p_dofs=m.con_p[0], con_p_offsets=m.con_p[1],
                            con_p_masters=m.con_p[2], con_p_weights=m.con_p[3], eta=eta, x=x,
                            stokes=ops.stokes(x), ablock=ops.ablock(x[:ops.nu]), bt=ops.bt(x[ops.nu:]),
                            mass=ops.mass(x[ops.nu:]), pressure_scaling=1.25)
        print(f"hanging2 {dim}D: {m.n_cells} cells, {len(m.hanging_u)} hanging velocity nodes", flush=True)


def ensemble48():
    cfgs = P.config_ensemble(48, seed=4)
    rng = np.random.default_rng(48)
    idx_all, val_all, lens, cheap = [], [], [], []
    for i, c in enumerate(cfgs):
        eta, rho, s = P.fields(c)
        p = O.make_problem(2, 2, c.coarse, c.n_levels, c.extent, c.bc, pressure_scaling=s)
        rhs = O.rhs_buoyancy(p, rho, c.gravity())
        sol, res = O.solve(p, harmonic(eta), rhs, params=O.default_params(tolerance=1e-8))
        assert res["converged"], i
        nu = O.sizes(p, c.n_levels - 1)["nu"]
        idx = np.sort(np.concatenate([rng.choice(nu, 300, replace=False),
                                      nu + rng.choice(sol.size - nu, 100, replace=False)]))
        idx_all.append(idx)
        val_all.append(sol[idx])
        lens.append(idx.size)
        cheap.append(res["cheap_iterations"])
        print(f"ensemble {i}: n={c.meta['n']} {res['cheap_iterations']} iterations", flush=True)
    np.savez_compressed(os.path.join(HERE, "ensemble48.npz"), idx=np.concatenate(idx_all),
                        values=np.concatenate(val_all), idx_len=np.array(lens), cheap=np.array(cheap))


if __name__ == "__main__":
    which = sys.argv[1:] or ["config3", "ensemble", "hanging"]
    if "hanging" in which:
        hanging2()
    if "ensemble" in which:
        ensemble48()
    if "config3" in which:
        config3_32()
