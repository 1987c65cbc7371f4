"""GPU parity of the general-mesh level operators (cell lists, geometry at the
quadrature points, hanging-node / Dirichlet / no-normal-flux lines: the level
operators of stokes_matrix_free_global_coarsening.cc:1455-1625) against the
matrix-based oracle (oracle/tria.py) on locally refined boxes and on the
spherical shell (source/geometry_model/spherical_shell.cc:361-374) with one
adaptive level.  FP64 bar: 1e-12 relative to the largest entry."""
import numpy as np
import pytest

from aspect_amd import TriaOperators
from aspect_amd import tria as T
from oracle import tria as OT

pytestmark = pytest.mark.gpu


def refined_box(dim, seed, bc):
    reps = (3, 2, 2)[:dim]
    t = T.Tria(T.BoxCoarse(dim, reps, (1.5, 1.0, 1.2)[:dim]), 1)
    t.refine(np.random.default_rng(seed).random(t.n_cells) < 0.35)
    return t.build(bc)


def shell(dim, g):
    t = T.Tria(T.ShellCoarse(dim), g)
    c = t.cell_centers()
    # one adaptive level on an outer wedge (a stand-in for the top 20% of |dT/dr|)
    t.refine((c[:, 0] > 0.2) & (np.linalg.norm(c, axis=1) > 0.75))
    return t.build({0: "tangential", 1: "tangential"})


def viscosity(mesh, seed):
    x = mesh.x_q.reshape(mesh.n_cells, -1, mesh.dim)
    r = np.linalg.norm(x, axis=2)
    rng = np.random.default_rng(seed)
    return np.exp(2.0 * np.sin(3.0 * r) + 0.3 * rng.normal(size=r.shape))


CASES = {
    "box3d": lambda: (refined_box(3, 1, {0: "zero", 2: "tangential", 3: "tangential", 5: "zero"}), 1.3, False),
    "box3d_compressible": lambda: (refined_box(3, 2, {1: "tangential", 4: "zero"}), 0.7, True),
    "box2d": lambda: (refined_box(2, 3, {0: "zero", 1: "tangential", 2: "tangential"}), 1.1, False),
    "shell3d": lambda: (shell(3, 0), 1.0, False),
    "shell2d": lambda: (shell(2, 2), 2.0, False),
}


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("case", list(CASES))
def test_tria_operators_match_oracle(case):
    mesh, s, comp = CASES[case]()
    print(f"\n[tria] {case}: cells {mesh.n_cells}, velocity DoFs {mesh.n_dofs_u}, pressure DoFs {mesh.n_dofs_p}, "
          f"hanging {len(mesh.hanging_u)}/{len(mesh.hanging_p)}, lines {len(mesh.con_u[0])}", flush=True)
    assert len(mesh.hanging_u) > 0
    eta = viscosity(mesh, 5)
    ref = OT.Operators(mesh, eta, pressure_scaling=s, compressible=comp)
    gpu = TriaOperators(mesh, pressure_scaling=s, compressible=comp)
    gpu.set_viscosity(eta)
    rng = np.random.default_rng(9)
    x = rng.normal(size=mesh.n_dofs_u + mesh.n_dofs_p)
    u, p = x[:mesh.n_dofs_u], x[mesh.n_dofs_u:]
    assert rel(gpu.vmult("stokes", x), ref.stokes(x)) <= 1e-12
    assert rel(gpu.vmult("ablock", u), ref.ablock(u)) <= 1e-12
    assert rel(gpu.vmult("bt", p), ref.bt(p)) <= 1e-12
    assert rel(gpu.vmult("mass", p), ref.mass(p)) <= 1e-12


def test_tria_apply_is_deterministic_and_symmetric():
    mesh, s, _ = CASES["shell3d"]()
    eta = viscosity(mesh, 6)
    gpu = TriaOperators(mesh, pressure_scaling=s)
    gpu.set_viscosity(eta)
    rng = np.random.default_rng(10)
    nu = mesh.n_dofs_u
    fixed = np.zeros(nu, bool)
    fixed[mesh.con_u[0]] = True
    x, y = rng.normal(size=nu), rng.normal(size=nu)
    x[fixed] = 0.0
    y[fixed] = 0.0
    ax = gpu.vmult("ablock", x)
    assert np.array_equal(ax, gpu.vmult("ablock", x))
    ay = gpu.vmult("ablock", y)
    assert abs(y @ ax - x @ ay) <= 1e-12 * np.sqrt((x @ ax) * (y @ ay))
    assert x @ ax > 0


def test_tria_refined_shell_larger():
    """Shell with one global refinement plus the adaptive level (several
    thousand cells: many full batches and a partial last one)."""
    mesh = shell(3, 1)
    print(f"\n[tria] shell3d g=1: cells {mesh.n_cells}", flush=True)
    eta = viscosity(mesh, 7)
    ref = OT.Operators(mesh, eta, pressure_scaling=1.0)
    gpu = TriaOperators(mesh, pressure_scaling=1.0)
    gpu.set_viscosity(eta)
    x = np.random.default_rng(11).normal(size=mesh.n_dofs_u + mesh.n_dofs_p)
    assert rel(gpu.vmult("stokes", x), ref.stokes(x)) <= 1e-12
