"""GPU parity of the free-surface stabilization term
(StokesOperator::local_apply_boundary_face, matrix_free_operators.cc:563-602)
against the oracle: the Stokes apply, the right-hand-side correction and a
solve with the term on."""
import numpy as np
import pytest

import oracle as O
from test_gpu_parity import FP64, rel, setup_pair

pytestmark = pytest.mark.gpu

# (dim, degree, coarse, levels, extent, bc, free-surface faces)
FS_CASES = [
    (3, 2, (2, 1, 1), 3, (2.0, 1.0, 1.0), (1, 1, 1, 1, 1, 0), (0, 0, 0, 0, 0, 1)),  # box, free top
    (3, 2, (3, 2, 1), 3, (1.5, 1.0, 0.5), (0, 1, 2, 0, 1, 0), (1, 0, 0, 1, 0, 1)),  # three faces, partial tiles
    (2, 2, (2, 1, 1), 5, (2.0, 1.0, 1.0), (1, 1, 1, 0, 0, 0), (0, 0, 0, 1, 0, 0)),  # 2D, free top
    (2, 3, (1, 2, 1), 4, (1.0, 2.0, 1.0), (0, 1, 2, 0, 0, 0), (1, 0, 0, 1, 0, 0)),  # 2D Q3
    (3, 3, (1, 1, 1), 3, (1.0, 1.0, 1.0), (1, 1, 1, 1, 1, 0), (0, 0, 0, 0, 0, 1)),  # 3D Q3
]


def _stab(p, faces, rng):
    return rng.normal(size=O.free_surface_stab_size(p, faces))


@pytest.mark.parametrize("case", FS_CASES)
def test_free_surface_stokes_vmult(case):
    *geo, faces = case
    p, h, eta, sz, rng = setup_pair(*geo, s=1.7)
    stab = _stab(p, faces, rng)
    h.set_free_surface(faces, stab)
    x = rng.normal(size=sz["nu"] + sz["np"])
    ref = O.stokes_vmult(p, geo[3] - 1, eta, x)
    ref[:sz["nu"]] += O.free_surface_apply(p, faces, stab, x[:sz["nu"]])
    assert rel(h.stokes_vmult(x), ref) < FP64
    # switched off again: the plain operator
    h.set_free_surface((0,) * 6)
    assert rel(h.stokes_vmult(x), O.stokes_vmult(p, geo[3] - 1, eta, x)) < FP64


@pytest.mark.parametrize("case", [FS_CASES[0], FS_CASES[2]])
def test_free_surface_rhs_correction(case):
    # correct_stokes_rhs with the face term (local_smoothing.cc:709-744):
    # rhs += -A u0 + (t . u0)(phi . n), u0 read as stored
    *geo, faces = case
    p, h, eta, sz, rng = setup_pair(*geo, s=1.3)
    stab = _stab(p, faces, rng)
    h.set_free_surface(faces, stab)
    con = O.constrained(p, geo[3] - 1)
    u0 = np.zeros(sz["nu"] + sz["np"])
    u0[:sz["nu"]][con] = rng.normal(size=con.sum())
    rhs = rng.normal(size=u0.size)
    ref = rhs + O.stokes_rhs_correction(p, eta, u0)
    ref[:sz["nu"]] += O.free_surface_apply(p, faces, stab, u0[:sz["nu"]], sign=-1.0, plain=True) * ~con
    assert rel(h.correct_stokes_rhs(rhs, u0), ref) < FP64


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_free_surface_solve(prec):
    # the solve with the term on reaches the tolerance under the oracle's
    # operator A + F (the term is not symmetric: GMRES, as in the reference)
    *geo, faces = FS_CASES[0]
    p, h, eta, sz, rng = setup_pair(*geo, s=1.0, prec=prec)
    t = np.zeros((O.free_surface_stab_size(p, faces) // 3, 3))
    t[:, 2] = -0.3  # rho dt theta |g| g_hat with g = -e_z
    h.set_free_surface(faces, t)
    h.build_preconditioner()
    rhs = O.rhs_buoyancy(p, 1.0 + rng.uniform(size=(sz["cells"], O.n_qp(p))), np.array([0.0, 0.0, -1.0]))
    sol, out = h.solve(rhs, tolerance=1e-8)
    assert out.converged
    nu = sz["nu"]
    x = sol.copy()
    x[nu:] = (x[nu:] - out.pressure_normalization_adjustment) / 1.0
    r = O.stokes_vmult(p, geo[3] - 1, eta, x)
    r[:nu] += O.free_surface_apply(p, faces, t.reshape(-1), x[:nu])
    assert np.linalg.norm(rhs - r) <= 1.05e-8 * out.r0_norm
