"""GPU parity of the mesh-deformation displacement solve
(MeshDeformationHandler::compute_mesh_displacements_gmg_for_degree,
source/mesh_deformation/interface.cc:1076-1360) against the oracle: the
vector Laplace level operator, the multigrid set-up and V-cycle, and the CG
solve with prescribed boundary displacements."""
import numpy as np
import pytest

import oracle as O
from aspect_amd import StokesMatrixFreeHandler, make_box
from test_gpu_parity import FP64, rel

pytestmark = pytest.mark.gpu

# (dim, degree, coarse, levels, extent, mesh bc: 0 none, 1 no normal displacement, 2 zero, 3 prescribed)
MESH_CASES = [
    (3, 2, (1, 1, 1), 4, (1.0, 1.0, 1.0), (1, 1, 1, 1, 2, 3)),   # box, free top surface
    (3, 2, (3, 2, 1), 3, (1.5, 1.0, 0.5), (0, 1, 2, 1, 2, 3)),   # partial tiles, an open face
    (2, 2, (2, 1, 1), 5, (2.0, 1.0, 1.0), (1, 1, 2, 3, 0, 0)),   # 2D
    (2, 3, (1, 2, 1), 4, (1.0, 2.0, 1.0), (1, 1, 2, 3, 0, 0)),   # 2D Q3
    (3, 3, (1, 1, 1), 3, (1.0, 1.0, 1.0), (1, 1, 1, 1, 2, 3)),   # 3D Q3
]


def _pair(dim, deg, coarse, nl, ext, mbc):
    box = make_box(dim, deg, coarse, nl, ext, (1,) * 6)
    h = StokesMatrixFreeHandler(box, "fp64")
    p = O.mesh_problem(O.make_problem(dim, deg, coarse, nl, ext, (1,) * 6), mbc)
    return h, p


def _u0(p, level, rng):
    """prescribed values on the constrained DoFs of the mesh problem (others zero)"""
    s = O.sizes(p, level)
    con = O.constrained(p, level)
    u0 = np.zeros(s["nu"])
    u0[con] = rng.normal(size=con.sum())
    return u0


@pytest.mark.parametrize("case", MESH_CASES)
def test_mesh_laplace_level_operator(case):
    h, p = _pair(*case)
    mbc = case[5]
    rng = np.random.default_rng(5)
    for level in (case[3] - 1, max(0, case[3] - 3)):
        x = rng.normal(size=O.sizes(p, level)["nu"])
        assert rel(h.mesh_laplace_vmult(mbc, level, x), O.laplace_vmult(p, level, x)) < FP64


@pytest.mark.parametrize("case", MESH_CASES)
def test_mesh_multigrid(case):
    # smoother set-up (degrees, eigenvalue-CG iterations, intervals) and one V-cycle
    h, p = _pair(*case)
    mbc = case[5]
    x = np.random.default_rng(6).normal(size=O.sizes(p, case[3] - 1)["nu"])
    y, info = h.mesh_vcycle(mbc, x)
    ref_info, ref_its = O.mesh_smoother_info(p)
    np.testing.assert_array_equal(info[:, 4], ref_info[:, 4])
    np.testing.assert_array_equal(info[:, 5], ref_its)
    np.testing.assert_allclose(info[:, :4], ref_info[:, :4], rtol=1e-9)
    assert rel(y, O.mesh_vcycle(p, x)) < 1e-10


@pytest.mark.parametrize("case", MESH_CASES)
def test_mesh_deformation_solve(case):
    h, p = _pair(*case)
    mbc = case[5]
    u0 = _u0(p, case[3] - 1, np.random.default_rng(7))
    sol, its, (res, req) = h.mesh_deformation_solve(mbc, u0, tolerance=1e-10)
    ref, rits, rres = O.mesh_solve(p, u0, tolerance=1e-10)
    assert res <= req
    assert abs(its - rits) <= max(1, 0.1 * rits)
    assert rel(sol, ref) < 1e-8
    con = O.constrained(p, case[3] - 1)
    assert np.array_equal(sol[con], u0[con])
