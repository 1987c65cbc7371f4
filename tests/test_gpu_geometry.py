"""GPU parity of the general-geometry path (per quadrature point J^{-1} and
JxW per level, the MappingQ data of MatrixFree: stokes_matrix_free_local_smoothing.cc:1643-1656
and :1699-1760) against the oracle on chunk meshes (annulus sectors and
spherical chunks, chart (r, longitude[, latitude]) of
source/geometry_model/chunk.cc:219-241) with zero-velocity or open faces."""
import numpy as np
import pytest

import oracle as O
from aspect_amd import StokesMatrixFreeHandler, make_box
from aspect_amd import problems as P
from test_gpu_parity import FP32, FP64, fp32_vcycle_bar, rel

pytestmark = pytest.mark.gpu

# (dim, degree, coarse, levels, bc); box extent 1: the chunk map takes xi in [0, 1]^dim
GEO_CASES = [
    (3, 2, (2, 2, 1), 3, (2, 2, 0, 0, 0, 0)),   # spherical chunk 8x8x4, walls at both radii
    (3, 2, (3, 2, 1), 3, (2, 0, 0, 2, 0, 0)),   # 12x8x4: partial tiles
    (2, 2, (2, 3), 4, (2, 2, 0, 0)),            # annulus sector 16x24
    (2, 3, (2, 3), 3, (2, 0, 2, 0)),            # Q3 x Q2
    (3, 3, (1, 1, 1), 3, (2, 2, 0, 0, 0, 0)),   # 3D Q3 x Q2
]


@pytest.fixture(autouse=True)
def _clear_oracle_geometry():
    O.set_geometry(None)
    yield
    O.set_geometry(None)


def _cells(coarse, dim, l):
    return tuple(coarse[d] << l for d in range(dim))


def geo_pair(dim, deg, coarse, nl, bc, s=1.0, prec="fp64", seed=0, mapping=None, averaging="none"):
    """Oracle problem and GPU handler on the same mapped mesh, per-qp viscosity
    from a radial profile with a random perturbation."""
    mapping = mapping or P.chunk_map(dim)
    ext = (1.0, 1.0, 1.0)
    p = O.make_problem(dim, deg, coarse, nl, ext, bc, pressure_scaling=s)
    tabs = [P.mapped_geometry(dim, deg, _cells(coarse, dim, l), mapping) for l in range(nl)]
    O.set_geometry(tabs)
    h = StokesMatrixFreeHandler(make_box(dim, deg, coarse, nl, ext, bc, pressure_scaling=s), prec)
    for l, t in enumerate(tabs):
        h.set_geometry(l, t)
    top = nl - 1
    sz = O.sizes(p, top)
    rng = np.random.default_rng(seed)
    x = P.mapped_qp_coordinates(dim, deg, _cells(coarse, dim, top), mapping)
    r = np.linalg.norm(x, axis=-1)
    eta = (1 + 30 * np.exp(-((r - 0.8) / 0.08) ** 2)) * np.exp(0.3 * rng.normal(size=r.shape))
    h.assemble(eta, averaging)
    eta_o = eta if averaging == "none" else {"arithmetic": lambda e: e.mean(axis=1),
                                            "harmonic": lambda e: 1 / (1 / e).mean(axis=1)}[averaging](eta)
    return p, h, eta_o, sz, rng, tabs, x


@pytest.mark.parametrize("case", GEO_CASES)
def test_geometry_operators_fp64(case):
    p, h, eta, sz, rng, tabs, _ = geo_pair(*case, s=1.7)
    nl = case[3]
    top = nl - 1
    x = rng.normal(size=sz["nu"] + sz["np"])
    assert rel(h.stokes_vmult(x), O.stokes_vmult(p, top, eta, x)) < FP64
    u, q = x[: sz["nu"]], x[sz["nu"]:]
    assert rel(h.ablock_vmult(-1, u), O.ablock_vmult(p, top, eta, u)) < FP64
    assert rel(h.mass_vmult(-1, q), O.mass_vmult(p, top, eta, q)) < FP64
    assert rel(h.bt_vmult(q), O.bt_vmult(p, top, q)) < FP64
    lv = O.level_viscosity(p, eta)
    for lev in range(nl):
        np.testing.assert_allclose(h.level_viscosity(lev), lv[lev], rtol=1e-13)
        s = O.sizes(p, lev)
        ul = rng.normal(size=s["nu"])
        ql = rng.normal(size=s["np"])
        assert rel(h.ablock_vmult(lev, ul), O.ablock_vmult(p, lev, lv[lev], ul)) < FP64
        assert rel(h.mass_vmult(lev, ql), O.mass_vmult(p, lev, lv[lev], ql)) < FP64
        assert rel(h.ablock_diagonal(lev), O.ablock_diagonal(p, lev, lv[lev])) < FP64
        assert rel(h.mass_diagonal(lev), O.mass_diagonal(p, lev, lv[lev])) < FP64
    rho = rng.uniform(0.5, 2.0, size=(sz["cells"], O.n_qp(p)))
    g = np.array([0.3, -1.0, 0.2])[: case[0]]
    assert rel(h.assemble_buoyancy_rhs(rho, g), O.rhs_buoyancy(p, rho, g)) < FP64


@pytest.mark.parametrize("case", GEO_CASES[:3])
def test_geometry_cellwise_viscosity(case):
    # one value per cell (harmonic averaging): the per-cell tables under geometry
    p, h, eta, sz, rng, _, _ = geo_pair(*case, s=1.1, averaging="harmonic")
    x = rng.normal(size=sz["nu"] + sz["np"])
    assert rel(h.stokes_vmult(x), O.stokes_vmult(p, case[3] - 1, eta, x)) < FP64
    lv = O.level_viscosity(p, eta)
    for lev in range(case[3]):
        np.testing.assert_allclose(h.level_viscosity(lev), lv[lev], rtol=1e-13)


@pytest.mark.parametrize("case", GEO_CASES[:3])
def test_geometry_level_operators_fp32(case):
    p, h, eta, sz, rng, _, _ = geo_pair(*case, s=1.3, prec="fp32")
    lv = O.level_viscosity(p, eta)
    for lev in range(case[3]):
        s = O.sizes(p, lev)
        ul = rng.normal(size=s["nu"])
        ql = rng.normal(size=s["np"])
        assert rel(h.ablock_vmult(lev, ul), O.ablock_vmult(p, lev, lv[lev], ul)) < FP32
        assert rel(h.mass_vmult(lev, ql), O.mass_vmult(p, lev, lv[lev], ql)) < FP32
        assert rel(h.ablock_diagonal(lev), O.ablock_diagonal(p, lev, lv[lev])) < FP32


def test_identity_geometry_matches_box_fast_path():
    # identity-map tables through the general path against the box fast path
    # (the plane kernel for 3D Q2, the tiled cell kernel otherwise) on the GPU
    for dim, deg, coarse, nl in ((3, 2, (3, 2, 1), 3), (2, 2, (2, 3), 4), (3, 3, (1, 1, 1), 3)):
        ext = (1.5, 1.0, 0.5)[:dim]
        bc = (2, 0, 0, 2, 0, 0)
        box = make_box(dim, deg, coarse, nl, ext, bc, pressure_scaling=1.3)
        hb = StokesMatrixFreeHandler(box, "fp64")
        hg = StokesMatrixFreeHandler(make_box(dim, deg, coarse, nl, ext, bc, pressure_scaling=1.3), "fp64")
        for l in range(nl):
            hg.set_geometry(l, P.mapped_geometry(dim, deg, _cells(coarse, dim, l), P.affine_map(np.diag(ext))))
        rng = np.random.default_rng(dim + deg)
        cells = int(np.prod(_cells(coarse, dim, nl - 1)))
        eta = np.exp(rng.normal(size=(cells, (deg + 1) ** dim)))
        hb.assemble(eta, "none")
        hg.assemble(eta, "none")
        s = hb.sizes(nl - 1)
        x = rng.normal(size=s["nu"] + s["np"])
        assert rel(hg.stokes_vmult(x), hb.stokes_vmult(x)) < FP64
        assert rel(hg.ablock_diagonal(nl - 1), hb.ablock_diagonal(nl - 1)) < FP64
        assert rel(hg.ablock_vmult(nl - 1, x[: s["nu"]]), hb.ablock_vmult(nl - 1, x[: s["nu"]])) < FP64


def oracle_spread(p, eta, xs, draws=3):
    """The oracle's own spread under last-bit changes: eta * (1 + 1e-15 N(0, 1))
    and other OpenMP thread counts (other reduction orders); per entry of the
    smoother set-up and per V-cycle kind.  The eigenvalue estimate runs CG
    until its own stopping rule; where a coarse level is nearly singular (two
    walls holding a 3x2x1 chunk: lambda_min / lambda_max ~ 2e-3, 74
    iterations) the smallest Ritz value, which sets the level-0 Chebyshev
    interval, moves by up to 1e-4 relative under such changes and the V-cycle
    by as much; elsewhere the spread is ~1e-15."""
    ref = O.smoother_setup(p, eta)
    v = {k: O.vcycle(p, eta, k, x) for k, x in xs.items()}
    s_info = np.zeros(ref[:, :, :4].shape)
    s_v = {k: 0.0 for k in xs}

    def take(e2):
        nonlocal s_info
        r2 = O.smoother_setup(p, e2)
        s_info = np.maximum(s_info, np.abs(r2[:, :, :4] - ref[:, :, :4]) / np.abs(ref[:, :, :4]))
        for k, x in xs.items():
            s_v[k] = max(s_v[k], rel(O.vcycle(p, e2, k, x), v[k]))
    for seed in range(draws):
        take(eta * (1 + 1e-15 * np.random.default_rng(100 + seed).normal(size=eta.shape)))
    nt = O.get_threads()
    try:
        for t in (1, 3):
            O.set_threads(t)
            take(eta)
    finally:
        O.set_threads(nt)
    return ref, v, s_info, s_v


@pytest.mark.parametrize("case", GEO_CASES)
def test_geometry_smoother_setup_and_vcycle(case):
    # bars: 1e-9 (set-up) and 1e-10 (V-cycle) as on the box, or ten times the
    # oracle's own spread under last-bit input perturbations where that is larger
    p, h, eta, sz, rng, _, _ = geo_pair(*case, s=1.9)
    h.build_preconditioner()
    info = h.smoother_info()
    xs = {0: rng.normal(size=sz["nu"]), 1: rng.normal(size=sz["np"])}
    ref, v, s_info, s_v = oracle_spread(p, eta, xs)
    np.testing.assert_allclose(info[:, :, 4], ref[:, :, 4])
    bar = np.maximum(1e-9, 10 * s_info)
    assert (np.abs(info[:, :, :4] - ref[:, :, :4]) <= bar * np.abs(ref[:, :, :4])).all()
    for kind, x in xs.items():
        assert rel(h.vcycle(kind, x), v[kind]) < max(1e-10, 10 * s_v[kind])


@pytest.mark.parametrize("case", GEO_CASES[:3])
def test_geometry_vcycle_fp32(case):
    # FP32 bar, or ten times the oracle's own spread where that is larger
    p, h, eta, sz, rng, _, _ = geo_pair(*case, s=1.9, prec="fp32")
    h.build_preconditioner()
    xs = {0: rng.normal(size=sz["nu"]), 1: rng.normal(size=sz["np"])}
    ref, v, _, s_v = oracle_spread(p, eta, xs)
    np.testing.assert_allclose(h.smoother_info()[:, :, 4], ref[:, :, 4])
    for kind, x in xs.items():
        assert rel(h.vcycle(kind, x), v[kind]) < max(fp32_vcycle_bar(p, eta, ref, kind, x), 10 * s_v[kind])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("case", [GEO_CASES[0], GEO_CASES[2], GEO_CASES[3]])
def test_geometry_solve(case, prec):
    # radial gravity on the chunk: solve to 1e-8, iterations within 10% of the oracle's
    p, h, eta, sz, rng, tabs, xq = geo_pair(*case, s=1.4, prec=prec)
    h.build_preconditioner()
    r = np.linalg.norm(xq, axis=-1)
    rho = 1 + 0.1 * np.sin(7 * xq[..., 0]) * np.cos(5 * xq[..., 1])
    rhs = np.zeros(sz["nu"] + sz["np"])
    for d in range(case[0]):
        e = np.zeros(case[0])
        e[d] = 1.0
        rhs += O.rhs_buoyancy(p, -rho * xq[..., d] / r, e)
    sol, out = h.solve(rhs, tolerance=1e-8)
    ref, res = O.solve(p, eta, rhs, params=O.default_params(tolerance=1e-8))
    assert out.converged and res["converged"]
    assert abs(out.cheap_iterations - res["cheap_iterations"]) <= max(1, 0.1 * res["cheap_iterations"])
    nu = sz["nu"]
    assert rel(sol[:nu], ref[:nu]) < 1e-6
    assert rel(sol[nu:], ref[nu:]) < 1e-6


@pytest.mark.parametrize("case", [GEO_CASES[0], GEO_CASES[2]])
def test_geometry_mesh_deformation(case):
    # the vector Laplace displacement solve on the mapped mesh: prescribed
    # displacement on the outer radius, zero on the inner radius
    dim, deg, coarse, nl, bc = case
    mbc = (2, 3, 0, 0, 0, 0)
    p0, h, _, _, rng, _, _ = geo_pair(*case)
    p = O.mesh_problem(p0, mbc)
    con = O.constrained(p, nl - 1)
    u0 = np.zeros(O.sizes(p, nl - 1)["nu"])
    u0[con] = rng.normal(size=con.sum())
    x = rng.normal(size=u0.size)
    assert rel(h.mesh_laplace_vmult(mbc, nl - 1, x), O.laplace_vmult(p, nl - 1, x)) < FP64
    y, info = h.mesh_vcycle(mbc, x)
    ref_info, ref_its = O.mesh_smoother_info(p)
    np.testing.assert_array_equal(info[:, 4], ref_info[:, 4])
    np.testing.assert_allclose(info[:, :4], ref_info[:, :4], rtol=1e-9)
    assert rel(y, O.mesh_vcycle(p, x)) < 1e-10
    sol, its, (res, req) = h.mesh_deformation_solve(mbc, u0, tolerance=1e-10)
    ref, rits, _ = O.mesh_solve(p, u0, tolerance=1e-10)
    assert res <= req and abs(its - rits) <= max(1, 0.1 * rits)
    assert rel(sol, ref) < 1e-8


def test_geometry_refusals():
    dim, deg, coarse, nl, bc = GEO_CASES[2]
    tab = P.mapped_geometry(dim, deg, _cells(coarse, dim, nl - 1), P.chunk_map(dim))
    h = StokesMatrixFreeHandler(make_box(dim, deg, coarse, nl, (1.0, 1.0), (1, 2, 0, 0)), "fp64")
    with pytest.raises(Exception, match="box geometry"):
        h.set_geometry(nl - 1, tab)
    with pytest.raises(ValueError):
        StokesMatrixFreeHandler(make_box(dim, deg, coarse, nl, (1.0, 1.0), bc), "fp64").set_geometry(nl - 1, tab[:-1])
    p, h, eta, *_ = geo_pair(*GEO_CASES[2])
    with pytest.raises(Exception, match="box geometry"):
        h.assemble(np.ones((tab.shape[0], tab.shape[1])), "project to Q1 only viscosity")
    with pytest.raises(Exception, match="box geometry"):
        h.set_free_surface((0, 0, 0, 1), np.zeros(O.free_surface_stab_size(p, (0, 0, 0, 1))))
    with pytest.raises(Exception, match="box geometry"):
        h.mesh_deformation_solve((2, 3, 1, 1), np.zeros(h.sizes(nl - 1)["nu"]))
