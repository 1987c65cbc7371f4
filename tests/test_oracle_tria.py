"""CPU tests that pin the general-mesh oracle (oracle/tria.py) and the mesh
builder (aspect_amd/tria.py: DoF maps, hanging-node, Dirichlet and
no-normal-flux lines, geometry at the quadrature points).

Each check is chosen so that a plausible slip fails it:
  * meshes without hanging nodes against the box oracle (oracle.cpp, pinned by
    brute-force element matrices and closed forms in test_oracle.py): plain
    boxes with every boundary type, and a curved (chunk) geometry, for the
    Stokes, A, B^T and mass operators — a transposed J^{-1}, a wrong DoF map or
    a wrong constraint semantic shows here;
  * textbook hanging-node counts (two cells, one refined: 2 / 1 hanging nodes
    in 2D, 16 / 5 in 3D);
  * the hanging-node lines reproduce every global Q2 (velocity) / Q1 (pressure)
    tensor polynomial at the hanging nodes (a wrong weight or a node matched to
    the wrong face fails);
  * on locally refined meshes the constrained operators reproduce closed-form
    integrals over the domain of polynomial fields: u^T A v = int 2 eta
    eps(u):eps(v) [- 2/3 eta div u div v], p^T B u = -s int p div u,
    p^T M p = s^2 int p^2 / eta (a missing C^T on the test side, a dropped
    hanging contribution or a cell-size error fails);
  * the shell's maps: Jacobians against finite differences of the points,
    node counts of the cubed-sphere / annulus lattice in closed form (a seam
    between panels that does not match adds nodes), the shell volume, and
    u . n = 0 on the no-normal-flux nodes after the constraints are applied.
"""
import numpy as np
import pytest

import oracle as O
from aspect_amd import problems as P
from aspect_amd import tria as T
from oracle import tria as OT

BCMAP = {0: None, 1: "tangential", 2: "zero"}


@pytest.fixture(autouse=True)
def _clear_geometry():
    O.set_geometry(None)
    yield
    O.set_geometry(None)


def box_mesh(dim, g, bc, extent):
    t = T.Tria(T.BoxCoarse(dim, (1,) * dim, extent), g)
    return t.build({f: BCMAP[b] for f, b in enumerate(bc[:2 * dim]) if BCMAP[b]})


def box_perm(mesh, dim, n, extent):
    """Box-vector index of every general-mesh DoF (velocity, pressure)."""
    h = np.array(extent[:dim], dtype=float) / n
    iu = np.round(mesh.x_u / (h / 2)).astype(np.int64)
    nv = 2 * n + 1
    lin_u = iu[:, 0] + nv * iu[:, 1] + (nv * nv * iu[:, 2] if dim == 3 else 0)
    ip = np.round(mesh.x_p / h).astype(np.int64)
    lin_p = ip[:, 0] + (n + 1) * ip[:, 1] + ((n + 1) ** 2 * ip[:, 2] if dim == 3 else 0)
    vn = nv ** dim
    return np.concatenate([c * vn + lin_u for c in range(dim)]), lin_p, dim * vn


@pytest.mark.parametrize("dim,bc,compressible", [(2, (1, 2, 1, 0, 0, 0), False), (3, (1, 1, 2, 1, 0, 1), True),
                                                 (3, (2, 2, 2, 2, 2, 2), False)])
def test_box_mesh_matches_box_oracle(dim, bc, compressible):
    g = 2
    n = 1 << g
    extent = (1.3, 0.9, 1.1)[:dim]
    mesh = box_mesh(dim, g, bc, extent)
    assert len(mesh.hanging_u) == 0 and len(mesh.hanging_p) == 0
    p = O.make_problem(dim, 2, (1,) * dim, g + 1, extent, bc, compressible=compressible, pressure_scaling=1.7)
    rng = np.random.default_rng(dim)
    eta = np.exp(rng.normal(size=(mesh.n_cells, 3 ** dim)))
    ops = OT.Operators(mesh, eta, pressure_scaling=1.7, compressible=compressible)
    pu, pp, nu_box = box_perm(mesh, dim, n, extent)
    x = rng.normal(size=nu_box + (n + 1) ** dim)
    xt = np.concatenate([x[pu], x[nu_box + pp]])
    y = O.stokes_vmult(p, g, eta, x)
    yt = ops.stokes(xt)
    ref = np.concatenate([y[pu], y[nu_box + pp]])
    assert np.max(np.abs(yt - ref)) <= 1e-12 * np.max(np.abs(ref))
    ya = O.ablock_vmult(p, g, eta, x[:nu_box])
    assert np.max(np.abs(ops.ablock(xt[:mesh.n_dofs_u]) - ya[pu])) <= 1e-12 * np.max(np.abs(ya))
    yb = O.bt_vmult(p, g, x[nu_box:])
    assert np.max(np.abs(ops.bt(xt[mesh.n_dofs_u:]) - yb[pu])) <= 1e-12 * np.max(np.abs(yb))
    ym = O.mass_vmult(p, g, eta, x[nu_box:])
    assert np.max(np.abs(ops.mass(xt[mesh.n_dofs_u:]) - ym[pp])) <= 1e-12 * np.max(np.abs(ym))


@pytest.mark.parametrize("dim", [2, 3])
def test_curved_mesh_matches_box_oracle(dim):
    g = 2
    n = 1 << g
    bc = (2, 0, 2, 0, 0, 2)
    mesh = box_mesh(dim, g, bc, (1.0, 1.0, 1.0))
    tabs = [P.mapped_geometry(dim, 2, (1 << l,) * dim, P.chunk_map(dim)) for l in range(g + 1)]
    mesh.geometry = tabs[-1]
    p = O.make_problem(dim, 2, (1,) * dim, g + 1, (1.0, 1.0, 1.0), bc, pressure_scaling=0.8)
    O.set_geometry(tabs)
    rng = np.random.default_rng(10 + dim)
    eta = np.exp(rng.normal(size=(mesh.n_cells, 3 ** dim)))
    ops = OT.Operators(mesh, eta, pressure_scaling=0.8)
    pu, pp, nu_box = box_perm(mesh, dim, n, (1.0, 1.0, 1.0))
    x = rng.normal(size=nu_box + (n + 1) ** dim)
    xt = np.concatenate([x[pu], x[nu_box + pp]])
    y = O.stokes_vmult(p, g, eta, x)
    ref = np.concatenate([y[pu], y[nu_box + pp]])
    assert np.max(np.abs(ops.stokes(xt) - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("dim,nu,np_", [(2, 2, 1), (3, 16, 5)])
def test_hanging_counts_two_cells(dim, nu, np_):
    t = T.Tria(T.BoxCoarse(dim, (2, 1, 1), (2.0, 1.0, 1.0)), 0)
    t.refine(np.array([True, False]))
    m = t.build()
    assert len(m.hanging_u) == nu and len(m.hanging_p) == np_
    assert m.n_cells == 2 ** dim + 1


# ------------------------------------------------------------------ polynomial fields
class Poly:
    """Tensor polynomial sum_a c_a x^a0 y^a1 z^a2 with exponents up to deg per variable."""

    def __init__(self, dim, deg, rng):
        self.dim, self.deg = dim, deg
        self.c = rng.normal(size=(deg + 1,) * dim)

    def _terms(self, x, dvar=None):
        out = np.zeros(len(x))
        for a in np.ndindex(*self.c.shape):
            t = np.full(len(x), self.c[a])
            for d in range(self.dim):
                e = a[d]
                if d == dvar:
                    t = t * (e * x[:, d] ** (e - 1) if e > 0 else 0.0)
                else:
                    t = t * x[:, d] ** e
            out += t
        return out

    def __call__(self, x):
        return self._terms(x)

    def grad(self, x):
        return np.stack([self._terms(x, d) for d in range(self.dim)], axis=1)


def domain_rule(dim, extent, npts=5):
    x, w = np.polynomial.legendre.leggauss(npts)
    x, w = 0.5 * (x + 1), 0.5 * w
    pts = T._lattice(dim, x) * np.array(extent[:dim])
    wt = np.prod(T._lattice(dim, w), axis=1) * np.prod(extent[:dim])
    return pts, wt


def refined_box(dim, seed):
    reps = (3, 2, 2)[:dim]
    extent = (1.5, 1.0, 1.2)[:dim]
    t = T.Tria(T.BoxCoarse(dim, reps, extent), 1)
    rng = np.random.default_rng(seed)
    t.refine(rng.random(t.n_cells) < 0.35)
    return t.build(), extent


@pytest.mark.parametrize("dim", [2, 3])
def test_hanging_lines_reproduce_polynomials(dim):
    mesh, _ = refined_box(dim, 3)
    assert len(mesh.hanging_u) > 0 and len(mesh.hanging_p) > 0
    rng = np.random.default_rng(4)
    for c in range(dim):
        f = Poly(dim, 2, rng)
        x = np.zeros(mesh.n_dofs_u)
        x[c * mesh.n_nodes_u:(c + 1) * mesh.n_nodes_u] = f(mesh.x_u)
        x0 = x.copy()
        x0[c * mesh.n_nodes_u + mesh.hanging_u] = 0.0
        xr = OT.resolve(mesh.con_u, x0)
        assert np.max(np.abs(xr - x)) <= 1e-12 * np.max(np.abs(x))
    q = Poly(dim, 1, rng)
    pv = q(mesh.x_p)
    p0 = pv.copy()
    p0[mesh.hanging_p] = 0.0
    assert np.max(np.abs(OT.resolve(mesh.con_p, p0) - pv)) <= 1e-12 * np.max(np.abs(pv))


@pytest.mark.parametrize("dim,compressible", [(2, False), (3, False), (3, True)])
def test_hanging_operators_closed_forms(dim, compressible):
    mesh, extent = refined_box(dim, 5 + dim)
    rng = np.random.default_rng(6)
    ea = rng.uniform(0.2, 0.6, size=dim)
    eta_f = lambda x: 1.0 + x @ ea                                   # linear viscosity
    s = 1.3
    ops = OT.Operators(mesh, eta_f(mesh.x_q.reshape(-1, dim)).reshape(mesh.n_cells, -1), pressure_scaling=s,
                       compressible=compressible)
    U = [Poly(dim, 2, rng) for _ in range(dim)]
    V = [Poly(dim, 2, rng) for _ in range(dim)]
    Pp = Poly(dim, 1, rng)

    def nodal(F):
        x = np.concatenate([f(mesh.x_u) for f in F])
        for c in range(dim):
            x[c * mesh.n_nodes_u + mesh.hanging_u] = 0.0
        return x

    xu, xv = nodal(U), nodal(V)
    pp = Pp(mesh.x_p)
    pp[mesh.hanging_p] = 0.0
    X, W = domain_rule(dim, extent)
    gu = np.stack([f.grad(X) for f in U], axis=1)                    # [pt, comp, d]
    gv = np.stack([f.grad(X) for f in V], axis=1)
    eu = 0.5 * (gu + gu.transpose(0, 2, 1))
    ev = 0.5 * (gv + gv.transpose(0, 2, 1))
    div_u, div_v = np.trace(gu, axis1=1, axis2=2), np.trace(gv, axis1=1, axis2=2)
    ex = eta_f(X)
    a_uv = np.sum(W * ex * (2 * np.sum(eu * ev, axis=(1, 2)) - (2.0 / 3.0 if compressible else 0.0) * div_u * div_v))
    a_uu = np.sum(W * ex * (2 * np.sum(eu * eu, axis=(1, 2)) - (2.0 / 3.0 if compressible else 0.0) * div_u ** 2))
    assert abs(xv @ ops.ablock(xu) - a_uv) <= 1e-11 * abs(a_uu)
    assert abs(xu @ ops.ablock(xu) - a_uu) <= 1e-11 * abs(a_uu)
    b = -s * np.sum(W * Pp(X) * div_u)
    assert abs(xu @ ops.bt(pp) - b) <= 1e-11 * np.sqrt(a_uu * np.sum(W * Pp(X) ** 2))
    st = ops.stokes(np.concatenate([xu, pp]))
    assert abs(pp @ st[mesh.n_dofs_u:] - b) <= 1e-11 * np.sqrt(a_uu * np.sum(W * Pp(X) ** 2))
    # mass with a constant viscosity (p^2 / eta stays polynomial)
    opc = OT.Operators(mesh, np.full((mesh.n_cells, 3 ** dim), 2.5), pressure_scaling=s)
    m = s * s * np.sum(W * Pp(X) ** 2) / 2.5
    assert abs(pp @ opc.mass(pp) - m) <= 1e-12 * m


# ------------------------------------------------------------------ the shell
@pytest.mark.parametrize("dim", [2, 3])
def test_shell_jacobians_against_differences(dim):
    sh = T.ShellCoarse(dim)
    rng = np.random.default_rng(7)
    root = rng.integers(0, sh.n_roots, size=50)
    xi = rng.uniform(0.05, 0.95, size=(50, dim))
    J = sh.jacobians(root, xi)
    h = 1e-6
    for e in range(dim):
        d = np.zeros(dim)
        d[e] = h
        fd = (sh.points(root, xi + d) - sh.points(root, xi - d)) / (2 * h)
        assert np.max(np.abs(fd - J[:, :, e])) <= 1e-8
    r = np.linalg.norm(sh.points(root, xi), axis=1)
    assert np.all((r >= sh.r0 - 1e-14) & (r <= sh.r1 + 1e-14))


@pytest.mark.parametrize("dim,g", [(2, 1), (3, 0), (3, 1)])
def test_shell_node_counts_and_volume(dim, g):
    m = T.Tria(T.ShellCoarse(dim), g).build()
    if dim == 3:
        M, Mp = 2 * 2 * (1 << g), 2 * (1 << g)
        assert m.n_nodes_u == (6 * M * M + 2) * (2 * 4 * (1 << g) + 1)
        assert m.n_nodes_p == (6 * Mp * Mp + 2) * (4 * (1 << g) + 1)
        vol = 4.0 / 3.0 * np.pi * (1.0 - 0.55 ** 3)
    else:
        assert m.n_nodes_u == 12 * (1 << g) * 2 * (2 * (1 << g) + 1)
        assert m.n_nodes_p == 12 * (1 << g) * ((1 << g) + 1)
        vol = np.pi * (1.0 - 0.55 ** 2)
    jxw = m.geometry[:, :, -1].sum()
    assert abs(jxw - vol) <= 1e-4 * vol
    assert np.all(m.geometry[:, :, -1] > 0)


@pytest.mark.parametrize("dim", [2, 3])
def test_shell_no_normal_flux_and_hanging_lines(dim):
    t = T.Tria(T.ShellCoarse(dim), 1 if dim == 2 else 0)
    c = t.cell_centers()
    t.refine((c[:, 0] > 0.25) & (np.linalg.norm(c, axis=1) > 0.7))
    m = t.build({0: "tangential", 1: "tangential"})
    assert len(m.hanging_u) > 0
    rng = np.random.default_rng(8)
    x = OT.resolve(m.con_u, rng.normal(size=m.n_dofs_u))
    nn = m.n_nodes_u
    hang = set(m.hanging_u.tolist())
    for bid in (0, 1):
        ids = np.array([i for i in m.boundary_nodes[bid] if i not in hang])
        u = np.stack([x[c * nn + ids] for c in range(dim)], axis=1)
        n = m.x_u[ids] / np.linalg.norm(m.x_u[ids], axis=1)[:, None]
        assert np.max(np.abs(np.sum(u * n, axis=1))) <= 1e-12 * np.max(np.abs(u))
    # hanging lines (pressure: no other lines) form a partition of unity
    dofs, offs, _, w = m.con_p
    sums = np.add.reduceat(w, offs[:-1])
    assert np.max(np.abs(sums - 1.0)) <= 1e-13


# ------------------------------------------------------------------ multigrid hierarchy
def box_hierarchy(dim, bc, seed=1):
    t = T.Tria(T.BoxCoarse(dim, (3, 2, 2)[:dim], (1.5, 1.0, 1.2)[:dim]), 1)
    t.refine(np.random.default_rng(seed).random(t.n_cells) < 0.35)
    return T.hierarchy(t, bc)


@pytest.mark.parametrize("dim", [2, 3])
def test_transfers_are_galerkin_on_boxes(dim):
    """On affine cells with a constant viscosity the spaces are nested and the
    quadrature exact, so P^T K_fine P = K_coarse on the free coarse DoFs for
    the constrained A block and mass (a wrong embedding weight, a dropped
    coarse constraint or a fine constrained row left in P fails)."""
    _, meshes, pu, pp = box_hierarchy(dim, {0: "zero", 2: "tangential", 3: "tangential"})
    for l in range(1, len(meshes)):
        of = OT.Operators(meshes[l], np.full((meshes[l].n_cells, 3 ** dim), 1.7), pressure_scaling=1.2)
        oc = OT.Operators(meshes[l - 1], np.full((meshes[l - 1].n_cells, 3 ** dim), 1.7), pressure_scaling=1.2)
        for P, K_f, C_f, fx_f, K_c, C_c, fx_c in ((pu[l], of.A, of.Cu, of.fixed_u, oc.A, oc.Cu, oc.fixed_u),
                                                   (pp[l], of.M, of.Cp, of.fixed_p, oc.M, oc.Cp, oc.fixed_p)):
            Af = OT.constrained_matrix(K_f, C_f, fx_f)
            Ac = OT.constrained_matrix(K_c, C_c, fx_c)
            G = (P.T @ Af @ P).toarray()
            free = np.ones(Ac.shape[0], bool)
            free[fx_c] = False
            ref = Ac.toarray()[np.ix_(free, free)]
            assert np.max(np.abs(G[np.ix_(free, free)] - ref)) <= 1e-12 * np.max(np.abs(ref))
            assert np.all(G[~free] == 0) and np.all(G[:, ~free] == 0)


def shell_hierarchy(dim, g, bc):
    t = T.Tria(T.ShellCoarse(dim), g)
    c = t.cell_centers()
    t.refine((c[:, 0] > 0.2) & (np.linalg.norm(c, axis=1) > 0.75))
    return T.hierarchy(t, bc)


def level_viscosities(meshes):
    return [np.exp(2.0 * np.sin(3.0 * np.linalg.norm(m.x_q, axis=2))) for m in meshes]


def test_shell_transfers_reproduce_constants():
    """Curved cells: constants (velocity per component with open boundaries,
    pressure) are prolongated exactly at the free fine DoFs."""
    _, meshes, pu, pp = shell_hierarchy(3, 0, None)
    mf, mc = meshes[1], meshes[0]
    for c in range(3):
        x = np.zeros(mc.n_dofs_u)
        x[c * mc.n_nodes_u:(c + 1) * mc.n_nodes_u] = 1.0
        y = pu[1] @ x
        free = np.ones(mf.n_dofs_u, bool)
        free[mf.con_u[0]] = False
        ref = np.zeros(mf.n_dofs_u)
        ref[c * mf.n_nodes_u:(c + 1) * mf.n_nodes_u] = 1.0
        assert np.max(np.abs(y[free] - ref[free])) <= 1e-13
    y = pp[1] @ np.ones(mc.n_dofs_p)
    free = np.ones(mf.n_dofs_p, bool)
    free[mf.con_p[0]] = False
    assert np.max(np.abs(y[free] - 1.0)) <= 1e-13


def test_oracle_vcycle_properties():
    """The V-cycle is symmetric (same pre- and post-smoother, R = P^T) and a
    contraction in the energy norm; the Chebyshev eigenvalue estimates bound
    the largest eigenvalue of D^{-1} A from below (Lanczos) before the 1.2
    safety factor."""
    from scipy.sparse.linalg import eigsh
    from scipy.sparse import diags
    _, meshes, pu, pp = box_hierarchy(3, {0: "zero", 2: "tangential", 4: "zero"})
    g = OT.GMG(meshes, pu, pp, level_viscosities(meshes), pressure_scaling=1.0)
    rng = np.random.default_rng(3)
    for kind in (0, 1):
        M = g.mats[kind][-1]
        n = M.shape[0]
        x, y = rng.normal(size=n), rng.normal(size=n)
        vx, vy = g.vcycle(kind, x), g.vcycle(kind, y)
        assert abs(y @ vx - x @ vy) <= 1e-10 * np.sqrt(abs((x @ vx) * (y @ vy)))
        e = rng.normal(size=n)
        e2 = e - g.vcycle(kind, M @ e)
        assert np.sqrt(e2 @ (M @ e2)) <= 0.6 * np.sqrt(e @ (M @ e))
        for l in range(1, len(meshes)):
            Ml = g.mats[kind][l]
            d = 1.0 / np.sqrt(g.diag[kind][l])
            lam = eigsh(diags(d) @ Ml @ diags(d), k=1, which="LA", return_eigenvectors=False)[0]
            est = g.cheb[kind][l].lmax / 1.2
            assert est <= lam * (1 + 1e-10) and est >= 0.85 * lam


def stokes_matrix(o):
    """The constrained block matrix [[A, B^T], [B, 0]] with identity rows on the
    constrained DoFs (the operator the solve inverts)."""
    from scipy.sparse import bmat, diags
    ku = np.ones(o.nu)
    ku[o.fixed_u] = 0.0
    kp = np.ones(o.np)
    kp[o.fixed_p] = 0.0
    Auu = OT.constrained_matrix(o.A, o.Cu, o.fixed_u)
    Bpu = diags(kp) @ (o.Cp.T @ o.B @ o.Cu) @ diags(ku)
    return bmat([[Auu, Bpu.T], [Bpu, diags(1.0 - kp)]]).tocsr()


def test_oracle_solve_matches_direct_solver():
    from scipy.sparse.linalg import spsolve
    _, meshes, pu, pp = box_hierarchy(3, {0: "zero", 2: "tangential", 4: "zero"})
    g = OT.GMG(meshes, pu, pp, level_viscosities(meshes), pressure_scaling=1.0)
    top = g.ops[-1]
    K = stokes_matrix(top)
    rng = np.random.default_rng(4)
    x0 = rng.normal(size=K.shape[0])
    assert np.max(np.abs(K @ x0 - top.stokes(x0))) <= 1e-12 * np.max(np.abs(K @ x0))
    b = rng.normal(size=K.shape[0])
    b[top.fixed_u] = 0.0
    b[top.nu + top.fixed_p] = 0.0
    x, its, res = OT.gmres(top.stokes, g.block_preconditioner, b, 1e-10 * np.linalg.norm(b), 300, 100)
    xd = spsolve(K.tocsc(), b)
    assert its < 60
    assert np.linalg.norm(x - xd) <= 1e-7 * np.linalg.norm(xd)


# ------------------------------------------------------------------ periodic boxes
def merged_mesh(mesh):
    """The same mesh with every pure periodic line (one master, weight 1, same
    component) replaced by identifying the two nodes in the cell lists."""
    import copy
    m = copy.copy(mesh)
    nn = mesh.n_nodes_u
    dofs, offs, masters, weights = mesh.con_u
    node_map = np.arange(nn)
    keep = np.ones(len(dofs), bool)
    for i, d in enumerate(dofs):
        if offs[i + 1] - offs[i] == 1 and weights[offs[i]] == 1.0 and masters[offs[i]] % nn != d % nn \
                and masters[offs[i]] // nn == d // nn:
            node_map[d % nn] = masters[offs[i]] % nn
            keep[i] = False
    m.cell_u = node_map[mesh.cell_u]
    ld = dofs[keep]
    lo = np.concatenate([[0], np.cumsum(np.diff(offs)[keep])])
    lm = np.concatenate([masters[offs[i]:offs[i + 1]] for i in np.nonzero(keep)[0]]) if keep.any() else masters[:0]
    lw = np.concatenate([weights[offs[i]:offs[i + 1]] for i in np.nonzero(keep)[0]]) if keep.any() else weights[:0]
    m.con_u = (ld, lo, lm, lw)
    dofs, offs, masters, weights = mesh.con_p
    pmap = np.arange(mesh.n_nodes_p)
    keep = np.diff(offs) != 1
    for i in np.nonzero(~keep)[0]:
        pmap[dofs[i]] = masters[offs[i]]
    m.cell_p = pmap[mesh.cell_p]
    m.con_p = (dofs[keep], np.concatenate([[0], np.cumsum(np.diff(offs)[keep])]), masters[:0], weights[:0])
    return m, node_map, pmap


@pytest.mark.parametrize("dim", [2, 3])
def test_periodic_lines_match_node_identification(dim):
    """A periodic box through constraint lines equals the box whose + face
    nodes are identified with the - face nodes in the cell lists."""
    t = T.Tria(T.BoxCoarse(dim, (3, 2, 2)[:dim], (1.5, 1.0, 1.2)[:dim], periodic=(0,)), 1)
    mesh = t.build({2: "zero", 3: "tangential"})
    merged, node_map, pmap = merged_mesh(mesh)
    rng = np.random.default_rng(21)
    eta = np.exp(rng.normal(size=(mesh.n_cells, 3 ** dim)))
    a = OT.Operators(mesh, eta, pressure_scaling=1.3)
    b = OT.Operators(merged, eta, pressure_scaling=1.3)
    x = rng.normal(size=a.nu + a.np)
    slave_u = np.zeros(a.nu, bool)
    for c in range(dim):
        slave_u[c * mesh.n_nodes_u + np.nonzero(node_map != np.arange(mesh.n_nodes_u))[0]] = True
    slave_p = pmap != np.arange(mesh.n_nodes_p)
    x[:a.nu][slave_u] = 0.0
    x[a.nu:][slave_p] = 0.0
    ya, yb = a.stokes(x), b.stokes(x)
    keep = np.concatenate([~slave_u, ~slave_p])
    assert slave_u.any() and slave_p.any()
    assert np.max(np.abs(ya[keep] - yb[keep])) <= 1e-12 * np.max(np.abs(yb))


@pytest.mark.parametrize("dim", [2, 3])
def test_periodic_seam_hanging_nodes(dim):
    """Hanging nodes across the periodic seam: fields of the non-periodic
    coordinates are reproduced there, and the energy of a quadratic field
    independent of x matches its closed form."""
    extent = (1.5, 1.0, 1.2)[:dim]
    t = T.Tria(T.BoxCoarse(dim, (3, 2, 2)[:dim], extent, periodic=(0,)), 1)
    c = t.cell_centers()
    t.refine((c[:, 0] < 0.3) & (c[:, 1] < 0.6))
    mesh = t.build()
    assert len(mesh.hanging_u) > 0
    seam = np.abs(mesh.x_u[mesh.hanging_u, 0]) < 1e-9
    assert seam.any()                       # hanging nodes found across the seam
    rng = np.random.default_rng(22)
    U = []
    for _ in range(dim):
        cf = rng.normal(size=(3,) * (dim - 1))
        U.append(cf)

    def val(cf, x):
        y = x[:, 1:]
        out = np.zeros(len(x))
        for a in np.ndindex(*cf.shape):
            out += cf[a] * np.prod([y[:, k] ** a[k] for k in range(dim - 1)], axis=0)
        return out

    def grad(cf, x):
        y = x[:, 1:]
        g = np.zeros((len(x), dim))
        for a in np.ndindex(*cf.shape):
            for k in range(dim - 1):
                if a[k] == 0:
                    continue
                t_ = cf[a] * a[k] * y[:, k] ** (a[k] - 1)
                for j in range(dim - 1):
                    if j != k:
                        t_ = t_ * y[:, j] ** a[j]
                g[:, k + 1] += t_
        return g

    x = np.concatenate([val(cf, mesh.x_u) for cf in U])
    x0 = x.copy()
    fixed = mesh.con_u[0]
    x0[fixed] = 0.0
    assert np.max(np.abs(OT.resolve(mesh.con_u, x0) - x)) <= 1e-12 * np.max(np.abs(x))
    ops = OT.Operators(mesh, np.full((mesh.n_cells, 3 ** dim), 1.9))
    X, W = domain_rule(dim, extent)
    gu = np.stack([grad(cf, X) for cf in U], axis=1)
    e = 0.5 * (gu + gu.transpose(0, 2, 1))
    ref = np.sum(W * 1.9 * 2 * np.sum(e * e, axis=(1, 2)))
    assert abs(x0 @ ops.ablock(x0) - ref) <= 1e-11 * ref


def test_fully_periodic_translations_in_kernel():
    t = T.Tria(T.BoxCoarse(2, (2, 3), (1.0, 1.5), periodic=(0, 1)), 1)
    c = t.cell_centers()
    t.refine(c[:, 0] + c[:, 1] < 0.8)
    mesh = t.build()
    ops = OT.Operators(mesh, np.exp(np.random.default_rng(23).normal(size=(mesh.n_cells, 9))))
    free = np.ones(ops.nu, bool)
    free[mesh.con_u[0]] = False
    for comp in range(2):
        u = np.zeros(ops.nu)
        u[comp * mesh.n_nodes_u:(comp + 1) * mesh.n_nodes_u] = 1.0
        u[~free] = 0.0
        assert np.max(np.abs(ops.ablock(u)[free])) <= 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_level_viscosities_of_a_linear_field(dim):
    """On affine cells the DGQ0 value of a linear field is its value at the
    cell centre, and the mean over the children of a cell is the value at the
    parent's centre: every level's table is the field at its cell centres."""
    t = T.Tria(T.BoxCoarse(dim, (3, 2, 2)[:dim], (1.5, 1.0, 1.2)[:dim]), 1)
    t.refine(np.random.default_rng(31).random(t.n_cells) < 0.35)
    trias, meshes, _, _ = T.hierarchy(t)
    parents = [None] + [T.parent_map(trias[l], trias[l - 1]) for l in range(1, len(trias))]
    a = np.array([0.7, -0.4, 0.9])[:dim]
    f = lambda x: 3.0 + x @ a
    tabs = OT.level_viscosities(meshes, parents, f(meshes[-1].x_q.reshape(-1, dim)).reshape(meshes[-1].n_cells, -1))
    for l in range(len(meshes) - 1):
        ref = f(trias[l].cell_centers())
        assert np.max(np.abs(tabs[l] - ref[:, None])) <= 1e-13 * np.max(np.abs(ref))
    # harmonic averaging of a cellwise constant table leaves it unchanged
    cst = np.repeat(np.exp(np.random.default_rng(32).normal(size=(meshes[-1].n_cells, 1))), 3 ** dim, axis=1)
    assert np.allclose(OT.level_viscosities(meshes, parents, cst, 2)[-1], cst, rtol=1e-14)


def test_rhs_correction_recovers_prescribed_translation():
    """Prescribed boundary velocity through the right-hand side: with u0 = a
    constant velocity on the zero-velocity boundary DoFs (prescribed values)
    and zero elsewhere, solving the homogeneous constrained system with
    rhs = correction(u0) gives x with x + u0 = the constant field (a
    translation with zero pressure solves the Stokes equations exactly)."""
    from scipy.sparse.linalg import spsolve
    t = T.Tria(T.BoxCoarse(3, (3, 2, 2), (1.5, 1.0, 1.2)), 1)
    c = t.cell_centers()
    t.refine((np.abs(c[:, 0] - 0.75) < 0.3) & (np.abs(c[:, 1] - 0.5) < 0.3) & (np.abs(c[:, 2] - 0.6) < 0.3))
    mesh = t.build({f: "zero" for f in range(6)})
    assert len(mesh.hanging_u) > 0
    ops = OT.Operators(mesh, np.exp(np.random.default_rng(33).normal(size=(mesh.n_cells, 27))),
                       pressure_scaling=1.4)
    vel = np.array([0.3, -1.1, 0.7])
    nn = mesh.n_nodes_u
    bnd = np.unique(np.concatenate([mesh.boundary_nodes[f] for f in range(6)]))
    u0 = np.zeros(ops.nu + ops.np)
    for cc in range(3):
        u0[cc * nn + bnd] = vel[cc]
    # hanging nodes lie inside (their masters are free), so u0 is consistent there
    rhs = OT.stokes_rhs_correction(ops, u0)
    K = stokes_matrix(ops).tolil()
    # enclosed flow (pressure up to a constant; the right-hand side is
    # consistent): pin one free pressure DoF to zero
    fp = np.ones(ops.np, bool)
    fp[mesh.con_p[0]] = False
    j = ops.nu + int(np.nonzero(fp)[0][0])
    K[j, :] = 0.0
    K[:, j] = 0.0
    K[j, j] = 1.0
    rhs[j] = 0.0
    x = spsolve(K.tocsc(), rhs)
    u = x[:ops.nu] + u0[:ops.nu]
    ur = OT.resolve(mesh.con_u, x[:ops.nu]) + u0[:ops.nu]
    free = np.ones(ops.nu, bool)
    free[mesh.con_u[0]] = False
    ref = np.concatenate([np.full(nn, v) for v in vel])
    assert np.max(np.abs(u[free] - ref[free])) <= 1e-8
    assert np.max(np.abs(ur[mesh.hanging_u] - ref[mesh.hanging_u])) <= 1e-8
    p = x[ops.nu:]
    assert np.max(np.abs(p[fp])) <= 1e-8


def test_rhs_correction_of_a_consistent_vector_is_the_operator():
    """u0 already satisfying the (homogeneous) lines: the plain read equals the
    constrained read, so the correction is -K u0 on the free rows."""
    mesh, _ = refined_box(3, 34)
    ops = OT.Operators(mesh, np.exp(np.random.default_rng(35).normal(size=(mesh.n_cells, 27))))
    rng = np.random.default_rng(36)
    x = rng.normal(size=ops.nu + ops.np)
    u0 = np.concatenate([OT.resolve(mesh.con_u, x[:ops.nu]), OT.resolve(mesh.con_p, x[ops.nu:])])
    y = -ops.stokes(u0)
    y[mesh.con_u[0]] = 0.0
    y[ops.nu + mesh.con_p[0]] = 0.0
    assert np.max(np.abs(OT.stokes_rhs_correction(ops, u0) - y)) <= 1e-12 * np.max(np.abs(y))


@pytest.mark.parametrize("dim", [2, 3])
def test_hanging2_golden_fixture(dim):
    """The committed two-cell hanging-node fixture (tests/golden/make_golden.py)
    is reproduced by the mesh builder and the oracle."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", f"hanging2_{dim}d.npz"))
    t = T.Tria(T.BoxCoarse(dim, (2, 1, 1)[:dim], (2.0, 1.0, 1.0)[:dim]), 0)
    t.refine(np.array([True, False]))
    bc = {0: "zero"}
    bc.update({f: "tangential" for f in range(1, 2 * dim)})
    m = t.build(bc)
    assert np.array_equal(m.cell_u, g["cell_u"]) and np.array_equal(m.cell_p, g["cell_p"])
    assert np.array_equal(m.con_u[0], g["con_u_dofs"]) and np.array_equal(m.con_u[2], g["con_u_masters"])
    assert np.allclose(m.con_u[3], g["con_u_weights"], rtol=0, atol=1e-15)
    assert np.allclose(m.geometry, g["geometry"], rtol=1e-15, atol=1e-15)
    ops = OT.Operators(m, g["eta"], pressure_scaling=float(g["pressure_scaling"]))
    assert np.max(np.abs(ops.stokes(g["x"]) - g["stokes"])) <= 1e-13 * np.max(np.abs(g["stokes"]))
