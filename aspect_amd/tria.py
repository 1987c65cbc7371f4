"""General level meshes for the matrix-free Stokes operators.

A mesh here is a set of coarse cells (quadrilaterals / hexahedra), each with
its own smooth map from the reference cell, refined globally and then once
more on flagged cells (one adaptive level, so neighbouring cells differ by at
most one level).  ``Tria.build`` produces what the operators of such a level
need, the host part of ``setup_dofs`` for general meshes
(source/simulator/solver/stokes_matrix_free_global_coarsening.cc:1481-1598):

* the Q2 velocity and Q1 pressure DoF maps (nodes identified by position;
  velocity DoF = component * n_nodes_u + node, component-major as on box
  levels);
* homogeneous Dirichlet lines for "Zero velocity boundary indicators"
  (:1537-1544), no-normal-flux lines with the manifold normal for the
  "Tangential velocity boundary indicators" (:1549-1562), and the Q2 / Q1
  hanging-node lines (:1588, :1616);
* the closed constraint lines (chains through constrained masters
  substituted), as after ``AffineConstraints::close``;
* J^{-1} and JxW at the QGauss(3) points of every cell (MatrixFree with
  update_gradients | update_JxW_values, :1651-1659), taken from the coarse
  cells' maps (an exact manifold description, not a polynomial mapping).

Choices where the reference's behaviour lives in deal.II (not vendored):
a hanging DoF on a boundary keeps its hanging-node line (its masters carry
the boundary lines); at a node where several tangential boundaries meet, only
axis-aligned normals are handled (every normal component is constrained);
a no-normal-flux line constrains the component with the largest normal
entry.

Coarse meshes provided: ``BoxCoarse`` (a box of repetitions, optionally
deformed by a smooth map) and ``ShellCoarse`` (the spherical shell of
source/geometry_model/spherical_shell.cc:361-374: 96 coarse cells in 3D,
six cube-sphere panels of 2 x 2 cells in 4 radial layers; 12 cells in 2D).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components
from scipy.spatial import cKDTree

__all__ = ["BoxCoarse", "ShellCoarse", "Tria", "TriaMesh", "coarsening_sequence", "gauss3", "hierarchy",
           "parent_map", "transfer"]


def gauss3():
    """3-point Gauss-Legendre rule on [0, 1] (QGauss(3) of a degree-2 velocity)."""
    x = np.array([0.5 - 0.5 * np.sqrt(0.6), 0.5, 0.5 + 0.5 * np.sqrt(0.6)])
    w = np.array([5.0, 8.0, 5.0]) / 18.0
    return x, w


def _lattice(dim, pts):
    """Tensor lattice of 1D points, x fastest: (len(pts)**dim, dim)."""
    g = np.meshgrid(*([pts] * dim), indexing="ij")
    return np.stack([a.ravel() for a in g][::-1], axis=1) if dim > 1 else np.asarray(pts)[:, None]


# --------------------------------------------------------------------- coarse meshes
class BoxCoarse:
    """Box [0, extent] with reps cells per direction; optional smooth map
    ``deform(x) -> (y, dy/dx)`` applied to the box points.  Boundary ids
    0..2*dim-1 are the faces x-, x+, y-, y+, z-, z+.  ``periodic``: axes whose
    two faces are identified (the + face's DoFs follow the - face's, as
    make_periodicity_constraints gives them; undeformed boxes only)."""

    def __init__(self, dim, reps, extent, deform=None, periodic=()):
        self.dim = dim
        self.reps = np.array(reps[:dim], dtype=np.int64)
        self.extent = np.array(extent[:dim], dtype=float)
        self.deform = deform
        self.periodic = tuple(int(a) for a in periodic)
        if self.periodic and deform is not None:
            raise ValueError("periodic boxes cannot be deformed")
        self.n_roots = int(np.prod(self.reps))
        r = np.arange(self.n_roots)
        idx = np.zeros((self.n_roots, dim), dtype=np.int64)
        rr = r.copy()
        for d in range(dim):
            idx[:, d] = rr % self.reps[d]
            rr //= self.reps[d]
        self.idx = idx
        self.root_boundary = -np.ones((self.n_roots, 2 * dim), dtype=np.int64)
        for d in range(dim):
            if d in self.periodic:
                continue
            self.root_boundary[idx[:, d] == 0, 2 * d] = 2 * d
            self.root_boundary[idx[:, d] == self.reps[d] - 1, 2 * d + 1] = 2 * d + 1

    def _box(self, root, xi):
        h = self.extent / self.reps
        return (self.idx[root] + xi) * h

    def points(self, root, xi):
        x = self._box(root, xi)
        return self.deform(x)[0] if self.deform else x

    def jacobians(self, root, xi):
        h = self.extent / self.reps
        J = np.zeros((len(root), self.dim, self.dim))
        J[:, np.arange(self.dim), np.arange(self.dim)] = h
        if self.deform:
            _, dy = self.deform(self._box(root, xi))
            J = np.einsum("mab,mbe->mae", dy, J)
        return J

    def normals(self, bid, x):
        """Unit normal of boundary bid at points x (box faces; the deformation
        must keep the boundary faces flat for tangential conditions)."""
        n = np.zeros((len(x), self.dim))
        n[:, bid // 2] = 1.0
        return n


class ShellCoarse:
    """Spherical shell R0 <= |x| <= R1.  3D: six cube-sphere panels, each 2 x 2
    cells (equiangular), 4 radial layers = 96 coarse cells; 2D: 12 cells around
    the annulus, 1 radial layer.  Boundary id 0 inner, 1 outer sphere."""

    PANELS = [((1, 0, 0), (0, 1, 0), (0, 0, 1)), ((-1, 0, 0), (0, 1, 0), (0, 0, 1)),
              ((0, 1, 0), (1, 0, 0), (0, 0, 1)), ((0, -1, 0), (1, 0, 0), (0, 0, 1)),
              ((0, 0, 1), (1, 0, 0), (0, 1, 0)), ((0, 0, -1), (1, 0, 0), (0, 1, 0))]

    def __init__(self, dim, r0=0.55, r1=1.0):
        self.dim = dim
        self.r0, self.r1 = float(r0), float(r1)
        if dim == 3:
            self.nlat, self.nrad = 2, 4
            self.n_roots = 6 * self.nlat * self.nlat * self.nrad
            r = np.arange(self.n_roots)
            self.panel = r // (self.nlat * self.nlat * self.nrad)
            rem = r % (self.nlat * self.nlat * self.nrad)
            self.li = rem % self.nlat
            self.lj = (rem // self.nlat) % self.nlat
            self.lk = rem // (self.nlat * self.nlat)
            P = np.array(self.PANELS, dtype=float)
            self.c, self.e1, self.e2 = P[:, 0], P[:, 1], P[:, 2]
        else:
            self.nang, self.nrad = 12, 1
            self.n_roots = self.nang
            r = np.arange(self.n_roots)
            self.li = r
            self.lk = np.zeros_like(r)
        self.root_boundary = -np.ones((self.n_roots, 2 * dim), dtype=np.int64)
        self.root_boundary[self.lk == 0, 2 * (dim - 1)] = 0
        self.root_boundary[self.lk == self.nrad - 1, 2 * (dim - 1) + 1] = 1

    def _radius(self, root, t):
        return self.r0 + (self.r1 - self.r0) * (self.lk[root] + t) / self.nrad

    def _panel(self, root, xi):
        th1 = 0.25 * np.pi * ((self.li[root] + xi[:, 0]) * (2.0 / self.nlat) - 1.0)
        th2 = 0.25 * np.pi * ((self.lj[root] + xi[:, 1]) * (2.0 / self.nlat) - 1.0)
        a, b = np.tan(th1), np.tan(th2)
        pnl = self.panel[root]
        p = self.c[pnl] + a[:, None] * self.e1[pnl] + b[:, None] * self.e2[pnl]
        rho = np.linalg.norm(p, axis=1)
        return p / rho[:, None], rho, a, b, pnl

    def points(self, root, xi):
        if self.dim == 2:
            th = 2 * np.pi * (self.li[root] + xi[:, 0]) / self.nang
            r = self._radius(root, xi[:, 1])
            return np.stack([r * np.cos(th), r * np.sin(th)], axis=1)
        n, _, _, _, _ = self._panel(root, xi)
        return self._radius(root, xi[:, 2])[:, None] * n

    def jacobians(self, root, xi):
        m = len(root)
        J = np.zeros((m, self.dim, self.dim))
        dr = (self.r1 - self.r0) / self.nrad
        if self.dim == 2:
            th = 2 * np.pi * (self.li[root] + xi[:, 0]) / self.nang
            r = self._radius(root, xi[:, 1])
            dth = 2 * np.pi / self.nang
            J[:, 0, 0] = -r * np.sin(th) * dth
            J[:, 1, 0] = r * np.cos(th) * dth
            J[:, 0, 1] = np.cos(th) * dr
            J[:, 1, 1] = np.sin(th) * dr
            return J
        n, rho, a, b, pnl = self._panel(root, xi)
        r = self._radius(root, xi[:, 2])
        k = 0.25 * np.pi * (2.0 / self.nlat)
        for col, (e, t) in enumerate(((self.e1[pnl], a), (self.e2[pnl], b))):
            dn = (e - n * np.sum(n * e, axis=1)[:, None]) / rho[:, None]  # d n / d (a or b)
            J[:, :, col] = (r * k * (1.0 + t * t))[:, None] * dn
        J[:, :, 2] = dr * n
        return J

    def normals(self, bid, x):
        return x / np.linalg.norm(x, axis=1)[:, None]


# --------------------------------------------------------------------- meshes
@dataclass
class TriaMesh:
    """Arrays of one general level (degree 2): see ``Tria.build``."""

    dim: int
    cell_u: np.ndarray        # (C, 3^dim) velocity node ids, lexicographic in the cell, x fastest
    cell_p: np.ndarray        # (C, 2^dim) pressure node ids
    n_nodes_u: int
    n_nodes_p: int
    x_u: np.ndarray           # (n_nodes_u, dim) node positions
    x_p: np.ndarray           # (n_nodes_p, dim)
    geometry: np.ndarray      # (C, 3^dim, dim*dim + 1): J^{-1} row-major (e*dim + d = d xi_e / d x_d), JxW
    x_q: np.ndarray           # (C, 3^dim, dim) quadrature points
    con_u: tuple              # (dofs, offsets, masters, weights) closed velocity lines
    con_p: tuple              # pressure lines
    cell_level: np.ndarray
    hanging_u: np.ndarray     # velocity nodes with hanging-node lines
    hanging_p: np.ndarray
    boundary_nodes: dict = field(default_factory=dict)  # boundary id -> velocity node ids

    @property
    def n_cells(self):
        return len(self.cell_u)

    @property
    def n_dofs_u(self):
        return self.dim * self.n_nodes_u

    @property
    def n_dofs_p(self):
        return self.n_nodes_p


def _cluster(x, tol):
    """Labels of points that coincide within tol, numbered by first appearance."""
    tree = cKDTree(x)
    pairs = tree.query_pairs(tol, output_type="ndarray")
    n = len(x)
    g = coo_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])), shape=(n, n)) if len(pairs) else \
        coo_matrix((n, n))
    _, lab = connected_components(g, directed=False)
    _, first = np.unique(lab, return_index=True)
    order = np.argsort(first)
    remap = np.empty(len(order), dtype=np.int64)
    remap[order] = np.arange(len(order))
    return remap[lab]


def _basis_1d(nodes, x):
    """Lagrange basis on the given 1D nodes at points x: (len(x), len(nodes))."""
    v = np.ones((len(x), len(nodes)))
    for i, xi in enumerate(nodes):
        for j, xj in enumerate(nodes):
            if i != j:
                v[:, i] *= (x - xj) / (xi - xj)
    return v


def _basis(dim, nodes, xi):
    """Tensor Lagrange basis at points xi (M, dim): (M, len(nodes)^dim), x fastest."""
    v = [_basis_1d(nodes, xi[:, d]) for d in range(dim)]
    out = v[0]
    for d in range(1, dim):
        out = (v[d][:, :, None] * out[:, None, :]).reshape(len(xi), -1)
    return out


class Tria:
    """Coarse cells refined globally, then (optionally) once more on flagged
    cells."""

    def __init__(self, coarse, global_refinement=0):
        self.coarse = coarse
        self.dim = coarse.dim
        self.g = int(global_refinement)
        n = 1 << self.g
        sub = _lattice(self.dim, np.arange(n)).astype(np.int64)
        self.root = np.repeat(np.arange(coarse.n_roots), len(sub))
        self.idx = np.tile(sub, (coarse.n_roots, 1))
        self.level = np.full(len(self.root), self.g, dtype=np.int64)
        self.adapted = False

    @property
    def n_cells(self):
        return len(self.root)

    def cell_centers(self):
        xi = (self.idx + 0.5) / (1 << self.level)[:, None]
        return self.coarse.points(self.root, xi)

    def refine(self, flags):
        """Refine the flagged cells once (allowed once, after the global
        refinement, so that neighbours differ by at most one level)."""
        if self.adapted:
            raise ValueError("one adaptive level only")
        flags = np.asarray(flags, dtype=bool)
        if flags.shape != (self.n_cells,):
            raise ValueError("one flag per cell")
        kids = _lattice(self.dim, np.arange(2)).astype(np.int64)
        root, idx, level = [], [], []
        for c in range(self.n_cells):
            if flags[c]:
                root.append(np.full(len(kids), self.root[c]))
                idx.append(2 * self.idx[c] + kids)
                level.append(np.full(len(kids), self.level[c] + 1))
            else:
                root.append([self.root[c]])
                idx.append(self.idx[c][None])
                level.append([self.level[c]])
        self.root = np.concatenate(root).astype(np.int64)
        self.idx = np.concatenate(idx).astype(np.int64)
        self.level = np.concatenate(level).astype(np.int64)
        self.adapted = True

    # ---------------------------------------------------------------- build
    def _periodic_shifts(self, P, tol):
        """Translations that carry points on a periodic face to the opposite
        face (one per combination of periodic axes the points can cross)."""
        axes = getattr(self.coarse, "periodic", ())
        out = []
        for mask in range(1, 1 << len(axes)):
            for signs in range(1 << bin(mask).count("1")):
                sh = np.zeros(self.dim)
                k = 0
                for i, ax in enumerate(axes):
                    if mask >> i & 1:
                        sh[ax] = self.coarse.extent[ax] * (1 if signs >> k & 1 else -1)
                        k += 1
                out.append(sh)
        return out

    def _map(self, cells, xi_local):
        """Physical points of local reference points (per cell, (P, dim)) of cells."""
        m = len(xi_local)
        sc = 1.0 / (1 << self.level[cells])
        xi = (self.idx[cells][:, None, :] + xi_local[None]) * sc[:, None, None]
        return self.coarse.points(np.repeat(self.root[cells], m), xi.reshape(-1, self.dim)).reshape(
            len(cells), m, self.dim)

    def build(self, bc=None):
        """Mesh arrays of the active cells.  bc: {boundary id: "zero" |
        "tangential"}; other boundary ids are open."""
        bc = bc or {}
        dim, C = self.dim, self.n_cells
        all_cells = np.arange(C)
        xu = _lattice(dim, np.array([0.0, 0.5, 1.0]))        # Q2 support points, x fastest
        X = self._map(all_cells, xu)                          # (C, 3^dim, dim)
        scale = float(np.max(np.abs(X)))
        tol = 1e-9 * max(scale, 1.0)
        lab = _cluster(X.reshape(-1, dim), tol)
        cell_u = lab.reshape(C, -1)
        nn = int(lab.max()) + 1
        x_u = np.zeros((nn, dim))
        x_u[lab] = X.reshape(-1, dim)
        corners = [int(np.dot(v, [3 ** d for d in range(dim)])) for v in _lattice(dim, np.array([0, 2]))]
        vert = cell_u[:, corners]
        pv, pinv = np.unique(vert.ravel(), return_inverse=True)
        # pressure nodes numbered by first appearance among the cell vertices
        _, first = np.unique(pinv, return_index=True)
        order = np.argsort(first)
        remap = np.empty(len(order), dtype=np.int64)
        remap[order] = np.arange(len(order))
        cell_p = remap[pinv].reshape(C, -1)
        npn = len(pv)
        x_p = np.zeros((npn, dim))
        x_p[cell_p.ravel()] = x_u[vert.ravel()]

        # ---- hanging nodes: nodes of finer cells at the quarter (velocity) /
        # half (pressure) points of the faces of a coarser neighbour
        lines_u, lines_p = {}, {}
        hang_u, hang_p = [], []
        if self.adapted:
            coarse_cells = np.nonzero(self.level == self.level.min())[0]
            tree_u, tree_p = cKDTree(x_u), cKDTree(x_p)
            for nodes1d, tree, lines, hang, cmap, nodes_q in (
                    (np.linspace(0, 1, 5), tree_u, lines_u, hang_u, cell_u, np.array([0.0, 0.5, 1.0])),
                    (np.linspace(0, 1, 3), tree_p, lines_p, hang_p, cell_p, np.array([0.0, 1.0]))):
                face_pts = []
                for d in range(dim):
                    for side in (0.0, 1.0):
                        f = _lattice(dim - 1, nodes1d) if dim > 1 else np.zeros((1, 0))
                        p = np.insert(f, d, side, axis=1)
                        face_pts.append(p)
                fp = np.unique(np.concatenate(face_pts), axis=0)
                # points that are not support points of the coarse cell itself
                step = nodes_q[1] - nodes_q[0]
                off = np.any(np.abs(fp / step - np.round(fp / step)) > 1e-12, axis=1)
                fp = fp[off]
                P = self._map(coarse_cells, fp).reshape(-1, dim)
                dist, hit = tree.query(P, distance_upper_bound=tol)
                for shift in self._periodic_shifts(P, tol):
                    d2, h2 = tree.query(P + shift, distance_upper_bound=tol)
                    better = np.isfinite(d2) & ~np.isfinite(dist)
                    dist[better], hit[better] = d2[better], h2[better]
                hit = hit.reshape(len(coarse_cells), len(fp))
                found = np.isfinite(dist.reshape(len(coarse_cells), len(fp)))
                W = _basis(dim, nodes_q, fp)                # (len(fp), n_local)
                for ci, c in enumerate(coarse_cells):
                    for k in np.nonzero(found[ci])[0]:
                        node = int(hit[ci, k])
                        if node in lines:
                            continue
                        w = W[k]
                        keep = np.abs(w) > 1e-12
                        lines[node] = (cmap[c][keep].astype(np.int64), w[keep])
                        hang.append(node)

        # ---- boundary nodes (cell faces on the faces of the coarse cells)
        bnodes = {}
        face_local = []
        lat = _lattice(dim, np.arange(3))
        for d in range(dim):
            for s, val in ((0, 0), (1, 2)):
                face_local.append(np.nonzero(lat[:, d] == val)[0])
        rb = self.coarse.root_boundary[self.root]                # (C, 2 dim)
        top = (1 << self.level) - 1
        for f in range(2 * dim):
            d, s = f // 2, f % 2
            on = (rb[:, f] >= 0) & ((self.idx[:, d] == 0) if s == 0 else (self.idx[:, d] == top))
            for bid in np.unique(rb[on, f]):
                sel = on & (rb[:, f] == bid)
                ids = np.unique(cell_u[sel][:, face_local[f]].ravel())
                bnodes[int(bid)] = np.union1d(bnodes.get(int(bid), np.zeros(0, np.int64)), ids)

        vel = {}  # constrained velocity dof -> (masters, weights)
        for node, (m, w) in lines_u.items():
            for comp in range(dim):
                vel[comp * nn + node] = (comp * nn + m, w)
        zero = set()
        for bid, kind in bc.items():
            if kind == "zero" and bid in bnodes:
                zero.update(int(n) for n in bnodes[bid])
        for n in zero:
            if n in lines_u:
                continue
            for comp in range(dim):
                vel[comp * nn + n] = (np.zeros(0, np.int64), np.zeros(0))
        normals = {}
        for bid, kind in bc.items():
            if kind != "tangential" or bid not in bnodes:
                continue
            ids = bnodes[bid]
            nv = self.coarse.normals(bid, x_u[ids])
            for n, v in zip(ids, nv):
                n = int(n)
                if n in lines_u or n in zero:
                    continue
                lst = normals.setdefault(n, [])
                if not any(abs(abs(np.dot(v, u)) - 1.0) < 1e-10 for u in lst):
                    lst.append(v)
        for n, lst in normals.items():
            if len(lst) == 1:
                v = lst[0]
                k = int(np.argmax(np.abs(v)))
                others = [j for j in range(dim) if j != k and abs(v[j]) > 1e-14]
                vel[k * nn + n] = (np.array([j * nn + n for j in others], dtype=np.int64),
                                   np.array([-v[j] / v[k] for j in others]))
            else:
                for v in lst:
                    k = int(np.argmax(np.abs(v)))
                    if np.sum(np.abs(v) > 1e-12) != 1:
                        raise NotImplementedError("non-axis-aligned normals meeting at a node")
                    vel[k * nn + n] = (np.zeros(0, np.int64), np.zeros(0))

        # periodic faces: nodes on the + face follow the matching - face node
        # (hanging-node and boundary lines take precedence)
        pl_p = dict(lines_p)
        for ax in getattr(self.coarse, "periodic", ()):
            L = self.coarse.extent[ax]
            for xs, nnodes, cons, comps in ((x_u, nn, vel, dim), (x_p, npn, pl_p, 1)):
                plus = np.nonzero(np.abs(xs[:, ax] - L) <= tol)[0]
                target = xs[plus].copy()
                target[:, ax] -= L
                dd, partner = cKDTree(xs).query(target, distance_upper_bound=tol)
                for n, ok, m in zip(plus, np.isfinite(dd), partner):
                    if not ok:
                        continue   # a hanging position of the + face (no node opposite)
                    for comp in range(comps):
                        key = comp * nnodes + int(n)
                        if key in cons:
                            continue
                        cons[key] = (np.array([comp * nnodes + int(m)], np.int64), np.array([1.0]))
        con_u = _close(vel)
        con_p = _close(pl_p)

        # ---- geometry at the QGauss(3) points
        xg, wg = gauss3()
        xq = _lattice(dim, xg)
        wq = np.prod(_lattice(dim, wg), axis=1)
        sc = 1.0 / (1 << self.level)
        xi = (self.idx[:, None, :] + xq[None]) * sc[:, None, None]
        rr = np.repeat(self.root, len(xq))
        J = self.coarse.jacobians(rr, xi.reshape(-1, dim)).reshape(C, len(xq), dim, dim)
        J = J * sc[:, None, None, None]
        Jinv = np.linalg.inv(J)
        jxw = np.abs(np.linalg.det(J)) * wq[None]
        geo = np.concatenate([Jinv.reshape(C, len(xq), dim * dim), jxw[..., None]], axis=2)
        x_q = self.coarse.points(rr, xi.reshape(-1, dim)).reshape(C, len(xq), dim)

        return TriaMesh(dim=dim, cell_u=cell_u.astype(np.int32), cell_p=cell_p.astype(np.int32),
                        n_nodes_u=nn, n_nodes_p=npn, x_u=x_u, x_p=x_p, geometry=geo, x_q=x_q,
                        con_u=con_u, con_p=con_p, cell_level=self.level.copy(),
                        hanging_u=np.array(sorted(hang_u), dtype=np.int64),
                        hanging_p=np.array(sorted(hang_p), dtype=np.int64), boundary_nodes=bnodes)


def _close(lines):
    """Substitute constrained masters until every master is unconstrained;
    returns (dofs, offsets, masters, weights) sorted by DoF."""
    lines = {int(k): (np.asarray(m, np.int64), np.asarray(w, float)) for k, (m, w) in lines.items()}
    for _ in range(8):
        changed = False
        for k, (m, w) in list(lines.items()):
            if not any(int(x) in lines for x in m):
                continue
            acc = {}
            for mm, ww in zip(m, w):
                mm = int(mm)
                if mm in lines:
                    for m2, w2 in zip(*lines[mm]):
                        acc[int(m2)] = acc.get(int(m2), 0.0) + ww * w2
                else:
                    acc[mm] = acc.get(mm, 0.0) + ww
            keys = sorted(a for a in acc if abs(acc[a]) > 1e-14)
            lines[k] = (np.array(keys, np.int64), np.array([acc[a] for a in keys]))
            changed = True
        if not changed:
            break
    else:
        raise RuntimeError("constraint chains did not close")
    dofs = np.array(sorted(lines), dtype=np.int64)
    offs = np.zeros(len(dofs) + 1, dtype=np.int64)
    ms, ws = [], []
    for i, d in enumerate(dofs):
        m, w = lines[int(d)]
        o = np.argsort(m)
        ms.append(m[o])
        ws.append(w[o])
        offs[i + 1] = offs[i] + len(m)
    masters = np.concatenate(ms) if ms else np.zeros(0, np.int64)
    weights = np.concatenate(ws) if ws else np.zeros(0)
    return dofs, offs, masters.astype(np.int64), weights.astype(float)


# --------------------------------------------------------------------- level hierarchy
def coarsening_sequence(tria):
    """Level meshes of the global-coarsening hierarchy, coarse first
    (MGTransferGlobalCoarseningTools::create_geometric_coarsening_sequence,
    used at stokes_matrix_free_global_coarsening.cc:1467): coarsening every
    cell whose siblings are all active turns the adaptive level (when there
    is one) into the uniform mesh of the global refinement, then each step
    removes one global refinement."""
    levels = [Tria(tria.coarse, k) for k in range(tria.g + 1)]
    if tria.adapted:
        levels.append(tria)
    return levels


def _cell_key(root, level, idx):
    return (int(root), int(level)) + tuple(int(i) for i in idx)


def transfer(fine, coarse, mf, mc):
    """Prolongation matrices (scipy CSR) from level mesh mc (of Tria coarse) to
    mf (of Tria fine), velocity and pressure: the fine interpolant of the
    coarse finite element function (the spaces are nested), as
    MGTwoLevelTransfer applies it (read the coarse vector with its
    constraints, embed cell by cell, write the fine vector with constrained
    entries zero).  Restriction is the transpose."""
    from scipy.sparse import csr_matrix
    dim = fine.dim
    lookup = {_cell_key(coarse.root[c], coarse.level[c], coarse.idx[c]): c for c in range(coarse.n_cells)}
    out = []
    for nodes1d, cells_f, cells_c, nn_f, nn_c, comps, con_f, con_c in (
            (np.array([0.0, 0.5, 1.0]), mf.cell_u, mc.cell_u, mf.n_nodes_u, mc.n_nodes_u, dim, mf.con_u, mc.con_u),
            (np.array([0.0, 1.0]), mf.cell_p, mc.cell_p, mf.n_nodes_p, mc.n_nodes_p, 1, mf.con_p, mc.con_p)):
        loc = _lattice(dim, nodes1d)
        seen = np.zeros(nn_f, dtype=bool)
        rows, cols, vals = [], [], []
        for c in range(fine.n_cells):
            lev, idx = fine.level[c], fine.idx[c]
            key = _cell_key(fine.root[c], lev, idx)
            if key in lookup:
                cc, shift, scale = lookup[key], np.zeros(dim), 1.0
            else:
                cc = lookup[_cell_key(fine.root[c], lev - 1, idx // 2)]
                shift, scale = (idx % 2) * 0.5, 0.5
            new = ~seen[cells_f[c]]
            if not np.any(new):
                continue
            xi = shift + scale * loc[new]
            w = _basis(dim, nodes1d, xi)
            w[np.abs(w) < 1e-14] = 0.0
            fn = cells_f[c][new]
            seen[fn] = True
            r, a = np.nonzero(w)
            rows.append(fn[r])
            cols.append(cells_c[cc][a])
            vals.append(w[r, a])
        rows, cols, vals = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
        R = np.concatenate([rows + k * nn_f for k in range(comps)])
        Cc = np.concatenate([cols + k * nn_c for k in range(comps)])
        V = np.tile(vals, comps)
        P = csr_matrix((V, (R, Cc)), shape=(comps * nn_f, comps * nn_c))
        # coarse constraints: constrained coarse DoFs through their masters
        dofs, offs, masters, weights = con_c
        free = np.ones(comps * nn_c, dtype=bool)
        free[dofs] = False
        fi = np.nonzero(free)[0]
        Cm = csr_matrix((np.concatenate([np.ones(len(fi)), weights]),
                         (np.concatenate([fi, np.repeat(dofs, np.diff(offs))]), np.concatenate([fi, masters]))),
                        shape=(comps * nn_c, comps * nn_c))
        P = (P @ Cm).tocsr()
        keep = np.ones(comps * nn_f)
        keep[con_f[0]] = 0.0
        P = csr_matrix(P.multiply(keep[:, None]))
        P.eliminate_zeros()
        P.sort_indices()
        out.append(P)
    return out[0], out[1]


def parent_map(fine, coarse):
    """For every cell of Tria fine, the cell of Tria coarse (the next level of
    the coarsening sequence) that contains it: its parent, or the same cell."""
    lookup = {_cell_key(coarse.root[c], coarse.level[c], coarse.idx[c]): c for c in range(coarse.n_cells)}
    out = np.empty(fine.n_cells, dtype=np.int64)
    for c in range(fine.n_cells):
        key = _cell_key(fine.root[c], fine.level[c], fine.idx[c])
        out[c] = lookup[key] if key in lookup else lookup[_cell_key(fine.root[c], fine.level[c] - 1,
                                                                   fine.idx[c] // 2)]
    return out


def hierarchy(tria, bc=None):
    """(level meshes coarse first, velocity prolongations, pressure
    prolongations); prolongation l maps level l-1 to level l."""
    trias = coarsening_sequence(tria)
    meshes = [t.build(bc) for t in trias]
    pu, pp = [None], [None]
    for l in range(1, len(trias)):
        a, b = transfer(trias[l], trias[l - 1], meshes[l], meshes[l - 1])
        pu.append(a)
        pp.append(b)
    return trias, meshes, pu, pp
