"""ORACLE — test infrastructure only.

Matrix-based CPU FP64 restatement of the Stokes level operators on a general
level mesh with affine constraints (hanging nodes, Dirichlet, no normal flux):
the element matrices are formed densely at the QGauss(3) points from the
per-point J^{-1} / JxW tables, assembled into global sparse matrices, and the
constraints applied as C^T K C with C the closed constraint map.  Only
``tests/`` imports this module.

Operators restated (ASPECT, source/simulator/solver/matrix_free_operators.cc):
  StokesOperator::local_apply        :376-547  (2 eta eps(u) - 2/3 eta div u I
                                                 [compressible] - s p I; -s div u)
  BTBlockOperator::local_apply       :668-701
  MassMatrixOperator (s^2 / eta)     :808-845
  ABlockOperator                     :974-1009
with the constraint handling of MatrixFree on a general mesh (deal.II, not
vendored): source values of constrained DoFs come from their masters
(read_dof_values), contributions to constrained DoFs go to their masters
(distribute_local_to_global), and constrained rows are identity rows
(MatrixFreeOperators::Base::vmult); the B^T block leaves constrained
velocity rows zero.

How it is pinned (tests/test_oracle_tria.py): on meshes without hanging nodes
against the box oracle (oracle.cpp, itself pinned by brute force and closed
forms), plain and with mapped geometry; on meshes with hanging nodes against
closed-form integrals of polynomial fields over the domain (2 eta eps:eps,
-s q div u, s^2 p q / eta), which the constrained operators reproduce
exactly for fields in the conforming space.
"""
from __future__ import annotations

import numpy as np
from scipy.sparse import coo_matrix, csr_matrix, identity

_XG = np.array([0.5 - 0.5 * np.sqrt(0.6), 0.5, 0.5 + 0.5 * np.sqrt(0.6)])


def _lag(nodes, x):
    v = np.ones((len(x), len(nodes)))
    d = np.zeros((len(x), len(nodes)))
    for i, xi in enumerate(nodes):
        for j, xj in enumerate(nodes):
            if i == j:
                continue
            v[:, i] *= (x - xj) / (xi - xj)
        for m, xm in enumerate(nodes):
            if m == i:
                continue
            t = np.ones(len(x)) / (xi - xm)
            for j, xj in enumerate(nodes):
                if j != i and j != m:
                    t *= (x - xj) / (xi - xj)
            d[:, i] += t
    return v, d


def _tensor(dim, nodes):
    """Values (NQ, n) and reference gradients (NQ, n, dim) of the tensor basis
    at the QGauss(3) points, both indices x fastest."""
    v1, d1 = _lag(nodes, _XG)
    n1, q1 = len(nodes), len(_XG)
    qi = np.stack(np.meshgrid(*([np.arange(q1)] * dim), indexing="ij")[::-1], -1).reshape(-1, dim)
    ai = np.stack(np.meshgrid(*([np.arange(n1)] * dim), indexing="ij")[::-1], -1).reshape(-1, dim)
    val = np.ones((len(qi), len(ai)))
    grad = np.ones((len(qi), len(ai), dim))
    for d in range(dim):
        vv = v1[qi[:, d]][:, ai[:, d]]
        dd = d1[qi[:, d]][:, ai[:, d]]
        val *= vv
        for e in range(dim):
            grad[:, :, e] *= dd if e == d else vv
    return val, grad


class Operators:
    """Global matrices of one mesh and viscosity (per quadrature point)."""

    def __init__(self, mesh, eta_qp, pressure_scaling=1.0, compressible=False):
        dim, C = mesh.dim, mesh.n_cells
        eta = np.asarray(eta_qp, float).reshape(C, -1)
        s = float(pressure_scaling)
        _, gu = _tensor(dim, [0.0, 0.5, 1.0])
        vp, _ = _tensor(dim, [0.0, 1.0])
        geo = mesh.geometry
        Jinv = geo[:, :, :dim * dim].reshape(C, -1, dim, dim)   # [e, d] = d xi_e / d x_d
        jxw = geo[:, :, dim * dim]
        g = np.einsum("cqed,qae->cqad", Jinv, gu)               # physical gradients (C, Q, nu, dim)
        W = jxw * eta
        nu, npl = gu.shape[1], vp.shape[1]
        # A[(c,a),(c',b)] = int eta (delta_cc' grad a . grad b + d_c' a d_c b) [- 2/3 eta d_c a d_c' b]
        lap = np.einsum("cq,cqad,cqbd->cab", W, g, g)
        cross = np.einsum("cq,cqae,cqbf->caebf", W, g, g)     # [cell, a, e, b, f] = int eta d_e a d_f b
        A = np.zeros((C, dim, nu, dim, nu))
        for ci in range(dim):
            A[:, ci, :, ci, :] += lap
            for cj in range(dim):
                A[:, ci, :, cj, :] += cross[:, :, cj, :, ci]
                if compressible:
                    A[:, ci, :, cj, :] -= (2.0 / 3.0) * cross[:, :, ci, :, cj]
        # B[i, (c,a)] = -s int psi_i d_c phi_a
        B = -s * np.einsum("kq,qi,kqae->kiea", jxw, vp, g)
        M = s * s * np.einsum("cq,qi,qj->cij", jxw / eta, vp, vp)
        nn, npn = mesh.n_nodes_u, mesh.n_nodes_p
        cu = mesh.cell_u.astype(np.int64)
        cp = mesh.cell_p.astype(np.int64)
        du = (np.arange(dim)[None, :, None] * nn + cu[:, None, :]).reshape(C, dim * nu)   # (C, dim*nu)
        r = np.broadcast_to(du[:, :, None], (C, dim * nu, dim * nu))
        cc = np.broadcast_to(du[:, None, :], (C, dim * nu, dim * nu))
        self.A = coo_matrix((A.reshape(C, dim * nu, dim * nu).ravel(), (r.ravel(), cc.ravel())),
                            shape=(dim * nn, dim * nn)).tocsr()
        r = np.broadcast_to(cp[:, :, None], (C, npl, dim * nu))
        cc = np.broadcast_to(du[:, None, :], (C, npl, dim * nu))
        self.B = coo_matrix((B.reshape(C, npl, dim * nu).ravel(), (r.ravel(), cc.ravel())),
                            shape=(npn, dim * nn)).tocsr()
        r = np.broadcast_to(cp[:, :, None], (C, npl, npl))
        cc = np.broadcast_to(cp[:, None, :], (C, npl, npl))
        self.M = coo_matrix((M.ravel(), (r.ravel(), cc.ravel())), shape=(npn, npn)).tocsr()
        self.Cu, self.fixed_u = constraint_map(mesh.con_u, dim * nn)
        self.Cp, self.fixed_p = constraint_map(mesh.con_p, npn)
        self.nu, self.np = dim * nn, npn

    def ablock(self, u):
        y = self.Cu.T @ (self.A @ (self.Cu @ u))
        y[self.fixed_u] = u[self.fixed_u]
        return y

    def bt(self, p):
        y = self.Cu.T @ (self.B.T @ (self.Cp @ p))
        y[self.fixed_u] = 0.0
        return y

    def mass(self, p):
        y = self.Cp.T @ (self.M @ (self.Cp @ p))
        y[self.fixed_p] = p[self.fixed_p]
        return y

    def stokes(self, x):
        u, p = x[:self.nu], x[self.nu:]
        uc, pc = self.Cu @ u, self.Cp @ p
        yu = self.Cu.T @ (self.A @ uc + self.B.T @ pc)
        yp = self.Cp.T @ (self.B @ uc)
        yu[self.fixed_u] = u[self.fixed_u]
        yp[self.fixed_p] = p[self.fixed_p]
        return np.concatenate([yu, yp])


def constraint_map(con, n):
    """C (n x n): free DoF -> itself, constrained DoF -> its masters' weights;
    and the list of constrained DoFs."""
    dofs, offs, masters, weights = con
    free = np.ones(n, dtype=bool)
    free[dofs] = False
    fi = np.nonzero(free)[0]
    rows = np.concatenate([fi, np.repeat(dofs, np.diff(offs))])
    cols = np.concatenate([fi, masters])
    vals = np.concatenate([np.ones(len(fi)), weights])
    return csr_matrix((vals, (rows, cols)), shape=(n, n)), np.asarray(dofs, np.int64)


def resolve(con, x):
    """x with constrained entries replaced by their masters' combination."""
    Cm, _ = constraint_map(con, len(x))
    return Cm @ x
