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
from scipy.sparse import coo_matrix, csr_matrix

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


def stokes_rhs_correction(ops, u0):
    """correct_stokes_rhs (stokes_matrix_free_local_smoothing.cc:617-754) on a
    general level: -C^T K u0 with u0 read as stored, constrained rows zero."""
    u, p = u0[:ops.nu], u0[ops.nu:]
    yu = -(ops.Cu.T @ (ops.A @ u + ops.B.T @ p))
    yp = -(ops.Cp.T @ (ops.B @ u))
    yu[ops.fixed_u] = 0.0
    yp[ops.fixed_p] = 0.0
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


# ---------------------------------------------------------------------- multigrid
def _tridiag_eigs(alphas, betas):
    m = len(alphas)
    d = np.empty(m)
    d[0] = 1.0 / alphas[0]
    for k in range(1, m):
        d[k] = 1.0 / alphas[k] + betas[k - 1] / alphas[k - 1]
    e = np.array([np.sqrt(betas[k]) / alphas[k] for k in range(m - 1)])
    T = np.diag(d) + (np.diag(e, 1) + np.diag(e, -1) if m > 1 else 0.0)
    w = np.linalg.eigvalsh(T)
    return w[0], w[-1]


class Chebyshev:
    """PreconditionChebyshev with the inverse diagonal (deal.II, not vendored;
    parameters of stokes_matrix_free_global_coarsening.cc:800-846): eigenvalue
    estimate by Jacobi-preconditioned CG from the mean-free pattern (i mod 11),
    stopped at |r| < 1e-2 or max_its, Lanczos tridiagonal from the CG
    coefficients, lambda_max scaled by 1.2; interval [lambda_max / range,
    lambda_max] (range > 1) or [min(0.9 lambda_max, lambda_min), lambda_max]."""

    def __init__(self, M, dinv, degree, max_its, rng_):
        self.M, self.dinv, self.degree = M, dinv, degree
        n = M.shape[0]
        pat = (np.arange(n) % 11).astype(float)
        r = pat - pat.mean()
        alphas, betas = [], []
        if np.sqrt(r @ r) >= 1e-2:
            z = dinv * r
            p = z.copy()
            rz = r @ z
            for it in range(1, max_its + 1):
                q = M @ p
                alpha = rz / (p @ q)
                r = r - alpha * q
                alphas.append(alpha)
                if np.sqrt(r @ r) < 1e-2 or it == max_its:
                    break
                z = dinv * r
                rz2 = r @ z
                beta = rz2 / rz
                rz = rz2
                p = z + beta * p
                betas.append(beta)
        if alphas:
            lmin, lmax = _tridiag_eigs(alphas, betas)
            lmax *= 1.2
        else:
            lmin = lmax = 1.0
        self.lmin, self.lmax, self.its = lmin, lmax, len(alphas)
        a = lmax / rng_ if rng_ > 1.0 else min(0.9 * lmax, lmin)
        self.delta = 0.5 * (lmax - a)
        self.theta = 0.5 * (lmax + a)

    def smooth(self, x, b, x_zero, steps):
        """Saad Alg. 12.1, `steps` applications (the first from zero when x_zero)."""
        M, dinv = self.M, self.dinv
        for s in range(steps):
            if x_zero and s == 0:
                d = dinv * b / self.theta
                x = d.copy()
            else:
                d = dinv * (b - M @ x) / self.theta
                x = x + d
            if self.degree < 2 or abs(self.delta) < 1e-40:
                continue
            sigma = self.theta / self.delta
            rho_old = self.delta / self.theta
            for _ in range(1, self.degree):
                rho = 1.0 / (2.0 * sigma - rho_old)
                d = rho * rho_old * d + 2.0 * rho / self.delta * dinv * (b - M @ x)
                x = x + d
                rho_old = rho
        return x


def constrained_matrix(K, Cm, fixed):
    """C^T K C with identity rows and columns on the constrained DoFs."""
    from scipy.sparse import diags
    n = K.shape[0]
    keep = np.ones(n)
    keep[fixed] = 0.0
    D = diags(keep)
    return (D @ (Cm.T @ K @ Cm) @ D + diags(1.0 - keep)).tocsr()


class GMG:
    """Global-coarsening V-cycles for the A block (kind 0) and the Schur
    complement approximation (kind 1) on general level meshes, coarse level
    first; prolongation l maps level l-1 to level l (restriction: its
    transpose).  Levels > 0: Chebyshev degree 4 over [lambda_max/15,
    lambda_max] from 10 iterations; level 0: the smoother as coarse solver,
    degree 8, range 1e-3, 100 iterations; A smoothed with 1 step, the
    Schur complement with 4 (stokes_matrix_free_global_coarsening.cc:800-883)."""

    def __init__(self, meshes, prolong_u, prolong_p, etas, pressure_scaling=1.0, compressible=False):
        self.ops = [Operators(m, e, pressure_scaling, compressible) for m, e in zip(meshes, etas)]
        self.P = [prolong_u, prolong_p]
        self.mats = [[], []]
        self.cheb = [[], []]
        self.diag = [[], []]
        for l, o in enumerate(self.ops):
            for kind, (K, Cm, fx) in enumerate(((o.A, o.Cu, o.fixed_u), (o.M, o.Cp, o.fixed_p))):
                M = constrained_matrix(K, Cm, fx)
                dg = M.diagonal()
                self.mats[kind].append(M)
                self.diag[kind].append(dg)
                coarse = l == 0
                self.cheb[kind].append(Chebyshev(M, 1.0 / dg, 8 if coarse else 4, 100 if coarse else 10,
                                                 1e-3 if coarse else 15.0))

    def vcycle(self, kind, b, l=None):
        l = len(self.ops) - 1 if l is None else l
        steps = 1 if kind == 0 else 4
        ch = self.cheb[kind][l]
        if l == 0:
            return ch.smooth(None, b, True, steps)
        x = ch.smooth(None, b, True, steps)
        r = b - self.mats[kind][l] @ x
        P = self.P[kind][l]
        x = x + P @ self.vcycle(kind, P.T @ r, l - 1)
        return ch.smooth(x, b, False, steps)

    def block_preconditioner(self, r):
        """BlockSchurPreconditioner::vmult, cheap variant
        (block_stokes_preconditioner.h:220-255): p = -S^{-1} r_p,
        u = A^{-1} (r_u - B^T p), one V-cycle each."""
        top = self.ops[-1]
        p = -self.vcycle(1, r[top.nu:])
        u = self.vcycle(0, r[:top.nu] - top.bt(p))
        return np.concatenate([u, p])

    def cg(self, kind, b, tol, maxit):
        """Preconditioned CG (one V-cycle) on the finest A block / mass from
        zero to |r| <= tol; returns (x, iterations)."""
        M = self.mats[kind][-1]
        x = np.zeros_like(b)
        r = b.copy()
        if np.sqrt(r @ r) <= tol:
            return x, 0
        z = self.vcycle(kind, r)
        rz = r @ z
        p = z.copy()
        for it in range(1, maxit + 1):
            q = M @ p
            alpha = rz / (p @ q)
            x = x + alpha * p
            r = r - alpha * q
            if np.sqrt(r @ r) <= tol:
                return x, it
            z = self.vcycle(kind, r)
            rz2 = r @ z
            p = z + (rz2 / rz) * p
            rz = rz2
        return x, maxit

    def block_preconditioner_expensive(self, r, a_tol=1e-2, s_tol=1e-6):
        """Expensive variant: S^{-1} by CG to s_tol |r_p| (at most 100
        iterations), A^{-1} by CG to a_tol |r_u - B^T p| (at most 10000)."""
        top = self.ops[-1]
        rp = r[top.nu:]
        nsp = np.linalg.norm(rp)
        p = -self.cg(1, rp, nsp * s_tol, 100)[0] if nsp > 1e-50 else np.zeros_like(rp)
        t = r[:top.nu] - top.bt(p)
        u = self.cg(0, t, np.linalg.norm(t) * a_tol, 10000)[0]
        return np.concatenate([u, p])

    def solve(self, b, tol, n_cheap, n_expensive, restart=100, a_tol=1e-2, s_tol=1e-6):
        """Cheap GMRES steps, then flexible GMRES with the expensive
        preconditioner (stokes_matrix_free_local_smoothing.cc:1259-1397).
        Returns (x, cheap iterations, expensive iterations, residual)."""
        top = self.ops[-1]
        x, its, res = gmres(top.stokes, self.block_preconditioner, b, tol, n_cheap, restart) if n_cheap > 0 \
            else (np.zeros_like(b), 0, np.linalg.norm(b))
        its_e = 0
        if res > tol and n_expensive > 0:
            x, its_e, res = gmres(top.stokes, lambda v: self.block_preconditioner_expensive(v, a_tol, s_tol), b, tol,
                                  n_expensive, restart, x0=x)
        return x, its, its_e, res


def gmres(apply, prec, b, tol, maxit, restart, x0=None):
    """Right-preconditioned restarted GMRES with modified Gram-Schmidt from x0
    (zero by default), the update built from the preconditioned vectors
    (flexible); stops when the residual estimate reaches tol.  Returns (x,
    iterations, residual estimate)."""
    x = np.zeros_like(b) if x0 is None else x0.copy()
    its = 0
    r = b - apply(x)
    beta = np.linalg.norm(r)
    while its < maxit and beta > tol:
        V = [r / beta]
        Z = []
        H = np.zeros((restart + 1, restart))
        cs, sn = np.zeros(restart), np.zeros(restart)
        g = np.zeros(restart + 1)
        g[0] = beta
        k = 0
        while k < restart and its < maxit:
            z = prec(V[k])
            Z.append(z)
            w = apply(z)
            for i in range(k + 1):
                H[i, k] = w @ V[i]
                w = w - H[i, k] * V[i]
            H[k + 1, k] = np.linalg.norm(w)
            V.append(w / H[k + 1, k] if H[k + 1, k] > 0 else w)
            for i in range(k):
                t = cs[i] * H[i, k] + sn[i] * H[i + 1, k]
                H[i + 1, k] = -sn[i] * H[i, k] + cs[i] * H[i + 1, k]
                H[i, k] = t
            den = np.hypot(H[k, k], H[k + 1, k])
            cs[k], sn[k] = H[k, k] / den, H[k + 1, k] / den
            H[k, k] = den
            H[k + 1, k] = 0.0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            k += 1
            its += 1
            if abs(g[k]) <= tol:
                break
        y = np.zeros(k)
        for i in range(k - 1, -1, -1):
            y[i] = (g[i] - H[i, i + 1:k] @ y[i + 1:k]) / H[i, i]
        for i in range(k):
            x = x + y[i] * Z[i]
        r = b - apply(x)
        beta = abs(g[k])
    return x, its, beta


def level_viscosities(meshes, parents, eta_qp, averaging=0):
    """Viscosity tables of every level (coarse first) from the finest level's
    values at its quadrature points: material averaging on the finest level
    (0 none, 1 arithmetic, 2 harmonic, 3 geometric mean of the cell's point
    values; material_model/interface.cc:556-680), its DGQ0 projection (JxW-
    weighted cell mean), and on coarser levels one value per cell, the mean of
    the contained cells of the next finer level (FE_DGQ(0) restriction in
    MGTwoLevelTransfer::interpolate, stokes_matrix_free_global_coarsening.cc:
    209-246)."""
    e = np.asarray(eta_qp, float).reshape(meshes[-1].n_cells, -1)
    nq = e.shape[1]
    if averaging == 1:
        e = np.repeat(e.mean(axis=1, keepdims=True), nq, axis=1)
    elif averaging == 2:
        e = np.repeat(1.0 / np.mean(1.0 / e, axis=1, keepdims=True), nq, axis=1)
    elif averaging == 3:
        e = np.repeat(np.exp(np.mean(np.log(e), axis=1, keepdims=True)), nq, axis=1)
    jxw = meshes[-1].geometry[:, :, -1]
    val = np.sum(jxw * e, axis=1) / np.sum(jxw, axis=1)
    out = [e]
    for l in range(len(meshes) - 2, -1, -1):
        par = parents[l + 1]
        cnt = np.bincount(par, minlength=meshes[l].n_cells)
        val = np.bincount(par, weights=val, minlength=meshes[l].n_cells) / cnt
        out.append(np.repeat(val[:, None], nq, axis=1))
    return out[::-1]
