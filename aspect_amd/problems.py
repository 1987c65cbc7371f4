"""Host-side problem definitions for the box configurations the solver is
measured and tested on: quadrature-point coordinates, material fields
(viscosity, density) and ASPECT's pressure scaling rule.

These play the role of ASPECT's material-model plugins and Simulator helpers
that sit above the Stokes handler; they produce the per-quadrature-point
inputs that ``StokesMatrixFreeHandler.assemble`` receives.

References (ASPECT tree):
  * nsinker material law        benchmarks/nsinker/nsinker.cc:61-76, 282-314
  * pressure scaling            source/simulator/helper_functions.cc:255-345
  * box length scale            source/geometry_model/box.cc:252-255
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .mfs import BC_TANGENTIAL, BC_ZERO, make_box


def gauss_points(n: int):
    """Gauss-Legendre points and weights on [0, 1]."""
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def qp_coordinates(dim, degree, n_cells, extent):
    """Coordinates of the QGauss(degree+1) points of every cell:
    array [cells, nq, dim], cells lexicographic (x fastest), q lexicographic (x fastest)."""
    xq, _ = gauss_points(degree + 1)
    h = [extent[d] / n_cells[d] for d in range(dim)]
    axes = []
    for d in range(dim):
        # [n_cells_d, Q] coordinates along d
        axes.append(((np.arange(n_cells[d])[:, None] + xq[None, :]) * h[d]))
    Q = degree + 1
    if dim == 2:
        X = axes[0][None, :, None, :]  # [cy, cx, qy, qx]
        Y = axes[1][:, None, :, None]
        X, Y = np.broadcast_arrays(X, Y)
        return np.stack([X, Y], axis=-1).reshape(n_cells[1] * n_cells[0], Q * Q, 2)
    X = axes[0][None, None, :, None, None, :]  # cz, cy, cx, qz, qy, qx
    Y = axes[1][None, :, None, None, :, None]
    Z = axes[2][:, None, None, :, None, None]
    X, Y, Z = np.broadcast_arrays(X, Y, Z)
    out = np.stack([X, Y, Z], axis=-1)
    return out.reshape(n_cells[2] * n_cells[1] * n_cells[0], Q ** 3, 3)


def cell_midpoints(dim, n_cells, extent):
    h = [extent[d] / n_cells[d] for d in range(dim)]
    grids = np.meshgrid(*[(np.arange(n_cells[d]) + 0.5) * h[d] for d in range(dim)][::-1], indexing="ij")
    return np.stack(grids[::-1], axis=-1).reshape(-1, dim)


def pressure_scaling(eta_at_midpoints: np.ndarray, length_scale: float) -> float:
    """compute_pressure_scaling_factor: exp(volume-weighted mean of log(eta)
    at cell midpoints) / length_scale (uniform cells: plain mean)."""
    return float(np.exp(np.mean(np.log(eta_at_midpoints))) / length_scale)


# ------------------------------------------------------------------ nsinker
NSINKER_CENTERS = np.array([  # first published sinker centres of the nsinker benchmark
    [2.4257829890e-01, 1.3469574514e-02, 3.8313885004e-01],
    [4.1465269048e-01, 6.7768972864e-02, 9.9312692973e-01],
    [4.8430804651e-01, 7.6533776604e-01, 3.1833815403e-02],
    [3.0935481671e-02, 9.3264044027e-01, 8.8787953411e-01],
])


def nsinker_chi(x, n_sinkers=4, delta=200.0, omega=0.1):
    """inside_outside_factor: product over sinkers of 1 - exp(-delta max(0, d - omega/2)^2)."""
    chi = np.ones(x.shape[:-1])
    for s in NSINKER_CENTERS[:n_sinkers]:
        d = np.linalg.norm(x - s[: x.shape[-1]], axis=-1)
        chi *= 1.0 - np.exp(-delta * np.maximum(0.0, d - omega / 2.0) ** 2)
    return chi


def nsinker_material(x, viscosity_ratio=1e4, sinker_density=10.0, n_sinkers=4):
    chi = nsinker_chi(x, n_sinkers)
    r = np.sqrt(viscosity_ratio)
    eta = (r - 1.0 / r) * (1.0 - chi) + 1.0 / r
    rho = sinker_density * (1.0 - chi)
    return eta, rho


# ------------------------------------------------------------------ seeded sinker configurations
def smooth_sinkers(x, centres, radius=0.05, width=0.01):
    """max over sinkers of 1/2 (1 - tanh((|x - c| - r) / w))."""
    chi = np.zeros(x.shape[:-1])
    for c in centres:
        d = np.linalg.norm(x - c, axis=-1)
        chi = np.maximum(chi, 0.5 * (1.0 - np.tanh((d - radius) / width)))
    return chi


@dataclass
class BoxConfig:
    name: str
    dim: int
    coarse: tuple
    n_levels: int
    bc: tuple
    degree: int = 2
    extent: tuple = (1.0, 1.0, 1.0)
    averaging: str = "none"
    gmg_precision: str = "fp64"
    tolerance: float = 1e-8
    meta: dict = field(default_factory=dict)

    @property
    def n_cells(self):
        return tuple(self.coarse[d] << (self.n_levels - 1) for d in range(self.dim))

    def box(self, pressure_scaling=1.0):
        return make_box(self.dim, self.degree, self.coarse, self.n_levels, self.extent, self.bc,
                        pressure_scaling=pressure_scaling)

    def gravity(self):
        return np.array([0.0, -1.0]) if self.dim == 2 else np.array([0.0, 0.0, -1.0])


def config_2d_32(**kw):
    """2D unit box, 32x32, eta = 1, rho = 1 + 0.01 sin(pi x) sin(pi y), free slip."""
    return BoxConfig("2d_box_32", 2, (1, 1), 6, (BC_TANGENTIAL,) * 4, **kw)


def config_sinker_64(**kw):
    """3D unit box, 64^3, 8 seeded sinkers, eta 1e3 inside / 1 outside, free slip."""
    return BoxConfig("sinker_64", 3, (1, 1, 1), 7, (BC_TANGENTIAL,) * 6,
                     meta=dict(n_sinkers=8, contrast=1e3, seed=2), **kw)


def config_sinker_128(**kw):
    """3D unit box, 128^3 (2x2x2 coarse box, 7 levels), 16 seeded sinkers, 1e6 contrast."""
    kw.setdefault("gmg_precision", "fp32")
    return BoxConfig("sinker_128", 3, (2, 2, 2), 7, (BC_TANGENTIAL,) * 6,
                     meta=dict(n_sinkers=16, contrast=1e6, seed=3), **kw)


def config_ensemble(n_problems=48, seed=4, **kw):
    """Ensemble of independent 2D problems: unit square, n x n cells with
    n = 16 m, m ~ U{1..16} (coarse mesh = odd part of n, so every problem has a
    multigrid hierarchy); 1-8 sinkers (radius 0.05, width 0.01), viscosity
    contrast 10^U(2,5), density 1 + a chi with a ~ U(0.001, 0.1); free slip;
    FP32 V-cycle."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_problems):
        n = 16 * int(rng.integers(1, 17))
        c = n
        while c % 2 == 0:
            c //= 2
        levels = int(np.log2(n // c)) + 1
        meta = dict(n_sinkers=int(rng.integers(1, 9)), contrast=float(10 ** rng.uniform(2, 5)),
                    amplitude=float(rng.uniform(0.001, 0.1)), seed=int(rng.integers(0, 2 ** 31)), n=n)
        kw2 = dict(gmg_precision="fp32", averaging="harmonic")
        kw2.update(kw)
        out.append(BoxConfig(f"ens_{i}", 2, (c, c), levels, (BC_TANGENTIAL,) * 4, meta=meta, **kw2))
    return out


def material(cfg: BoxConfig, x):
    """(eta, rho) at points x [..., dim] for a configuration."""
    if cfg.name.startswith("2d_box"):
        eta = np.ones(x.shape[:-1])
        rho = 1.0 + 0.01 * np.sin(np.pi * x[..., 0]) * np.sin(np.pi * x[..., 1])
        return eta, rho
    if cfg.name.startswith("sinker"):
        rng = np.random.default_rng(cfg.meta["seed"])
        centres = rng.uniform(0.1, 0.9, size=(cfg.meta["n_sinkers"], cfg.dim))
        chi = smooth_sinkers(x, centres)
        eta = 1.0 + (cfg.meta["contrast"] - 1.0) * chi
        rho = 1.0 + chi
        return eta, rho
    if cfg.name.startswith("nsinker"):
        return nsinker_material(x, cfg.meta.get("viscosity_ratio", 1e4), cfg.meta.get("sinker_density", 10.0))
    if cfg.name.startswith("ens_"):
        rng = np.random.default_rng(cfg.meta["seed"])
        centres = rng.uniform(0.1, 0.9, size=(cfg.meta["n_sinkers"], cfg.dim))
        chi = smooth_sinkers(x, centres)
        eta = 1.0 + (cfg.meta["contrast"] - 1.0) * chi
        rho = 1.0 + cfg.meta["amplitude"] * chi
        return eta, rho
    raise ValueError(cfg.name)


def config_nsinker(nref=2, averaging="none", **kw):
    """ASPECT's nsinker benchmark (benchmarks/nsinker/nsinker_gmg.prm): unit cube,
    zero velocity on all faces, 4 sinkers, ratio 1e4, pressure scaling 1."""
    return BoxConfig(f"nsinker_{nref}", 3, (1, 1, 1), nref + 1, (BC_ZERO,) * 6, averaging=averaging,
                     tolerance=1e-6, meta=dict(viscosity_ratio=1e4, sinker_density=10.0), **kw)


def fields(cfg: BoxConfig):
    """eta and rho at every quadrature point ([cells, nq]) and the pressure scaling."""
    x = qp_coordinates(cfg.dim, cfg.degree, cfg.n_cells, cfg.extent)
    eta, rho = material(cfg, x)
    if cfg.name.startswith("nsinker"):
        s = 1.0  # nsinker's pressure-scaling signal returns 1
    else:
        mid = cell_midpoints(cfg.dim, cfg.n_cells, cfg.extent)
        eta_mid, _ = material(cfg, mid)
        s = pressure_scaling(eta_mid, 0.01 * cfg.extent[0])
    return eta, rho, s


# ------------------------------------------------------------------ general geometry (mapped boxes)
def _lobatto(k):
    """Support points of FE_Q(k) on [0, 1] (Gauss-Lobatto)."""
    if k == 1:
        return np.array([0.0, 1.0])
    inner = np.sort(np.polynomial.legendre.Legendre.basis(k).deriv().roots())
    return np.concatenate([[0.0], 0.5 * (inner + 1.0), [1.0]])


def _lagrange_tables(k):
    """S[q][i] and G[q][i]: FE_Q(k) basis values and derivatives at the QGauss(k+1) points on [0, 1]."""
    pts = _lobatto(k)
    xq, _ = gauss_points(k + 1)
    S = np.ones((k + 1, k + 1))
    G = np.zeros((k + 1, k + 1))
    for i in range(k + 1):
        for q in range(k + 1):
            v = 1.0
            for j in range(k + 1):
                if j != i:
                    v *= (xq[q] - pts[j]) / (pts[i] - pts[j])
            S[q, i] = v
            d = 0.0
            for m in range(k + 1):
                if m == i:
                    continue
                t = 1.0 / (pts[i] - pts[m])
                for j in range(k + 1):
                    if j != i and j != m:
                        t *= (xq[q] - pts[j]) / (pts[i] - pts[j])
                d += t
            G[q, i] = d
    return S, G


def chunk_map(dim, r_range=(0.55, 1.0), lon_range=(0.0, 0.8), lat_range=(-0.4, 0.4)):
    """Map of box coordinates xi in [0, 1]^dim onto a chunk in the chart
    (radius, longitude[, latitude]) of ASPECT's chunk geometry
    (source/geometry_model/chunk.cc:219-241, push_forward_sphere): 2D annulus
    sector, 3D spherical chunk; positively oriented."""
    def f(xi):
        r = r_range[0] + (r_range[1] - r_range[0]) * xi[..., 0]
        lon = lon_range[0] + (lon_range[1] - lon_range[0]) * xi[..., 1]
        if dim == 2:
            return np.stack([r * np.cos(lon), r * np.sin(lon)], axis=-1)
        lat = lat_range[0] + (lat_range[1] - lat_range[0]) * xi[..., 2]
        return np.stack([r * np.cos(lat) * np.cos(lon), r * np.cos(lat) * np.sin(lon), r * np.sin(lat)], axis=-1)
    return f


def affine_map(B, c=None):
    """x = B xi + c."""
    B = np.asarray(B, dtype=float)
    c = np.zeros(B.shape[0]) if c is None else np.asarray(c, dtype=float)
    return lambda xi: xi @ B.T + c


def mapped_nodes(dim, degree, n_cells, mapping):
    """Physical positions of the FE_Q(k) nodes of every cell under the
    iso-parametric map with nodes on mapping(xi): [cells, nv, dim]."""
    pts = _lobatto(degree)
    ax = [(np.arange(n_cells[d])[:, None] + pts[None, :]) / n_cells[d] for d in range(dim)]
    N = degree + 1
    if dim == 2:
        X = ax[0][None, :, None, :]
        Y = ax[1][:, None, :, None]
        X, Y = np.broadcast_arrays(X, Y)
        xi = np.stack([X, Y], axis=-1).reshape(n_cells[1] * n_cells[0], N * N, 2)
    else:
        X = ax[0][None, None, :, None, None, :]
        Y = ax[1][None, :, None, None, :, None]
        Z = ax[2][:, None, None, :, None, None]
        X, Y, Z = np.broadcast_arrays(X, Y, Z)
        xi = np.stack([X, Y, Z], axis=-1).reshape(n_cells[2] * n_cells[1] * n_cells[0], N ** 3, 3)
    return mapping(xi)


def mapped_geometry(dim, degree, n_cells, mapping):
    """Per-quadrature-point geometry of the iso-parametric FE_Q(k) map with
    nodes on mapping(xi), the data MatrixFree stores for a MappingQ:
    [cells, nq, dim*dim + 1] = J^{-1} row-major (d xi_e / d x_d) and JxW."""
    S, G = _lagrange_tables(degree)
    _, w1 = gauss_points(degree + 1)
    Xn = mapped_nodes(dim, degree, n_cells, mapping)  # [cells, nv, dim]
    # D_j[q, a] = d phi_a / d xi_j at q (tensor products, x fastest)
    mats = []
    for j in range(dim):
        M = np.ones((1, 1))
        for d in reversed(range(dim)):
            M = np.kron(M, G if d == j else S)
        mats.append(M)
    wq = np.ones(1)
    for d in reversed(range(dim)):
        wq = np.kron(wq, w1)
    J = np.stack([np.einsum("qa,cai->cqi", Dj, Xn) for Dj in mats], axis=-1)  # [c, q, i(x), j(xi)]
    Jinv = np.linalg.inv(J)  # [c, q, e(xi), d(x)]
    jxw = np.linalg.det(J) * wq[None, :]
    cells, nq = jxw.shape
    return np.concatenate([Jinv.reshape(cells, nq, dim * dim), jxw[..., None]], axis=-1)


def mapped_qp_coordinates(dim, degree, n_cells, mapping):
    """Physical coordinates of the quadrature points under the iso-parametric map: [cells, nq, dim]."""
    S, _ = _lagrange_tables(degree)
    M = np.ones((1, 1))
    for d in reversed(range(dim)):
        M = np.kron(M, S)
    return np.einsum("qa,cai->cqi", M, mapped_nodes(dim, degree, n_cells, mapping))


def shell_temperature(x, seed=0, r0=0.55, r1=1.0, amplitude=0.01):
    """Conductive profile of a spherical shell, (1/r - 1/R1) / (1/R0 - 1/R1),
    plus a seeded perturbation of spherical-harmonic degree <= 8 (a sum of
    powers up to 8 of x/|x| . d_i for random directions d_i) under a radial
    envelope vanishing at both surfaces."""
    rng = np.random.default_rng(seed)
    r = np.linalg.norm(x, axis=-1)
    n = x / r[..., None]
    dim = x.shape[-1]
    pert = np.zeros_like(r)
    for _ in range(6):
        d = rng.normal(size=dim)
        d /= np.linalg.norm(d)
        pert += rng.normal() * (n @ d) ** rng.integers(1, 9)
    env = np.sin(np.pi * (r - r0) / (r1 - r0))
    return (1.0 / r - 1.0 / r1) / (1.0 / r0 - 1.0 / r1) + amplitude * env * pert


def shell_viscosity(x, seed=0):
    """eta = exp(-ln(1e4) T), times 30 below r = 0.9."""
    T = shell_temperature(x, seed)
    r = np.linalg.norm(x, axis=-1)
    return np.exp(-np.log(1e4) * T) * np.where(r < 0.9, 30.0, 1.0)


def config_shell(dim=3, global_refinement=2, seed=0, fraction=0.2):
    """Spherical shell R = [0.55, 1] (96 coarse cells in 3D), global refinement,
    then one adaptive level on the cells with the largest |dT/dr| (the given
    fraction, by cell count).  Returns the refined Tria."""
    from .tria import ShellCoarse, Tria
    t = Tria(ShellCoarse(dim), global_refinement)
    c = t.cell_centers()
    r = np.linalg.norm(c, axis=1)
    h = 1e-4
    dTdr = (shell_temperature(c * ((r + h) / r)[:, None], seed) -
            shell_temperature(c * ((r - h) / r)[:, None], seed)) / (2 * h)
    k = int(round(fraction * t.n_cells))
    flags = np.zeros(t.n_cells, dtype=bool)
    flags[np.argsort(-np.abs(dTdr), kind="stable")[:k]] = True
    t.refine(flags)
    return t
