"""ORACLE — test infrastructure only.

ctypes bindings to ``oracle/liboracle.so``, the CPU FP64 restatement of
ASPECT's matrix-free Stokes solve (see the header of ``oracle/oracle.cpp`` for
the reference files it follows).  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` import this package; the product
in ``aspect_amd/`` never does.

How each function is pinned: the brute-force dense element-matrix path and
the sum-factorized path are checked against each other, both against closed
forms, and the solves against the iteration counts of the reference's own
test outputs (``tests/test_oracle.py``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build(force: bool = False) -> str:
    """Compile the oracle with g++ (OpenMP).  Returns the library path."""
    srcs = [os.path.join(_HERE, f) for f in ("oracle.cpp", "cpu_port.cpp")]
    if force or not os.path.exists(_LIB_PATH) or max(map(os.path.getmtime, srcs)) > os.path.getmtime(_LIB_PATH):
        cmd = ["g++", "-O3", "-march=x86-64-v2", "-fopenmp", "-shared", "-fPIC", "-std=c++17",
               "-o", _LIB_PATH + ".tmp", *srcs]
        subprocess.run(cmd, check=True)
        os.replace(_LIB_PATH + ".tmp", _LIB_PATH)
    return _LIB_PATH


class Problem(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("degree", C.c_int),
        ("coarse", C.c_int * 3),
        ("n_levels", C.c_int),
        ("extent", C.c_double * 3),
        ("bc", C.c_int * 6),
        ("compressible", C.c_int),
        ("pressure_scaling", C.c_double),
    ]


class SolveParams(C.Structure):
    _fields_ = [
        ("tolerance", C.c_double),
        ("n_cheap_steps", C.c_int),
        ("n_expensive_steps", C.c_int),
        ("restart", C.c_int),
        ("a_block_tolerance", C.c_double),
        ("s_block_tolerance", C.c_double),
        ("level_eta_mode", C.c_int),
        ("cheap_krylov", C.c_int),   # 0 GMRES, 1 IDR(s)
        ("idr_s", C.c_int),
        ("solve_newton_system", C.c_int),
    ]


class SolveResult(C.Structure):
    _fields_ = [
        ("cheap_iterations", C.c_int),
        ("expensive_iterations", C.c_int),
        ("a_preconditioner_applications", C.c_int),
        ("s_preconditioner_applications", C.c_int),
        ("converged", C.c_int),
        ("initial_nonlinear_residual", C.c_double),
        ("final_linear_residual", C.c_double),
        ("solver_tolerance", C.c_double),
        ("r0_norm", C.c_double),
        ("pressure_normalization_adjustment", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = C.CDLL(_LIB_PATH)
        _lib.orc_last_error.restype = C.c_char_p
        _lib.orc_port_last_error.restype = C.c_char_p
        _lib.orc_port_isa.restype = C.c_char_p
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


def make_problem(dim=3, degree=2, coarse=(1, 1, 1), n_levels=3, extent=(1.0, 1.0, 1.0),
                 bc=(1, 1, 1, 1, 1, 1), compressible=False, pressure_scaling=1.0) -> Problem:
    p = Problem()
    p.dim = dim
    p.degree = degree
    for i in range(3):
        p.coarse[i] = int(coarse[i]) if i < len(coarse) else 1
        p.extent[i] = float(extent[i]) if i < len(extent) else 1.0
    p.n_levels = n_levels
    for i in range(6):
        p.bc[i] = int(bc[i]) if i < len(bc) else 0
    p.compressible = int(compressible)
    p.pressure_scaling = pressure_scaling
    return p


def sizes(p: Problem, level: int):
    out = np.zeros(4, dtype=np.int64)
    _check(lib().orc_sizes(C.byref(p), level, _p(out)))
    return dict(cells=int(out[0]), nu=int(out[1]), np=int(out[2]), vnodes=int(out[3]))


def n_qp(p: Problem) -> int:
    return (p.degree + 1) ** p.dim


def constrained(p, level):
    s = sizes(p, level)
    out = np.zeros(s["nu"], dtype=np.uint8)
    _check(lib().orc_constrained(C.byref(p), level, _p(out)))
    return out.astype(bool)


def qp_points(p, level):
    s = sizes(p, level)
    out = np.zeros((s["cells"], n_qp(p), p.dim))
    _check(lib().orc_qp_points(C.byref(p), level, _p(out)))
    return out


def quadrature_weights(p, level):
    out = np.zeros(n_qp(p))
    _check(lib().orc_quadrature_weights(C.byref(p), level, _p(out)))
    return out


def velocity_support_points(p, level):
    s = sizes(p, level)
    out = np.zeros((s["vnodes"], p.dim))
    _check(lib().orc_velocity_support_points(C.byref(p), level, _p(out)))
    return out


def pressure_support_points(p, level):
    s = sizes(p, level)
    out = np.zeros((s["np"], p.dim))
    _check(lib().orc_pressure_support_points(C.byref(p), level, _p(out)))
    return out


def _eta(eta):
    eta = np.ascontiguousarray(eta, dtype=np.float64)
    per_qp = 1 if eta.ndim == 2 else 0
    return eta, per_qp


def stokes_vmult(p, level, eta, src, out=None):
    eta, per_qp = _eta(eta)
    src = np.ascontiguousarray(src, dtype=np.float64)
    dst = np.zeros_like(src) if out is None else out
    _check(lib().orc_stokes_vmult(C.byref(p), level, _p(eta), per_qp, _p(src), _p(dst)))
    return dst


def ablock_vmult(p, level, eta, src):
    eta, per_qp = _eta(eta)
    src = np.ascontiguousarray(src, dtype=np.float64)
    dst = np.zeros_like(src)
    _check(lib().orc_ablock_vmult(C.byref(p), level, _p(eta), per_qp, _p(src), _p(dst)))
    return dst


def bt_vmult(p, level, src_p):
    s = sizes(p, level)
    src_p = np.ascontiguousarray(src_p, dtype=np.float64)
    dst = np.zeros(s["nu"])
    _check(lib().orc_bt_vmult(C.byref(p), level, _p(src_p), _p(dst)))
    return dst


def mass_vmult(p, level, eta, src):
    eta, per_qp = _eta(eta)
    src = np.ascontiguousarray(src, dtype=np.float64)
    dst = np.zeros_like(src)
    _check(lib().orc_mass_vmult(C.byref(p), level, _p(eta), per_qp, _p(src), _p(dst)))
    return dst


def ablock_diagonal(p, level, eta):
    eta, per_qp = _eta(eta)
    out = np.zeros(sizes(p, level)["nu"])
    _check(lib().orc_ablock_diagonal(C.byref(p), level, _p(eta), per_qp, _p(out)))
    return out


def mass_diagonal(p, level, eta):
    eta, per_qp = _eta(eta)
    out = np.zeros(sizes(p, level)["np"])
    _check(lib().orc_mass_diagonal(C.byref(p), level, _p(eta), per_qp, _p(out)))
    return out


def cell_matrices(p, level, eta):
    eta, per_qp = _eta(eta)
    s = sizes(p, level)
    nv = (p.degree + 1) ** p.dim
    npl = p.degree ** p.dim
    nloc = p.dim * nv + npl
    K = np.zeros((s["cells"], nloc, nloc))
    M = np.zeros((s["cells"], npl, npl))
    D = np.zeros((s["cells"], nloc), dtype=np.int64)
    _check(lib().orc_cell_matrices(C.byref(p), level, _p(eta), per_qp, _p(K), _p(M), _p(D)))
    return K, M, D


def prolongate(p, fine_level, kind, src):
    s = sizes(p, fine_level)
    src = np.ascontiguousarray(src, dtype=np.float64)
    dst = np.zeros(s["nu"] if kind == 0 else s["np"])
    _check(lib().orc_prolongate(C.byref(p), fine_level, kind, _p(src), _p(dst)))
    return dst


def restrict(p, fine_level, kind, src):
    s = sizes(p, fine_level - 1)
    src = np.ascontiguousarray(src, dtype=np.float64)
    dst = np.zeros(s["nu"] if kind == 0 else s["np"])
    _check(lib().orc_restrict(C.byref(p), fine_level, kind, _p(src), _p(dst)))
    return dst


def level_viscosity(p, eta):
    eta, per_qp = _eta(eta)
    counts = [sizes(p, l)["cells"] for l in range(p.n_levels)]
    out = np.zeros(sum(counts))
    _check(lib().orc_level_viscosity(C.byref(p), _p(eta), per_qp, _p(out)))
    res, off = [], 0
    for c in counts:
        res.append(out[off:off + c].copy())
        off += c
    return res


def stokes_rhs_correction(p, eta, u0):
    """-A(u0) with constrained values of u0 read as stored (correct_stokes_rhs)."""
    eta, per_qp = _eta(eta)
    u0 = np.ascontiguousarray(u0, dtype=np.float64)
    out = np.zeros_like(u0)
    _check(lib().orc_stokes_rhs_correction(C.byref(p), _p(eta), per_qp, _p(u0), _p(out)))
    return out


def level_viscosity_q1(p, eta_qp):
    """Per-quadrature-point viscosity tables of every level (level 0 first) for
    "project to Q1 only viscosity"; the last entry is the active table."""
    eta_qp = np.ascontiguousarray(eta_qp, dtype=np.float64)
    nq = n_qp(p)
    counts = [sizes(p, l)["cells"] for l in range(p.n_levels)]
    out = np.zeros(sum(counts) * nq)
    _check(lib().orc_level_viscosity_q1(C.byref(p), _p(eta_qp), _p(out)))
    res, off = [], 0
    for c in counts:
        res.append(out[off:off + c * nq].reshape(c, nq).copy())
        off += c * nq
    return res


def smoother_setup(p, eta):
    eta, per_qp = _eta(eta)
    out = np.zeros((p.n_levels, 2, 5))
    _check(lib().orc_smoother_setup(C.byref(p), _p(eta), per_qp, _p(out)))
    return out


def vcycle(p, eta, kind, src):
    eta, per_qp = _eta(eta)
    src = np.ascontiguousarray(src, dtype=np.float64)
    dst = np.zeros_like(src)
    _check(lib().orc_vcycle(C.byref(p), _p(eta), per_qp, kind, _p(src), _p(dst)))
    return dst


def free_surface_stab_size(p, faces):
    """Number of values of the stabilization table for the marked faces."""
    nq = (p.degree + 1) ** (p.dim - 1)
    n = [p.coarse[d] << (p.n_levels - 1) for d in range(p.dim)]
    tot = 0
    for f in range(2 * p.dim):
        if faces[f]:
            d = f // 2
            cells = int(np.prod([n[e] for e in range(p.dim) if e != d]))
            tot += cells * nq * p.dim
    return tot


def free_surface_apply(p, faces, stab, src_u, sign=1.0, plain=False, out=None):
    """The free-surface face term on the active level added to out (N_u values)."""
    faces = np.ascontiguousarray(np.resize(np.asarray(faces, dtype=np.int32), 6))
    stab = np.ascontiguousarray(stab, dtype=np.float64)
    src_u = np.ascontiguousarray(src_u, dtype=np.float64)
    dst = np.zeros_like(src_u) if out is None else out
    _check(lib().orc_free_surface_apply(C.byref(p), _p(faces), _p(stab), _p(src_u), _p(dst), C.c_double(sign),
                                        int(plain)))
    return dst


def time_vcycle(p, eta, kind, reps=1):
    """(set-up seconds, seconds per V-cycle) of the oracle's multigrid on the CPU."""
    eta, per_qp = _eta(eta)
    out = np.zeros(2)
    _check(lib().orc_time_vcycle(C.byref(p), _p(eta), per_qp, kind, int(reps), _p(out)))
    return float(out[0]), float(out[1])


def chebyshev(p, eta, kind, level, src):
    eta, per_qp = _eta(eta)
    src = np.ascontiguousarray(src, dtype=np.float64)
    dst = np.zeros_like(src)
    _check(lib().orc_chebyshev(C.byref(p), _p(eta), per_qp, kind, level, _p(src), _p(dst)))
    return dst


def rhs_buoyancy(p, rho_qp, g):
    s = sizes(p, p.n_levels - 1)
    rho_qp = np.ascontiguousarray(rho_qp, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    out = np.zeros(s["nu"] + s["np"])
    _check(lib().orc_rhs_buoyancy(C.byref(p), _p(rho_qp), _p(g), _p(out)))
    return out


def default_params(tolerance=1e-7, n_cheap=1000, n_expensive=1000, restart=100,
                   a_tol=1e-2, s_tol=1e-6, krylov="GMRES", idr_s=2, level_eta="DGQ0",
                   newton=False) -> SolveParams:
    prm = SolveParams()
    prm.tolerance = tolerance
    prm.n_cheap_steps = n_cheap
    prm.n_expensive_steps = n_expensive
    prm.restart = restart
    prm.a_block_tolerance = a_tol
    prm.s_block_tolerance = s_tol
    prm.level_eta_mode = {"DGQ0": 0, "DGQ1": 1}[level_eta]
    prm.cheap_krylov = {"GMRES": 0, "IDR(s)": 1}[krylov]
    prm.idr_s = idr_s
    prm.solve_newton_system = int(bool(newton))
    return prm


def solve(p, eta, rhs, x0=None, params: SolveParams | None = None):
    eta, per_qp = _eta(eta)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    x0 = np.zeros_like(rhs) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
    params = params or default_params()
    sol = np.zeros_like(rhs)
    res = SolveResult()
    _check(lib().orc_solve(C.byref(p), _p(eta), per_qp, C.byref(params), _p(rhs), _p(x0), _p(sol),
                           C.byref(res)))
    return sol, {f: getattr(res, f) for f, _ in SolveResult._fields_}


class CpuPort:
    """The SIMD-across-cells CPU Stokes apply (``cpu_port.cpp``) for the finest
    level of a 3D degree-2 box problem: the CPU baseline's timed operator."""

    def __init__(self, p: Problem):
        self._plan = C.c_void_p()
        if lib().orc_port_create(C.byref(p), C.byref(self._plan)) != 0:
            raise RuntimeError(lib().orc_port_last_error().decode())

    def stokes_vmult(self, eta, src, out=None, threads=0):
        eta, per_qp = _eta(eta)
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.empty_like(src) if out is None else out
        lib().orc_port_stokes_vmult(self._plan, _p(eta), per_qp, _p(src), _p(dst), int(threads))
        return dst

    def __del__(self):
        if getattr(self, "_plan", None):
            lib().orc_port_destroy(self._plan)
            self._plan = None


def simd_width_name() -> str:
    """Instruction set the CPU port's batch kernel runs with on this host."""
    return lib().orc_port_isa().decode()


def set_threads(n: int) -> None:
    """OpenMP thread count of the oracle's parallel loops."""
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    """OpenMP thread count the oracle's parallel loops run with."""
    return int(lib().orc_get_threads())


# ------------------------------------------------------------------ mesh deformation
def mesh_problem(p: Problem, mesh_bc) -> Problem:
    """The problem with the mesh-deformation constraints: per face 0 none,
    1 no normal displacement, 2 zero displacement, 3 prescribed displacement
    (constrained like 2, values from u0)."""
    q = Problem.from_buffer_copy(p)
    for i in range(6):
        b = int(mesh_bc[i]) if i < len(mesh_bc) else 0
        q.bc[i] = 2 if b == 3 else b
    return q


def laplace_vmult(p, level, src):
    src = np.ascontiguousarray(src, dtype=np.float64)
    dst = np.zeros_like(src)
    _check(lib().orc_laplace_vmult(C.byref(p), level, _p(src), _p(dst)))
    return dst


def laplace_diagonal(p, level):
    out = np.zeros(sizes(p, level)["nu"])
    _check(lib().orc_laplace_diagonal(C.byref(p), level, _p(out)))
    return out


def mesh_smoother_info(p):
    """[level][lambda_min, lambda_max, theta, delta, degree] and eigen-CG iterations per level."""
    nl = p.n_levels
    out = np.zeros(nl * 6)
    _check(lib().orc_mesh_smoother_info(C.byref(p), _p(out)))
    return out[:nl * 5].reshape(nl, 5), out[nl * 5:]


def mesh_vcycle(p, src):
    src = np.ascontiguousarray(src, dtype=np.float64)
    dst = np.zeros_like(src)
    _check(lib().orc_mesh_vcycle(C.byref(p), _p(src), _p(dst)))
    return dst


def mesh_solve(p, u0, tolerance=1e-10):
    u0 = np.ascontiguousarray(u0, dtype=np.float64)
    sol = np.zeros_like(u0)
    its = C.c_int()
    res = np.zeros(2)
    _check(lib().orc_mesh_solve(C.byref(p), _p(u0), C.c_double(tolerance), _p(sol), C.byref(its), _p(res)))
    return sol, its.value, res


# ------------------------------------------------------------------ general geometry
def set_geometry(tables):
    """Geometry tables of the levels for every following call: a list (level 0
    first) of [cells, nq, dim*dim + 1] arrays (J^{-1} row-major, JxW), or
    None to return to the box."""
    _check(lib().orc_clear_geometry())
    if tables is None:
        return
    for l, t in enumerate(tables):
        t = np.ascontiguousarray(t, dtype=np.float64).reshape(-1)
        _check(lib().orc_set_geometry(l, C.c_longlong(t.size), _p(t)))
