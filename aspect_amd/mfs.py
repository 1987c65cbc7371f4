"""Python mirror of ASPECT's ``StokesMatrixFreeHandler`` over the C-ABI of
``libaspect_mfs.so`` (declared in ``include/aspect_mfs.h``).

The class keeps the reference interface's method names and meaning
(include/aspect/simulator/solver/stokes_matrix_free.h:47-98 and
include/aspect/simulator/solver/interface.h:36-101): ``setup_dofs``,
``assemble``, ``build_preconditioner``, ``solve`` returning ``SolverOutputs``,
``name``, and the memory getters.  All arithmetic runs in the HIP library;
this module only marshals arrays.  Loading fails loudly when the library is
missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaspect_mfs.so")

MFS_OK, MFS_ERROR, MFS_NO_CONVERGENCE = 0, 1, 2
GMG_FP64, GMG_FP32 = 0, 1
AVERAGE = {"none": 0, "arithmetic average only viscosity": 1, "harmonic average only viscosity": 2,
           "geometric average only viscosity": 3, "project to Q1 only viscosity": 4,
           "arithmetic": 1, "harmonic": 2, "geometric": 3, "q1": 4}
BC_OPEN, BC_TANGENTIAL, BC_ZERO = 0, 1, 2

# every symbol include/aspect_mfs.h declares
EXPORTED = [
    "mfs_last_error", "mfs_create", "mfs_destroy", "mfs_name", "mfs_sizes", "mfs_set_viscosity",
    "mfs_set_pressure_scaling", "mfs_build_preconditioner", "mfs_stokes_vmult", "mfs_ablock_vmult",
    "mfs_mass_vmult", "mfs_bt_vmult", "mfs_ablock_diagonal", "mfs_mass_diagonal", "mfs_prolongate",
    "mfs_restrict", "mfs_level_viscosity", "mfs_smoother_info", "mfs_chebyshev", "mfs_vcycle",
    "mfs_block_preconditioner", "mfs_solve", "mfs_assemble_buoyancy_rhs", "mfs_stokes_vmult_device",
    "mfs_ablock_vmult_device", "mfs_solve_device", "mfs_assemble_buoyancy_rhs_device", "mfs_synchronize",
    "mfs_memory_consumption", "mfs_time_operator", "mfs_set_graph_capture", "mfs_correct_stokes_rhs",
    "mfs_ensemble_create", "mfs_ensemble_destroy", "mfs_ensemble_member", "mfs_ensemble_setup", "mfs_ensemble_solve",
    "mfs_validate_operator_arguments", "mfs_ablock_vmult_device_noalloc", "mfs_set_free_surface",
    "mfs_mesh_deformation_solve", "mfs_mesh_laplace_vmult", "mfs_mesh_vcycle", "mfs_set_geometry",
    "mfs_box_dof_index", "mfs_tria_create", "mfs_tria_destroy", "mfs_tria_sizes", "mfs_tria_set_viscosity",
    "mfs_tria_vmult", "mfs_tria_vmult_device", "mfs_tria_time", "mfs_tria_get_viscosity", "mfs_tria_correct_stokes_rhs", "mfs_tria_gmg_create", "mfs_tria_gmg_destroy", "mfs_tria_gmg_set_viscosity", "mfs_tria_gmg_build",
    "mfs_tria_gmg_set_graph_capture", "mfs_probe_contraction",
    "mfs_tria_gmg_info", "mfs_tria_gmg_vcycle", "mfs_tria_solve",
]


class Box(C.Structure):
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
        ("cheap_krylov", C.c_int),
        ("idr_s", C.c_int),
        ("solve_newton_system", C.c_int),
        ("failure_strategy", C.c_int),
        ("history_file", C.c_char_p),
    ]


# "Krylov method for cheap solver steps"
KRYLOV = {"GMRES": 0, "IDR(s)": 1}
# "Linear solver failure strategy"
FAILURE = {"abort": 0, "continue with nonlinear solver": 1}
# operator codes and modes of mfs_validate_operator_arguments
OPS = {"stokes": 0, "ablock": 1, "bt": 2, "mass": 3, "adiag": 4, "mdiag": 5, "rhs": 6, "laplace": 7, "ldiag": 8}
MODES = {"set": 0, "resid": 1, "add": 2}


def _params(tolerance, n_cheap_steps, n_expensive_steps, restart, a_block_tolerance, s_block_tolerance, krylov,
            idr_s, solve_newton_system=False, failure_strategy="abort", history_file=None):
    hf = None if history_file is None else os.fsencode(history_file)
    return SolveParams(tolerance, n_cheap_steps, n_expensive_steps, restart, a_block_tolerance, s_block_tolerance,
                       KRYLOV[krylov], idr_s, int(bool(solve_newton_system)), FAILURE[failure_strategy], hf)


def _vec(a, n, what):
    """Contiguous FP64 copy of a host vector, checked against the length the library reads or writes."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 1 or a.size != n:
        raise ValueError(f"{what} must be a vector of {n} values (got shape {a.shape})")
    return a


def box_dof_index(box, points, component):
    """Position of each DoF (support point [n, dim] or [n, 3], vector component
    0..dim-1 velocity, dim pressure) in the library's block vector of the
    box's finest level: the permutation from a deal.II DoF numbering (no
    device needed).  Raises SolverError for a point off the box's nodes or two
    DoFs at one node."""
    lib = load_library()
    pts = np.asarray(points, dtype=np.float64)
    n = pts.shape[0]
    p3 = np.zeros((n, 3))
    p3[:, : pts.shape[1]] = pts
    comp = np.ascontiguousarray(component, dtype=np.int32)
    if comp.shape != (n,):
        raise ValueError("component must hold one entry per support point")
    out = np.zeros(n, dtype=np.int64)
    _check(lib.mfs_box_dof_index(C.byref(box), C.c_longlong(n), _ptr(np.ascontiguousarray(p3)),
                                 comp.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p)))
    return out


def validate_operator_arguments(op, mode, src_u=None, src_p=None, eta=None, rhs_u=None, rhs_p=None, dst_u=None,
                                dst_p=None):
    """Host-side pointer check the library runs before every operator launch
    (no device needed); arguments are addresses (ints) or None.  Raises
    SolverError naming the first null pointer the (op, mode) kernel uses."""
    lib = load_library()
    ptrs = [C.c_void_p(x) if x else None for x in (src_u, src_p, eta, rhs_u, rhs_p, dst_u, dst_p)]
    _check(lib.mfs_validate_operator_arguments(OPS[op], MODES[mode], *ptrs))


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
        ("solve_seconds", C.c_double),
        ("cheap_final_residual", C.c_double),
        ("cheap_seconds", C.c_double),
    ]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load the HIP library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise ImportError(f"{path} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                              " (no CPU fallback exists)")
        lib = C.CDLL(path)
        lib.mfs_last_error.restype = C.c_char_p
        for name in EXPORTED:
            getattr(lib, name)  # raises AttributeError if a declared symbol is missing
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class SolverError(RuntimeError):
    pass


class NoConvergence(SolverError):
    """Mirror of SolverControl::NoConvergence raised by the handler's solve."""

    def __init__(self, msg, outputs):
        super().__init__(msg)
        self.outputs = outputs


def _check(rc):
    if rc == MFS_ERROR:
        raise SolverError(load_library().mfs_last_error().decode())
    return rc


@dataclass
class SolverOutputs:
    """StokesSolver::SolverOutputs (interface.h:36-58) plus solver counters."""
    initial_nonlinear_residual: float
    final_linear_residual: float
    pressure_normalization_adjustment: float
    cheap_iterations: int
    expensive_iterations: int
    a_preconditioner_applications: int
    s_preconditioner_applications: int
    converged: bool
    solver_tolerance: float
    r0_norm: float
    solve_seconds: float
    cheap_final_residual: float = 0.0
    cheap_seconds: float = 0.0


def make_box(dim=3, degree=2, coarse=(1, 1, 1), n_levels=3, extent=(1.0, 1.0, 1.0), bc=(1,) * 6,
             compressible=False, pressure_scaling=1.0) -> Box:
    b = Box()
    b.dim, b.degree, b.n_levels = dim, degree, n_levels
    for i in range(3):
        b.coarse[i] = int(coarse[i]) if i < len(coarse) else 1
        b.extent[i] = float(extent[i]) if i < len(extent) else 1.0
    for i in range(6):
        b.bc[i] = int(bc[i]) if i < len(bc) else 0
    b.compressible = int(compressible)
    b.pressure_scaling = float(pressure_scaling)
    return b


class StokesMatrixFreeHandler:
    """GPU implementation of the matrix-free Stokes handler on box meshes.

    ``gmg_precision`` selects GMGNumberType of the V-cycle: "fp64" (the
    reference's ``using GMGNumberType = double``) or "fp32".
    """

    def __init__(self, box: Box, gmg_precision: str = "fp64"):
        self.lib = load_library()
        self.box = box
        self.gmg_precision = gmg_precision
        self._h = C.c_void_p()
        self._averaging = 0
        self.setup_dofs()

    # -- StokesMatrixFreeHandler interface ---------------------------------
    def setup_dofs(self):
        if self._h:
            self.lib.mfs_destroy(self._h)
            self._h = C.c_void_p()
        prec = GMG_FP32 if self.gmg_precision == "fp32" else GMG_FP64
        _check(self.lib.mfs_create(C.byref(self.box), prec, C.byref(self._h)))

    def assemble(self, eta_qp, averaging="none", pressure_scaling=None):
        """evaluate_material_model(): viscosity at every quadrature point."""
        if pressure_scaling is not None:
            _check(self.lib.mfs_set_pressure_scaling(self._h, C.c_double(pressure_scaling)))
        eta = np.ascontiguousarray(eta_qp, dtype=np.float64)
        s = self.sizes(self.top)
        if eta.size != s["cells"] * self.n_qp:
            raise ValueError(f"viscosity table must hold {s['cells']} x {self.n_qp} values")
        mode = AVERAGE[averaging] if isinstance(averaging, str) else int(averaging)
        _check(self.lib.mfs_set_viscosity(self._h, _ptr(eta), mode))
        self._averaging = mode

    def set_geometry(self, level, geo):
        """Per-quadrature-point geometry of a level: [cells, nq, dim*dim + 1]
        (J^{-1} row-major, then JxW)."""
        s = self.sizes(level)
        d = self.box.dim
        g = np.ascontiguousarray(geo, dtype=np.float64).reshape(-1)
        if g.size != s["cells"] * self.n_qp * (d * d + 1):
            raise ValueError(f"geometry table must hold {s['cells']} x {self.n_qp} x {d * d + 1} values")
        _check(self.lib.mfs_set_geometry(self._h, level, _ptr(g)))

    def set_free_surface(self, faces, stab=None):
        """Free-surface stabilization on the marked faces (x-, x+, y-, y+, z-, z+):
        stab holds dim values per face quadrature point of the marked faces."""
        f = (C.c_int * 6)(*[int(v) for v in list(faces) + [0] * (6 - len(faces))])
        if stab is None:
            _check(self.lib.mfs_set_free_surface(self._h, f, None))
            return
        st = np.ascontiguousarray(stab, dtype=np.float64).reshape(-1)
        d = self.box.dim
        nq = (self.box.degree + 1) ** (d - 1)
        n = [self.box.coarse[e] << (self.box.n_levels - 1) for e in range(d)]
        need = sum(int(np.prod([n[e] for e in range(d) if e != fi // 2])) * nq * d for fi in range(2 * d) if f[fi])
        if st.size != need:
            raise ValueError(f"stabilization table must hold {need} values")
        _check(self.lib.mfs_set_free_surface(self._h, f, _ptr(st)))

    # -- mesh deformation (MeshDeformationHandler::compute_mesh_displacements_gmg)
    def _mesh_bc(self, bc):
        return (C.c_int * 6)(*[int(v) for v in list(bc) + [0] * (6 - len(bc))])

    def mesh_deformation_solve(self, bc, u0, tolerance=1e-10):
        """Vector Laplace displacement solve; bc per face: 0 none, 1 no normal
        displacement, 2 zero, 3 prescribed (from u0).  Returns (solution,
        iterations, (residual, required residual))."""
        nu = self.sizes(self.top)["nu"]
        u0 = _vec(u0, nu, "u0")
        sol = np.zeros(nu)
        its = C.c_int()
        res = np.zeros(2)
        rc = _check(self.lib.mfs_mesh_deformation_solve(self._h, self._mesh_bc(bc), _ptr(u0), C.c_double(tolerance),
                                                        _ptr(sol), C.byref(its), _ptr(res)))
        if rc == MFS_NO_CONVERGENCE:
            raise NoConvergence(self.lib.mfs_last_error().decode(), (its.value, res))
        return sol, its.value, (res[0], res[1])

    def mesh_laplace_vmult(self, bc, level, src):
        n = self.sizes(level)["nu"]
        return self._unary(self.lib.mfs_mesh_laplace_vmult, (self._mesh_bc(bc), level), src, n)

    def mesh_vcycle(self, bc, src):
        """One V-cycle of the mesh multigrid and the smoother set-up
        ([level][lambda_min, lambda_max, theta, delta, degree, eigen-CG its])."""
        n = self.sizes(self.top)["nu"]
        src = _vec(src, n, "source vector")
        dst = np.zeros(n)
        info = np.zeros(6 * (self.top + 1))
        _check(self.lib.mfs_mesh_vcycle(self._h, self._mesh_bc(bc), _ptr(src), _ptr(dst), _ptr(info)))
        return dst, info.reshape(self.top + 1, 6)

    def build_preconditioner(self):
        _check(self.lib.mfs_build_preconditioner(self._h))

    def solve(self, rhs, x0=None, tolerance=1e-7, n_cheap_steps=1000, n_expensive_steps=1000, restart=100,
              a_block_tolerance=1e-2, s_block_tolerance=1e-6, raise_on_failure=True, krylov="GMRES", idr_s=2,
              solve_newton_system=False, failure_strategy="abort", history_file=None):
        """solve() (interface.h:89-94): rhs and x0 hold N_u + N_p values.  A
        failed solve raises NoConvergence (with the library's message) unless
        raise_on_failure is False or failure_strategy is "continue with
        nonlinear solver"."""
        s = self.sizes(self.top)
        n = s["nu"] + s["np"]
        rhs = _vec(rhs, n, "rhs")
        sol = np.zeros_like(rhs)
        prm = _params(tolerance, n_cheap_steps, n_expensive_steps, restart, a_block_tolerance, s_block_tolerance,
                      krylov, idr_s, solve_newton_system, failure_strategy, history_file)
        res = SolveResult()
        xp = None
        if x0 is not None:
            xp = _vec(x0, n, "x0")
        rc = _check(self.lib.mfs_solve(self._h, _ptr(rhs), _ptr(xp) if xp is not None else None, _ptr(sol),
                                       C.byref(prm), C.byref(res)))
        out = SolverOutputs(res.initial_nonlinear_residual, res.final_linear_residual,
                            res.pressure_normalization_adjustment, res.cheap_iterations, res.expensive_iterations,
                            res.a_preconditioner_applications, res.s_preconditioner_applications,
                            bool(res.converged), res.solver_tolerance, res.r0_norm, res.solve_seconds,
                            res.cheap_final_residual, res.cheap_seconds)
        if rc == MFS_NO_CONVERGENCE and raise_on_failure:
            raise NoConvergence(self.lib.mfs_last_error().decode(), out)
        return sol, out

    def name(self) -> str:
        buf = C.create_string_buffer(64)
        _check(self.lib.mfs_name(self._h, buf, 64))
        return buf.value.decode()

    def memory_consumption(self):
        out = (C.c_longlong * 3)()
        _check(self.lib.mfs_memory_consumption(self._h, out))
        return dict(vectors=out[0], cell_data=out[1], krylov=out[2])

    # -- operators and pieces (parity entry points) ------------------------
    @property
    def top(self):
        return self.box.n_levels - 1

    @property
    def n_qp(self):
        return (self.box.degree + 1) ** self.box.dim

    def sizes(self, level):
        out = (C.c_longlong * 4)()
        _check(self.lib.mfs_sizes(self._h, level, out))
        return dict(cells=out[0], nu=out[1], np=out[2], vnodes=out[3])

    def _unary(self, fn, level_args, src, nout, nin=None):
        src = _vec(src, nout if nin is None else nin, "source vector")
        dst = np.zeros(nout)
        _check(fn(self._h, *level_args, _ptr(src), _ptr(dst)))
        return dst

    def stokes_vmult(self, src):
        s = self.sizes(self.top)
        return self._unary(self.lib.mfs_stokes_vmult, (), src, s["nu"] + s["np"])

    def ablock_vmult(self, level, src):
        s = self.sizes(self.top if level == -1 else level)
        return self._unary(self.lib.mfs_ablock_vmult, (level,), src, s["nu"])

    def mass_vmult(self, level, src):
        s = self.sizes(self.top if level == -1 else level)
        return self._unary(self.lib.mfs_mass_vmult, (level,), src, s["np"])

    def bt_vmult(self, src_p):
        s = self.sizes(self.top)
        return self._unary(self.lib.mfs_bt_vmult, (), src_p, s["nu"], s["np"])

    def ablock_diagonal(self, level):
        out = np.zeros(self.sizes(level)["nu"])
        _check(self.lib.mfs_ablock_diagonal(self._h, level, _ptr(out)))
        return out

    def mass_diagonal(self, level):
        out = np.zeros(self.sizes(level)["np"])
        _check(self.lib.mfs_mass_diagonal(self._h, level, _ptr(out)))
        return out

    def prolongate(self, fine_level, kind, src):
        s, c = self.sizes(fine_level), self.sizes(fine_level - 1)
        key = "nu" if kind == 0 else "np"
        return self._unary(self.lib.mfs_prolongate, (fine_level, kind), src, s[key], c[key])

    def restrict(self, fine_level, kind, src):
        s, f = self.sizes(fine_level - 1), self.sizes(fine_level)
        key = "nu" if kind == 0 else "np"
        return self._unary(self.lib.mfs_restrict, (fine_level, kind), src, s[key], f[key])

    def level_viscosity(self, level):
        """Level table: one value per cell, or [cells, nq] with "project to Q1"."""
        cells = self.sizes(level)["cells"]
        nq = (self.box.degree + 1) ** self.box.dim
        out = np.zeros(cells * nq if self._averaging == 4 else cells)
        _check(self.lib.mfs_level_viscosity(self._h, level, _ptr(out)))
        return out.reshape(cells, nq) if self._averaging == 4 else out

    def smoother_info(self):
        out = np.zeros((self.box.n_levels, 2, 5))
        _check(self.lib.mfs_smoother_info(self._h, _ptr(out)))
        return out

    def chebyshev(self, kind, level, src):
        s = self.sizes(level)
        return self._unary(self.lib.mfs_chebyshev, (kind, level), src, s["nu"] if kind == 0 else s["np"])

    def vcycle(self, kind, src):
        s = self.sizes(self.top)
        return self._unary(self.lib.mfs_vcycle, (kind,), src, s["nu"] if kind == 0 else s["np"])

    def block_preconditioner(self, src):
        s = self.sizes(self.top)
        return self._unary(self.lib.mfs_block_preconditioner, (), src, s["nu"] + s["np"])

    def assemble_buoyancy_rhs(self, rho_qp, gravity):
        rho = np.ascontiguousarray(rho_qp, dtype=np.float64)
        g = np.ascontiguousarray(np.resize(np.asarray(gravity, dtype=np.float64), 3))
        s = self.sizes(self.top)
        if rho.size != s["cells"] * self.n_qp:
            raise ValueError(f"density table must hold {s['cells']} x {self.n_qp} values")
        out = np.zeros(s["nu"] + s["np"])
        _check(self.lib.mfs_assemble_buoyancy_rhs(self._h, _ptr(rho), _ptr(g), _ptr(out)))
        return out

    # device-pointer entry points (torch tensors or raw ints)
    def stokes_vmult_device(self, d_src: int, d_dst: int):
        _check(self.lib.mfs_stokes_vmult_device(self._h, C.c_void_p(d_src), C.c_void_p(d_dst)))

    def ablock_vmult_device(self, level: int, d_src: int, d_dst: int):
        _check(self.lib.mfs_ablock_vmult_device(self._h, level, C.c_void_p(d_src), C.c_void_p(d_dst)))

    def ablock_vmult_device_noalloc(self, level: int, d_src: int, d_dst: int):
        """The same apply without any device allocation: the work array is the
        library's fixed per-GPU scratch array (calls on one GPU must not overlap)."""
        _check(self.lib.mfs_ablock_vmult_device_noalloc(self._h, level, C.c_void_p(d_src), C.c_void_p(d_dst)))

    def assemble_buoyancy_rhs_device(self, d_rho: int, gravity, d_rhs: int):
        g = np.ascontiguousarray(np.resize(np.asarray(gravity, dtype=np.float64), 3))
        _check(self.lib.mfs_assemble_buoyancy_rhs_device(self._h, C.c_void_p(d_rho), _ptr(g), C.c_void_p(d_rhs)))

    def solve_device(self, d_rhs: int, d_x0, d_sol: int, tolerance=1e-7, n_cheap_steps=1000,
                     n_expensive_steps=1000, restart=100, a_block_tolerance=1e-2, s_block_tolerance=1e-6,
                     krylov="GMRES", idr_s=2, solve_newton_system=False, failure_strategy="abort",
                     history_file=None):
        prm = _params(tolerance, n_cheap_steps, n_expensive_steps, restart, a_block_tolerance, s_block_tolerance,
                      krylov, idr_s, solve_newton_system, failure_strategy, history_file)
        res = SolveResult()
        rc = _check(self.lib.mfs_solve_device(self._h, C.c_void_p(d_rhs), C.c_void_p(d_x0) if d_x0 else None,
                                              C.c_void_p(d_sol), C.byref(prm), C.byref(res)))
        return rc, {f: getattr(res, f) for f, _ in SolveResult._fields_}

    def time_operator(self, which: int, level: int, d_src: int, d_dst: int, reps: int) -> float:
        """Average device seconds of reps applications (hipEvents on the handle's stream)."""
        out = C.c_double()
        _check(self.lib.mfs_time_operator(self._h, which, level, C.c_void_p(d_src), C.c_void_p(d_dst), reps,
                                          C.byref(out)))
        return out.value

    def synchronize(self):
        _check(self.lib.mfs_synchronize(self._h))

    def correct_stokes_rhs(self, rhs, u0):
        """rhs - A(u0) for boundary values u0 on the constrained DoFs (correct_stokes_rhs)."""
        s = self.sizes(self.top)
        n = s["nu"] + s["np"]
        out = _vec(rhs, n, "rhs").copy()
        u0 = _vec(u0, n, "u0")
        _check(self.lib.mfs_correct_stokes_rhs(self._h, _ptr(u0), _ptr(out)))
        return out

    def set_graph_capture(self, enable: bool):
        """Replay V-cycles from a captured hipGraph (default) or launch them eagerly."""
        _check(self.lib.mfs_set_graph_capture(self._h, int(bool(enable))))

    @classmethod
    def _member(cls, h, box: Box, gmg_precision: str):
        """Non-owning view of an ensemble member's handle."""
        obj = cls.__new__(cls)
        obj.lib = load_library()
        obj.box = box
        obj.gmg_precision = gmg_precision
        obj._h = h
        obj._averaging = 0
        obj._owned = False
        return obj

    def __del__(self):
        try:
            if self._h and getattr(self, "_owned", True):
                self.lib.mfs_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass


class Ensemble:
    """Independent problems of uneven sizes solved together (one handle and
    HIP stream per problem, host worker threads)."""

    def __init__(self, boxes, gmg_precision: str = "fp32"):
        self.lib = load_library()
        self.boxes = list(boxes)
        self.gmg_precision = gmg_precision
        arr = (Box * len(self.boxes))(*self.boxes)
        self._e = C.c_void_p()
        prec = GMG_FP32 if gmg_precision == "fp32" else GMG_FP64
        _check(self.lib.mfs_ensemble_create(len(self.boxes), arr, prec, C.byref(self._e)))
        self.members = []
        for i, b in enumerate(self.boxes):
            h = C.c_void_p()
            _check(self.lib.mfs_ensemble_member(self._e, i, C.byref(h)))
            self.members.append(StokesMatrixFreeHandler._member(h, b, gmg_precision))

    def setup(self, etas, averaging="harmonic", n_threads=16):
        if len(etas) != len(self.members):
            raise ValueError(f"ensemble of {len(self.members)} problems: got {len(etas)} viscosity tables")
        etas = [np.ascontiguousarray(e, dtype=np.float64) for e in etas]
        for i, (e, m) in enumerate(zip(etas, self.members)):
            if e.size != m.sizes(m.top)["cells"] * m.n_qp:
                raise ValueError(f"viscosity table of member {i} has the wrong size")
        ptrs = (C.c_void_p * len(etas))(*[e.ctypes.data for e in etas])
        mode = AVERAGE[averaging] if isinstance(averaging, str) else int(averaging)
        _check(self.lib.mfs_ensemble_setup(self._e, ptrs, mode, n_threads))
        for m in self.members:
            m._averaging = mode

    def solve(self, rhss, tolerance=1e-8, n_cheap_steps=1000, n_expensive_steps=1000, restart=100,
              a_block_tolerance=1e-2, s_block_tolerance=1e-6, krylov="GMRES", idr_s=2, n_threads=16):
        if len(rhss) != len(self.members):
            raise ValueError(f"ensemble of {len(self.members)} problems: got {len(rhss)} right-hand sides")
        rhss = [_vec(r, m.sizes(m.top)["nu"] + m.sizes(m.top)["np"], f"rhs of member {i}")
                for i, (r, m) in enumerate(zip(rhss, self.members))]
        sols = [np.zeros_like(r) for r in rhss]
        rp = (C.c_void_p * len(rhss))(*[r.ctypes.data for r in rhss])
        sp = (C.c_void_p * len(sols))(*[x.ctypes.data for x in sols])
        prm = _params(tolerance, n_cheap_steps, n_expensive_steps, restart, a_block_tolerance, s_block_tolerance,
                      krylov, idr_s)
        res = (SolveResult * len(rhss))()
        secs = C.c_double(0)
        _check(self.lib.mfs_ensemble_solve(self._e, rp, sp, C.byref(prm), res, n_threads, C.byref(secs)))
        outs = [SolverOutputs(r.initial_nonlinear_residual, r.final_linear_residual,
                              r.pressure_normalization_adjustment, r.cheap_iterations, r.expensive_iterations,
                              r.a_preconditioner_applications, r.s_preconditioner_applications, bool(r.converged),
                              r.solver_tolerance, r.r0_norm, r.solve_seconds, r.cheap_final_residual,
                              r.cheap_seconds) for r in res]
        return sols, outs, secs.value

    def __del__(self):
        try:
            if self._e:
                self.members = []
                self.lib.mfs_ensemble_destroy(self._e)
                self._e = C.c_void_p()
        except Exception:
            pass


class TriaDesc(C.Structure):
    _fields_ = [("dim", C.c_int), ("degree", C.c_int), ("n_cells", C.c_longlong), ("n_nodes_u", C.c_longlong),
                ("n_nodes_p", C.c_longlong), ("cell_nodes_u", C.c_void_p), ("cell_nodes_p", C.c_void_p),
                ("geometry", C.c_void_p), ("compressible", C.c_int), ("pressure_scaling", C.c_double)]


class Constraints(C.Structure):
    _fields_ = [("n_lines", C.c_longlong), ("dofs", C.c_void_p), ("offsets", C.c_void_p), ("masters", C.c_void_p),
                ("weights", C.c_void_p)]


class TriaOperators:
    """Stokes level operators on a general mesh with hanging nodes (the level
    operators of the global-coarsening handler,
    source/simulator/solver/stokes_matrix_free_global_coarsening.cc:1455-1625).

    ``mesh`` carries the arrays ``aspect_amd.tria.Tria.build`` returns: cell
    node lists, geometry at the quadrature points and closed constraint lines
    ``(dofs, offsets, masters, weights)`` of the velocity and pressure blocks.
    """

    OPS = {"stokes": 0, "ablock": 1, "bt": 2, "mass": 3}

    def __init__(self, mesh, pressure_scaling=1.0, compressible=False):
        self.lib = load_library()
        self.dim = mesh.dim
        self._keep = []

        def arr(a, t):
            a = np.ascontiguousarray(a, dtype=t)
            self._keep.append(a)
            return a.ctypes.data

        d = TriaDesc()
        d.dim, d.degree = mesh.dim, 2
        d.n_cells, d.n_nodes_u, d.n_nodes_p = mesh.n_cells, mesh.n_nodes_u, mesh.n_nodes_p
        d.cell_nodes_u = arr(mesh.cell_u, np.int32)
        d.cell_nodes_p = arr(mesh.cell_p, np.int32)
        d.geometry = arr(mesh.geometry, np.float64)
        d.compressible = int(bool(compressible))
        d.pressure_scaling = float(pressure_scaling)

        def con(c):
            dofs, offs, masters, weights = c
            k = Constraints()
            k.n_lines = len(dofs)
            k.dofs, k.offsets = arr(dofs, np.int64), arr(offs, np.int64)
            k.masters, k.weights = arr(masters, np.int64), arr(weights, np.float64)
            return k

        cu, cp = con(mesh.con_u), con(mesh.con_p)
        self._t = C.c_void_p()
        _check(self.lib.mfs_tria_create(C.byref(d), C.byref(cu), C.byref(cp), C.byref(self._t)))
        self._keep = []
        out = (C.c_longlong * 3)()
        _check(self.lib.mfs_tria_sizes(self._t, out))
        self.n_cells, self.nu, self.np = out[0], out[1], out[2]

    def set_viscosity(self, eta_qp):
        e = np.ascontiguousarray(eta_qp, dtype=np.float64)
        if e.size != self.n_cells * 3 ** self.dim:
            raise ValueError("one viscosity per cell and quadrature point expected")
        _check(self.lib.mfs_tria_set_viscosity(self._t, _ptr(e)))

    def correct_stokes_rhs(self, u0, rhs):
        """rhs - K u0 with u0 read as stored (prescribed values on constrained
        DoFs) and nothing added to constrained rows."""
        u = _vec(u0, self.nu + self.np, "u0")
        r = _vec(rhs, self.nu + self.np, "rhs").copy()
        _check(self.lib.mfs_tria_correct_stokes_rhs(self._t, _ptr(u), _ptr(r)))
        return r

    def viscosity(self):
        """The object's viscosity table [cell][q]."""
        out = np.empty(self.n_cells * 3 ** self.dim)
        _check(self.lib.mfs_tria_get_viscosity(self._t, _ptr(out)))
        return out.reshape(self.n_cells, -1)

    def vmult(self, op, src):
        o = self.OPS[op]
        n_in = {0: self.nu + self.np, 1: self.nu, 2: self.np, 3: self.np}[o]
        n_out = {0: self.nu + self.np, 1: self.nu, 2: self.nu, 3: self.np}[o]
        x = _vec(src, n_in, "source vector")
        y = np.empty(n_out)
        _check(self.lib.mfs_tria_vmult(self._t, o, _ptr(x), _ptr(y)))
        return y

    def vmult_device(self, op, d_src, d_dst):
        _check(self.lib.mfs_tria_vmult_device(self._t, self.OPS[op], C.c_void_p(d_src), C.c_void_p(d_dst)))

    def time(self, op, d_src, d_dst, reps):
        """Average device seconds of reps applies (hipEvents on the object's stream)."""
        out = C.c_double(0)
        _check(self.lib.mfs_tria_time(self._t, self.OPS[op], C.c_void_p(d_src), C.c_void_p(d_dst), int(reps),
                                      C.byref(out)))
        return out.value

    def __del__(self):
        try:
            if self._t:
                self.lib.mfs_tria_destroy(self._t)
                self._t = C.c_void_p()
        except Exception:
            pass


class Csr(C.Structure):
    _fields_ = [("rows", C.c_longlong), ("cols", C.c_longlong), ("row_ptr", C.c_void_p), ("col_idx", C.c_void_p),
                ("values", C.c_void_p)]


class TriaMultigrid:
    """Global-coarsening multigrid over general levels (coarse first) and the
    cheap Stokes solve on the finest level
    (source/simulator/solver/stokes_matrix_free_global_coarsening.cc:800-883,
    :1259-1313).  ``levels``: ``TriaOperators`` with their viscosity set;
    ``prolong_u`` / ``prolong_p``: scipy CSR prolongations, entry l mapping
    level l-1 to level l (entry 0 unused), as ``aspect_amd.tria.hierarchy``
    returns them; ``parents`` (optional): entry l maps the cells of level l to
    the cells of level l-1 (``aspect_amd.tria.parent_map``), needed by
    ``set_viscosity``."""

    def __init__(self, levels, prolong_u, prolong_p, parents=None):
        self.lib = load_library()
        self.levels = list(levels)   # kept alive: the multigrid object refers to them
        n = len(self.levels)
        keep = []

        def csr(P):
            P = P.tocsr()
            a = [np.ascontiguousarray(P.indptr, np.int64), np.ascontiguousarray(P.indices, np.int64),
                 np.ascontiguousarray(P.data, np.float64)]
            keep.extend(a)
            m = Csr()
            m.rows, m.cols = P.shape
            m.row_ptr, m.col_idx, m.values = a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data
            return m

        cu = (Csr * max(n - 1, 1))(*[csr(prolong_u[l]) for l in range(1, n)])
        cp = (Csr * max(n - 1, 1))(*[csr(prolong_p[l]) for l in range(1, n)])
        lv = (C.c_void_p * n)(*[t._t.value for t in self.levels])
        par = None
        if parents is not None:
            arrs = [np.ascontiguousarray(parents[l], np.int64) for l in range(1, n)]
            keep.extend(arrs)
            par = (C.c_void_p * max(n - 1, 1))(*[a.ctypes.data for a in arrs])
        self._g = C.c_void_p()
        _check(self.lib.mfs_tria_gmg_create(n, lv, cu, cp, par, C.byref(self._g)))

    def set_viscosity(self, eta_qp, averaging="none"):
        """Viscosity of every level from the finest level's values at its
        quadrature points (material averaging on the finest level, DGQ0 cell
        means transferred down)."""
        e = np.ascontiguousarray(eta_qp, dtype=np.float64)
        t = self.levels[-1]
        if e.size != t.n_cells * 3 ** t.dim:
            raise ValueError("one viscosity per cell and quadrature point of the finest level expected")
        mode = AVERAGE[averaging] if isinstance(averaging, str) else int(averaging)
        _check(self.lib.mfs_tria_gmg_set_viscosity(self._g, _ptr(e), mode))

    def build(self):
        _check(self.lib.mfs_tria_gmg_build(self._g))

    def set_graph_capture(self, enable):
        """Replay each V-cycle from a captured hipGraph (default) or launch it kernel by kernel."""
        _check(self.lib.mfs_tria_gmg_set_graph_capture(self._g, int(bool(enable))))

    def info(self, level, kind):
        out = (C.c_double * 4)()
        t = self.levels[level]
        diag = np.empty(t.nu if kind == 0 else t.np)
        _check(self.lib.mfs_tria_gmg_info(self._g, level, kind, out, _ptr(diag)))
        return {"lmin": out[0], "lmax": out[1], "its": int(out[2]), "colours": int(out[3])}, diag

    def vcycle(self, kind, src):
        t = self.levels[-1]
        x = _vec(src, t.nu if kind == 0 else t.np, "source vector")
        y = np.empty_like(x)
        _check(self.lib.mfs_tria_gmg_vcycle(self._g, kind, _ptr(x), _ptr(y)))
        return y

    def solve(self, rhs, tolerance=1e-8, n_cheap_steps=1000, n_expensive_steps=1000, restart=100,
              a_block_tolerance=1e-2, s_block_tolerance=1e-6, x0=None, solve_newton_system=False,
              failure_strategy="abort"):
        t = self.levels[-1]
        b = _vec(rhs, t.nu + t.np, "right-hand side")
        xg = None if x0 is None else _vec(x0, t.nu + t.np, "initial guess")
        x = np.empty_like(b)
        prm = _params(tolerance, n_cheap_steps, n_expensive_steps, restart, a_block_tolerance, s_block_tolerance,
                      "GMRES", 2, solve_newton_system, failure_strategy)
        res = SolveResult()
        rc = _check(self.lib.mfs_tria_solve(self._g, _ptr(b), _ptr(xg) if xg is not None else None, _ptr(x),
                                            C.byref(prm), C.byref(res)))
        return x, SolverOutputs(res.initial_nonlinear_residual, res.final_linear_residual,
                                res.pressure_normalization_adjustment, res.cheap_iterations,
                                res.expensive_iterations, res.a_preconditioner_applications,
                                res.s_preconditioner_applications, bool(res.converged), res.solver_tolerance,
                                res.r0_norm, res.solve_seconds, res.cheap_final_residual, res.cheap_seconds), rc

    def __del__(self):
        try:
            if self._g:
                self.lib.mfs_tria_gmg_destroy(self._g)
                self._g = C.c_void_p()
        except Exception:
            pass
