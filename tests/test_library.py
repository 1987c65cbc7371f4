"""CPU checks of the product boundary: the HIP library loads and exports every
symbol include/aspect_mfs.h declares (no compute calls, no GPU needed), and the
host-side problem definitions agree with the oracle's geometry."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
from aspect_amd import mfs, problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "aspect_mfs.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mfs_\w+)\s*\(", txt)))


def test_header_and_python_binding_agree():
    assert declared_symbols() == sorted(mfs.EXPORTED)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(mfs.LIB_PATH):
        pytest.skip("libaspect_mfs.so not built")
    lib = ctypes.CDLL(mfs.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_library_refuses_bad_box_without_gpu_work():
    if not os.path.exists(mfs.LIB_PATH):
        pytest.skip("libaspect_mfs.so not built")
    lib = mfs.load_library()
    box = mfs.make_box(dim=4)
    h = ctypes.c_void_p()
    assert lib.mfs_create(ctypes.byref(box), 0, ctypes.byref(h)) == mfs.MFS_ERROR
    assert b"dim" in lib.mfs_last_error()


@pytest.mark.parametrize("dim,deg", [(2, 2), (3, 2), (3, 3)])
def test_qp_coordinates_match_oracle(dim, deg):
    p = O.make_problem(dim, deg, (2, 1, 1), 3, (2.0, 1.0, 0.5), (1,) * 6)
    n = [2 << 2, 1 << 2, 1 << 2][:dim]
    ours = P.qp_coordinates(dim, deg, n, (2.0, 1.0, 0.5))
    np.testing.assert_allclose(ours, O.qp_points(p, 2), atol=1e-15)


def test_pressure_scaling_rule():
    # geometric mean of midpoint viscosities over the box length scale 0.01 * x extent
    eta = np.array([1.0, 100.0])
    assert abs(P.pressure_scaling(eta, 0.01) - 10.0 / 0.01) < 1e-9


# pointers each (op, mode) kernel dereferences: src_u, src_p, eta, rhs_u, rhs_p, dst_u, dst_p
_NEEDS = {
    "stokes": ("src_u", "src_p", "eta", "dst_u", "dst_p"),
    "ablock": ("src_u", "eta", "dst_u"),
    "bt": ("src_p", "dst_u"),
    "mass": ("src_p", "eta", "dst_p"),
    "adiag": ("eta", "dst_u"),
    "mdiag": ("eta", "dst_p"),
    "rhs": ("eta", "dst_u"),
    "laplace": ("src_u", "dst_u"),
    "ldiag": ("dst_u",),
}
_RESID = {"stokes": ("rhs_u", "rhs_p"), "ablock": ("rhs_u",), "bt": ("rhs_u",), "mass": ("rhs_p",),
          "laplace": ("rhs_u",)}
_ALL = ("src_u", "src_p", "eta", "rhs_u", "rhs_p", "dst_u", "dst_p")


@pytest.mark.parametrize("op", sorted(_NEEDS))
@pytest.mark.parametrize("mode", ["set", "resid", "add"])
def test_operator_argument_validation(op, mode):
    # the host-side check every operator launch runs (no device involved)
    if not os.path.exists(mfs.LIB_PATH):
        pytest.skip("libaspect_mfs.so not built")
    need = set(_NEEDS[op]) | (set(_RESID.get(op, ())) if mode == "resid" else set())
    # exactly the needed pointers: accepted
    mfs.validate_operator_arguments(op, mode, **{k: 4096 for k in need})
    for k in need:  # each needed pointer missing: rejected, naming it
        args = {j: 4096 for j in need if j != k}
        with pytest.raises(mfs.SolverError, match=f"null {k} pointer"):
            mfs.validate_operator_arguments(op, mode, **args)
    for k in set(_ALL) - need:  # pointers the kernel does not use may be null
        mfs.validate_operator_arguments(op, mode, **{j: 4096 for j in need})


def test_null_handle_is_an_error_not_a_launch():
    if not os.path.exists(mfs.LIB_PATH):
        pytest.skip("libaspect_mfs.so not built")
    lib = mfs.load_library()
    buf = (ctypes.c_double * 4)()
    out = (ctypes.c_longlong * 4)()
    calls = [
        lambda: lib.mfs_stokes_vmult_device(None, buf, buf),
        lambda: lib.mfs_ablock_vmult_device(None, 0, buf, buf),
        lambda: lib.mfs_stokes_vmult(None, buf, buf),
        lambda: lib.mfs_sizes(None, 0, out),
        lambda: lib.mfs_set_viscosity(None, buf, 0),
        lambda: lib.mfs_build_preconditioner(None),
        lambda: lib.mfs_vcycle(None, 0, buf, buf),
        lambda: lib.mfs_solve_device(None, buf, None, buf, None, None),
        lambda: lib.mfs_synchronize(None),
        lambda: lib.mfs_bt_vmult(None, buf, buf),
        lambda: lib.mfs_prolongate(None, 1, 0, buf, buf),
    ]
    for c in calls:
        assert c() == mfs.MFS_ERROR
        assert b"null" in lib.mfs_last_error()


def test_vector_length_check():
    with pytest.raises(ValueError):
        mfs._vec(np.zeros(5), 6, "rhs")
    with pytest.raises(ValueError):
        mfs._vec(np.zeros((2, 3)), 6, "rhs")
    assert mfs._vec([1, 2, 3], 3, "rhs").dtype == np.float64


@pytest.mark.parametrize("dim,deg,coarse,nl", [(2, 2, (2, 3), 3), (3, 2, (2, 1, 1), 2), (2, 3, (1, 2), 3),
                                               (3, 3, (1, 1, 1), 2)])
def test_box_dof_index_recovers_permutation(dim, deg, coarse, nl):
    # deal.II-style DoF numbering (a random order of support points, the FE_Q
    # Gauss-Lobatto points of every cell) mapped back to the library's layout
    import oracle as O
    from aspect_amd import make_box
    from aspect_amd.mfs import SolverError, box_dof_index
    ext = (1.5, 0.75, 2.0)[:dim]
    box = make_box(dim, deg, coarse, nl, ext, (0,) * 6)
    p = O.make_problem(dim, deg, coarse, nl, ext, (0,) * 6)
    s = O.sizes(p, nl - 1)
    vp = O.velocity_support_points(p, nl - 1)
    pp = O.pressure_support_points(p, nl - 1)
    pts = np.concatenate([np.tile(vp, (dim, 1)), pp])
    comp = np.concatenate([np.repeat(np.arange(dim), s["vnodes"]), np.full(s["np"], dim)])
    perm = np.random.default_rng(dim * deg).permutation(len(comp))
    idx = box_dof_index(box, pts[perm], comp[perm])
    np.testing.assert_array_equal(idx, perm)
    bad = pts[perm].copy()
    bad[5, 0] += 0.1 * ext[0] / (coarse[0] << (nl - 1))
    with pytest.raises(SolverError, match="not a node"):
        box_dof_index(box, bad, comp[perm])
    dup = perm.copy()
    dup[1] = dup[0]
    with pytest.raises(SolverError, match="two DoFs"):
        box_dof_index(box, pts[dup], comp[dup])


def test_tria_refuses_bad_meshes_without_gpu_work():
    """mfs_tria_create checks the mesh and the constraint lines on the host
    before any device call: node ids in range, degree 2, finite geometry,
    every DoF constrained once, closed lines (no constrained master)."""
    if not os.path.exists(mfs.LIB_PATH):
        pytest.skip("libaspect_mfs.so not built")
    from aspect_amd import tria as T
    lib = mfs.load_library()
    t = T.Tria(T.BoxCoarse(2, (2, 1), (2.0, 1.0)), 0)
    t.refine(np.array([True, False]))
    mesh = t.build({0: "zero"})
    keep = []

    def arr(a, ty):
        a = np.ascontiguousarray(a, dtype=ty)
        keep.append(a)
        return a.ctypes.data

    def desc(cell_u=None, degree=2, geometry=None):
        d = mfs.TriaDesc()
        d.dim, d.degree = 2, degree
        d.n_cells, d.n_nodes_u, d.n_nodes_p = mesh.n_cells, mesh.n_nodes_u, mesh.n_nodes_p
        d.cell_nodes_u = arr(mesh.cell_u if cell_u is None else cell_u, np.int32)
        d.cell_nodes_p = arr(mesh.cell_p, np.int32)
        d.geometry = arr(mesh.geometry if geometry is None else geometry, np.float64)
        d.pressure_scaling = 1.0
        return d

    def con(lines):
        dofs, offs, masters, weights = lines
        k = mfs.Constraints()
        k.n_lines = len(dofs)
        k.dofs, k.offsets = arr(dofs, np.int64), arr(offs, np.int64)
        k.masters, k.weights = arr(masters, np.int64), arr(weights, np.float64)
        return k

    def create(d, cu):
        h = ctypes.c_void_p()
        rc = lib.mfs_tria_create(ctypes.byref(d), ctypes.byref(cu), None, ctypes.byref(h))
        return rc, lib.mfs_last_error().decode()

    good = con(mesh.con_u)
    bad_u = mesh.cell_u.copy()
    bad_u[1, 3] = mesh.n_nodes_u
    rc, msg = create(desc(cell_u=bad_u), good)
    assert rc == mfs.MFS_ERROR and "node id out of range" in msg
    rc, msg = create(desc(degree=3), good)
    assert rc == mfs.MFS_ERROR and "degree 2" in msg
    geo = mesh.geometry.copy()
    geo[0, 0, 0] = np.nan
    rc, msg = create(desc(geometry=geo), good)
    assert rc == mfs.MFS_ERROR and "geometry not finite" in msg
    dofs, offs, masters, weights = mesh.con_u
    # a hanging line whose master is a Dirichlet DoF: not closed
    hang = int(np.nonzero(np.diff(offs) > 0)[0][0])
    open_masters = masters.copy()
    open_masters[offs[hang]] = dofs[np.nonzero(np.diff(offs) == 0)[0][0]]
    rc, msg = create(desc(), con((dofs, offs, open_masters, weights)))
    assert rc == mfs.MFS_ERROR and "close the constraints" in msg
    twice = np.concatenate([dofs, dofs[:1]])
    rc, msg = create(desc(), con((twice, np.concatenate([offs, offs[-1:]]), masters, weights)))
    assert rc == mfs.MFS_ERROR and "constrained twice" in msg
