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
