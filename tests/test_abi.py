"""The C-ABI library (libcmm.so) loads without a GPU and exports every function include/cmm.h
declares; argument errors are reported with CMM_ERR_ARG before any device work."""
import ctypes as C
import os
import re

import pytest

import cmm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "cmm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cmm_[a-z0-9_]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert declared() == sorted(cmm.SYMBOLS)


def test_library_exports_every_symbol():
    lib = cmm.lib()
    missing = [s for s in declared() if not hasattr(lib, s)]
    assert not missing, missing
    assert lib.cmm_version() >= 1


def test_argument_errors_without_device():
    lib = cmm.lib()
    c = cmm.AdvectCoeffs()
    import numpy as np
    t = np.zeros(4)
    # l_order 0 is invalid
    rc = lib.cmm_advect_coeffs_fill(t.ctypes.data_as(cmm.DP), t.ctypes.data_as(cmm.DP), 0, 0, 30, 1, 1e-3, -1,
                                    C.byref(c))
    assert rc == -1 and b"bad argument" in lib.cmm_last_error()
    # a non power-of-two fine grid is rejected before any allocation
    p = cmm.EulerParams()
    p.g_coarse, p.g_fine, p.g_psi, p.g_vort = cmm.grid(64), cmm.grid(100), cmm.grid(64), cmm.grid(128)
    p.dt, p.l_order, p.map_update_order_num = 0.1, 3, 1
    h = C.c_void_p()
    assert lib.cmm_euler_create(C.byref(p), C.byref(h)) == -1
    assert h.value is None
    # the direction is -1 (backward map) or +1 (forward map)
    c.direction = 0
    c.l_order = 3
    g = cmm.grid(32)
    arr = (C.c_void_p * 4)(1, 1, 1, 1)
    assert lib.cmm_map_advect(g, g, C.c_void_p(8), C.c_void_p(16), arr, C.byref(c), C.c_void_p(24),
                              C.c_void_p(32), None, 0, None) == -1
    assert b"direction" in lib.cmm_last_error()
    # workspace sizes: the combined fields C1 (all), C12 (RK3/RK4 variants) and a copy of level 0
    # (all but Euler), none for AB2 and RK3-mod (they read level 1 on its own) or a single level
    for tin, L, fields in ((30, 3, 3), (40, 3, 3), (41, 4, 3), (10, 3, 1), (21, 2, 2), (20, 3, 0), (31, 3, 0),
                           (30, 1, 0)):
        c.time_integration_num, c.l_order = tin, L
        assert lib.cmm_advect_workspace_bytes(g, g, C.byref(c)) == (fields * 4 * 8 * 32 * 32 + 64 * 32 * 32 if fields else 0)


def test_lagrange_coefficients_match_closed_form():
    import numpy as np
    c = cmm.advect_coeffs(np.array([-0.2, -0.1, 0.0, 0.1]), np.full(4, 0.1), 0, 3, 30, 1, 1e-3)
    assert np.allclose(list(c.L1)[:3], [3, -3, 1], atol=1e-13)
    assert np.allclose(list(c.L12)[:3], [1.875, -1.25, 0.375], atol=1e-13)


def test_uniform_particles_deterministic():
    import numpy as np
    a = cmm.uniform_particles(1000, 0)
    b = cmm.uniform_particles(1000, 0)
    assert np.array_equal(a, b)
    assert a.min() >= 0 and a.max() < cmm.TWO_PI
    assert abs(a.mean() - np.pi) < 0.2
    assert not np.array_equal(a, cmm.uniform_particles(1000, 1))
