"""Shared test inputs: the reference's test-supply grids and basis, fixture loaders."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def tiny_grid():
    """GridController__TEST_SUPPLY TINY: 4 points, unit weights (GridController__TEST_SUPPLY.cpp:31-38)."""
    pts = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [-2.0, 1.0, 1.0], [3.0, 0.5, -1.0]])
    return pts, np.ones(4)


def very_small_grid():
    """VERY_SMALL: 5 points (GridController__TEST_SUPPLY.cpp:39-46)."""
    pts = np.array([[-1.0, 0.0, 0.0], [0.0, 0.5, 0.4], [0.0, -0.4, -0.5], [0.8, 0.1, -0.1], [0.1, -0.1, 0.1]])
    return pts, np.array([0.1, 0.2, 0.15, 0.15, 0.4])


def small_mixed_basis():
    """BasisController__TEST_SUPPLY SMALL_MIXED: Cartesian s, p, d shells on three centres
    (BasisController__TEST_SUPPLY.cpp:59-81), contraction coefficients libint2-renormalised."""
    from serenity_amd._lib import lib
    import ctypes as C
    from serenity_amd.xc import BasisController
    shells = [(0, [1.0, 2.0], [1.0, 0.1], [0.0, 0.0, 0.0]),
              (1, [1.0, 2.0], [1.0, 0.1], [1.0, 0.0, 0.0]),
              (2, [0.5, 1.2], [0.8, 0.4], [-1.0, 2.5, 1.0])]
    alpha, coef, l, nprim, center = [], [], [], [], []
    for am, a, c, A in shells:
        a = np.array(a, dtype=np.float64)
        c = np.array(c, dtype=np.float64)
        rc = lib().xc_normalize_contraction(am, len(a), a.ctypes.data_as(C.POINTER(C.c_double)),
                                            c.ctypes.data_as(C.POINTER(C.c_double)))
        assert rc == 0
        alpha += list(a)
        coef += list(c)
        l.append(am)
        nprim.append(len(a))
        center += A
    return BasisController(l, [0, 0, 0], nprim, center, alpha, coef)


def load_json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def fixtures():
    return np.load(os.path.join(GOLDEN, "fixtures.npz"))


def h2_fde_system():
    """Active and environment H2 of TestSystem_H2_6_31Gs_{ACTIVE,ENVIRONMENT}_FDE on the
    active system's grid (SSF, accuracy 4, weight threshold 1e-14)."""
    from serenity_amd import xc
    fx = fixtures()
    out = {}
    for tag in ("act", "env"):
        Z, X = fx["h2fde_%s_Z" % tag], fx["h2fde_%s_xyz_angstrom" % tag] * xc.ANGSTROM_TO_BOHR
        out[tag] = (xc.BasisController.from_file(GOLDEN + "/H-6-31GS", Z, X, spherical=False),
                    fx["h2fde_%s_P" % tag])
        if tag == "act":
            out["grid"] = xc.molecular_grid(Z, X, grid_type="SSF", accuracy=4, weight_threshold=1e-14)
    return out
