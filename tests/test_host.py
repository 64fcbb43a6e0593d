"""Host-side parts of the library (grid, basis, overlap, workloads, C ABI surface).
CPU only: nothing here launches a kernel."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import common
from serenity_amd import _lib, xc
from serenity_amd import workloads as W

# Lebedev-Laikov rule -> algebraic degree integrated exactly.
DEGREE = {6: 3, 14: 5, 26: 7, 38: 9, 50: 11, 74: 13, 86: 15, 110: 17, 146: 19, 170: 21, 194: 23, 230: 25, 266: 27,
          302: 29, 350: 31, 434: 35, 590: 41, 770: 47, 974: 53, 1202: 59, 1454: 65, 1730: 71, 2030: 77, 2354: 83,
          2702: 89, 3074: 95, 3470: 101, 3890: 107, 4334: 113, 4802: 119, 5294: 125, 5810: 131}


def dfact(k):
    r = 1.0
    while k > 1:
        r *= k
        k -= 2
    return r


@pytest.mark.parametrize("npts", sorted(DEGREE))
def test_lebedev_rule_exactness(npts):
    """Unit points, weights summing to 1, and exact sphere averages of x^a y^b z^c up to the
    rule's degree: <x^a y^b z^c> = (a-1)!!(b-1)!!(c-1)!!/(a+b+c+1)!! for even a,b,c."""
    x, y, z, w = xc.lebedev(npts)
    assert np.allclose(x * x + y * y + z * z, 1.0, atol=1e-14)
    assert abs(w.sum() - 1.0) < 1e-13
    deg = min(DEGREE[npts], 24)
    for a in range(0, deg + 1, 2):
        for b in range(0, deg + 1 - a, 2):
            c = deg - a - b if (deg - a - b) % 2 == 0 else deg - a - b - 1
            for cc in (c, max(c - 2, 0)):
                val = np.sum(w * x ** a * y ** b * z ** cc)
                exact = dfact(a - 1) * dfact(b - 1) * dfact(cc - 1) / dfact(a + b + cc + 1)
                assert abs(val - exact) < 1e-13, (npts, a, b, cc)
    # odd moments vanish
    assert abs(np.sum(w * x * y * z)) < 1e-14


def test_grid_factory_volume_windows():
    """GridFactory_test.cpp: volume of the radius-7 sphere around C60 on Becke/Voronoi/SSF grids."""
    d = common.load_json("grid_factory.json")
    X = np.array(d["c60_bohr"])
    assert len(d["cases"]) == 8
    for c in d["cases"]:
        g = xc.molecular_grid(np.full(60, 6), X, grid_type=c["grid_type"], radial_type=c["radial"],
                              accuracy=c["accuracy"], smoothing=c["smoothing"], weight_threshold=c["weight_threshold"],
                              sort=False)
        r = np.linalg.norm(g.points, axis=1)
        s = g.weights[r <= d["sphere_radius"]].sum() / (d["sphere_radius"] ** 3 * 4.0 / 3.0 * np.pi)
        assert abs(s - 1.0) < c["near"], c
        if c["greater"] is not None:
            assert s > 1.0 + c["greater"], c
        if c["less"] is not None:
            assert s < 1.0 - c["less"], c


def test_hilbert_sort_2x2x2():
    """HilbertRTreeSorting_test.cpp:32-54: exact order of the eight cube corners."""
    h = common.load_json("hilbert.json")
    # drive the sort through a single-atom-free path: build a grid whose points are the corners
    pts = np.array(h["coords"])
    w = np.array(h["weights"])
    import ctypes
    lib = _lib.lib()
    # the molecular grid API sorts; reproduce the sort alone through a tiny helper grid
    from serenity_amd.xc import _d  # noqa
    out_p = pts.copy()
    out_w = w.copy()
    sort_points(out_p, out_w)
    assert np.array_equal(out_w, np.array(h["wref"]))
    assert np.array_equal(out_p, np.array(h["pref"]))


def sort_points(p, w):
    lib = _lib.lib()
    f = lib.xc_hilbert_sort
    f.restype = C.c_int
    f.argtypes = [C.c_int64, _lib.dp, _lib.dp]
    pc = np.ascontiguousarray(p)
    wc = np.ascontiguousarray(w)
    _lib.check(f(len(w), pc.ctypes.data_as(_lib.dp), wc.ctypes.data_as(_lib.dp)))
    p[...] = pc
    w[...] = wc


def test_radial_quadrature_single_atom():
    """A free atom's grid integrates a Gaussian exactly enough: int exp(-a r^2) = (pi/a)^(3/2)."""
    for Z in (1, 6, 8):
        g = xc.molecular_grid([Z], np.zeros((1, 3)), accuracy=4, weight_threshold=0.0)
        r2 = (g.points ** 2).sum(1)
        for a in (0.3, 1.0, 5.0):
            assert abs(np.sum(g.weights * np.exp(-a * r2)) / (np.pi / a) ** 1.5 - 1.0) < 1e-6


def test_ssf_weights_partition_unity():
    """SSF cell functions partition space: the molecular grid integrates a smooth density the
    same as the sum of its free-atom integrals."""
    Z, X = W.water_cluster(1, 1, 2)
    g = xc.molecular_grid(Z, X, accuracy=4)
    f = lambda p: sum(np.exp(-0.8 * ((p - a) ** 2).sum(1)) for a in X)
    assert abs(np.sum(g.weights * f(g.points)) / (len(X) * (np.pi / 0.8) ** 1.5) - 1.0) < 1e-6


@pytest.mark.parametrize("cfg,nao", [(1, 192), (2, 490), (3, 2752)])
def test_config_basis_sizes(cfg, nao):
    c = W.CONFIGS[cfg]
    Z, X = W.geometry(c["system"])
    b = xc.BasisController.from_file(c["basis"], Z, X)
    assert b.nao == nao


def test_overlap_diagonal_and_grid_quadrature():
    """Analytic overlap: unit diagonal (renormalised contractions, solid harmonics and Cartesian
    factors), and agreement with a grid quadrature of the AO products."""
    from oracle import oracle as O
    Z, X = W.water_cluster(1, 1, 1)
    for spherical in (True, False):
        b = xc.BasisController.from_file("DEF2-TZVP", Z, X, spherical=spherical)
        S = b.overlap()
        assert np.allclose(np.diag(S), 1.0, atol=1e-12)
        assert np.allclose(S, S.T, atol=1e-15)
        g = xc.molecular_grid(Z, X, accuracy=6, weight_threshold=0.0)
        ao = O.ao_block(b, g.points, thr=0.0, deriv=0)["values"]
        Sg = ao.T @ (ao * g.weights[:, None])
        assert np.abs(Sg - S).max() < 2e-5


def test_workload_geometries():
    Z, X = W.water_cluster(2, 2, 2)
    Xa = X / xc.ANGSTROM_TO_BOHR
    for m in range(8):
        o, h1, h2 = Xa[3 * m:3 * m + 3]
        assert abs(np.linalg.norm(h1 - o) - 0.9572) < 1e-12
        ang = np.degrees(np.arccos(np.dot(h1 - o, h2 - o) / (np.linalg.norm(h1 - o) * np.linalg.norm(h2 - o))))
        assert abs(ang - 104.52) < 1e-9
    Z2, X2 = W.water_cluster(2, 2, 2)
    assert np.array_equal(X, X2)
    Z, X = W.alkane(20)
    assert (Z == 6).sum() == 20 and (Z == 1).sum() == 42
    Xa = X / xc.ANGSTROM_TO_BOHR
    cc = np.linalg.norm(Xa[1:20] - Xa[:19], axis=1)
    assert np.all(np.abs(cc - 1.53) < 0.07)
    Z, X, sub = W.benzene_water(24)
    assert len(Z) == 12 + 72 and sub.max() == 24
    d = np.linalg.norm(X[:, None] - X[None], axis=2) / xc.ANGSTROM_TO_BOHR
    np.fill_diagonal(d, 9.9)
    assert d.min() > 0.9


def test_synthetic_density_is_idempotent():
    Z, X = W.alkane(3)
    b = xc.BasisController.from_file("DEF2-SVP", Z, X)
    S = b.overlap()
    P = W.synthetic_density(S, W.n_electrons(Z) // 2)
    assert abs(np.sum(P * S) - W.n_electrons(Z)) < 1e-10
    assert np.allclose(P @ S @ P, 2 * P, atol=1e-9)


def test_library_exports_every_declared_symbol():
    names = _lib.exported_symbols_declared_in_header()
    assert len(names) >= 30
    lib = _lib.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_host_error_paths():
    """Errors come back as status codes with a message (SerenityError counterpart)."""
    with pytest.raises(xc.XCError, match="duplicated atoms"):
        xc.molecular_grid([8, 8], np.zeros((2, 3)))
    # closer than 1e-5 bohr is a duplicate (Geometry.cpp:540-549), regardless of the element
    with pytest.raises(xc.XCError, match="duplicated atoms"):
        xc.molecular_grid([8, 1], np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 5e-6]]), accuracy=1)
    assert xc.molecular_grid([8, 1], np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.8]]), accuracy=1).getNGridPoints() > 0
    with pytest.raises(xc.XCError, match="no entry for element"):
        xc.BasisController.from_file("DEF2-SVP", [19], np.zeros((1, 3)))  # the tables hold H..Ar
    with pytest.raises(xc.XCError, match="accuracy"):
        xc.molecular_grid([8], np.zeros((1, 3)), accuracy=9)
