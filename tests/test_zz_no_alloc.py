"""xc_functional_device_noalloc (no device allocation, static per-device workspace) against
xc_functional on the same densities: equal outputs and energy, on two calls in a row with the
second one smaller, for an LDA-type and a GGA functional; sizes beyond the workspace are refused
before any launch."""
import ctypes as C

import numpy as np
import pytest

from serenity_amd import _lib, xc
from serenity_amd import workloads as W

pytestmark = pytest.mark.gpu

NOALLOC_MAX_BLOCKS = 131072


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _density(ctx):
    Z, X = W.water_cluster(2, 1, 1)
    b = xc.BasisController.from_file("DEF2-SVP", Z, X)
    g = xc.molecular_grid(Z, X, accuracy=3)
    P = W.synthetic_density(b.overlap(), W.n_electrons(Z) // 2)
    bf = xc.BasisFunctionOnGridController(ctx, b, g, 128, 1e-9, 1)
    rho, grad = xc.DensityOnGridCalculator(bf).calcDensityAndGradientOnGrid(P)
    bf._release()
    return np.ascontiguousarray(rho), np.ascontiguousarray(np.asarray(grad).reshape(3, -1)), g.weights.copy()


def _reference(ctx, fid, rho, grad, w, bs):
    N = len(rho)
    eps, v, g = np.zeros(N), np.zeros(N), np.zeros(3 * N)
    E = C.c_double()
    _lib.check(_lib.lib().xc_functional(ctx.handle, fid, N, bs, rho.ctypes.data_as(_lib.dp),
                                        grad.ctypes.data_as(_lib.dp), w.ctypes.data_as(_lib.dp),
                                        eps.ctypes.data_as(_lib.dp), v.ctypes.data_as(_lib.dp),
                                        g.ctypes.data_as(_lib.dp), C.byref(E)))
    return eps, v, g, E.value


def _noalloc(ctx, fid, rho, grad, w, bs, buf):
    N = len(rho)
    ctx.upload(buf["rho"], rho)
    ctx.upload(buf["grad"], grad.ravel())
    ctx.upload(buf["w"], w)
    # stale values in the outputs must be overwritten
    for k in ("eps", "v", "g"):
        ctx.upload(buf[k], np.full(3 * N if k == "g" else N, 7.0))
    E = C.c_double()
    _lib.check(_lib.lib().xc_functional_device_noalloc(ctx.handle, fid, N, bs, buf["rho"], buf["grad"], buf["w"],
                                                       buf["eps"], buf["v"], buf["g"], C.byref(E)))
    out = [np.zeros(N), np.zeros(N), np.zeros(3 * N)]
    for a, k in zip(out, ("eps", "v", "g")):
        ctx.download(a, buf[k])
    return out[0], out[1], out[2], E.value


@pytest.mark.parametrize("functional", ["PBE", "LDA"])
def test_noalloc_functional_equals_allocating_entry_point(functional):
    ctx = xc.Context(0)
    try:
        rho, grad, w = _density(ctx)
        N = len(rho)
        fid = _lib.FUNCTIONALS[functional]
        buf = {k: ctx.alloc(8 * (3 * N if k in ("grad", "g") else N)) for k in ("rho", "grad", "w", "eps", "v", "g")}
        # two calls in a row, the second on a smaller input (a prefix of other points, other block size)
        for n, bs in ((N, 128), (N // 3 + 17, 96)):
            sl = slice(N - n, N)
            r, gr, ww = rho[sl].copy(), np.ascontiguousarray(grad[:, sl]), w[sl].copy()
            ref = _reference(ctx, fid, r, gr, ww, bs)
            got = _noalloc(ctx, fid, r, gr, ww, bs, buf)
            say("%s noalloc vs xc_functional: %d points, blocks of %d, E %.12f vs %.12f" %
                (functional, n, bs, got[3], ref[3]))
            for a, b in zip(got[:3], ref[:3]):
                assert np.array_equal(a, b)
            assert got[3] == ref[3]
        # beyond the workspace: refused on the host, nothing launched
        big = NOALLOC_MAX_BLOCKS * 128 + 1
        rc = _lib.lib().xc_functional_device_noalloc(ctx.handle, fid, big, 128, buf["rho"], buf["grad"], buf["w"],
                                                     buf["eps"], buf["v"], buf["g"], None)
        assert rc == 1 and b"workspace" in _lib.lib().xc_last_error()
        for p in buf.values():
            ctx.free(p)
    finally:
        ctx.close()
