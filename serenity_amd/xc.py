"""Python host side of the XC Fock build, shaped like Serenity's grid classes.

Class and method names follow the reference so that call sites read alike:

=============================================  ==================================================
this module                                    Serenity (src/...)
=============================================  ==================================================
``GridController``                             grid/GridController.h (points 3 x N, weights N)
``BasisController``                            basis/BasisController.h (shell list, AO order)
``BasisFunctionOnGridController``              data/grid/BasisFunctionOnGridController.h:105-131
``DensityOnGridCalculator``                    data/grid/DensityOnGridCalculator.h:45-134
``FunctionalLibrary``                          dft/functionals/FunctionalLibrary.h:51-64
``ScalarOperatorToMatrixAdder``                data/grid/ScalarOperatorToMatrixAdder.h:48-196
``FuncPotential``                              potentials/FuncPotential.h:48-145
=============================================  ==================================================

Every call goes through the C ABI of ``libserenity_xc.so``; errors raise
``XCError`` (the counterpart of ``SerenityError``).  Matrices are numpy
arrays, addressed in column-major (Fortran) order like Eigen's.
"""
import ctypes as C
import os
import weakref

import numpy as np

from . import _lib
from ._lib import FUNCTIONALS, GGA, PRECISIONS, GRID_TYPES, RADIAL_TYPES, XC_BUSY, XCError, check

BASIS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "basis")
ANGSTROM_TO_BOHR = 1.0 / 0.529177210544

__all__ = [
    "Context", "BasisController", "GridController", "BasisFunctionOnGridController",
    "DensityOnGridCalculator", "FunctionalLibrary", "FunctionalData", "ScalarOperatorToMatrixAdder",
    "FuncPotential", "molecular_grid", "lebedev", "overlap_matrix", "XCError", "ANGSTROM_TO_BOHR",
    "DensityFittedCoulomb", "fock_build_packed",
]


def _d(a):
    return a.ctypes.data_as(_lib.dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_lib.ip) if a is not None else None


def _f64(a):
    return np.require(np.asarray(a, dtype=np.float64), requirements=["C", "A"])


class Context:
    """One HIP device and one stream (all kernels of this context go there)."""

    def __init__(self, device=0):
        L = _lib.lib()
        h = C.c_void_p()
        check(L.xc_ctx_create(int(device), C.byref(h)))
        self.handle = h
        self.device = device
        self._children = weakref.WeakSet()   # device handles created on this context

    def _register(self, obj):
        self._children.add(obj)

    @property
    def stream(self):
        return _lib.lib().xc_ctx_stream(self.handle)

    def synchronize(self):
        check(_lib.lib().xc_ctx_synchronize(self.handle))

    def memory_in_use(self):
        v = C.c_int64()
        check(_lib.lib().xc_ctx_memory(self.handle, C.byref(v)))
        return v.value

    def alloc(self, nbytes):
        p = C.c_void_p()
        check(_lib.lib().xc_device_alloc(self.handle, C.c_size_t(nbytes), C.byref(p)))
        return p

    def free(self, ptr):
        check(_lib.lib().xc_device_free(self.handle, ptr))

    def upload(self, dptr, array):
        a = np.ascontiguousarray(array)
        check(_lib.lib().xc_memcpy_h2d(self.handle, dptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes)))

    def download(self, array, dptr):
        check(_lib.lib().xc_memcpy_d2h(self.handle, array.ctypes.data_as(C.c_void_p), dptr,
                                       C.c_size_t(array.nbytes)))

    def copy(self, dst, src, nbytes):
        """Device-to-device copy by the library's streaming copy kernel."""
        check(_lib.lib().xc_device_copy(self.handle, dst, src, C.c_size_t(nbytes)))

    def set_deterministic(self, enable=True):
        """Sum every Fock element's terms in exact integer parts, so that F is the same bit for bit
        on every run; only before any basis, grid or controller of the context exists."""
        check(_lib.lib().xc_ctx_set_deterministic(self.handle, int(bool(enable))))

    def profile(self, enable=True):
        """Record HIP events between the build stages of every following build."""
        check(_lib.lib().xc_profile(self.handle, int(bool(enable))))

    def profile_collect(self):
        """dict stage -> summed milliseconds since the last collect, and the build count."""
        ms = np.zeros(len(_lib.STAGES))
        nb = C.c_int()
        check(_lib.lib().xc_profile_collect(self.handle, _d(ms), C.byref(nb)))
        return dict(zip(_lib.STAGES, ms.tolist())), nb.value

    def close(self):
        """Release every handle created on this context (servers, then controllers, then the rest),
        then the context."""
        if not self.handle:
            return
        kids = list(self._children)
        order = lambda o: (0 if isinstance(o, Server) else
                           1 if isinstance(o, (BasisFunctionOnGridController, DensityFittedCoulomb)) else 2)
        for k in sorted(kids, key=order):
            k._release()
        check(_lib.lib().xc_ctx_destroy(self.handle))
        self.handle = None


class BasisController:
    """Contracted shells in Serenity's AO order (atoms in input order, shells in
    file order).  ``coef`` holds libint2-renormalised contraction coefficients."""

    def __init__(self, l, pure, nprim, center, alpha, coef, atom=None):
        self.l = np.ascontiguousarray(l, dtype=np.int32)
        self.pure = np.ascontiguousarray(pure, dtype=np.int32)
        self.nprim = np.ascontiguousarray(nprim, dtype=np.int32)
        self.center = _f64(np.asarray(center, dtype=np.float64).reshape(-1, 3))
        self.alpha = _f64(alpha)
        self.coef = _f64(coef)
        self.atom = None if atom is None else np.ascontiguousarray(atom, dtype=np.int32)
        nf = np.where(self.pure != 0, 2 * self.l + 1, (self.l + 1) * (self.l + 2) // 2)
        self.nfunc = nf.astype(np.int32)
        self.first = np.concatenate([[0], np.cumsum(nf)[:-1]]).astype(np.int32)
        self.nao = int(nf.sum())
        self.nshell = len(self.l)

    @classmethod
    def from_file(cls, basis_name, Z, coords_bohr, spherical=True):
        """Load a Turbomole-format basis ("DEF2-SVP", "DEF2-TZVP" or a path)."""
        path = basis_name if os.path.exists(basis_name) else os.path.join(BASIS_DIR, basis_name.upper())
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = _f64(np.asarray(coords_bohr, dtype=np.float64).reshape(-1, 3))
        L = _lib.lib()
        ns = C.c_int()
        pl, pn, pa = _lib.ip(), _lib.ip(), _lib.ip()
        pc, pal, pco = _lib.dp(), _lib.dp(), _lib.dp()
        check(L.xc_load_basis(path.encode(), len(Z), _i(Z), _d(xyz), int(bool(spherical)), C.byref(ns),
                              C.byref(pl), C.byref(pn), C.byref(pa), C.byref(pc), C.byref(pal), C.byref(pco)))
        try:
            n = ns.value
            l = np.ctypeslib.as_array(pl, shape=(n,)).copy()
            nprim = np.ctypeslib.as_array(pn, shape=(n,)).copy()
            atom = np.ctypeslib.as_array(pa, shape=(n,)).copy()
            center = np.ctypeslib.as_array(pc, shape=(n * 3,)).copy()
            npr = int(nprim.sum())
            alpha = np.ctypeslib.as_array(pal, shape=(npr,)).copy()
            coef = np.ctypeslib.as_array(pco, shape=(npr,)).copy()
        finally:
            for p in (pl, pn, pa, pc, pal, pco):
                L.xc_free_host(C.cast(p, C.c_void_p))
        return cls(l, np.full(n, int(bool(spherical))), nprim, center, alpha, coef, atom)

    def as_args(self):
        """(nshell, l, pure, nprim, center, alpha, coef) as ctypes pointers."""
        return (self.nshell, _i(self.l), _i(self.pure), _i(self.nprim), _d(self.center), _d(self.alpha),
                _d(self.coef))

    def overlap(self):
        return overlap_matrix(self)

    def getAtomIndicesOfBasis(self):
        """Atom of every AO (BasisController::getAtomIndicesOfBasis); needs shells loaded per atom."""
        if self.atom is None:
            raise XCError(4, "BasisController: shells carry no atom indices")
        return np.repeat(self.atom, self.nfunc).astype(np.int32)


def overlap_matrix(basis):
    S = np.zeros((basis.nao, basis.nao), order="F")
    check(_lib.lib().xc_overlap(*basis.as_args(), _d(S.reshape(-1, order="F"))))
    return S


class GridController:
    """Grid points (N x 3 here; 3 x N column-major in memory) and weights."""

    def __init__(self, points, weights):
        self.points = _f64(np.asarray(points, dtype=np.float64).reshape(-1, 3))
        self.weights = _f64(weights)
        if len(self.weights) != len(self.points):
            raise XCError(1, "GridController: points and weights differ in length")

    def getNGridPoints(self):
        return len(self.weights)

    def _device(self, ctx):
        """The grid's resident copy on ``ctx``, shared by every controller built on this grid."""
        cache = self.__dict__.setdefault("_device_grids", {})
        dg = cache.get(id(ctx))
        if dg is None or dg.handle is None or dg.ctx_ref() is not ctx:
            dg = _DeviceGrid(ctx, self)
            cache[id(ctx)] = dg
        return dg


def molecular_grid(Z, coords_bohr, grid_type="SSF", radial_type="AHLRICHS", accuracy=4, smoothing=3,
                   weight_threshold=1e-14, sort=True, return_before_cut=False, ctx=None):
    """Atom-centred integration grid (GridFactory::produce + Hilbert sort).  With ``ctx`` the
    fuzzy-cell partition runs on that context's GPU (same grid as the host build)."""
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    xyz = _f64(np.asarray(coords_bohr, dtype=np.float64).reshape(-1, 3))
    L = _lib.lib()
    n = C.c_int64()
    nb = C.c_int64()
    pp, pw = _lib.dp(), _lib.dp()
    args = (len(Z), _i(Z), _d(xyz), GRID_TYPES[grid_type], RADIAL_TYPES[radial_type], int(accuracy),
            int(smoothing), float(weight_threshold), int(bool(sort)), C.byref(n), C.byref(pp), C.byref(pw),
            C.byref(nb))
    if ctx is None:
        check(L.xc_molecular_grid(*args))
    else:
        check(L.xc_molecular_grid_device(ctx.handle, *args))
    try:
        N = n.value
        pts = np.ctypeslib.as_array(pp, shape=(N * 3,)).copy().reshape(N, 3)
        w = np.ctypeslib.as_array(pw, shape=(N,)).copy()
    finally:
        L.xc_free_host(C.cast(pp, C.c_void_p))
        L.xc_free_host(C.cast(pw, C.c_void_p))
    g = GridController(pts, w)
    if return_before_cut:
        return g, nb.value
    return g


def lebedev(npoints):
    x, y, z, w = (np.zeros(npoints) for _ in range(4))
    check(_lib.lib().xc_lebedev(int(npoints), _d(x), _d(y), _d(z), _d(w)))
    return x, y, z, w


class _Handle:
    """A device handle owned by a Context; released explicitly or when collected."""
    _destroy = None

    def _release(self):
        h, self.handle = getattr(self, "handle", None), None
        if h:
            getattr(_lib.lib(), self._destroy)(h)

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass


class _DeviceBasis(_Handle):
    _destroy = "xc_basis_destroy"

    def __init__(self, ctx, basis):
        h = C.c_void_p()
        check(_lib.lib().xc_basis_create(ctx.handle, basis.nshell, _i(basis.l), _i(basis.pure), _i(basis.nprim),
                                         _d(basis.center), _d(basis.alpha), _d(basis.coef), C.byref(h)))
        self.handle = h
        ctx._register(self)


class DensityFittedCoulomb(_Handle):
    """Density-fitted Coulomb matrix on the GPU (the J engine behind Serenity's densFitJ = RI):
    ``coulomb(P)`` returns J[P] = B^T M^-1 B p, with the 3-centre contractions on the device
    (no 3-centre integral stored) and the metric M = (P|Q) factorised once on the host."""
    _destroy = "xc_coulomb_destroy"

    def __init__(self, ctx, basis, aux, threshold=1e-14):
        import scipy.linalg
        self._db = _DeviceBasis(ctx, basis)
        self._da = _DeviceBasis(ctx, aux)
        h = C.c_void_p()
        check(_lib.lib().xc_coulomb_create(self._db.handle, self._da.handle, C.c_double(threshold), C.byref(h)))
        self.handle = h
        ctx._register(self)
        self.n, self.naux = basis.nao, aux.nao
        M = np.zeros((aux.nao, aux.nao))
        check(_lib.lib().xc_ri_metric(*aux.as_args(), _d(M)))
        self.metric = M
        self._cho = scipy.linalg.cho_factor(M)
        np_, nh = C.c_int64(), C.c_int64()
        check(_lib.lib().xc_coulomb_info(h, C.byref(np_), C.byref(nh)))
        self.npairs, self.hermite_terms = np_.value, nh.value

    def rhs(self, P):
        """b_Q = sum P_mu,nu (mu nu|Q)."""
        b = np.zeros(self.naux)
        Pf = np.asfortranarray(P, dtype=np.float64)
        check(_lib.lib().xc_coulomb_rhs(self.handle, Pf.ctypes.data_as(_lib.dp), _d(b)))
        return b

    def matrix(self, c):
        """J_mu,nu = sum_Q c_Q (mu nu|Q)."""
        J = np.zeros((self.n, self.n), order="F")
        check(_lib.lib().xc_coulomb_matrix(self.handle, _d(np.ascontiguousarray(c, dtype=np.float64)),
                                           J.ctypes.data_as(_lib.dp)))
        return np.array(J)

    def coulomb(self, P):
        import scipy.linalg
        c = scipy.linalg.cho_solve(self._cho, self.rhs(P))
        return self.matrix(c)


class _DeviceGrid(_Handle):
    _destroy = "xc_grid_destroy"

    def __init__(self, ctx, grid):
        h = C.c_void_p()
        check(_lib.lib().xc_grid_create(ctx.handle, grid.getNGridPoints(), _d(grid.points), _d(grid.weights),
                                        C.byref(h)))
        self.handle = h
        self.ctx_ref = weakref.ref(ctx)
        ctx._register(self)


class BasisFunctionOnGridController(_Handle):
    """Screened AO values (and gradients) of every grid block, resident in HBM.

    ``BasisFunctionOnGridController(basis, grid, maxBlockSize, radialThreshold,
    highestDerivative)`` as in the reference; ``highestDerivative`` is 0, 1 or 2
    (2 adds the second derivatives the GGA nuclear gradient needs).  ``keep_fp32`` keeps an FP32 copy of the tiles for the mixed-precision
    contractions; ``keep_fp32="only"`` keeps FP32 tiles alone (FP32 builds only,
    a third of the footprint)."""

    def __init__(self, ctx, basis, grid, max_block_size=128, radial_threshold=1e-9, highest_derivative=1,
                 keep_fp32=False):
        self.ctx = ctx
        self.basis = basis
        self.grid = grid
        self._db = _DeviceBasis(ctx, basis)
        self._dg = grid._device(ctx)
        h = C.c_void_p()
        check(_lib.lib().xc_bfogc_create(self._db.handle, self._dg.handle, int(max_block_size),
                                         float(radial_threshold), int(highest_derivative),
                                         2 if keep_fp32 == "only" else int(bool(keep_fp32)),
                                         C.byref(h)))
        self.handle = h
        ctx._register(self)
        self.highest_derivative = highest_derivative
        self.keep_fp32 = keep_fp32
        nb, npts, nao, mbs = C.c_int(), C.c_int64(), C.c_int(), C.c_int()
        check(_lib.lib().xc_bfogc_info(h, C.byref(nb), C.byref(npts), C.byref(nao), C.byref(mbs)))
        self._nblocks, self._npts, self._nao, self._mbs = nb.value, npts.value, nao.value, mbs.value

    def getNBlocks(self):
        return self._nblocks

    def getNGridPoints(self):
        return self._npts

    def getNBasisFunctions(self):
        return self._nao

    def getMaxBlockSize(self):
        return self._mbs

    def getFirstIndexOfBlock(self, block):
        return block * self._mbs

    def getHighestDerivative(self):
        return self.highest_derivative

    def stats(self):
        """(sum_b n_b, sum_b B_b n_b^2, resident tile bytes)."""
        a, b, c = C.c_int64(), C.c_double(), C.c_int64()
        check(_lib.lib().xc_bfogc_stats(self.handle, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def update_density(self, dP, precision="FP64", with_gradient=True):
        """Compute and keep resident the grid density of the device density matrix ``dP``
        (the cache read by NAddFuncPotential.build_cached)."""
        check(_lib.lib().xc_bfogc_update_density(self.handle, PRECISIONS[precision], dP, int(bool(with_gradient))))

    def getBlockOnGridData(self, block):
        """dict(functionValues B x n, derivativeValues (x, y, z) or None, negligible n)."""
        n = self._nao
        B = min(self._mbs, self._npts - block * self._mbs)
        vals = np.zeros(B * n)
        grads = [np.zeros(B * n) for _ in range(3)] if self.highest_derivative >= 1 else None
        neg = np.zeros(n, dtype=np.int32)
        bs = C.c_int()
        check(_lib.lib().xc_bfogc_block_data(self.handle, int(block), _d(vals),
                                             *([_d(g) for g in grads] if grads else [None, None, None]),
                                             _i(neg), C.byref(bs)))
        out = {"functionValues": vals.reshape(n, B).T, "negligible": neg}
        out["derivativeValues"] = tuple(g.reshape(n, B).T for g in grads) if grads else None
        if self.highest_derivative >= 2:
            h = np.zeros(6 * n * B)
            check(_lib.lib().xc_bfogc_block_hessian(self.handle, int(block), _d(h)))
            out["secondDerivativeValues"] = tuple(h.reshape(6, n, B)[c].T for c in range(6))
        return out

    _destroy = "xc_bfogc_destroy"


class DensityOnGridCalculator:
    """rho and grad rho of a (restricted) density matrix on the grid."""

    def __init__(self, bfogc, block_average_threshold=1e-11):
        self.bfogc = bfogc
        self.blockAverageThreshold = block_average_threshold
        self._nonneg = None

    def _run(self, P, want_grad):
        n = self.bfogc.getNBasisFunctions()
        N = self.bfogc.getNGridPoints()
        P = np.asarray(P, dtype=np.float64)
        if P.shape != (n, n):
            raise XCError(1, "DensityOnGridCalculator: density matrix is %s, basis has %d functions" % (P.shape, n))
        Pf = np.asfortranarray(P)
        rho = np.zeros(N)
        grad = np.zeros(3 * N) if want_grad else None
        nonneg = np.zeros(self.bfogc.getNBlocks(), dtype=np.int32)
        check(_lib.lib().xc_density(self.bfogc.handle, Pf.ctypes.data_as(_lib.dp), _d(rho), _d(grad), _i(nonneg)))
        self._nonneg = nonneg
        return rho, (grad.reshape(3, N) if want_grad else None)

    def calcDensityOnGrid(self, P):
        return self._run(P, False)[0]

    def calcDensityAndGradientOnGrid(self, P):
        return self._run(P, True)

    def getNonNegligibleBlocks(self):
        return self._nonneg


class FunctionalData:
    def __init__(self, epuv, dFdRho, dFdGradRho, energy):
        self.epuv = epuv
        self.dFdRho = dFdRho
        self.dFdGradRho = dFdGradRho
        self.energy = energy


class FunctionalLibrary:
    """FunctionalLibrary(maxBlockSize).calcData(GRADIENTS, functional, rho[, grad], weights)."""

    def __init__(self, ctx, max_block_size=128):
        self.ctx = ctx
        self.max_block_size = max_block_size

    def calcData(self, functional, rho, grad, weights):
        N = len(rho)
        f = FUNCTIONALS[functional]
        gga = functional in GGA
        rho = _f64(rho)
        w = _f64(weights)
        gr = _f64(np.asarray(grad).reshape(-1)) if (gga and grad is not None) else None
        eps, v = np.zeros(N), np.zeros(N)
        g = np.zeros(3 * N) if gga else None
        E = C.c_double()
        check(_lib.lib().xc_functional(self.ctx.handle, f, N, self.max_block_size, _d(rho), _d(gr), _d(w),
                                       _d(eps), _d(v), _d(g), C.byref(E)))
        return FunctionalData(eps, v, g.reshape(3, N) if gga else None, E.value)


class ScalarOperatorToMatrixAdder:
    """F += integral of a scalar (and gradient) potential over the grid."""

    def __init__(self, bfogc, block_ave_threshold=1e-11):
        self.bfogc = bfogc
        self.blockAveThreshold = block_ave_threshold

    def addScalarOperatorToMatrix(self, F, scalar, gradient=None):
        n = self.bfogc.getNBasisFunctions()
        N = self.bfogc.getNGridPoints()
        if F.shape != (n, n):
            raise XCError(1, "ScalarOperatorToMatrixAdder: matrix is %s, basis has %d functions" % (F.shape, n))
        if len(scalar) != N:
            raise XCError(1, "ScalarOperatorToMatrixAdder: the scalar operator and the grid differ")
        Ff = np.asfortranarray(F, dtype=np.float64)
        g = _f64(np.asarray(gradient).reshape(-1)) if gradient is not None else None
        if g is not None and len(g) != 3 * N:
            raise XCError(1, "ScalarOperatorToMatrixAdder: the gradient operator and the grid differ")
        check(_lib.lib().xc_fock_add(self.bfogc.handle, float(self.blockAveThreshold), _d(_f64(scalar)), _d(g),
                                     Ff.ctypes.data_as(_lib.dp)))
        F[...] = Ff
        return F


class FuncPotential:
    """XC potential of a restricted density matrix: getMatrix(P) -> F, getEnergy() -> E_xc.

    ``precision`` "FP64" runs both contractions on FP64 matrix cores; "FP32" runs
    them on FP32 matrix cores (needs ``keep_fp32`` tiles).  The functional and
    the E_xc / F accumulations are FP64 in both."""

    def __init__(self, bfogc, functional="PBE", precision="FP64", block_ave_threshold=1e-11):
        if functional not in FUNCTIONALS:
            raise XCError(1, "unknown functional " + functional)
        self.bfogc = bfogc
        self.functional = functional
        self.precision = precision
        self.blockAveThreshold = block_ave_threshold
        self._energy = None

    def getMatrix(self, P):
        n = self.bfogc.getNBasisFunctions()
        Pf = np.asfortranarray(P, dtype=np.float64)
        if Pf.shape != (n, n):
            raise XCError(1, "FuncPotential: density matrix is %s, basis has %d functions" % (Pf.shape, n))
        F = np.zeros((n, n), order="F")
        E = C.c_double()
        check(_lib.lib().xc_fock_build(self.bfogc.handle, FUNCTIONALS[self.functional], PRECISIONS[self.precision],
                                       float(self.blockAveThreshold), Pf.ctypes.data_as(_lib.dp),
                                       F.ctypes.data_as(_lib.dp), C.byref(E)))
        self._energy = E.value
        return F

    def getEnergy(self):
        if self._energy is None:
            raise XCError(4, "FuncPotential: getMatrix has not been called")
        return self._energy

    def getMatrixUnrestricted(self, Pa, Pb):
        """FuncPotential<UNRESTRICTED>::getMatrix: (F_alpha, F_beta) of the spin density matrices."""
        n = self.bfogc.getNBasisFunctions()
        Pa = np.asfortranarray(Pa, dtype=np.float64)
        Pb = np.asfortranarray(Pb, dtype=np.float64)
        if Pa.shape != (n, n) or Pb.shape != (n, n):
            raise XCError(1, "FuncPotential: spin density matrices are %s and %s, basis has %d functions"
                          % (Pa.shape, Pb.shape, n))
        Fa = np.zeros((n, n), order="F")
        Fb = np.zeros((n, n), order="F")
        E = C.c_double()
        check(_lib.lib().xc_fock_build_unrestricted(self.bfogc.handle, FUNCTIONALS[self.functional],
                                                    PRECISIONS[self.precision], float(self.blockAveThreshold),
                                                    Pa.ctypes.data_as(_lib.dp), Pb.ctypes.data_as(_lib.dp),
                                                    Fa.ctypes.data_as(_lib.dp), Fb.ctypes.data_as(_lib.dp),
                                                    C.byref(E)))
        self._energy = E.value
        return Fa, Fb

    def getGeomGradients(self, P, ao_atom=None, natoms=None):
        """XC part of the nuclear gradient (natoms x 3) of the density P at the fixed grid
        (FuncPotential::getGeomGradients).  ``P`` is a matrix (restricted) or a pair
        (P_alpha, P_beta) (unrestricted).  The controller needs highestDerivative 1 (LDA)
        or 2 (GGA).  ``ao_atom`` defaults to the basis' getAtomIndicesOfBasis()."""
        n = self.bfogc.getNBasisFunctions()
        pair = isinstance(P, (tuple, list))
        Ps = [np.asfortranarray(p, dtype=np.float64) for p in (P if pair else [P])]
        for Pf in Ps:
            if Pf.shape != (n, n):
                raise XCError(1, "FuncPotential: density matrix is %s, basis has %d functions" % (Pf.shape, n))
        atom = np.ascontiguousarray(self.bfogc.basis.getAtomIndicesOfBasis() if ao_atom is None else ao_atom,
                                    dtype=np.int32)
        na = int(atom.max()) + 1 if natoms is None else int(natoms)
        G = np.zeros((na, 3), order="F")
        E = C.c_double()
        f = FUNCTIONALS[self.functional]
        if pair:
            check(_lib.lib().xc_geom_gradient_unrestricted(self.bfogc.handle, f, Ps[0].ctypes.data_as(_lib.dp),
                                                           Ps[1].ctypes.data_as(_lib.dp), na, _i(atom),
                                                           G.ctypes.data_as(_lib.dp), C.byref(E)))
        else:
            check(_lib.lib().xc_geom_gradient(self.bfogc.handle, f, Ps[0].ctypes.data_as(_lib.dp), na, _i(atom),
                                              G.ctypes.data_as(_lib.dp), C.byref(E)))
        self._energy = E.value
        return np.array(G)

    def build_device(self, dP, dF, sync=True):
        """Resident form: dP, dF are device pointers (ctypes c_void_p)."""
        E = C.c_double()
        check(_lib.lib().xc_fock_build_device(self.bfogc.handle, FUNCTIONALS[self.functional],
                                              PRECISIONS[self.precision], float(self.blockAveThreshold), dP, dF,
                                              C.byref(E) if sync else None))
        if sync:
            self._energy = E.value
            return E.value
        return None


def fock_build_packed(potentials, dPs, dFs):
    """XC matrices and energies of several controllers in one packed launch per stage
    (xc_fock_build_packed): ``potentials`` are FuncPotential objects of distinct controllers
    sharing one functional, precision and block-average threshold; ``dPs``/``dFs`` device
    pointers of their density and XC matrices.  Returns the energies (also kept on each
    potential)."""
    k = len(potentials)
    if k == 0:
        return []
    p0 = potentials[0]
    for p in potentials:
        if (p.functional, p.precision, p.blockAveThreshold) != (p0.functional, p0.precision, p0.blockAveThreshold):
            raise XCError(1, "fock_build_packed: the potentials differ in functional, precision or threshold")
    hs = (C.c_void_p * k)(*[p.bfogc.handle for p in potentials])
    Ps = (C.c_void_p * k)(*[C.c_void_p(x.value if isinstance(x, C.c_void_p) else x) for x in dPs])
    Fs = (C.c_void_p * k)(*[C.c_void_p(x.value if isinstance(x, C.c_void_p) else x) for x in dFs])
    E = np.zeros(k)
    check(_lib.lib().xc_fock_build_packed(k, hs, FUNCTIONALS[p0.functional], PRECISIONS[p0.precision],
                                          float(p0.blockAveThreshold), Ps, Fs, _d(E)))
    for p, e in zip(potentials, E):
        p._energy = float(e)
    return E.tolist()


class NAddFuncPotential:
    """Non-additive functional potential of the active subsystem in an embedding
    (NAddFuncPotential::getMatrix/getEnergy, restricted, NAddFuncPotential.cpp:192-301).

    ``active`` and every controller in ``environment`` are BasisFunctionOnGridControllers of
    the subsystem bases on ONE shared (supersystem) grid.  getMatrix(P_act, [P_env...]) returns
    the integral of v[rho_sup] - v[rho_act] in the active basis; getEnergy() the energy
    E[rho_sup] - E[rho_act] - sum_env E[rho_env] of that build."""

    def __init__(self, active, environment, functional="PBE", precision="FP64", block_ave_threshold=1e-11):
        if functional not in FUNCTIONALS:
            raise XCError(1, "unknown functional " + functional)
        self.active = active
        self.environment = list(environment)
        self.functional = functional
        self.precision = precision
        self.blockAveThreshold = block_ave_threshold
        self._energy = None

    def _env_handles(self):
        return (C.c_void_p * max(1, len(self.environment)))(*[e.handle for e in self.environment])

    def getMatrix(self, P_active, P_environment):
        n = self.active.getNBasisFunctions()
        Pa = np.asfortranarray(P_active, dtype=np.float64)
        if Pa.shape != (n, n):
            raise XCError(1, "NAddFuncPotential: active density matrix is %s, basis has %d functions" % (Pa.shape, n))
        if len(P_environment) != len(self.environment):
            raise XCError(1, "NAddFuncPotential: %d environment densities for %d environment controllers"
                          % (len(P_environment), len(self.environment)))
        Pe = []
        for ctl, P in zip(self.environment, P_environment):
            m = ctl.getNBasisFunctions()
            Pf = np.asfortranarray(P, dtype=np.float64)
            if Pf.shape != (m, m):
                raise XCError(1, "NAddFuncPotential: environment density matrix is %s, basis has %d functions"
                              % (Pf.shape, m))
            Pe.append(Pf)
        ptrs = (_lib.dp * max(1, len(Pe)))(*[p.ctypes.data_as(_lib.dp) for p in Pe])
        F = np.zeros((n, n), order="F")
        E = C.c_double()
        check(_lib.lib().xc_nadd_fock_build(self.active.handle, len(Pe), self._env_handles(),
                                            FUNCTIONALS[self.functional], PRECISIONS[self.precision],
                                            float(self.blockAveThreshold), Pa.ctypes.data_as(_lib.dp), ptrs,
                                            F.ctypes.data_as(_lib.dp), C.byref(E)))
        self._energy = E.value
        return F

    def getMatrixUnrestricted(self, Pa_active, Pb_active, Pa_environment, Pb_environment):
        """NAddFuncPotential<UNRESTRICTED>::getMatrix: spin density matrices of the active and
        environment subsystems in, (F_alpha, F_beta) of the active subsystem out
        (xc_nadd_fock_build_unrestricted); getEnergy() then returns the non-additive energy."""
        n = self.active.getNBasisFunctions()
        if len(Pa_environment) != len(self.environment) or len(Pb_environment) != len(self.environment):
            raise XCError(1, "NAddFuncPotential: environment spin densities do not match the environment controllers")

        def mat(P, m, what):
            Pf = np.asfortranarray(P, dtype=np.float64)
            if Pf.shape != (m, m):
                raise XCError(1, "NAddFuncPotential: %s density matrix is %s, basis has %d functions"
                              % (what, Pf.shape, m))
            return Pf
        Pa, Pb = mat(Pa_active, n, "active alpha"), mat(Pb_active, n, "active beta")
        Pea = [mat(P, c.getNBasisFunctions(), "environment alpha") for c, P in zip(self.environment, Pa_environment)]
        Peb = [mat(P, c.getNBasisFunctions(), "environment beta") for c, P in zip(self.environment, Pb_environment)]
        pa = (_lib.dp * max(1, len(Pea)))(*[p.ctypes.data_as(_lib.dp) for p in Pea])
        pb = (_lib.dp * max(1, len(Peb)))(*[p.ctypes.data_as(_lib.dp) for p in Peb])
        Fa, Fb = np.zeros((n, n), order="F"), np.zeros((n, n), order="F")
        E = C.c_double()
        check(_lib.lib().xc_nadd_fock_build_unrestricted(
            self.active.handle, len(Pea), self._env_handles(), FUNCTIONALS[self.functional],
            PRECISIONS[self.precision], float(self.blockAveThreshold), Pa.ctypes.data_as(_lib.dp),
            Pb.ctypes.data_as(_lib.dp), pa, pb, Fa.ctypes.data_as(_lib.dp), Fb.ctypes.data_as(_lib.dp), C.byref(E)))
        self._energy = E.value
        return Fa, Fb

    def getEnergy(self):
        if self._energy is None:
            raise XCError(4, "NAddFuncPotential: getMatrix has not been called")
        return self._energy

    def build_device(self, dP_active, dP_environment, dF, sync=True):
        """Resident form: device pointers (ctypes c_void_p) for every density matrix and F."""
        if len(dP_environment) != len(self.environment):
            raise XCError(1, "NAddFuncPotential: %d environment densities for %d environment controllers"
                          % (len(dP_environment), len(self.environment)))
        ptrs = (C.c_void_p * max(1, len(dP_environment)))(*dP_environment)
        E = C.c_double()
        check(_lib.lib().xc_nadd_fock_build_device(self.active.handle, len(dP_environment), self._env_handles(),
                                                   FUNCTIONALS[self.functional], PRECISIONS[self.precision],
                                                   float(self.blockAveThreshold), dP_active, ptrs, dF,
                                                   C.byref(E) if sync else None))
        if sync:
            self._energy = E.value
            return E.value
        return None

    def build_cached(self, dF, sync=True):
        """Build from the densities cached in every controller (update_density); dF: device pointer."""
        E = C.c_double()
        check(_lib.lib().xc_nadd_fock_build_cached(self.active.handle, len(self.environment), self._env_handles(),
                                                   FUNCTIONALS[self.functional], PRECISIONS[self.precision],
                                                   float(self.blockAveThreshold), dF,
                                                   C.byref(E) if sync else None))
        if sync:
            self._energy = E.value
            return E.value
        return None


class Server(_Handle):
    """Asynchronous XC Fock builds on resident controllers: ``nslots`` builds overlap on the
    device, each on its own stream; with ``use_graphs`` a controller's build is captured once
    per (functional, precision, threshold) and replayed.  submit() returns a ticket, or None
    when no slot is free or the controller already has a build in flight."""

    _destroy = "xc_server_destroy"

    def __init__(self, ctx, nslots=4, use_graphs=True):
        h = C.c_void_p()
        check(_lib.lib().xc_server_create(ctx.handle, int(nslots), int(bool(use_graphs)), C.byref(h)))
        self.handle = h
        self.ctx = ctx
        self._out = {}  # ticket -> (F, energy holder) kept alive until finished
        ctx._register(self)

    def submit(self, bfogc, functional, precision, P, block_ave_threshold=1e-11):
        if functional not in FUNCTIONALS or precision not in PRECISIONS:
            raise XCError(1, "Server.submit: unknown functional or precision %s / %s" % (functional, precision))
        n = bfogc.getNBasisFunctions()
        Pf = np.asfortranarray(P, dtype=np.float64)
        if Pf.shape != (n, n):
            raise XCError(1, "Server.submit: density matrix is %s, basis has %d functions" % (Pf.shape, n))
        F = np.zeros((n, n), order="F")
        E = (C.c_double * 1)()
        t = C.c_int64()
        rc = _lib.lib().xc_server_submit(self.handle, bfogc.handle, FUNCTIONALS[functional], PRECISIONS[precision],
                                         float(block_ave_threshold), Pf.ctypes.data_as(_lib.dp),
                                         F.ctypes.data_as(_lib.dp), E, C.byref(t))
        if rc == XC_BUSY:
            return None
        check(rc)
        self._out[t.value] = (F, E)
        return t.value

    def poll(self, ticket):
        """(F, E) once the build of ``ticket`` has finished, else None."""
        done = C.c_int()
        check(_lib.lib().xc_server_poll(self.handle, C.c_int64(ticket), C.byref(done)))
        if not done.value:
            return None
        F, E = self._out.pop(ticket)
        return F, E[0]

    def wait(self, ticket):
        check(_lib.lib().xc_server_wait(self.handle, C.c_int64(ticket)))
        F, E = self._out.pop(ticket)
        return F, E[0]

    def stats(self):
        b, s, g = C.c_int(), C.c_int64(), C.c_int64()
        check(_lib.lib().xc_server_stats(self.handle, C.byref(b), C.byref(s), C.byref(g)))
        return dict(busy_slots=b.value, submitted=s.value, graph_launches=g.value)
