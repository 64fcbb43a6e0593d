"""TEST INFRASTRUCTURE -- ctypes front end of the CPU oracle (libxc_oracle.so).

Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of bench.py
use this module, as the checker; the product library never does.  Every
function mirrors a Serenity routine (see the header of xc_oracle.cpp for the
file:line of each).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxc_oracle.so")
FUNCTIONALS = {"LDA": 0, "PBE": 1, "TF": 2, "SLATER": 3, "VWN5": 4, "PBEX": 5, "PBEC": 6, "PBE_LIBXC": 7, "BP86": 8}
GGA = {"PBE", "PBEX", "PBEC", "PBE_LIBXC", "BP86"}

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.oracle_max_threads.restype = C.c_int
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(ip)


def _basis_args(b):
    return (C.c_int(b.nshell), _i(b.l), _i(b.pure), _i(b.nprim), _d(b.center), _d(b.alpha), _d(b.coef))


def _pts(points):
    return np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))


def ao_block(basis, points, thr=1e-9, deriv=1):
    """Dense AO data of one block: dict(values, grad (3,B,n), hess (6,B,n), negligible)."""
    pts = _pts(points)
    B, n = len(pts), basis.nao
    val = np.zeros(B * n)
    g = [np.zeros(B * n) for _ in range(3)]
    h = np.zeros(6 * B * n)
    neg = np.zeros(n, dtype=np.int32)
    lib().oracle_ao_block(*_basis_args(basis), _d(pts), C.c_int(B), C.c_double(thr), C.c_int(deriv), _d(val),
                          *[_d(x) for x in g], _d(h), _i(neg))
    return dict(values=val.reshape(n, B).T, grad=np.stack([x.reshape(n, B).T for x in g]),
                hess=np.stack([h[c * B * n:(c + 1) * B * n].reshape(n, B).T for c in range(6)]), negligible=neg)


def density(basis, points, P, block_size=128, thr=1e-9, grad=True, hess=False):
    pts = _pts(points)
    N = len(pts)
    Pf = np.asfortranarray(P, dtype=np.float64)
    rho = np.zeros(N)
    gr = np.zeros(3 * N) if (grad or hess) else None
    he = np.zeros(6 * N) if hess else None
    nb = (N + block_size - 1) // block_size
    nonneg = np.zeros(nb, dtype=np.int32)
    lib().oracle_density(*_basis_args(basis), _d(pts), C.c_int64(N), C.c_int(block_size), C.c_double(thr),
                         Pf.ctypes.data_as(dp), _d(rho), _d(gr), _d(he), _i(nonneg))
    out = dict(rho=rho, nonneg=nonneg)
    if gr is not None:
        out["grad"] = gr.reshape(3, N)
    if he is not None:
        out["hess"] = he.reshape(6, N)
    return out


def functional(name, rho, grad, weights, block_size=128, params=None):
    """eps, v, g (3 x N or None), E for a restricted density."""
    N = len(rho)
    gga = name in GGA
    if gga and grad is None:
        raise ValueError("GGA functional needs the density gradient")
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    gr = np.ascontiguousarray(np.asarray(grad, dtype=np.float64).reshape(-1)) if gga else None
    if gga and gr.size != 3 * N:
        raise ValueError("gradient must hold 3 x N values")
    eps, v = np.zeros(N), np.zeros(N)
    g = np.zeros(3 * N) if gga else None
    E = C.c_double()
    prm = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
    lib().oracle_functional(C.c_int(FUNCTIONALS[name]), C.c_int64(N), C.c_int(block_size), _d(rho), _d(gr), _d(w),
                            _d(prm), _d(eps), _d(v), _d(g), C.byref(E))
    return eps, v, (g.reshape(3, N) if gga else None), E.value


def fock_add(basis, points, weights, v, g=None, F=None, block_size=128, thr=1e-9, block_ave_thr=1e-11,
             block_range=(0, -1)):
    pts = _pts(points)
    N = len(pts)
    n = basis.nao
    Fo = np.zeros((n, n), order="F") if F is None else np.asfortranarray(F, dtype=np.float64).copy(order="F")
    gg = None if g is None else np.ascontiguousarray(np.asarray(g, dtype=np.float64).reshape(-1))
    lib().oracle_fock_add(*_basis_args(basis), _d(pts), _d(np.ascontiguousarray(weights, dtype=np.float64)),
                          C.c_int64(N), C.c_int(block_size), C.c_double(thr), C.c_double(block_ave_thr),
                          _d(np.ascontiguousarray(v, dtype=np.float64)), _d(gg), Fo.ctypes.data_as(dp),
                          C.c_int64(block_range[0]), C.c_int64(block_range[1]))
    return Fo


def fock_build(basis, points, weights, P, name, block_size=128, thr=1e-9, block_ave_thr=1e-11, params=None):
    """FuncPotential::getMatrix/getEnergy: (F, E_xc, rho, grad)."""
    pts = _pts(points)
    N = len(pts)
    n = basis.nao
    Pf = np.asfortranarray(P, dtype=np.float64)
    F = np.zeros((n, n), order="F")
    E = C.c_double()
    rho = np.zeros(N)
    grad = np.zeros(3 * N)
    prm = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
    lib().oracle_fock_build(*_basis_args(basis), _d(pts), _d(np.ascontiguousarray(weights, dtype=np.float64)),
                            C.c_int64(N), C.c_int(block_size), C.c_double(thr), C.c_double(block_ave_thr),
                            C.c_int(FUNCTIONALS[name]), _d(prm), Pf.ctypes.data_as(dp), F.ctypes.data_as(dp),
                            C.byref(E), _d(rho), _d(grad))
    return F, E.value, rho, grad.reshape(3, N)


def nadd_fock_build(active_basis, env_bases, points, weights, P_active, P_envs, name, block_size=128, thr=1e-9,
                    block_ave_thr=1e-11):
    """NAddFuncPotential::getMatrix/getEnergy, restricted (NAddFuncPotential.cpp:192-301).

    rho_sup = sum of the subsystem densities on the shared grid
    (SupersystemDensityOnGridController.cpp:157-190); v = v[rho_sup] - v[rho_act]
    (likewise d eps/d grad rho) is integrated with the active basis;
    E = E[rho_sup] - E[rho_act] - sum_env E[rho_env] (NAddEnergyHelper::getEnergy,
    NAddFuncPotential.cpp:507-514).  Returns (F, E)."""
    gga = name in GGA
    dens = [density(active_basis, points, P_active, block_size=block_size, thr=thr, grad=gga)]
    for b, P in zip(env_bases, P_envs):
        dens.append(density(b, points, P, block_size=block_size, thr=thr, grad=gga))
    rho_sup = np.sum([d["rho"] for d in dens], axis=0)
    grad_sup = np.sum([d["grad"] for d in dens], axis=0) if gga else None
    _, v_sup, g_sup, e_sup = functional(name, rho_sup, grad_sup, weights, block_size=block_size)
    _, v_act, g_act, e_act = functional(name, dens[0]["rho"], dens[0]["grad"] if gga else None, weights,
                                        block_size=block_size)
    E = e_sup - e_act
    for d in dens[1:]:
        E -= functional(name, d["rho"], d["grad"] if gga else None, weights, block_size=block_size)[3]
    F = fock_add(active_basis, points, weights, v_sup - v_act, (g_sup - g_act) if gga else None,
                 block_size=block_size, thr=thr, block_ave_thr=block_ave_thr)
    return F, E


def functional_pol(name, rho_a, rho_b, grad_a, grad_b, weights, block_size=128, params=None):
    """Unrestricted calcData: (eps, v_a, v_b, g_a, g_b, E) with g_s = d eps / d grad rho_s (3 x N)."""
    N = len(rho_a)
    gga = name in GGA
    eps, va, vb = np.zeros(N), np.zeros(N), np.zeros(N)
    ga, gb = np.zeros(3 * N), np.zeros(3 * N)
    E = C.c_double()
    prm = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    lib().oracle_functional_pol(C.c_int(FUNCTIONALS[name]), C.c_int64(N), C.c_int(block_size), _d(f(rho_a)),
                                _d(f(rho_b)), _d(f(grad_a)) if gga else None, _d(f(grad_b)) if gga else None,
                                _d(f(weights)), _d(prm), _d(eps), _d(va), _d(vb), _d(ga), _d(gb), C.byref(E))
    return eps, va, vb, ga.reshape(3, N), gb.reshape(3, N), E.value


def fock_build_pol(basis, points, weights, Pa, Pb, name, block_size=128, thr=1e-9, block_ave_thr=1e-11,
                   params=None):
    """FuncPotential<UNRESTRICTED>::getMatrix/getEnergy: (F_a, F_b, E_xc)."""
    pts = _pts(points)
    N = len(pts)
    n = basis.nao
    Fa = np.zeros((n, n), order="F")
    Fb = np.zeros((n, n), order="F")
    E = C.c_double()
    prm = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
    lib().oracle_fock_build_pol(*_basis_args(basis), _d(pts), _d(np.ascontiguousarray(weights, dtype=np.float64)),
                                C.c_int64(N), C.c_int(block_size), C.c_double(thr), C.c_double(block_ave_thr),
                                C.c_int(FUNCTIONALS[name]), _d(prm),
                                np.asfortranarray(Pa, dtype=np.float64).ctypes.data_as(dp),
                                np.asfortranarray(Pb, dtype=np.float64).ctypes.data_as(dp),
                                Fa.ctypes.data_as(dp), Fb.ctypes.data_as(dp), C.byref(E))
    return Fa, Fb, E.value


def xc_gradient(basis, points, weights, P, name, ao_atom, natoms, block_size=128, thr=1e-9, params=None):
    """FuncPotential::getGeomGradients: (natoms x 3 gradient, E_xc)."""
    pts = _pts(points)
    N = len(pts)
    Pf = np.asfortranarray(P, dtype=np.float64)
    G = np.zeros((natoms, 3), order="F")
    E = C.c_double()
    prm = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
    atom = np.ascontiguousarray(ao_atom, dtype=np.int32)
    lib().oracle_xc_gradient(*_basis_args(basis), _d(pts), _d(np.ascontiguousarray(weights, dtype=np.float64)),
                             C.c_int64(N), C.c_int(block_size), C.c_double(thr), C.c_int(FUNCTIONALS[name]),
                             _d(prm), Pf.ctypes.data_as(dp), C.c_int(natoms), _i(atom), G.ctypes.data_as(dp),
                             C.byref(E))
    return np.array(G), E.value


def max_threads():
    return lib().oracle_max_threads()
