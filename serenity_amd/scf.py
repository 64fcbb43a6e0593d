"""Restricted Kohn-Sham SCF around the XC Fock build.

Mirrors the callers of FuncPotential in Serenity:

* ``ScfIntegrals`` -- the one-electron Hamiltonian h = T + V, the overlap S
  and the Coulomb matrix J (exact four-centre integrals or density fitting,
  Serenity's default ``densFitJ = RI`` with the RI_J_WEIGEND auxiliary basis,
  Settings.h:128-138; RI_J_IntegralController.cpp:41-121).
* ``perform`` / ``perform_unrestricted`` -- ``Scf<SCFMode>::perform`` (src/scf/Scf.cpp:38-148): build
  F = h + J[P] + V_xc[P] from the current density, extrapolate it by DIIS,
  diagonalise in the canonically orthogonalised basis, occupy the lowest
  orbitals, repeat until the energy, the density and the DIIS error are
  converged.
* ``ScfTask.run`` -- ``ScfTask<RESTRICTED>::run`` (src/tasks/ScfTask.cpp:368-)
  with Serenity's grid protocol: the SCF iterations use the SMALL grid
  (accuracy 2, ScfTask.cpp:358-359) and the final energy is evaluated once on
  the DEFAULT grid (accuracy 4) from the converged density
  (finalDFTEnergyEvaluation, ScfTask.cpp:127-153).

Precision modes of the XC build inside the SCF: "FP64" runs every build on
FP64 matrix cores.  "MIXED" iterates with FP32-MFMA contractions until the
FP32-level convergence criteria hold, then restarts the DIIS and finishes with
FP64 builds from that density (iterative refinement); the energies before and
after the FP64 finish are both reported.

The linear algebra (eigh, DIIS, the density-fitted J contractions) is numpy on
the host.  At 2752 AOs ((H2O)64/def2-TZVP) an iteration then spends ~0.95 s in
the GPU J engine, ~0.67 s in host linear algebra (three n^3 products and an
n x n eigendecomposition) and ~0.03 s in the XC build; the per-phase times are
returned.  (torch, tried for the products, reported no GPU when loaded in the
same process after this library, so they stay on the host.)
"""
import ctypes as C
import json
import os
import time

import numpy as np

from . import _lib
from . import xc
from .xc import check

_d = lambda a: a.ctypes.data_as(_lib.dp)
_i = lambda a: a.ctypes.data_as(_lib.ip)


def nuclear_repulsion(Z, coords_bohr):
    Z = np.asarray(Z, dtype=np.float64)
    X = np.asarray(coords_bohr, dtype=np.float64).reshape(-1, 3)
    e = 0.0
    for a in range(len(Z)):
        for b in range(a):
            e += Z[a] * Z[b] / np.linalg.norm(X[a] - X[b])
    return e


def one_electron(basis, Z, coords_bohr):
    """(T, V): kinetic energy and nuclear attraction, n x n."""
    n = basis.nao
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    X = np.ascontiguousarray(np.asarray(coords_bohr, dtype=np.float64).reshape(-1, 3))
    T = np.zeros((n, n), order="F")
    V = np.zeros((n, n), order="F")
    check(_lib.lib().xc_one_electron(*basis.as_args(), len(Z), _i(Z), _d(X), _d(T.reshape(-1, order="F")),
                                     _d(V.reshape(-1, order="F"))))
    return T, V


def eri_packed(basis, schwarz=1e-14):
    """(mu nu|la si) over packed AO pairs pair(mu, nu) = mu (mu + 1) / 2 + nu, mu >= nu."""
    npair = basis.nao * (basis.nao + 1) // 2
    E = np.zeros((npair, npair))
    check(_lib.lib().xc_eri_packed(*basis.as_args(), C.c_double(schwarz), _d(E)))
    return E


def ri_j_integrals(basis, aux, schwarz=1e-14):
    """(B3, M): (P|mu nu) as naux x npair and the metric (P|Q)."""
    npair = basis.nao * (basis.nao + 1) // 2
    B3 = np.zeros((aux.nao, npair))
    M = np.zeros((aux.nao, aux.nao))
    check(_lib.lib().xc_ri_j_integrals(*basis.as_args(), *aux.as_args(), C.c_double(schwarz), _d(B3), _d(M)))
    return B3, M


def pack_density(P):
    """Packed pair vector of a symmetric matrix with the off-diagonal pairs doubled."""
    n = P.shape[0]
    r, c = np.tril_indices(n)
    v = P[r, c].copy()
    v[r != c] *= 2.0
    return v


def unpack_symmetric(v, n):
    out = np.zeros((n, n))
    r, c = np.tril_indices(n)
    out[r, c] = v
    out[c, r] = v
    return out


class ScfIntegrals:
    """S, h, E_nuc and the Coulomb matrix of one system (restricted).

    ``aux`` = None: exact four-centre J (Serenity's densFitJ NONE);
    otherwise density-fitted J over the auxiliary basis (densFitJ RI)."""

    def __init__(self, basis, Z, coords_bohr, aux=None, schwarz=1e-14, coulomb=None):
        t0 = time.perf_counter()
        self.basis, self.aux = basis, aux
        self.n = basis.nao
        self.S = basis.overlap()
        self.T, self.V = one_electron(basis, Z, coords_bohr)
        self.h = self.T + self.V
        self.e_nuc = nuclear_repulsion(Z, coords_bohr)
        self.gpu_j = coulomb  # a DensityFittedCoulomb: J on the GPU instead of the host integrals
        if coulomb is not None:
            self.eri = self.D = None
        elif aux is None:
            self.eri = eri_packed(basis, schwarz)
            self.D = None
        else:
            B3, M = ri_j_integrals(basis, aux, schwarz)
            Lm = np.linalg.cholesky(M)
            # D = L^-1 B3, so that J = D^T (D p) = B3^T M^-1 B3 p
            self.D = np.linalg.solve(Lm, B3)
            self.eri = None
        self.setup_s = time.perf_counter() - t0

    def coulomb(self, P):
        if self.gpu_j is not None:
            return self.gpu_j.coulomb(P)
        p = pack_density(P)
        if self.D is not None:
            j = self.D.T @ (self.D @ p)
        else:
            j = self.eri @ p
        return unpack_symmetric(j, self.n)

    def exchange(self, P):
        """K_mu,nu = sum_la,si (mu la|nu si) P_la,si from the exact four-centre integrals
        (held as an n^4 array after the first call; small systems only)."""
        if self.eri is None:
            raise xc.XCError(4, "ScfIntegrals: exchange needs the exact four-centre integrals (aux=None)")
        if getattr(self, "_eri4", None) is None:
            n = self.n
            r, c = np.tril_indices(n)
            pair = np.zeros((n, n), dtype=np.int64)
            pair[r, c] = np.arange(len(r))
            pair[c, r] = pair[r, c]
            self._eri4 = self.eri[pair[:, :, None, None], pair[None, None, :, :]]  # (mu nu|la si)
        return np.einsum("mlns,ls->mn", self._eri4, P, optimize=True)


def canonical_orthogonalization(S, thr=1e-7):
    """X with X^T S X = 1 (eigenvectors of S with eigenvalue above thr; canOrthThreshold)."""
    e, U = np.linalg.eigh(S)
    keep = e > thr
    return U[:, keep] / np.sqrt(e[keep])


class Diis:
    """Pulay DIIS on Fock matrices with the orthonormal-basis commutator error."""

    def __init__(self, max_store=10):
        self.max_store = max_store
        self.F, self.E = [], []

    def reset(self):
        self.F, self.E = [], []

    def extrapolate(self, F, err):
        self.F.append(F)
        self.E.append(err)
        if len(self.F) > self.max_store:
            self.F.pop(0)
            self.E.pop(0)
        m = len(self.F)
        if m < 2:
            return F
        B = -np.ones((m + 1, m + 1))
        B[m, m] = 0.0
        for i in range(m):
            for j in range(i + 1):
                B[i, j] = B[j, i] = float(np.vdot(self.E[i], self.E[j]))
        rhs = np.zeros(m + 1)
        rhs[m] = -1.0
        try:
            c = np.linalg.solve(B, rhs)[:m]
        except np.linalg.LinAlgError:
            return F
        return sum(ci * Fi for ci, Fi in zip(c, self.F))


def perform(ints, nocc, xc_build, P0=None, max_cycles=100, energy_thr=1e-10, rmsd_thr=1e-9, diis_thr=1e-8,
            diis_store=10, log=None, exchange=0.0):
    """Scf<RESTRICTED>::perform.  xc_build(P) -> (E_xc, V_xc), or None for Hartree-Fock;
    ``exchange`` is the fraction of exact exchange (F -= x K / 2, E -= x tr(P K) / 4).  Returns a dict
    with the energy of the last density, P, orbitals, orbital energies, cycles and the energy history."""
    say = log or (lambda *a: None)
    X = canonical_orthogonalization(ints.S)
    S, h = ints.S, ints.h
    if P0 is None:
        e, Cp = np.linalg.eigh(X.T @ h @ X)
        Co = X @ Cp[:, :nocc]
        P = 2.0 * Co @ Co.T
    else:
        P = np.array(P0, dtype=np.float64)
    diis = Diis(diis_store)
    e_old = None
    hist = []
    t_xc = t_j = t_alg = 0.0
    for cycle in range(1, max_cycles + 1):
        t0 = time.perf_counter()
        J = ints.coulomb(P)
        t1 = time.perf_counter()
        e_xc, Vxc = xc_build(P) if xc_build is not None else (0.0, 0.0)
        t2 = time.perf_counter()
        t_j += t1 - t0
        t_xc += t2 - t1
        F = h + J + Vxc
        energy = ints.e_nuc + float(np.vdot(P, h)) + 0.5 * float(np.vdot(P, J)) + e_xc
        if exchange:
            K = ints.exchange(P)
            F = F - 0.5 * exchange * K
            energy -= 0.25 * exchange * float(np.vdot(P, K))
        t3 = time.perf_counter()
        FPS = F @ P @ S
        err = X.T @ (FPS - FPS.T) @ X
        err_norm = float(np.sqrt(np.mean(err * err)))
        Fx = diis.extrapolate(F, err)
        eps, Cp = np.linalg.eigh(X.T @ Fx @ X)
        Cmo = X @ Cp
        Co = Cmo[:, :nocc]
        P_new = 2.0 * Co @ Co.T
        rmsd = float(np.sqrt(np.mean((P_new - P) ** 2)))
        t_alg += time.perf_counter() - t3
        de = None if e_old is None else energy - e_old
        hist.append(dict(cycle=cycle, energy=energy, dE=de, rmsd=rmsd, diis_error=err_norm, e_xc=e_xc))
        say("cycle %3d  E = %.12f  dE = %s  rmsd = %.2e  diis = %.2e" %
            (cycle, energy, "      -     " if de is None else "%.2e" % de, rmsd, err_norm))
        if de is not None and abs(de) < energy_thr and rmsd < rmsd_thr and err_norm < diis_thr:
            return dict(energy=energy, P=P, C=Cmo, eps=eps, cycles=cycle, converged=True, history=hist,
                        e_xc=e_xc, t_xc=t_xc, t_j=t_j, t_alg=t_alg)
        e_old = energy
        P = P_new
    return dict(energy=energy, P=P, C=Cmo, eps=eps, cycles=max_cycles, converged=False, history=hist, e_xc=e_xc,
                t_xc=t_xc, t_j=t_j, t_alg=t_alg)


def perform_unrestricted(ints, nalpha, nbeta, xc_build, P0=None, max_cycles=150, energy_thr=1e-10, rmsd_thr=1e-9,
                         diis_thr=1e-8, diis_store=10, log=None):
    """Scf<UNRESTRICTED>::perform.  xc_build(Pa, Pb) -> (E_xc, Va, Vb).  F_s = h + J[Pa + Pb] + V_s;
    one DIIS over both spins (stacked commutator errors, common coefficients)."""
    say = log or (lambda *a: None)
    X = canonical_orthogonalization(ints.S)
    S, h = ints.S, ints.h
    if P0 is None:
        _, Cp = np.linalg.eigh(X.T @ h @ X)
        C0 = X @ Cp
        Pa, Pb = C0[:, :nalpha] @ C0[:, :nalpha].T, C0[:, :nbeta] @ C0[:, :nbeta].T
    else:
        Pa, Pb = (np.array(p, dtype=np.float64) for p in P0)
    diis = Diis(diis_store)
    e_old = None
    hist = []
    for cycle in range(1, max_cycles + 1):
        P = Pa + Pb
        J = ints.coulomb(P)
        e_xc, Va, Vb = xc_build(Pa, Pb)
        Fa, Fb = h + J + Va, h + J + Vb
        energy = ints.e_nuc + float(np.vdot(P, h)) + 0.5 * float(np.vdot(P, J)) + e_xc
        errs = []
        for F, Ps in ((Fa, Pa), (Fb, Pb)):
            FPS = F @ Ps @ S
            errs.append(X.T @ (FPS - FPS.T) @ X)
        err = np.concatenate([e.ravel() for e in errs])
        err_norm = float(np.sqrt(np.mean(err * err)))
        Fx = diis.extrapolate(np.stack([Fa, Fb]), err)
        new = []
        for F, no in ((Fx[0], nalpha), (Fx[1], nbeta)):
            _, Cp = np.linalg.eigh(X.T @ F @ X)
            Co = (X @ Cp)[:, :no]
            new.append(Co @ Co.T)
        rmsd = float(np.sqrt(0.5 * (np.mean((new[0] - Pa) ** 2) + np.mean((new[1] - Pb) ** 2))))
        de = None if e_old is None else energy - e_old
        hist.append(dict(cycle=cycle, energy=energy, dE=de, rmsd=rmsd, diis_error=err_norm, e_xc=e_xc))
        say("cycle %3d  E = %.12f  dE = %s  rmsd = %.2e  diis = %.2e" %
            (cycle, energy, "      -     " if de is None else "%.2e" % de, rmsd, err_norm))
        if de is not None and abs(de) < energy_thr and rmsd < rmsd_thr and err_norm < diis_thr:
            return dict(energy=energy, Pa=Pa, Pb=Pb, cycles=cycle, converged=True, history=hist, e_xc=e_xc)
        e_old = energy
        Pa, Pb = new
    return dict(energy=energy, Pa=Pa, Pb=Pb, cycles=max_cycles, converged=False, history=hist, e_xc=e_xc)


def gpu_xc_build_unrestricted(pot):
    def build(Pa, Pb):
        Fa, Fb = pot.getMatrixUnrestricted(Pa, Pb)
        return pot.getEnergy(), Fa, Fb
    return build


def total_energy(ints, P, e_xc):
    """E = E_nuc + tr(P h) + 1/2 tr(P J[P]) + E_xc[P]."""
    J = ints.coulomb(P)
    return ints.e_nuc + float(np.vdot(P, ints.h)) + 0.5 * float(np.vdot(P, J)) + e_xc


# FP32-level convergence of the mixed-precision iterations (the FP32 contractions leave
# about 1e-7 relative noise in F, below which the DIIS error does not fall).
MIXED_THRESHOLDS = dict(energy_thr=1e-6, rmsd_thr=1e-5, diis_thr=1e-4)


def gpu_xc_build(pot):
    def build(P):
        F = pot.getMatrix(P)
        return pot.getEnergy(), F
    return build


GUESS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "initial_guess.json")
ELEMENT_SYMBOLS = {z + 1: s for z, s in enumerate(("H", "HE", "LI", "BE", "B", "C", "N", "O", "F", "NE", "NA", "MG", "AL",
                                                     "SI", "P", "S", "CL", "AR"))}


def atomic_density_guess(basis_name, Z, spherical=True):
    """Serenity's ATOM_SCF initial guess (AtomicDensityGuessCalculator.cpp:84-181): for every element the
    stored spherically averaged free-atom density matrix in the MINAO basis is projected into the
    atom's target basis, P_t = X P_m X^T with X the least-squares solution of S_tt X = S_tm (SVD,
    singular values below 1e-6 of the largest dropped, as the reference's JacobiSVD threshold), and
    the atom blocks are placed on the diagonal in atom order."""
    table = json.load(open(GUESS_FILE))["spherical" if spherical else "cartesian"]
    origin = np.zeros((1, 3))
    per_element, blocks = {}, []
    for z in np.asarray(Z, dtype=int):
        sym = ELEMENT_SYMBOLS.get(int(z))
        if sym is None or sym not in table:
            raise xc.XCError(1, "atomic_density_guess: no stored free-atom density for Z = %d" % z)
        if sym not in per_element:
            bt = xc.BasisController.from_file(basis_name, [z], origin, spherical=spherical)
            bm = xc.BasisController.from_file("MINAO", [z], origin, spherical=spherical)
            both = xc.BasisController(np.concatenate([bt.l, bm.l]), np.concatenate([bt.pure, bm.pure]),
                                      np.concatenate([bt.nprim, bm.nprim]), np.vstack([bt.center, bm.center]),
                                      np.concatenate([bt.alpha, bm.alpha]), np.concatenate([bt.coef, bm.coef]))
            S = both.overlap()
            nt = bt.nao
            U, sv, Vt = np.linalg.svd(S[:nt, :nt])
            keep = sv > 1e-6 * sv[0]
            X = Vt[keep].T @ ((U[:, keep].T @ S[:nt, nt:]) / sv[keep][:, None])
            d = table[sym]
            Pm = np.asarray(d["P"], dtype=np.float64).reshape(d["n"], d["n"])
            if Pm.shape[0] != bm.nao:
                raise xc.XCError(1, "atomic_density_guess: stored %s density has %d functions, MINAO %d"
                                 % (sym, Pm.shape[0], bm.nao))
            per_element[sym] = X @ Pm @ X.T
        blocks.append(per_element[sym])
    n = sum(b.shape[0] for b in blocks)
    P = np.zeros((n, n))
    o = 0
    for b in blocks:
        k = b.shape[0]
        P[o:o + k, o:o + k] = b
        o += k
    return P


class ScfTask:
    """ScfTask<RESTRICTED>::run for a KS-DFT system on the GPU XC build.

    Grids: SCF iterations on the SMALL grid (accuracy ``small_accuracy``), the final
    energy on the DEFAULT grid (accuracy ``accuracy``) from the converged density,
    as Serenity does (ScfTask.cpp:127-153, 358-359)."""

    def __init__(self, ctx, Z, coords_bohr, basis="DEF2-SVP", functional="LDA", ri=True, aux="RI_J_WEIGEND",
                 accuracy=4, small_accuracy=2, spherical=True, schwarz=1e-14, fp32_tiles=True, log=None,
                 aux_spherical=None, radial_threshold=1e-9, block_ave_threshold=1e-11, highest_derivative=1,
                 spin=0, unrestricted=False, gpu_coulomb=False, guess="ATOM_SCF"):
        self.ctx = ctx
        self.Z = np.ascontiguousarray(Z, dtype=np.int32)
        self.X = np.asarray(coords_bohr, dtype=np.float64).reshape(-1, 3)
        self.functional = functional
        self.say = log or (lambda *a: None)
        t0 = time.perf_counter()
        self.basis = xc.BasisController.from_file(basis, self.Z, self.X, spherical=spherical)
        self.auxbasis = (xc.BasisController.from_file(aux, self.Z, self.X,
                                                      spherical=spherical if aux_spherical is None else aux_spherical)
                         if ri else None)
        self.radial_threshold = radial_threshold
        self.block_ave_threshold = block_ave_threshold
        self.highest_derivative = highest_derivative
        jgpu = xc.DensityFittedCoulomb(ctx, self.basis, self.auxbasis) if (gpu_coulomb and ri) else None
        self.ints = ScfIntegrals(self.basis, self.Z, self.X, self.auxbasis, schwarz, coulomb=jgpu)
        nel = int(self.Z.sum())
        if (nel - spin) % 2 or spin < 0 or spin > nel:
            raise xc.XCError(1, "ScfTask: %d electrons cannot have %d unpaired" % (nel, spin))
        self.nocc = nel // 2
        self.nalpha, self.nbeta = (nel + spin) // 2, (nel - spin) // 2
        self.unrestricted = bool(unrestricted or spin)
        self.grid_small = xc.molecular_grid(self.Z, self.X, accuracy=small_accuracy, ctx=ctx)
        self.grid_default = (self.grid_small if accuracy == small_accuracy
                             else xc.molecular_grid(self.Z, self.X, accuracy=accuracy, ctx=ctx))
        # initial guess (Settings::scf.initialguess, default ATOM_SCF); H_CORE: core Hamiltonian orbitals
        if guess not in ("ATOM_SCF", "H_CORE"):
            raise xc.XCError(1, "ScfTask: unknown initial guess " + str(guess))
        self.guess = guess
        self.P_guess = atomic_density_guess(basis, self.Z, spherical) if guess == "ATOM_SCF" else None
        if self.P_guess is not None and self.P_guess.shape[0] != self.basis.nao:
            raise xc.XCError(1, "ScfTask: guess density does not match the basis")
        self.setup_s = time.perf_counter() - t0
        self.fp32_tiles = fp32_tiles
        self._bf = {}

    def _potential(self, which, precision):
        """FuncPotential on the small or default grid; one controller per grid serves both precisions
        (FP64 tiles plus, with fp32_tiles, their FP32 copy)."""
        grid = self.grid_small if which == "small" else self.grid_default
        if which not in self._bf:
            self._bf[which] = xc.BasisFunctionOnGridController(self.ctx, self.basis, grid, 128,
                                                               self.radial_threshold, self.highest_derivative,
                                                               keep_fp32=self.fp32_tiles)
        return xc.FuncPotential(self._bf[which], self.functional, precision, self.block_ave_threshold)

    def final_energy(self, P, precision="FP64"):
        """finalDFTEnergyEvaluation: E_xc on the default grid from the converged P."""
        return self.final_evaluation(P, precision, update=False)[0]

    def final_evaluation(self, P, precision="FP64", update=True):
        """finalDFTEnergyEvaluation (ScfTask.cpp:127-153): the energy of P with V_xc on the default
        grid, and the density of the orbitals updated with that Fock matrix (as the reference leaves
        it in the electronic structure).  With equal grids the SCF density is kept."""
        pot = self._potential("default", precision)
        Vxc = pot.getMatrix(P)
        J = self.ints.coulomb(P)
        e = self.ints.e_nuc + float(np.vdot(P, self.ints.h)) + 0.5 * float(np.vdot(P, J)) + pot.getEnergy()
        if self.grid_default is self.grid_small or not update:
            return e, P
        X = canonical_orthogonalization(self.ints.S)
        _, Cp = np.linalg.eigh(X.T @ (self.ints.h + J + Vxc) @ X)
        Co = (X @ Cp)[:, :self.nocc]
        return e, 2.0 * Co @ Co.T

    def run_unrestricted(self, precision="FP64", max_cycles=150):
        """Scf<UNRESTRICTED> on the small grid, final energy of the converged spin densities on the
        default grid (no orbital update is reported)."""
        t0 = time.perf_counter()
        r = perform_unrestricted(self.ints, self.nalpha, self.nbeta,
                                 gpu_xc_build_unrestricted(self._potential("small", precision)),
                                 max_cycles=max_cycles, log=self.say,
                                 **(MIXED_THRESHOLDS if precision == "FP32" else {}))
        pot = self._potential("default", precision)
        pot.getMatrixUnrestricted(r["Pa"], r["Pb"])
        energy = total_energy(self.ints, r["Pa"] + r["Pb"], pot.getEnergy())
        return dict(mode="UNRESTRICTED " + precision, energy=energy, energy_small_grid=r["energy"],
                    cycles=r["cycles"], converged=r["converged"], Pa=r["Pa"], Pb=r["Pb"],
                    wall_s=round(time.perf_counter() - t0, 3))

    def run(self, mode="FP64", max_cycles=100, P0=None):
        """mode "FP64" or "MIXED" (iterations with the mixed-precision builds, then an FP64 finish from
        their converged density); returns energies, cycle counts and timings."""
        if self.unrestricted:
            if mode not in ("FP64", "FP32"):
                raise xc.XCError(1, "ScfTask: unrestricted runs take the precision FP64 or FP32")
            return self.run_unrestricted(mode, max_cycles)
        t0 = time.perf_counter()
        out = dict(mode=mode, nao=self.basis.nao, naux=None if self.auxbasis is None else self.auxbasis.nao,
                   npts_small=self.grid_small.getNGridPoints(), npts_default=self.grid_default.getNGridPoints())
        if P0 is None:
            P0 = self.P_guess
        if mode == "FP64":
            r = perform(self.ints, self.nocc, gpu_xc_build(self._potential("small", "FP64")), P0=P0,
                        max_cycles=max_cycles, log=self.say)
            out.update(cycles=r["cycles"], converged=r["converged"], energy_small_grid=r["energy"])
            out["energy"], out["P_final"] = self.final_evaluation(r["P"], "FP64")
        elif mode == "MIXED":
            # iterations with FP32 contractions (the dense blocks' density refined in FP64), the
            # energy of their converged density with the same builds, then the FP64 finish
            r32 = perform(self.ints, self.nocc, gpu_xc_build(self._potential("small", "MIXED")), P0=P0,
                          max_cycles=max_cycles, log=self.say, **MIXED_THRESHOLDS)
            out["cycles_fp32"] = r32["cycles"]
            out["energy_before_fp64_finish"] = self.final_energy(r32["P"], "MIXED")
            out["P_fp32"] = r32["P"]
            r = perform(self.ints, self.nocc, gpu_xc_build(self._potential("small", "FP64")), P0=r32["P"],
                        max_cycles=max_cycles, log=self.say)
            out.update(cycles_fp64_finish=r["cycles"], converged=r32["converged"] and r["converged"],
                       energy_small_grid=r["energy"])
            out["cycles"] = r32["cycles"] + r["cycles"]
            out["energy"], out["P_final"] = self.final_evaluation(r["P"], "FP64")
        else:
            raise xc.XCError(1, "ScfTask: unknown mode " + str(mode))
        out["P"] = r["P"]
        out["t_xc_s"] = round(r["t_xc"], 3)
        out["t_j_s"] = round(r["t_j"], 3)
        out["t_alg_s"] = round(r["t_alg"], 3)
        out["wall_s"] = round(time.perf_counter() - t0, 3)
        return out

    def geom_gradient(self, P):
        """XC part of the nuclear gradient of P on the default grid (FuncPotential::getGeomGradients)."""
        pot = self._potential("default", "FP64")
        return pot.getGeomGradients(P, natoms=len(self.Z))

    def close(self):
        for bf in self._bf.values():
            bf._release()
        self._bf = {}
        if self.ints.gpu_j is not None:
            self.ints.gpu_j._release()
