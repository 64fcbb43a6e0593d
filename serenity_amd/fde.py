"""Freeze-and-thaw subsystem DFT on the GPU XC builds (restricted, embeddingMode NADD_FUNC).

FreezeAndThawTask<RESTRICTED>::run (src/tasks/FreezeAndThawTask.cpp:151-312) visits the
subsystems in turn; each visit is an FDETask (src/tasks/FDETask.cpp:60-140): the active
subsystem's KS equations in its own basis with every other subsystem frozen,

    F_act = h_act + J[rho_act + rho_env] + V_xc[rho_act]
            + V_xc^nadd[rho_act, rho_env] + V_T^nadd[rho_act, rho_env],

where h_act holds the attraction of every nucleus of the supersystem, J is the Coulomb matrix of
the supersystem density (density fitted over the supersystem's auxiliary basis, on the GPU J
engine), and the non-additive XC and kinetic potentials are NAddFuncPotential builds on the
supersystem grid (NAddFuncPotential.cpp:192-301).  A cycle ends when every subsystem's density
matrix changed by less than ``conv_thresh`` (Frobenius norm, FreezeAndThawTask.cpp:250-265, default
1e-6); the supersystem energy is

    E = E_nuc + sum_i tr(P_i h_i) + 1/2 tr(P J[P]) + sum_i E_xc[rho_i] + E_xc^nadd + T^nadd,

with P the block-diagonal supersystem density matrix (so sum_i E_xc[rho_i] + E_xc^nadd =
E_xc[rho_sup]).  Every subsystem's XC and non-additive builds run on the one supersystem grid, as
the reference does when it keeps the supersystem grid for all subsystems (gridCutOff < 0).
The environment subsystems' grid densities stay cached between builds
(xc_bfogc_update_density, the reference's DensityOnGridControllers of the passive systems).
"""
import time

import numpy as np

from . import scf as SCF
from . import xc


class _EmbeddedIntegrals:
    """The integral interface of scf.perform for one active subsystem with a frozen environment:
    S and h of the subsystem, h including the Coulomb potential of the frozen environment density,
    and J of the active density alone (both from the supersystem J engine)."""

    def __init__(self, task, i, J_env):
        sl = task.slices[i]
        self.S = task.S[i]
        self.h = task.h[i] + J_env[sl, sl]
        self.e_nuc = 0.0
        self._task, self._sl = task, sl

    def coulomb(self, P):
        return self._task.coulomb_of(self._sl, P)[self._sl, self._sl]


class FreezeAndThawTask:
    """Restricted freeze-and-thaw over closed-shell subsystems given as (Z, coords_bohr) pairs."""

    def __init__(self, ctx, subsystems, basis="DEF2-SVP", functional="PBE", kinetic="TF", aux="RI_J_WEIGEND",
                 accuracy=4, spherical=True, precision="FP64", log=None):
        self.ctx = ctx
        self.functional, self.kinetic, self.precision = functional, kinetic, precision
        self.say = log or (lambda *a: None)
        t0 = time.perf_counter()
        self.Zs = [np.ascontiguousarray(Z, dtype=np.int32) for Z, _ in subsystems]
        self.Xs = [np.asarray(X, dtype=np.float64).reshape(-1, 3) for _, X in subsystems]
        self.Z = np.concatenate(self.Zs)
        self.X = np.vstack(self.Xs)
        self.bases = [xc.BasisController.from_file(basis, Z, X, spherical=spherical) for Z, X in zip(self.Zs, self.Xs)]
        # the supersystem basis is the subsystem bases in subsystem order (atoms in that order)
        self.basis_sup = xc.BasisController.from_file(basis, self.Z, self.X, spherical=spherical)
        offs = np.cumsum([0] + [b.nao for b in self.bases])
        if offs[-1] != self.basis_sup.nao:
            raise xc.XCError(1, "FreezeAndThawTask: subsystem bases do not add up to the supersystem basis")
        self.slices = [slice(int(offs[i]), int(offs[i + 1])) for i in range(len(self.bases))]
        self.aux_sup = xc.BasisController.from_file(aux, self.Z, self.X, spherical=spherical)
        self.J_engine = xc.DensityFittedCoulomb(ctx, self.basis_sup, self.aux_sup)
        self.grid = xc.molecular_grid(self.Z, self.X, accuracy=accuracy, ctx=ctx)
        keep = precision != "FP64"
        self.bf = [xc.BasisFunctionOnGridController(ctx, b, self.grid, 128, 1e-9, 1, keep_fp32=keep)
                   for b in self.bases]
        self.S = [b.overlap() for b in self.bases]
        self.h = []
        for b in self.bases:
            T, V = SCF.one_electron(b, self.Z, self.X)  # attraction of every nucleus of the supersystem
            self.h.append(T + V)
        self.e_nuc = SCF.nuclear_repulsion(self.Z, self.X)
        self.nocc = []
        for Z in self.Zs:
            if int(Z.sum()) % 2:
                raise xc.XCError(1, "FreezeAndThawTask: restricted subsystems need an even electron count")
            self.nocc.append(int(Z.sum()) // 2)
        self.P = [SCF.atomic_density_guess(basis, Z, spherical) for Z in self.Zs]
        n = self.basis_sup.nao
        self._dP = [ctx.alloc(8 * b.nao ** 2) for b in self.bases]
        self._Psup = np.zeros((n, n))
        self.setup_s = time.perf_counter() - t0

    def _embed(self, sl, P):
        Ps = self._Psup
        Ps[:] = 0.0
        Ps[sl, sl] = P
        return Ps

    def coulomb_of(self, sl, P):
        """J[P] in the supersystem basis for a density matrix P of the subsystem at AO slice sl."""
        return self.J_engine.coulomb(self._embed(sl, P))

    def supersystem_density(self):
        n = self.basis_sup.nao
        P = np.zeros((n, n))
        for sl, Pi in zip(self.slices, self.P):
            P[sl, sl] = Pi
        return P

    def _cache_density(self, i):
        self.ctx.upload(self._dP[i], np.asfortranarray(self.P[i]))
        self.bf[i].update_density(self._dP[i], "FP64")

    def _embedded_build(self, i):
        """xc_build(P_act) of scf.perform: E and V of the active subsystem's own XC functional plus the
        non-additive XC and kinetic potentials against the cached environment densities."""
        env = [self.bf[j] for j in range(len(self.bf)) if j != i]
        own = xc.FuncPotential(self.bf[i], self.functional, self.precision)
        nadd = [xc.NAddFuncPotential(self.bf[i], env, f, "FP64") for f in (self.functional, self.kinetic)]

        def build(P):
            V = own.getMatrix(P)
            E = own.getEnergy()
            self.ctx.upload(self._dP[i], np.asfortranarray(P))
            self.bf[i].update_density(self._dP[i], "FP64")
            n = self.bases[i].nao
            dF = self.ctx.alloc(8 * n * n)
            try:
                for pot in nadd:
                    E += pot.build_cached(dF)
                    Vn = np.zeros((n, n), order="F")
                    self.ctx.download(Vn, dF)
                    V = V + Vn
            finally:
                self.ctx.free(dF)
            return E, V
        return build

    def energy(self):
        """Supersystem energy of the current subsystem densities."""
        Psup = self.supersystem_density()
        J = self.J_engine.coulomb(Psup)
        e = self.e_nuc + 0.5 * float(np.vdot(Psup, J))
        e_xc_sub = 0.0
        for i in range(len(self.bf)):
            e += float(np.vdot(self.P[i], self.h[i]))
            pot = xc.FuncPotential(self.bf[i], self.functional, "FP64")
            pot.getMatrix(self.P[i])
            e_xc_sub += pot.getEnergy()
        for i in range(len(self.bf)):
            self._cache_density(i)
        # E_xc^nadd and T^nadd of the full partition: the same for any choice of active subsystem
        env = self.bf[1:]
        n = self.bases[0].nao
        dF = self.ctx.alloc(8 * n * n)
        try:
            e_nadd = [xc.NAddFuncPotential(self.bf[0], env, f, "FP64").build_cached(dF)
                      for f in (self.functional, self.kinetic)]
        finally:
            self.ctx.free(dF)
        return e + e_xc_sub + sum(e_nadd), dict(e_xc_subsystems=e_xc_sub, e_xc_nadd=e_nadd[0], t_nadd=e_nadd[1])

    def run(self, max_cycles=50, conv_thresh=1e-6, scf_kwargs=None):
        scf_kwargs = dict(scf_kwargs or {})
        t0 = time.perf_counter()
        for i in range(len(self.bf)):
            self._cache_density(i)
        history, scf_cycles = [], 0
        converged = False
        for cycle in range(1, max_cycles + 1):
            changes = []
            for i in range(len(self.bf)):
                sl = self.slices[i]
                Penv = self.supersystem_density()
                Penv[sl, sl] = 0.0
                J_env = self.J_engine.coulomb(Penv)
                ints = _EmbeddedIntegrals(self, i, J_env)
                r = SCF.perform(ints, self.nocc[i], self._embedded_build(i), P0=self.P[i], **scf_kwargs)
                scf_cycles += r["cycles"]
                changes.append(float(np.linalg.norm(r["P"] - self.P[i])))
                self.P[i] = r["P"]
                self._cache_density(i)
            e, parts = self.energy()
            history.append(dict(cycle=cycle, energy=e, max_density_change=max(changes)))
            self.say("FaT cycle %2d  E = %.12f  max |dP| = %.2e" % (cycle, e, max(changes)))
            if max(changes) < conv_thresh:
                converged = True
                break
        e, parts = self.energy()
        return dict(energy=e, cycles=len(history), converged=converged, history=history, scf_cycles=scf_cycles,
                    wall_s=time.perf_counter() - t0, **parts)

    def close(self):
        for d in self._dP:
            self.ctx.free(d)
        self._dP = []
        for b in self.bf:
            b._release()
        self.bf = []
        self.J_engine._release()
