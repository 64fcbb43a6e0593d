/*
 * serenity_xc.h -- C ABI of the MI355X-native XC Fock build.
 *
 * The library replaces the hot path of Serenity's FuncPotential
 * (src/potentials/FuncPotential.cpp:73-111): AO values/gradients on grid
 * blocks, density and density gradient on the grid, the XC functional, and the
 * grid-to-matrix Fock assembly.  Every entry point returns an int status
 * (XC_OK = 0); on failure xc_last_error() returns a thread-local message.  Host
 * pointers are borrowed for the duration of the call; handles own their device
 * memory.  Matrices are n x n column-major doubles (Eigen's default layout),
 * grid points are 3 x N column-major doubles (x, y, z contiguous per point,
 * Eigen::Matrix3Xd), all lengths in bohr and energies in hartree.
 *
 * Threading: calls on one xc_ctx are serialised internally; different
 * contexts are independent.
 */
#ifndef SERENITY_XC_H
#define SERENITY_XC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  XC_OK = 0,
  XC_ERR_ARG = 1,      /* invalid argument (null handle, size mismatch, ...) */
  XC_ERR_DEVICE = 2,   /* HIP runtime error */
  XC_ERR_MEMORY = 3,   /* device footprint above the admitted budget */
  XC_ERR_STATE = 4,    /* call order / missing data */
  XC_ERR_INPUT = 5,    /* bad input data (parse error, duplicated atoms, ...) */
  XC_BUSY = 6          /* server: no free slot, or the controller has a build in flight (retry) */
};

/* Functionals available on the device.  LDA = Slater + VWN5, PBE = PBE
 * exchange + PBE correlation (PW92 LDA part), TF = Thomas-Fermi kinetic,
 * BP86 = Becke 88 exchange + Perdew 86 correlation (on PZ81).
 * Serenity: functional_definitions.dat:44,47,48,80.  XC_FUNC_PBE carries the
 * constants of XCFun (Serenity's default evaluator); XC_FUNC_PBE_LIBXC those of
 * LibXC (customFunc.impl = LIBXC), which differ in the exchange mu. */
enum { XC_FUNC_LDA = 0, XC_FUNC_PBE = 1, XC_FUNC_TF = 2, XC_FUNC_SLATER = 3, XC_FUNC_VWN5 = 4,
       XC_FUNC_PBEX = 5, XC_FUNC_PBEC = 6, XC_FUNC_PBE_LIBXC = 7, XC_FUNC_BP86 = 8 };

/* Precision of the two dense contractions (phi.P and phi^T.(w v phi)).  The
 * functional, the E_xc reduction and the F accumulation are FP64 in every mode.
 * XC_PREC_FP32 contracts with fl32(P) and adds the FP64 term tr(F (P - fl32(P)))
 * to E_xc (dE_xc/dP = F), which removes the first-order error of rounding P.
 * XC_PREC_MIXED is XC_PREC_FP32 plus an FP64 refinement of the density of the
 * blocks whose largest density after the FP32 pass exceeds 2 bohr^-3 (the
 * heavy-atom cores, about a tenth of the density work of a water cluster),
 * contracted again with FP64 tiles and fl32(P) in FP64: it removes the FP32
 * rounding bias of E_xc that grows with the number of heavy atoms.  It needs a
 * controller with both tile sets (keep_fp32 = 1) and is available for the
 * restricted single-system builds (xc_fock_build[_device], serving). */
enum { XC_PREC_FP64 = 0, XC_PREC_FP32 = 1, XC_PREC_MIXED = 2 };

/* Grid partition and radial quadrature (Serenity Options::GRID_TYPES / RADIAL_GRID_TYPES). */
enum { XC_GRID_BECKE = 0, XC_GRID_VORONOI = 1, XC_GRID_SSF = 2 };
enum { XC_RAD_BECKE = 0, XC_RAD_HANDY = 1, XC_RAD_AHLRICHS = 2, XC_RAD_KNOWLES = 3, XC_RAD_EQUI = 4 };

typedef struct xc_ctx xc_ctx;       /* one device + one stream */
typedef struct xc_basis xc_basis;   /* contracted shells, resident on the device */
typedef struct xc_grid xc_grid;     /* grid points and weights, resident on the device */
typedef struct xc_bfogc xc_bfogc;   /* basis functions on grid blocks: screening lists + resident AO tiles */
typedef struct xc_server xc_server; /* asynchronous builds on resident controllers (serving) */

/* ---- errors, version ---------------------------------------------------- */
const char* xc_last_error(void);
int xc_version(void);

/* ---- context ------------------------------------------------------------ */
/* Creates a context on HIP device `device` with its own (default-priority) stream. */
int xc_ctx_create(int device, xc_ctx** out);
/* Fails with XC_ERR_STATE while basis, grid or controller handles of the
 * context are alive: destroy those first. */
int xc_ctx_destroy(xc_ctx* ctx);
int xc_ctx_synchronize(xc_ctx* ctx);
/* The hipStream_t every kernel of this context is launched on. */
void* xc_ctx_stream(xc_ctx* ctx);
/* Device bytes currently held by handles of this context (in 2 MiB steps for blocks of
 * 1 MiB and more).  Released blocks stay with the context for reuse and return to the
 * device when an allocation fails and when the context is destroyed. */
int xc_ctx_memory(xc_ctx* ctx, int64_t* bytes_in_use);
/* Bit-for-bit repeatable Fock matrices (no counterpart in the reference, whose OpenMP
 * dynamic schedule also sums the blocks' contributions in a varying order:
 * src/data/grid/ScalarOperatorToMatrixAdder.cpp:66-110).  enable != 0:
 * every row reduce of the context adds each element's terms as exact 64-bit integer parts
 * (2^-32 and 2^-72 units) instead of FP64 additions in a varying order, so F is the same on
 * every run, packed or alone; the element is exact up to 2^-73 per term.  Costs ~0.14 ms of
 * the 7.8 ms mixed build of (H2O)64/def2-TZVP.  Valid for terms below 2^19 in magnitude in
 * columns of fewer than 2^23 terms whose absolute values sum to less than 2^30 (a Fock
 * element's terms are far smaller).  Fails with XC_ERR_STATE once a basis,
 * grid or controller of the context exists: set it right after xc_ctx_create. */
int xc_ctx_set_deterministic(xc_ctx* ctx, int enable);

/* ---- host-side helpers (no device work) ------------------------------- */
/* libint2-style renormalisation of one contraction, in place
 * (Serenity Shell ctor, src/basis/Shell.cpp:30-47). */
int xc_normalize_contraction(int l, int nprim, const double* alpha, double* coef);

/* Parses a Turbomole-format basis file for every atom of a geometry and
 * returns the shell list in Serenity's AO order (atoms in input order, shells
 * in file order; AtomCenteredBasisController.cpp:60-75,
 * BasisFunctionProvider.cpp:60-130).  Output arrays are allocated by the
 * library and released with xc_free_host.  coef is renormalised. */
int xc_load_basis(const char* path, int natoms, const int* Z, const double* coords_bohr, int pure,
                  int* nshell, int** shell_l, int** shell_nprim, int** shell_atom, double** shell_center,
                  double** alpha, double** coef);
void xc_free_host(void* p);

/* Molecular grid (Serenity GridFactory::produce + AtomCenteredGridController
 * sort).  points: 3 x N (xyz per point), weights: N; released with xc_free_host. */
int xc_molecular_grid(int natoms, const int* Z, const double* coords_bohr, int grid_type, int radial_type,
                      int accuracy, int smoothing, double weight_threshold, int sort, int64_t* npoints,
                      double** points, double** weights, int64_t* npoints_before_cut);

/* Hilbert-curve ordering of npoints points (3 x N, xyz per point) and their
 * weights, in place (HilbertRTreeSorting::sort, src/grid/HilbertRTreeSorting.cpp
 * :27-216; equal keys keep their input order). */
int xc_hilbert_sort(int64_t npoints, double* points, double* weights);
/* xc_molecular_grid built on the context's device: the host makes each
 * element's reference atom grid once; the candidates, the fuzzy-cell partition
 * (the O(N natoms^2) part), the weight cut and the Hilbert order run on the
 * device with the host build's arithmetic (shared functions, a stable sort on
 * the same keys), so the result is the same grid.  No counterpart in the
 * reference, whose grid is built on the CPU (GridFactory.cpp:51-300,
 * HilbertRTreeSorting.cpp:27-216). */
int xc_molecular_grid_device(xc_ctx* ctx, int natoms, const int* Z, const double* coords, int grid_type,
                             int radial_type, int accuracy, int smoothing, double weight_threshold, int sort,
                             int64_t* npoints, double** points, double** weights, int64_t* npoints_before_cut);

/* Lebedev-Laikov rule with `npoints` points: x,y,z,w arrays of length npoints
 * (weights sum to 1). */
int xc_lebedev(int npoints, double* x, double* y, double* z, double* w);

/* Analytic overlap matrix S (n x n column-major) of a shell list. */
int xc_overlap(int nshell, const int* shell_l, const int* shell_pure, const int* shell_nprim,
               const double* shell_center, const double* alpha, const double* coef, double* S);

/* ---- integrals of the SCF driver (host, OpenMP) --------------------------
 * The full SCF around the XC build needs the one-electron Hamiltonian and the
 * Coulomb matrix, which Serenity takes from libint2 through
 * OneElectronIntegralController (src/integrals/OneElectronIntegralController.cpp)
 * and ERIPotential / RI_J_IntegralController
 * (src/integrals/RI_J_IntegralController.cpp:41-121).  Shells as in
 * xc_basis_create, l <= 4.  AO pairs are packed as
 * pair(mu, nu) = mu (mu + 1) / 2 + nu, mu >= nu; npair = n (n + 1) / 2. */
/* Boys function F_m(T), 0 <= m <= 17. */
int xc_boys(int m, double T, double* F);
/* Kinetic energy T and nuclear attraction V of point nuclei (charges Z, coords
 * 3 x natoms in bohr), n x n column-major. */
int xc_one_electron(int nshell, const int* shell_l, const int* shell_pure, const int* shell_nprim,
                    const double* shell_center, const double* alpha, const double* coef, int natoms, const int* Z,
                    const double* coords, double* T, double* V);
/* Four-centre Coulomb integrals (mu nu|la si): eri is npair x npair; quartets
 * whose Schwarz bound is below schwarz_threshold are zero.  At most 2^31
 * entries (XC_ERR_INPUT beyond). */
int xc_eri_packed(int nshell, const int* shell_l, const int* shell_pure, const int* shell_nprim,
                  const double* shell_center, const double* alpha, const double* coef, double schwarz_threshold,
                  double* eri);
/* The metric (P|Q) of an auxiliary shell list alone (naux x naux). */
int xc_ri_metric(int naux_shell, const int* aux_l, const int* aux_pure, const int* aux_nprim,
                 const double* aux_center, const double* aux_alpha, const double* aux_coef, double* M);

/* ---- density-fitted Coulomb matrix on the GPU (J engine) -----------------
 * The Coulomb part of Serenity's DFT Fock matrix with densFitJ = RI
 * (RI_J_IntegralController.cpp:41-121 over libint2 3-centre integrals) without
 * storing 3-centre integrals: every significant primitive pair of the orbital
 * basis (|K| (pi/p)^1.5 >= threshold) is contracted with the density into
 * Hermite form and meets the auxiliary functions through Hermite Coulomb
 * integrals.  Orbital shells up to f, auxiliary up to g.
 *   xc_coulomb_rhs:    b_P = sum_mu,nu P_mu,nu (mu nu|P)     (P host n x n, b host naux)
 *   xc_coulomb_matrix: J_mu,nu = sum_P c_P (mu nu|P)         (c host naux, J host n x n)
 * with c = M^-1 b between the two (M from xc_ri_metric). */
typedef struct xc_coulomb xc_coulomb;
int xc_coulomb_create(xc_basis* basis, xc_basis* aux, double threshold, xc_coulomb** out);
int xc_coulomb_info(const xc_coulomb* jc, int64_t* npairs, int64_t* hermite_terms);
int xc_coulomb_rhs(xc_coulomb* jc, const double* P, double* b);
int xc_coulomb_matrix(xc_coulomb* jc, const double* c, double* J);
int xc_coulomb_destroy(xc_coulomb* jc);

/* Density-fitted Coulomb integrals over an auxiliary shell list: B3 is
 * naux x npair (row P holds (P|mu nu)), M the naux x naux metric (P|Q). */
int xc_ri_j_integrals(int nshell, const int* shell_l, const int* shell_pure, const int* shell_nprim,
                      const double* shell_center, const double* alpha, const double* coef, int naux_shell,
                      const int* aux_l, const int* aux_pure, const int* aux_nprim, const double* aux_center,
                      const double* aux_alpha, const double* aux_coef, double schwarz_threshold, double* B3,
                      double* M);

/* ---- resident objects --------------------------------------------------- */
/* Shells in AO order.  shell_center: 3 x nshell; alpha/coef: concatenated
 * primitives (coef already renormalised).  Replaces BasisController's shell
 * list as seen by BasisFunctionOnGridController (BasisController.h:157-193). */
int xc_basis_create(xc_ctx* ctx, int nshell, const int* shell_l, const int* shell_pure, const int* shell_nprim,
                    const double* shell_center, const double* alpha, const double* coef, xc_basis** out);
int xc_basis_destroy(xc_basis* basis);
int xc_basis_nao(const xc_basis* basis, int* nao);

/* Grid of npoints points; replaces GridController::getGridPoints/getWeights
 * (src/grid/GridController.cpp:30-50). */
int xc_grid_create(xc_ctx* ctx, int64_t npoints, const double* points, const double* weights, xc_grid** out);
int xc_grid_destroy(xc_grid* grid);

/* BasisFunctionOnGridController(basis, grid, maxBlockSize, radialThreshold,
 * highestDerivative) (BasisFunctionOnGridController.h:105-107).  Runs the
 * block screening and evaluates the AO values (gradients if
 * highest_derivative >= 1, second derivatives if 2) of every block into
 * resident HBM tiles.  highest_derivative is 0, 1 or 2 (the Fock build needs
 * at most 1; the GGA nuclear gradient needs 2, as getGeomGradients sets it,
 * FuncPotential.cpp:123-128).
 * keep_fp32: 0 = FP64 tiles only, 1 = FP64 and FP32 tiles (XC_PREC_FP32 builds),
 * 2 = FP32 tiles only (XC_PREC_FP32 builds; a third of the footprint, for
 * serving many resident systems). */
int xc_bfogc_create(xc_basis* basis, xc_grid* grid, int max_block_size, double radial_threshold,
                    int highest_derivative, int keep_fp32, xc_bfogc** out);
int xc_bfogc_destroy(xc_bfogc* bfogc);
/* getNBlocks / getNGridPoints / getNBasisFunctions / getFirstIndexOfBlock. */
int xc_bfogc_info(const xc_bfogc* bfogc, int* nblocks, int64_t* npoints, int* nao, int* max_block_size);
/* Sum over blocks of the number of non-negligible AOs, and of B_b * n_b^2
 * (the algorithmic work of both contractions is 4 * sum B_b n_b^2 flops). */
int xc_bfogc_stats(const xc_bfogc* bfogc, int64_t* sum_nb, double* sum_B_nb2, int64_t* tile_bytes);
/* getBlockOnGridData(block) (BasisFunctionOnGridController.cpp:122-131):
 * copies the block's dense B x nao column-major value matrix (and gradient
 * components if requested and available; pass NULL to skip) and the
 * negligible flags (1 = screened out) to the host. */
int xc_bfogc_block_data(const xc_bfogc* bfogc, int block, double* values, double* grad_x, double* grad_y,
                        double* grad_z, int* negligible, int* block_size);
/* Second derivatives of the block's AOs (blockOnGridData->secondDerivativeValues,
 * BasisFunctionOnGridController.h:73-79): hess is 6 x n x B (planes xx, xy, xz,
 * yy, yz, zz; each n x B with B the block size), zero for negligible AOs.
 * Needs highest_derivative 2. */
int xc_bfogc_block_hessian(const xc_bfogc* bfogc, int block, double* hess);

/* ---- the path ----------------------------------------------------------- */
/* DensityOnGridCalculator::calcDensityAndGradientOnGrid(P, rho, grad)
 * (DensityOnGridCalculator.cpp:54-65, MatrixOperatorToGridTransformer.cpp
 * :37-198).  P: host n x n.  rho: host N; grad: host 3 x N SoA (x[N], y[N],
 * z[N]) or NULL for values only.  nonnegligible_blocks: host int[nblocks]
 * (1 where the block has at least one non-negligible AO) or NULL. */
int xc_density(xc_bfogc* bfogc, const double* P, double* rho, double* grad, int* nonnegligible_blocks);

/* FunctionalLibrary::calcData(GRADIENTS, functional, density)
 * (FunctionalLibrary.cpp:38-71, XCFun.cpp:38-158) on host arrays of npoints:
 * eps = energy density per volume, v = d eps/d rho, g = d eps/d grad rho
 * (3 x N SoA, may be NULL for LDA-type functionals).  grad may be NULL for
 * LDA-type functionals.  Blocks of block_size points whose sum |rho| is below
 * block_size * 1e-12 are skipped (outputs zero).  energy = sum_p w_p eps_p. */
int xc_functional(xc_ctx* ctx, int functional, int64_t npoints, int block_size, const double* rho,
                  const double* grad, const double* weights, double* eps, double* v, double* g, double* energy);

/* No counterpart in the reference: xc_functional on device arrays that allocates
 * nothing, for callers that cannot call hipMalloc (code under HIP graph capture,
 * serving loops and harnesses that forbid allocation).  d_rho (N), d_grad (3 x N
 * SoA; may be NULL for LDA-type functionals), d_weights (N) in; d_eps, d_v (N)
 * and d_g (3 x N, may be NULL; zero-filled for LDA-type functionals) out, all
 * device buffers of the context's device; energy (host, may be NULL: then the
 * call only enqueues) = sum_p w_p eps_p.  Results equal xc_functional's.  The
 * per-block energies live in one static device workspace per device of at most
 * 131072 blocks (1 MiB), so at most 131072 * block_size points (XC_ERR_ARG
 * beyond); calls that overlap on one device share that workspace and race, so
 * serialise them (calls on one context are). */
int xc_functional_device_noalloc(xc_ctx* ctx, int functional, int64_t npoints, int block_size, const double* d_rho,
                                 const double* d_grad, const double* d_weights, double* d_eps, double* d_v,
                                 double* d_g, double* energy);

/* ScalarOperatorToMatrixAdder::addScalarOperatorToMatrix(F, v[, g])
 * (ScalarOperatorToMatrixAdder.cpp:51-113, addBlock :179-303): F += integral
 * of the scalar potential v (and gradient potential g, 3 x N SoA, or NULL).
 * F: host n x n, updated in place. */
int xc_fock_add(xc_bfogc* bfogc, double block_ave_threshold, const double* v, const double* g, double* F);

/* FuncPotential::getMatrix + getEnergy (FuncPotential.cpp:66-111) fused:
 * F = V_xc[P] (overwritten), energy = E_xc[P].  Host arrays. */
int xc_fock_build(xc_bfogc* bfogc, int functional, int precision, double block_ave_threshold, const double* P,
                  double* F, double* energy);

/* Same on device-resident buffers (P, F: device n x n; energy: host),
 * enqueued on the context stream; synchronises before returning the energy
 * unless energy is NULL (then the value stays in the context and can be read
 * with xc_last_energy after xc_ctx_synchronize). */
int xc_fock_build_device(xc_bfogc* bfogc, int functional, int precision, double block_ave_threshold,
                         const double* dP, double* dF, double* energy);

/* Packed builds: FuncPotential::getMatrix + getEnergy (FuncPotential.cpp:73-111) of
 * nsys controllers at once, for callers with many small systems (serving, embedding
 * subsystems).  One functional and precision for all; P_i, F_i device n_i x n_i
 * matrices, energies[i] host.  The work of every system shares one launch per stage
 * and work class (merged block and work lists with a per-system table); each F_i and
 * E_xc,i equal those of xc_fock_build_device on the same controller.  The controllers
 * must be distinct, of one context, with 128-point blocks.  Synchronises before
 * returning. */
int xc_fock_build_packed(int nsys, xc_bfogc* const* bfogcs, int functional, int precision,
                         double block_ave_threshold, const double* const* dP, double* const* dF, double* energies);

/* FuncPotential<UNRESTRICTED>::getMatrix / getEnergy (FuncPotential.cpp:73-111 with
 * SCFMode = UNRESTRICTED): spin density matrices P_a, P_b in, the alpha and beta
 * XC matrices F_a, F_b (overwritten) and E_xc out.  The functional takes the spin
 * densities (XCFun's XC_A_B_AX_AY_AZ_BX_BY_BZ variables); a block is skipped only
 * when both spin densities are negligible, and each spin's Fock pass has its own
 * block-average skip (ScalarOperatorToMatrixAdder.cpp:200-206, 251-268). */
int xc_fock_build_unrestricted(xc_bfogc* bfogc, int functional, int precision, double block_ave_threshold,
                               const double* P_alpha, const double* P_beta, double* F_alpha, double* F_beta,
                               double* energy);
/* Same on device buffers (xc_device_alloc). */
int xc_fock_build_unrestricted_device(xc_bfogc* bfogc, int functional, int precision, double block_ave_threshold,
                                      const double* dP_alpha, const double* dP_beta, double* dF_alpha,
                                      double* dF_beta, double* energy);

/* FuncPotential::getGeomGradients (FuncPotential.cpp:113-239): the XC part of
 * the nuclear gradient of the restricted density P (host, n x n column-major)
 * at a fixed grid (no weight derivatives, as the reference), written to grad
 * as natoms x 3 column-major (Eigen's nAtoms x 3 layout).  ao_atom[nu] is the
 * atom of AO nu (BasisController::getAtomIndicesOfBasis).  The controller
 * needs highest_derivative 1 (LDA) or 2 (GGA) and FP64 tiles; energy gets E_xc
 * of P (may be NULL). */
int xc_geom_gradient(xc_bfogc* bfogc, int functional, const double* P, int natoms, const int* ao_atom,
                     double* grad, double* energy);
/* The unrestricted form (FuncPotential<UNRESTRICTED>::getGeomGradients): the same sum
 * per spin, 2 P^s_mu,nu with the spin's potential w v_s, w g_s. */
int xc_geom_gradient_unrestricted(xc_bfogc* bfogc, int functional, const double* P_alpha, const double* P_beta,
                                  int natoms, const int* ao_atom, double* grad, double* energy);
int xc_last_energy(xc_ctx* ctx, double* energy);

/* NAddFuncPotential::getMatrix + getEnergy, restricted (src/potentials/
 * NAddFuncPotential.cpp:192-301): the non-additive potential of the active
 * subsystem in an embedding.  rho_sup = rho_act + sum of the environment
 * densities on the shared supersystem grid (SupersystemDensityOnGridController
 * .cpp:157-190); F = integral of v[rho_sup] - v[rho_act] (and the gradient
 * terms) in the active basis, overwritten; energy = E[rho_sup] - E[rho_act]
 * - sum_env E[rho_env] (NAddEnergyHelper::getEnergy, NAddFuncPotential.cpp
 * :507-514).  Every controller must be built on the same grid handle.
 * P_active: n_act x n_act; P_env[i]: n_i x n_i; F: n_act x n_act (host). */
int xc_nadd_fock_build(xc_bfogc* active, int nenv, xc_bfogc* const* env, int functional, int precision,
                       double block_ave_threshold, const double* P_active, const double* const* P_env, double* F,
                       double* energy);
/* NAddFuncPotential<UNRESTRICTED>::getMatrix / getEnergy (NAddFuncPotential.cpp:192-301
 * with SCFMode = UNRESTRICTED): spin density matrices of the active (Pa_active,
 * Pb_active) and environment subsystems in; F_a, F_b = the integrals of
 * v_s[rho_sup] - v_s[rho_act] in the active basis and the energy
 * E[sup] - E[act] - sum_env E[env] of the spin-polarised functional out (host
 * buffers).  Same grid and controller rules as xc_nadd_fock_build. */
int xc_nadd_fock_build_unrestricted(xc_bfogc* active, int nenv, xc_bfogc* const* env, int functional, int precision,
                                    double block_ave_threshold, const double* Pa_active, const double* Pb_active,
                                    const double* const* Pa_env, const double* const* Pb_env, double* Fa,
                                    double* Fb, double* energy);
/* DensityMatrixDensityOnGridController::calculateData (DensityMatrixDensityOnGrid
 * Controller.cpp:134-170), explicit form of its lazy cache: computes rho (and
 * grad rho if with_gradient) of the device density matrix dP and keeps them
 * resident in the controller until the next update.  Clears the controller's
 * cached energies. */
int xc_bfogc_update_density(xc_bfogc* bfogc, int precision, const double* dP, int with_gradient);
/* NAddFuncPotential::getMatrix on the cached densities of every subsystem (the
 * reference's DensityOnGridControllers and NAddEnergyHelper keep theirs until
 * the density matrix changes, NAddFuncPotential.cpp:507-514).  E[rho_env] is
 * computed once per environment controller, functional and density update.
 * Results equal xc_nadd_fock_build_device on the same densities. */
int xc_nadd_fock_build_cached(xc_bfogc* active, int nenv, xc_bfogc* const* env, int functional, int precision,
                              double block_ave_threshold, double* dF, double* energy);
/* Same on device-resident P and F (energy as for xc_fock_build_device). */
int xc_nadd_fock_build_device(xc_bfogc* active, int nenv, xc_bfogc* const* env, int functional, int precision,
                              double block_ave_threshold, const double* dP_active, const double* const* dP_env,
                              double* dF, double* energy);

/* ---- serving ------------------------------------------------------------ */
/* Asynchronous XC Fock builds on resident controllers (no counterpart in the
 * reference, whose builds are synchronous calls of FuncPotential::getMatrix):
 * nslots builds overlap on the device, each on its own stream with pinned
 * staging of P and F.  With use_graphs, the first build of a controller for a
 * (functional, precision, threshold) is captured into a HIP graph and later
 * ones replay it.  A controller may have one build in flight at a time. */
int xc_server_create(xc_ctx* ctx, int nslots, int use_graphs, xc_server** out);
int xc_server_destroy(xc_server* server);
/* Starts F = V_xc[P], energy = E_xc[P] (host P, F and energy, written when the
 * ticket completes).  Returns XC_BUSY without side effects when no slot is free
 * or the controller has a build in flight. */
int xc_server_submit(xc_server* server, xc_bfogc* bfogc, int functional, int precision,
                     double block_ave_threshold, const double* P, double* F, double* energy, int64_t* ticket);
/* done = 1 once the build has finished and F / energy are written. */
int xc_server_poll(xc_server* server, int64_t ticket, int* done);
int xc_server_wait(xc_server* server, int64_t ticket);
int xc_server_stats(const xc_server* server, int* busy_slots, int64_t* submitted, int64_t* graph_launches);

/* Stage timing with HIP events on the context stream.  While enabled, every
 * build records an event between its stages; xc_profile_collect waits for
 * them, returns the summed milliseconds of each stage over the builds since
 * the last collect (index 0 P to FP32, 1 density, 2 functional, 3 fock (the
 * blocks' sub-tiles), 4 reduce of the sub-tiles into F's rows, 5 finalize +
 * energy sum) and resets. */
enum { XC_STAGE_CONVERT = 0, XC_STAGE_DENSITY = 1, XC_STAGE_FUNCTIONAL = 2, XC_STAGE_FOCK = 3, XC_STAGE_REDUCE = 4,
       XC_STAGE_FINALIZE = 5, XC_NSTAGES = 6 };
int xc_profile(xc_ctx* ctx, int enable);
int xc_profile_collect(xc_ctx* ctx, double* stage_ms, int* nbuilds);

/* Device buffer helpers so that callers without their own allocator (ctypes,
 * tests, benchmarks) can hold resident P and F. */
int xc_device_alloc(xc_ctx* ctx, size_t bytes, void** dptr);
int xc_device_free(xc_ctx* ctx, void* dptr);
int xc_memcpy_h2d(xc_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel of the library (reads and
 * writes `bytes` once; used to calibrate HBM byte counters). */
int xc_device_copy(xc_ctx* ctx, void* dst, const void* src, size_t bytes);
int xc_memcpy_d2h(xc_ctx* ctx, void* dst, const void* src, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* SERENITY_XC_H */
