/*
 * serenity_xc.hpp -- C++ host classes of the XC Fock build, shaped like
 * Serenity's grid-path classes, over the C ABI in serenity_xc.h.
 *
 * Names, constructor arguments and method names follow the reference so that
 * Serenity code (or a test written like Serenity's gtest cases) reads alike:
 *
 *   serenity_amd::BasisFunctionOnGridController   BasisFunctionOnGridController.h:105-131
 *   serenity_amd::DensityOnGridCalculator         DensityOnGridCalculator.h:45-134
 *   serenity_amd::DensityMatrixController         DensityMatrixController.h (setDensityMatrix,
 *                                                 getDensityMatrix, notifying its dependents)
 *   serenity_amd::DensityMatrixDensityOnGridController
 *                                                 DensityMatrixDensityOnGridController.h:45-120
 *   serenity_amd::ScalarOperatorToMatrixAdder     ScalarOperatorToMatrixAdder.h:48-196
 *   serenity_amd::FunctionalLibrary               FunctionalLibrary.h:51-64
 *   serenity_amd::SystemController                the parts of SystemController.h the XC path reads
 *                                                 (geometry, basis, settings.dft.functional)
 *   serenity_amd::FuncPotential                   FuncPotential.h:48-145
 *
 * The reference's forms are there (FuncPotential(system, dMat, grid, functional) with
 * getMatrix(), getEnergy(P) and notify(); FunctionalLibrary::calcData(type, functional,
 * densityOnGridController, order)); the explicit forms that take P as an argument are kept beside
 * them for callers without a density-matrix controller.
 *
 * Matrices are column-major (Eigen's default): element (i, j) of an n x n
 * matrix at data[i + n * j].  Grid points are 3 x N column-major.  Every
 * failure of the library throws serenity_amd::SerenityError (the reference's
 * SerenityError counterpart) with the library's message.  Header only; link
 * against libserenity_xc.so.
 */
#ifndef SERENITY_XC_HPP
#define SERENITY_XC_HPP

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "serenity_xc.h"

namespace serenity_amd {

struct SerenityError : std::runtime_error {
  int code;
  SerenityError(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

inline void check(int rc) {
  if (rc != XC_OK) throw SerenityError(rc, xc_last_error());
}

/* Column-major dense matrix (the shape Serenity's Eigen matrices have on this path). */
struct Matrix {
  int rows = 0, cols = 0;
  std::vector<double> data;
  Matrix() = default;
  Matrix(int r, int c) : rows(r), cols(c), data((size_t)r * c, 0.0) {}
  double& operator()(int i, int j) { return data[(size_t)i + (size_t)rows * j]; }
  double operator()(int i, int j) const { return data[(size_t)i + (size_t)rows * j]; }
  void setZero() { std::fill(data.begin(), data.end(), 0.0); }
};

enum class Functional {
  LDA = XC_FUNC_LDA,
  PBE = XC_FUNC_PBE,
  TF = XC_FUNC_TF,
  PBE_LIBXC = XC_FUNC_PBE_LIBXC,
  BP86 = XC_FUNC_BP86
};
inline bool isGGA(Functional f) { return f == Functional::PBE || f == Functional::PBE_LIBXC || f == Functional::BP86; }
enum class Precision { FP64 = XC_PREC_FP64, FP32 = XC_PREC_FP32, MIXED = XC_PREC_MIXED };
/* PartialDerivatives.h:931: derivatives with respect to the gradient invariant sigma = |grad rho|^2,
 * to the gradient components, or the potential alone. */
enum class FUNCTIONAL_DATA_TYPE { GRADIENT_INVARIANTS = 0, GRADIENTS = 1, POTENTIAL = 2 };

/* ObjectSensitiveClass / NotifyingClass: lazily evaluated objects register with what they depend on
 * and are told when it changes. */
class ObjectSensitiveClass {
 public:
  virtual ~ObjectSensitiveClass() = default;
  virtual void notify() = 0;
};

class NotifyingClass {
 public:
  void addSensitiveObject(std::weak_ptr<ObjectSensitiveClass> o) { _objects.push_back(std::move(o)); }
  void notifyObjects() {
    std::vector<std::weak_ptr<ObjectSensitiveClass>> alive;
    for (auto& w : _objects)
      if (auto o = w.lock()) {
        o->notify();
        alive.push_back(w);
      }
    _objects.swap(alive);
  }

 private:
  std::vector<std::weak_ptr<ObjectSensitiveClass>> _objects;
};

/* One HIP device and its stream. */
class Context {
 public:
  explicit Context(int device = 0) { check(xc_ctx_create(device, &_ctx)); }
  ~Context() { xc_ctx_destroy(_ctx); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  xc_ctx* get() const { return _ctx; }
  void synchronize() { check(xc_ctx_synchronize(_ctx)); }

 private:
  xc_ctx* _ctx = nullptr;
};

/* Contracted shells in AO order (Serenity's BasisController as seen by the grid path). */
class BasisController {
 public:
  struct Shells {
    std::vector<int> l, pure, nprim;
    std::vector<double> center, alpha, coef;  // center 3 x nshell; coef renormalised
  };
  BasisController(std::shared_ptr<Context> ctx, const Shells& s) : _ctx(std::move(ctx)), _shells(s) {
    check(xc_basis_create(_ctx->get(), (int)s.l.size(), s.l.data(), s.pure.data(), s.nprim.data(), s.center.data(),
                          s.alpha.data(), s.coef.data(), &_basis));
    check(xc_basis_nao(_basis, &_nao));
  }
  /* Loads a Turbomole-format basis file for a geometry (Z, coordinates in bohr). */
  static std::shared_ptr<BasisController> fromFile(std::shared_ptr<Context> ctx, const std::string& path,
                                                   const std::vector<int>& Z, const std::vector<double>& xyz,
                                                   bool spherical) {
    int ns = 0;
    int *l, *np, *atom;
    double *c, *a, *k;
    check(xc_load_basis(path.c_str(), (int)Z.size(), Z.data(), xyz.data(), spherical ? 1 : 0, &ns, &l, &np, &atom, &c,
                        &a, &k));
    Shells s;
    int nprim = 0;
    for (int i = 0; i < ns; ++i) nprim += np[i];
    s.l.assign(l, l + ns);
    s.nprim.assign(np, np + ns);
    s.pure.assign(ns, spherical ? 1 : 0);
    s.center.assign(c, c + 3 * ns);
    s.alpha.assign(a, a + nprim);
    s.coef.assign(k, k + nprim);
    for (void* p : {(void*)l, (void*)np, (void*)atom, (void*)c, (void*)a, (void*)k}) xc_free_host(p);
    return std::make_shared<BasisController>(std::move(ctx), s);
  }
  ~BasisController() { xc_basis_destroy(_basis); }
  BasisController(const BasisController&) = delete;
  BasisController& operator=(const BasisController&) = delete;
  unsigned int getNBasisFunctions() const { return (unsigned int)_nao; }
  const Shells& getShells() const { return _shells; }
  xc_basis* get() const { return _basis; }

 private:
  std::shared_ptr<Context> _ctx;
  Shells _shells;
  xc_basis* _basis = nullptr;
  int _nao = 0;
};

/* Grid points (3 x N) and weights (GridController::getGridPoints / getWeights). */
class GridController {
 public:
  GridController(std::shared_ptr<Context> ctx, std::vector<double> points, std::vector<double> weights)
      : _ctx(std::move(ctx)), _points(std::move(points)), _weights(std::move(weights)) {
    if (_points.size() != 3 * _weights.size())
      throw SerenityError(XC_ERR_ARG, "GridController: points and weights differ in length");
    check(xc_grid_create(_ctx->get(), (int64_t)_weights.size(), _points.data(), _weights.data(), &_grid));
  }
  /* Atom-centred molecular grid (GridFactory::produce + Hilbert sort). */
  static std::shared_ptr<GridController> molecular(std::shared_ptr<Context> ctx, const std::vector<int>& Z,
                                                   const std::vector<double>& xyz, int gridType = XC_GRID_SSF,
                                                   int accuracy = 4, double weightThreshold = 1e-14,
                                                   int smoothing = 3) {
    int64_t n = 0, nb = 0;
    double *p, *w;
    check(xc_molecular_grid((int)Z.size(), Z.data(), xyz.data(), gridType, XC_RAD_AHLRICHS, accuracy, smoothing,
                            weightThreshold, 1, &n, &p, &w, &nb));
    std::vector<double> pts(p, p + 3 * n), wts(w, w + n);
    xc_free_host(p);
    xc_free_host(w);
    return std::make_shared<GridController>(std::move(ctx), std::move(pts), std::move(wts));
  }
  ~GridController() { xc_grid_destroy(_grid); }
  GridController(const GridController&) = delete;
  GridController& operator=(const GridController&) = delete;
  unsigned int getNGridPoints() const { return (unsigned int)_weights.size(); }
  const std::vector<double>& getGridPoints() const { return _points; }
  const std::vector<double>& getWeights() const { return _weights; }
  xc_grid* get() const { return _grid; }

 private:
  std::shared_ptr<Context> _ctx;
  std::vector<double> _points, _weights;
  xc_grid* _grid = nullptr;
};

/* BasisFunctionOnGridController(basis, grid, maxBlockSize, radialThreshold, highestDerivative). */
class BasisFunctionOnGridController {
 public:
  struct BasisFunctionBlockOnGridData {
    Matrix functionValues;                     // B x n
    std::vector<Matrix> derivativeValues;      // x, y, z (B x n each) when highestDerivative >= 1
    std::vector<int> negligible;               // n flags
  };
  BasisFunctionOnGridController(std::shared_ptr<BasisController> basis, std::shared_ptr<GridController> grid,
                                unsigned int maxBlockSize, double radialThreshold, unsigned int highestDerivative,
                                bool keepFP32 = false)
      : _basis(std::move(basis)), _grid(std::move(grid)), _highestDerivative(highestDerivative) {
    if (highestDerivative > 2)
      throw SerenityError(XC_ERR_ARG, "Third derivatives of basis functions on the grid are not implemented!");
    check(xc_bfogc_create(_basis->get(), _grid->get(), (int)maxBlockSize, radialThreshold, (int)highestDerivative,
                          keepFP32 ? 1 : 0, &_bf));
    int64_t np = 0;
    check(xc_bfogc_info(_bf, &_nBlocks, &np, &_nao, &_maxBlockSize));
  }
  ~BasisFunctionOnGridController() { xc_bfogc_destroy(_bf); }
  BasisFunctionOnGridController(const BasisFunctionOnGridController&) = delete;
  BasisFunctionOnGridController& operator=(const BasisFunctionOnGridController&) = delete;
  std::shared_ptr<BasisFunctionBlockOnGridData> getBlockOnGridData(unsigned int blockIndex) const {
    if ((int)blockIndex >= _nBlocks) throw SerenityError(XC_ERR_ARG, "block index out of range");
    const unsigned int first = getFirstIndexOfBlock(blockIndex);
    const int B = (int)std::min<unsigned int>(_maxBlockSize, getNGridPoints() - first);
    auto d = std::make_shared<BasisFunctionBlockOnGridData>();
    d->functionValues = Matrix(B, _nao);
    d->negligible.assign(_nao, 0);
    double* g[3] = {nullptr, nullptr, nullptr};
    if (_highestDerivative >= 1) {
      d->derivativeValues.assign(3, Matrix(B, _nao));
      for (int c = 0; c < 3; ++c) g[c] = d->derivativeValues[c].data.data();
    }
    int bs = 0;
    check(xc_bfogc_block_data(_bf, (int)blockIndex, d->functionValues.data.data(), g[0], g[1], g[2],
                              d->negligible.data(), &bs));
    return d;
  }
  unsigned int getNBlocks() const { return (unsigned int)_nBlocks; }
  unsigned int getFirstIndexOfBlock(unsigned int blockIndex) const { return blockIndex * (unsigned int)_maxBlockSize; }
  unsigned int getNGridPoints() const { return _grid->getNGridPoints(); }
  unsigned int getNBasisFunctions() const { return (unsigned int)_nao; }
  unsigned int getMaxBlockSize() const { return (unsigned int)_maxBlockSize; }
  unsigned int getHighestDerivative() const { return _highestDerivative; }
  std::shared_ptr<BasisController> getBasisController() const { return _basis; }
  std::shared_ptr<GridController> getGridController() const { return _grid; }
  xc_bfogc* get() const { return _bf; }

 private:
  std::shared_ptr<BasisController> _basis;
  std::shared_ptr<GridController> _grid;
  xc_bfogc* _bf = nullptr;
  int _nBlocks = 0, _nao = 0, _maxBlockSize = 0;
  unsigned int _highestDerivative;
};

/* DensityOnGridCalculator(bfogc, blockAverageThreshold): rho and grad rho of a restricted P. */
class DensityOnGridCalculator {
 public:
  DensityOnGridCalculator(std::shared_ptr<BasisFunctionOnGridController> bf, double blockAverageThreshold)
      : _bf(std::move(bf)), _blockAverageThreshold(blockAverageThreshold) {}
  std::vector<double> calcDensityOnGrid(const Matrix& P) {
    std::vector<double> rho(_bf->getNGridPoints());
    run(P, rho, nullptr);
    return rho;
  }
  /* rho (N) and gradient (3 x N, x[N] y[N] z[N]) are resized and overwritten. */
  void calcDensityAndGradientOnGrid(const Matrix& P, std::vector<double>& rho, std::vector<double>& grad) {
    rho.assign(_bf->getNGridPoints(), 0.0);
    grad.assign(3 * (size_t)_bf->getNGridPoints(), 0.0);
    run(P, rho, &grad);
  }
  const std::vector<int>& getNonNegligibleBlocks() const { return _nonNegligible; }
  std::shared_ptr<BasisFunctionOnGridController> getBasisFunctionOnGridController() const { return _bf; }

 private:
  void run(const Matrix& P, std::vector<double>& rho, std::vector<double>* grad) {
    const int n = (int)_bf->getNBasisFunctions();
    if (P.rows != n || P.cols != n)
      throw SerenityError(XC_ERR_ARG, "DensityOnGridCalculator: density matrix and basis differ in size");
    _nonNegligible.assign(_bf->getNBlocks(), 0);
    check(xc_density(_bf->get(), P.data.data(), rho.data(), grad ? grad->data() : nullptr, _nonNegligible.data()));
  }
  std::shared_ptr<BasisFunctionOnGridController> _bf;
  double _blockAverageThreshold;
  std::vector<int> _nonNegligible;
};

/* DensityMatrixController (restricted): holds the density matrix and tells its dependents when it
 * is replaced (setDensityMatrix -> notifyObjects). */
class DensityMatrixController : public NotifyingClass {
 public:
  explicit DensityMatrixController(Matrix P) : _P(std::move(P)) {}
  const Matrix& getDensityMatrix() const { return _P; }
  void setDensityMatrix(Matrix P) {
    _P = std::move(P);
    notifyObjects();
  }

 private:
  Matrix _P;
};

/* DensityMatrixDensityOnGridController: rho and grad rho of the controller's density matrix on the
 * grid, computed on first use and kept until the density matrix changes. */
class DensityMatrixDensityOnGridController : public ObjectSensitiveClass,
                                             public std::enable_shared_from_this<DensityMatrixDensityOnGridController> {
 public:
  static std::shared_ptr<DensityMatrixDensityOnGridController> produce(
      std::shared_ptr<DensityOnGridCalculator> calc, std::shared_ptr<DensityMatrixController> dMat) {
    auto d = std::shared_ptr<DensityMatrixDensityOnGridController>(
        new DensityMatrixDensityOnGridController(std::move(calc), dMat));
    dMat->addSensitiveObject(d);
    return d;
  }
  const std::vector<double>& getDensityOnGrid() {
    update();
    return _rho;
  }
  const std::vector<double>& getDensityGradientOnGrid() {
    update();
    return _grad;
  }
  std::shared_ptr<BasisFunctionOnGridController> getBasisFunctionOnGridController() const {
    return _calc->getBasisFunctionOnGridController();
  }
  void notify() override { _valid = false; }

 private:
  DensityMatrixDensityOnGridController(std::shared_ptr<DensityOnGridCalculator> calc,
                                       std::shared_ptr<DensityMatrixController> dMat)
      : _calc(std::move(calc)), _dMat(std::move(dMat)) {}
  void update() {
    if (_valid) return;
    _calc->calcDensityAndGradientOnGrid(_dMat->getDensityMatrix(), _rho, _grad);
    _valid = true;
  }
  std::shared_ptr<DensityOnGridCalculator> _calc;
  std::shared_ptr<DensityMatrixController> _dMat;
  std::vector<double> _rho, _grad;
  bool _valid = false;
};

/* FunctionalData of calcData: epuv, dFdRho, and dFdGradRho (3 x N, GRADIENTS) or dFdSigma (N,
 * GRADIENT_INVARIANTS), energy. */
struct FunctionalData {
  std::vector<double> epuv, dFdRho, dFdGradRho, dFdSigma;
  double energy = 0.0;
};

class FunctionalLibrary {
 public:
  FunctionalLibrary(std::shared_ptr<Context> ctx, unsigned int maxBlockSize) : _ctx(std::move(ctx)), _bs(maxBlockSize) {}
  /* calcData(type, functional, densityOnGridController, order) (FunctionalLibrary.h:61-63): the
   * functional on the controller's (cached) density, with the weights of its grid.  order 1. */
  FunctionalData calcData(FUNCTIONAL_DATA_TYPE type, Functional f,
                          std::shared_ptr<DensityMatrixDensityOnGridController> densityOnGridController,
                          unsigned int order = 1) {
    if (order != 1)
      throw SerenityError(XC_ERR_ARG, "FunctionalLibrary::calcData: derivatives of order 1 only");
    if (type == FUNCTIONAL_DATA_TYPE::POTENTIAL)
      throw SerenityError(XC_ERR_ARG, "FunctionalLibrary::calcData: POTENTIAL data are not available");
    const auto& rho = densityOnGridController->getDensityOnGrid();
    const auto& grad = densityOnGridController->getDensityGradientOnGrid();
    const auto& w = densityOnGridController->getBasisFunctionOnGridController()->getGridController()->getWeights();
    FunctionalData d = calcData(f, rho, &grad, w);
    if (type == FUNCTIONAL_DATA_TYPE::GRADIENT_INVARIANTS && !d.dFdGradRho.empty()) {
      // d eps / d grad rho = 2 (d eps / d sigma) grad rho
      const size_t N = rho.size();
      d.dFdSigma.assign(N, 0.0);
      for (size_t p = 0; p < N; ++p) {
        const double gx = grad[p], gy = grad[N + p], gz = grad[2 * N + p];
        const double s = gx * gx + gy * gy + gz * gz;
        if (s > 0.0)
          d.dFdSigma[p] = (d.dFdGradRho[p] * gx + d.dFdGradRho[N + p] * gy + d.dFdGradRho[2 * N + p] * gz) / (2.0 * s);
      }
      d.dFdGradRho.clear();
    }
    return d;
  }
  /* The explicit form on given arrays (GRADIENTS data). */
  FunctionalData calcData(Functional f, const std::vector<double>& rho, const std::vector<double>* grad,
                          const std::vector<double>& weights) {
    FunctionalData d;
    const int64_t N = (int64_t)rho.size();
    d.epuv.resize(N);
    d.dFdRho.resize(N);
    const bool gga = isGGA(f);
    if (gga) d.dFdGradRho.resize(3 * N);
    check(xc_functional(_ctx->get(), (int)f, N, (int)_bs, rho.data(), grad ? grad->data() : nullptr, weights.data(),
                        d.epuv.data(), d.dFdRho.data(), gga ? d.dFdGradRho.data() : nullptr, &d.energy));
    return d;
  }

 private:
  std::shared_ptr<Context> _ctx;
  unsigned int _bs;
};

/* The parts of Serenity's SystemController the XC path reads: the context, the geometry (atomic
 * numbers, coordinates in bohr), the basis and the functional of the system's DFT settings. */
class SystemController {
 public:
  SystemController(std::shared_ptr<Context> ctx, std::vector<int> Z, std::vector<double> xyz,
                   std::shared_ptr<BasisController> basis, Functional functional = Functional::PBE)
      : _ctx(std::move(ctx)), _Z(std::move(Z)), _xyz(std::move(xyz)), _basis(std::move(basis)), _f(functional) {}
  std::shared_ptr<Context> getContext() const { return _ctx; }
  const std::vector<int>& getAtomicNumbers() const { return _Z; }
  const std::vector<double>& getCoordinates() const { return _xyz; }
  std::shared_ptr<BasisController> getBasisController() const { return _basis; }
  Functional getFunctional() const { return _f; }

 private:
  std::shared_ptr<Context> _ctx;
  std::vector<int> _Z;
  std::vector<double> _xyz;
  std::shared_ptr<BasisController> _basis;
  Functional _f;
};

/* ScalarOperatorToMatrixAdder(bfogc, blockAveThreshold): F += integral of v (and g). */
class ScalarOperatorToMatrixAdder {
 public:
  ScalarOperatorToMatrixAdder(std::shared_ptr<BasisFunctionOnGridController> bf, double blockAveThreshold)
      : _bf(std::move(bf)), _thr(blockAveThreshold) {}
  void addScalarOperatorToMatrix(Matrix& F, const std::vector<double>& scalar) { add(F, scalar, nullptr); }
  void addScalarOperatorToMatrix(Matrix& F, const std::vector<double>& scalar, const std::vector<double>& gradient) {
    add(F, scalar, &gradient);
  }

 private:
  void add(Matrix& F, const std::vector<double>& v, const std::vector<double>* g) {
    const unsigned int N = _bf->getNGridPoints();
    if (v.size() != N || (g && g->size() != 3 * (size_t)N))
      throw SerenityError(XC_ERR_ARG,
                          "The scalar operator and the BasisFunctionOnGridController passed to "
                          "addScalarOperatorToMatrix are not defined on the same grid!");
    if (F.rows != (int)_bf->getNBasisFunctions() || F.cols != F.rows)
      throw SerenityError(XC_ERR_ARG, "ScalarOperatorToMatrixAdder: matrix and basis differ in size");
    check(xc_fock_add(_bf->get(), _thr, v.data(), g ? g->data() : nullptr, F.data.data()));
  }
  std::shared_ptr<BasisFunctionOnGridController> _bf;
  double _thr;
};

/* FuncPotential (FuncPotential.h:48-145).  The reference's form: FuncPotential(system, dMat, grid,
 * functional), created with std::make_shared; getMatrix() builds V_xc of the controller's current
 * density matrix on first use and keeps it until the density matrix changes (notify), getEnergy(P)
 * returns the E_xc of that build (P is not read, as in the reference).  The explicit form
 * FuncPotential(bfogc, functional, precision) has getMatrix(P) and getEnergy(). */
class FuncPotential : public ObjectSensitiveClass, public std::enable_shared_from_this<FuncPotential> {
 public:
  FuncPotential(std::shared_ptr<SystemController> system, std::shared_ptr<DensityMatrixController> dMat,
                std::shared_ptr<GridController> grid, Functional functional, Precision p = Precision::FP64,
                double blockAveThreshold = 1e-11)
      : _system(std::move(system)), _dMat(std::move(dMat)), _f(functional), _p(p), _thr(blockAveThreshold) {
    _bf = std::make_shared<BasisFunctionOnGridController>(_system->getBasisController(), std::move(grid), 128, 1e-9,
                                                          1, p != Precision::FP64);
  }
  FuncPotential(std::shared_ptr<BasisFunctionOnGridController> bf, Functional f, Precision p = Precision::FP64,
                double blockAveThreshold = 1e-11)
      : _bf(std::move(bf)), _f(f), _p(p), _thr(blockAveThreshold) {}
  const Matrix& getMatrix() {
    if (!_dMat) throw SerenityError(XC_ERR_STATE, "FuncPotential: constructed without a density matrix controller");
    if (!_registered) {
      std::shared_ptr<FuncPotential> self;
      try {
        self = shared_from_this();
      } catch (const std::bad_weak_ptr&) {
        throw SerenityError(XC_ERR_STATE, "FuncPotential: create it with std::make_shared to use getMatrix()");
      }
      _dMat->addSensitiveObject(self);
      _registered = true;
    }
    if (!_cacheValid) {
      _F = getMatrix(_dMat->getDensityMatrix());
      _cacheValid = true;
    }
    return _F;
  }
  double getEnergy(const Matrix& /*P*/) {
    getMatrix();
    return _energy;
  }
  void notify() override { _cacheValid = false; }
  Functional getFunctional() const { return _f; }
  Matrix getMatrix(const Matrix& P) {
    const int n = (int)_bf->getNBasisFunctions();
    if (P.rows != n || P.cols != n)
      throw SerenityError(XC_ERR_ARG, "FuncPotential: density matrix and basis differ in size");
    Matrix F(n, n);
    check(xc_fock_build(_bf->get(), (int)_f, (int)_p, _thr, P.data.data(), F.data.data(), &_energy));
    _built = true;
    return F;
  }
  /* FuncPotential<UNRESTRICTED>::getMatrix: the alpha and beta XC matrices. */
  std::pair<Matrix, Matrix> getMatrix(const Matrix& Palpha, const Matrix& Pbeta) {
    const int n = (int)_bf->getNBasisFunctions();
    if (Palpha.rows != n || Palpha.cols != n || Pbeta.rows != n || Pbeta.cols != n)
      throw SerenityError(XC_ERR_ARG, "FuncPotential: spin density matrices and basis differ in size");
    Matrix Fa(n, n), Fb(n, n);
    check(xc_fock_build_unrestricted(_bf->get(), (int)_f, (int)_p, _thr, Palpha.data.data(), Pbeta.data.data(),
                                     Fa.data.data(), Fb.data.data(), &_energy));
    _built = true;
    return {std::move(Fa), std::move(Fb)};
  }
  double getEnergy() const {
    if (!_built) throw SerenityError(XC_ERR_STATE, "FuncPotential: getMatrix has not been called");
    return _energy;
  }

 private:
  std::shared_ptr<SystemController> _system;
  std::shared_ptr<DensityMatrixController> _dMat;
  std::shared_ptr<BasisFunctionOnGridController> _bf;
  Functional _f;
  Precision _p;
  double _thr;
  double _energy = 0.0;
  bool _built = false;
  bool _registered = false, _cacheValid = false;
  Matrix _F;
};

/* NAddFuncPotential (NAddFuncPotential.cpp:192-301), restricted: the non-additive potential of the active
 * subsystem, F = integral of v[rho_sup] - v[rho_act] in the active basis, getEnergy() =
 * E[rho_sup] - E[rho_act] - sum_env E[rho_env].  Every controller must share one GridController. */
class NAddFuncPotential {
 public:
  NAddFuncPotential(std::shared_ptr<BasisFunctionOnGridController> active,
                    std::vector<std::shared_ptr<BasisFunctionOnGridController>> environment, Functional f,
                    Precision p = Precision::FP64, double blockAveThreshold = 1e-11)
      : _act(std::move(active)), _env(std::move(environment)), _f(f), _p(p), _thr(blockAveThreshold) {}
  Matrix getMatrix(const Matrix& Pact, const std::vector<Matrix>& Penv) {
    const int n = (int)_act->getNBasisFunctions();
    if (Pact.rows != n || Pact.cols != n)
      throw SerenityError(XC_ERR_ARG, "NAddFuncPotential: active density matrix and basis differ in size");
    if (Penv.size() != _env.size())
      throw SerenityError(XC_ERR_ARG, "NAddFuncPotential: one density matrix per environment subsystem is needed");
    std::vector<xc_bfogc*> env;
    std::vector<const double*> P;
    for (size_t i = 0; i < _env.size(); ++i) {
      const int m = (int)_env[i]->getNBasisFunctions();
      if (Penv[i].rows != m || Penv[i].cols != m)
        throw SerenityError(XC_ERR_ARG, "NAddFuncPotential: environment density matrix and basis differ in size");
      env.push_back(_env[i]->get());
      P.push_back(Penv[i].data.data());
    }
    Matrix F(n, n);
    check(xc_nadd_fock_build(_act->get(), (int)env.size(), env.data(), (int)_f, (int)_p, _thr, Pact.data.data(),
                             P.data(), F.data.data(), &_energy));
    _built = true;
    return F;
  }
  /* NAddFuncPotential<UNRESTRICTED>::getMatrix: alpha and beta density matrices of the active
   * and environment subsystems in, the alpha and beta non-additive potentials out. */
  std::pair<Matrix, Matrix> getMatrix(const Matrix& PactA, const Matrix& PactB, const std::vector<Matrix>& PenvA,
                                      const std::vector<Matrix>& PenvB) {
    const int n = (int)_act->getNBasisFunctions();
    if (PactA.rows != n || PactA.cols != n || PactB.rows != n || PactB.cols != n)
      throw SerenityError(XC_ERR_ARG, "NAddFuncPotential: active density matrices and basis differ in size");
    if (PenvA.size() != _env.size() || PenvB.size() != _env.size())
      throw SerenityError(XC_ERR_ARG, "NAddFuncPotential: two density matrices per environment subsystem are needed");
    std::vector<xc_bfogc*> env;
    std::vector<const double*> Pa, Pb;
    for (size_t i = 0; i < _env.size(); ++i) {
      const int m = (int)_env[i]->getNBasisFunctions();
      if (PenvA[i].rows != m || PenvA[i].cols != m || PenvB[i].rows != m || PenvB[i].cols != m)
        throw SerenityError(XC_ERR_ARG, "NAddFuncPotential: environment density matrix and basis differ in size");
      env.push_back(_env[i]->get());
      Pa.push_back(PenvA[i].data.data());
      Pb.push_back(PenvB[i].data.data());
    }
    Matrix Fa(n, n), Fb(n, n);
    check(xc_nadd_fock_build_unrestricted(_act->get(), (int)env.size(), env.data(), (int)_f, (int)_p, _thr,
                                          PactA.data.data(), PactB.data.data(), Pa.data(), Pb.data(),
                                          Fa.data.data(), Fb.data.data(), &_energy));
    _built = true;
    return {std::move(Fa), std::move(Fb)};
  }
  double getEnergy() const {
    if (!_built) throw SerenityError(XC_ERR_STATE, "NAddFuncPotential: getMatrix has not been called");
    return _energy;
  }

 private:
  std::shared_ptr<BasisFunctionOnGridController> _act;
  std::vector<std::shared_ptr<BasisFunctionOnGridController>> _env;
  Functional _f;
  Precision _p;
  double _thr;
  double _energy = 0.0;
  bool _built = false;
};

}  // namespace serenity_amd

#endif /* SERENITY_XC_HPP */
