// C++ driver of the parity tests: exercises the Serenity-shaped C++ classes of
// include/serenity_xc.hpp on inputs written by tests/test_cpp_host.py and
// writes the results back as text.  Input format: whitespace-separated
// sections "name count values...".
//
//   cpp_parity --version                   (no device needed)
//   cpp_parity <case> <input> <output>     case: bfogc | density | fock | funcpot | nadd
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "serenity_xc.hpp"

using namespace serenity_amd;

static std::map<std::string, std::vector<double>> read_sections(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open " + path);
  std::map<std::string, std::vector<double>> out;
  std::string name;
  size_t count;
  while (in >> name >> count) {
    std::vector<double> v(count);
    for (size_t i = 0; i < count; ++i) in >> v[i];
    out[name] = v;
  }
  return out;
}

static std::vector<int> ints(const std::vector<double>& v) { return std::vector<int>(v.begin(), v.end()); }

static void write(std::ofstream& out, const std::string& name, const std::vector<double>& v) {
  out << name << " " << v.size() << "\n" << std::setprecision(17);
  for (double x : v) out << x << "\n";
}

int main(int argc, char** argv) {
  if (argc == 2 && std::string(argv[1]) == "--version") {
    std::cout << "cpp_parity xc_version " << xc_version() << std::endl;
    return 0;
  }
  if (argc != 4) {
    std::cerr << "usage: cpp_parity <case> <input> <output>" << std::endl;
    return 2;
  }
  const std::string cs = argv[1];
  try {
    auto in = read_sections(argv[2]);
    std::ofstream out(argv[3]);
    auto ctx = std::make_shared<Context>(0);
    std::shared_ptr<BasisController> basis;
    if (in.count("shell_l")) {
      BasisController::Shells s;
      s.l = ints(in["shell_l"]);
      s.pure = ints(in["shell_pure"]);
      s.nprim = ints(in["shell_nprim"]);
      s.center = in["shell_center"];
      s.alpha = in["alpha"];
      s.coef = in["coef"];
      basis = std::make_shared<BasisController>(ctx, s);
    }
    std::shared_ptr<GridController> grid;
    if (in.count("points")) grid = std::make_shared<GridController>(ctx, in["points"], in["weights"]);
    const unsigned bs = in.count("block_size") ? (unsigned)in["block_size"][0] : 128;
    const double thr = in.count("radial_threshold") ? in["radial_threshold"][0] : 1e-9;
    if (cs == "bfogc") {
      auto bf = std::make_shared<BasisFunctionOnGridController>(basis, grid, bs, thr, 1);
      auto d = bf->getBlockOnGridData(0);
      write(out, "values", d->functionValues.data);
      for (int c = 0; c < 3; ++c) write(out, std::string("grad") + "xyz"[c], d->derivativeValues[c].data);
      write(out, "nblocks", {(double)bf->getNBlocks()});
    } else if (cs == "density") {
      auto bf = std::make_shared<BasisFunctionOnGridController>(basis, grid, bs, thr, 1);
      DensityOnGridCalculator calc(bf, 0.0);
      const int n = (int)basis->getNBasisFunctions();
      Matrix P(n, n);
      P.data = in["P"];
      std::vector<double> rho, grad;
      calc.calcDensityAndGradientOnGrid(P, rho, grad);
      write(out, "rho", rho);
      write(out, "grad", grad);
    } else if (cs == "fock") {
      auto bf = std::make_shared<BasisFunctionOnGridController>(basis, grid, bs, thr, 1);
      ScalarOperatorToMatrixAdder adder(bf, in["block_ave_threshold"][0]);
      const int n = (int)basis->getNBasisFunctions();
      Matrix F(n, n);
      adder.addScalarOperatorToMatrix(F, in["potential"], in["gradient"]);
      write(out, "F", F.data);
    } else if (cs == "funcpot") {
      // basis file + geometry: BasisController::fromFile, molecular grid, FuncPotential
      std::ifstream pf(argv[2]);
      std::string token, basis_path;
      while (pf >> token)
        if (token == "basis_path") pf >> basis_path;
      auto Z = ints(in["Z"]);
      auto b = BasisController::fromFile(ctx, basis_path, Z, in["xyz"], in["spherical"][0] != 0);
      auto g = GridController::molecular(ctx, Z, in["xyz"], (int)in["grid_type"][0], (int)in["accuracy"][0],
                                         in["weight_threshold"][0]);
      auto bf = std::make_shared<BasisFunctionOnGridController>(b, g, 128, 1e-9, 1);
      const int n = (int)b->getNBasisFunctions();
      Matrix P(n, n);
      P.data = in["P"];
      FuncPotential pot(bf, (Functional)(int)in["functional"][0]);
      Matrix F = pot.getMatrix(P);
      write(out, "F", F.data);
      write(out, "energy", {pot.getEnergy()});
      // error behaviour: a wrong-size density matrix throws SerenityError
      bool threw = false;
      try {
        pot.getMatrix(Matrix(n - 1, n - 1));
      } catch (const SerenityError&) {
        threw = true;
      }
      write(out, "threw_on_bad_P", {threw ? 1.0 : 0.0});
    } else if (cs == "funcpot_ref") {
      // the reference's form: FuncPotential(system, dMat, grid, functional), getMatrix(), notify
      std::ifstream pf(argv[2]);
      std::string token, basis_path;
      while (pf >> token)
        if (token == "basis_path") pf >> basis_path;
      auto Z = ints(in["Z"]);
      auto b = BasisController::fromFile(ctx, basis_path, Z, in["xyz"], true);
      auto g = GridController::molecular(ctx, Z, in["xyz"], XC_GRID_SSF, (int)in["accuracy"][0]);
      const int n = (int)b->getNBasisFunctions();
      Matrix P(n, n);
      P.data = in["P"];
      const auto f = (Functional)(int)in["functional"][0];
      auto system = std::make_shared<SystemController>(ctx, Z, in["xyz"], b, f);
      auto dMat = std::make_shared<DensityMatrixController>(P);
      auto pot = std::make_shared<FuncPotential>(system, dMat, g, f);
      const Matrix& F1 = pot->getMatrix();
      write(out, "F1", F1.data);
      write(out, "E1", {pot->getEnergy(P)});
      // a new density matrix reaches the potential through notify
      Matrix P2 = P;
      for (double& x : P2.data) x *= 0.9;
      dMat->setDensityMatrix(P2);
      write(out, "F2", pot->getMatrix().data);
      write(out, "E2", {pot->getEnergy(P2)});
      // FunctionalLibrary::calcData on a density-on-grid controller of the same density matrix
      auto bf = std::make_shared<BasisFunctionOnGridController>(b, g, 128, 1e-9, 1);
      auto calc = std::make_shared<DensityOnGridCalculator>(bf, 1e-11);
      auto dog = DensityMatrixDensityOnGridController::produce(calc, dMat);
      FunctionalLibrary lib(ctx, 128);
      auto d = lib.calcData(FUNCTIONAL_DATA_TYPE::GRADIENTS, f, dog);
      write(out, "calc_energy", {d.energy});
      write(out, "calc_dFdRho", d.dFdRho);
      auto d2 = lib.calcData(FUNCTIONAL_DATA_TYPE::GRADIENT_INVARIANTS, f, dog);
      write(out, "calc_dFdSigma", d2.dFdSigma);
      write(out, "rho", dog->getDensityOnGrid());
      write(out, "grad", dog->getDensityGradientOnGrid());
      // the explicit form without shared ownership cannot register for notify
      bool threw = false;
      try {
        FuncPotential loose(system, dMat, g, f);
        loose.getMatrix();
      } catch (const SerenityError&) {
        threw = true;
      }
      write(out, "threw_without_make_shared", {threw ? 1.0 : 0.0});
    } else if (cs == "nadd") {
      // active and environment subsystems (sections prefixed act_ / env_) on one grid
      auto subsystem = [&](const std::string& pre) {
        BasisController::Shells s;
        s.l = ints(in[pre + "shell_l"]);
        s.pure = ints(in[pre + "shell_pure"]);
        s.nprim = ints(in[pre + "shell_nprim"]);
        s.center = in[pre + "shell_center"];
        s.alpha = in[pre + "alpha"];
        s.coef = in[pre + "coef"];
        auto b = std::make_shared<BasisController>(ctx, s);
        return std::make_shared<BasisFunctionOnGridController>(b, grid, 128, 1e-9, 1);
      };
      auto act = subsystem("act_"), env = subsystem("env_");
      const int na = (int)act->getNBasisFunctions(), ne = (int)env->getNBasisFunctions();
      Matrix Pa(na, na), Pe(ne, ne);
      Pa.data = in["P_act"];
      Pe.data = in["P_env"];
      NAddFuncPotential pot(act, {env}, (Functional)(int)in["functional"][0]);
      Matrix F = pot.getMatrix(Pa, {Pe});
      write(out, "F", F.data);
      write(out, "energy", {pot.getEnergy()});
      bool threw = false;
      try {
        pot.getMatrix(Pa, {});
      } catch (const SerenityError&) {
        threw = true;
      }
      write(out, "threw_on_missing_env", {threw ? 1.0 : 0.0});
      // NAddFuncPotential<UNRESTRICTED>: the closed-shell dimer with both spins at P / 2
      Matrix Pa2 = Pa, Pe2 = Pe;
      for (double& x : Pa2.data) x *= 0.5;
      for (double& x : Pe2.data) x *= 0.5;
      auto FaFb = pot.getMatrix(Pa2, Pa2, {Pe2}, {Pe2});
      write(out, "F_alpha", FaFb.first.data);
      write(out, "F_beta", FaFb.second.data);
      write(out, "energy_unrestricted", {pot.getEnergy()});
    } else {
      std::cerr << "unknown case " << cs << std::endl;
      return 2;
    }
  } catch (const SerenityError& e) {
    std::cerr << "SerenityError(" << e.code << "): " << e.what() << std::endl;
    return 1;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
