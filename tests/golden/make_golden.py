"""Write the golden vectors of the XC path from the Serenity tree.

Run once in a container that has the reference mounted (default
``/root/reference``); the outputs in this directory are committed data and
nothing reads the reference at test time.  What is extracted:

* ``kat_ao.json``       -- expected AO values / gradients / Hessians of the
  SMALL_MIXED basis on the TINY grid (Mathematica known answers,
  src/data/grid/BasisFunctionOnGridController_test.cpp:43-503, tol 1e-8).
* ``kat_density.json``  -- the 10x10 density matrix and expected rho, grad rho,
  Hessian of rho on TINY (src/data/grid/DensityOnGridCalculator_test.cpp:38-256,
  tol 5e-8).
* ``kat_fock.json``     -- scalar and gradient potentials on VERY_SMALL and the
  expected Fock matrix, block size 3 (src/data/grid/ScalarOperatorToMatrixAdder
  _test.cpp:40-147, tol 1e-8).
* ``kat_funcpot_h2.json`` -- the 12x12 LDA and BP86 Fock matrices of H2/def2-TZVP
  (src/potentials/FuncPotential_test.cpp:59-100, 148-232, tol 1e-6).
* ``fixtures.npz``      -- density matrices (HDF5 datasets read with h5dump)
  and stored energies of the H2/def2-TZVP, He/def2-TZVP (PBE) and
  water/def2-TZVP (PBE) test systems, with their geometries.
* ``grid_factory.json`` -- the C60 test geometry and the sphere-volume windows
  of src/grid/construction/GridFactory_test.cpp (Becke/Voronoi/SSF grids).
* ``hilbert.json``      -- the 2x2x2 Hilbert-sort case (src/grid/HilbertRTreeSorting_test.cpp:32-54).
* ``SC-CC-PV5Z``        -- the Sc entry of data/basis/CC-PV5Z (spherical
  normalisation invariant, BasisFunctionOnGridController_test.cpp:505-701).
"""
import ast
import json
import operator
import os
import re
import subprocess
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
H5DUMP = "/opt/conda/bin/h5dump"

_OPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: operator.truediv}


def num(expr):
    """Evaluate a numeric literal expression with + - * / and sqrt()."""
    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant):
            return float(node.value)
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
            return -ev(node.operand)
        if isinstance(node, ast.BinOp) and type(node.op) in _OPS:
            return _OPS[type(node.op)](ev(node.left), ev(node.right))
        if isinstance(node, ast.Call) and getattr(node.func, "id", "") == "sqrt":
            return float(np.sqrt(ev(node.args[0])))
        raise ValueError("unsupported expression: " + expr)
    return ev(ast.parse(expr.strip(), mode="eval"))


def read(rel):
    return open(os.path.join(REF, rel)).read()


def kat_ao():
    text = read("src/data/grid/BasisFunctionOnGridController_test.cpp")
    body = text.split("TEST_F(BasisFunctionOnGridControllerTest, TestBasisFunctionValuesAndDerivatives)")[1]
    body = body.split("TEST_F(")[0]
    out = []
    pat = re.compile(r"EXPECT_NEAR\((.+?),\s*(values|derivatives\.[xyz]|hessian\.[xyz]{2})\((\d+),\s*(\d+)\),\s*expectedPrecision\)")
    for m in pat.finditer(body):
        out.append(dict(kind=m.group(2), point=int(m.group(3)), ao=int(m.group(4)), expected=num(m.group(1))))
    # one entry is written as EXPECT_NEAR(0, value / sqrt(c), tol): the value itself must vanish
    pat0 = re.compile(r"EXPECT_NEAR\(0,\s*(values|derivatives\.[xyz]|hessian\.[xyz]{2})\((\d+),\s*(\d+)\)\s*/\s*sqrt")
    for m in pat0.finditer(body):
        out.append(dict(kind=m.group(1), point=int(m.group(2)), ao=int(m.group(3)), expected=0.0))
    return dict(tolerance=1e-8, radial_threshold=1e-11, block_size=10, entries=out)


def kat_density():
    text = read("src/data/grid/DensityOnGridCalculator_test.cpp")
    P = np.zeros((10, 10))
    for m in re.finditer(r"densityMatrix\((\d+),\s*(\d+)\)\s*=\s*(.+?);", text):
        P[int(m.group(1)), int(m.group(2))] = num(m.group(3))
    exp = []
    pat = re.compile(r"EXPECT_NEAR\(([-0-9.eE]+),\s*(densOnGrid|gradOnGrid\.[xyz]|hessOnGrid\.[xyz]{2})\[(\d+)\],\s*expectedPrecision\)")
    seen = set()
    for m in pat.finditer(text):
        key = (m.group(2), int(m.group(3)))
        if key in seen:
            continue
        seen.add(key)
        exp.append(dict(kind=m.group(2), point=int(m.group(3)), expected=float(m.group(1))))
    return dict(tolerance=5e-8, block_size=3, radial_threshold=0.0, P=P.tolist(), entries=exp)


def kat_fock():
    text = read("src/data/grid/ScalarOperatorToMatrixAdder_test.cpp")
    pot = [0.0] * 5
    grad = [[0.0] * 5 for _ in range(3)]
    for m in re.finditer(r"\bpot\[(\d+)\]\s*=\s*(.+?);", text):
        pot[int(m.group(1))] = num(m.group(2))
    for m in re.finditer(r"gradPot\.([xyz])\[(\d+)\]\s*=\s*(.+?);", text):
        grad["xyz".index(m.group(1))][int(m.group(2))] = num(m.group(3))
    F = []
    for m in re.finditer(r"EXPECT_NEAR\(result\((\d+),\s*(\d+)\),\s*([-0-9.eE]+),\s*expectedPrecision\)", text):
        F.append(dict(row=int(m.group(1)), col=int(m.group(2)), expected=float(m.group(3))))
    return dict(tolerance=1e-8, block_size=3, radial_threshold=1e-11, block_ave_threshold=1e-11, potential=pot,
                gradient=grad, entries=F)


def kat_funcpot():
    text = read("src/potentials/FuncPotential_test.cpp")

    def fref(test):
        body = text.split("TEST_F(FuncPotentialTest, %s)" % test)[1].split("TEST_F(")[0]
        vals = re.search(r"Fref <<(.+?);", body, re.S).group(1)
        return np.array([float(x) for x in vals.replace("\n", " ").split(",")]).reshape(12, 12).tolist()
    return dict(tolerance=1e-6, functional="LDA", F=fref("H2_FockMatrix_LDA"), F_BP86=fref("H2_FockMatrix_GGA"))


def kat_nadd():
    """Restricted LDA NAdd Fock entries of the H2 dimer (NAddFuncPotential_test.cpp:58-82)."""
    text = read("src/potentials/NAddFuncPotential_test.cpp")
    body = text.split("TEST_F(NAddFuncPotentialTest, H2Dimer_FockMatrix_LDA)")[1].split("TEST_F(")[0]
    F = [dict(row=int(m.group(1)), col=int(m.group(2)), expected=float(m.group(3)))
         for m in re.finditer(r"EXPECT_NEAR\(F\((\d+),\s*(\d+)\),\s*([-0-9.eE]+),\s*([-0-9.eE]+)\)", body)]
    tol = float(re.search(r"EXPECT_NEAR\(F\(0, 0\),\s*[-0-9.eE]+,\s*([-0-9.eE]+)\)", body).group(1))
    return dict(tolerance=tol, functional="LDA", basis="6-31GS", grid="SSF accuracy 4 of the active system",
                entries=F)


def h5_dataset(path, name):
    """Dataset `name` of an HDF5 file as a float64 array (h5dump binary export)."""
    import tempfile
    head = subprocess.run([H5DUMP, "-H", "-d", name, path], check=True, capture_output=True, text=True).stdout
    dims = re.search(r"DATASPACE\s+SIMPLE\s+\{\s*\(\s*([0-9, ]+)\)", head).group(1)
    shape = tuple(int(x) for x in dims.split(","))
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "d.bin")
        subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", out, path], check=True, capture_output=True)
        arr = np.fromfile(out, dtype="<f8")
    assert arr.size == int(np.prod(shape)), (path, name, arr.size, shape)
    return arr.reshape(shape)


def energies(path):
    out = {}
    for line in open(path):
        parts = line.split()
        if len(parts) == 2 and parts[0].isdigit():
            out[int(parts[0])] = float(parts[1])
    return out


def xyz(path):
    lines = open(path).read().splitlines()
    n = int(lines[0])
    sym, pos = [], []
    for line in lines[2:2 + n]:
        p = line.split()
        sym.append(p[0].capitalize())
        pos.append([float(x) for x in p[1:4]])
    return sym, np.array(pos)


def fixtures():
    base = os.path.join(REF, "data/testresources")
    out = {}
    for tag, d in (("h2", "TestSystem_H2_DEF2_TZVP"), ("he", "TestSystem_He_Def2-TZVP_DFT"),
                   ("water", "TestSystem_Water_Def2-TZVP_DFT"), ("h2fde_act", "TestSystem_H2_6_31Gs_ACTIVE_FDE"),
                   ("h2fde_env", "TestSystem_H2_6_31Gs_ENVIRONMENT_FDE")):
        P = h5_dataset(os.path.join(base, d, d + ".dmat.res.h5"), "densityMatrix")
        sym, pos = xyz(os.path.join(base, d, d + ".xyz"))
        e = energies(os.path.join(base, d, d + ".energies.res"))
        out[tag + "_P"] = P
        out[tag + "_Z"] = np.array([{"H": 1, "He": 2, "O": 8}[s] for s in sym], dtype=np.int32)
        out[tag + "_xyz_angstrom"] = pos
        out[tag + "_energy_ids"] = np.array(sorted(e), dtype=np.int32)
        out[tag + "_energy_values"] = np.array([e[k] for k in sorted(e)])
    # BP86: H2 / 6-31G* (Cartesian), RI-J, Becke grid; density and Fock matrices
    d = "TestSystem_H2_6_31Gs_BP86"
    out["h2bp86_P"] = h5_dataset(os.path.join(base, d, d + ".dmat.res.h5"), "densityMatrix")
    out["h2bp86_F"] = h5_dataset(os.path.join(base, d, d + ".FockMatrix.res.h5"), "FockMatrix")
    sym, pos = xyz(os.path.join(base, d, d + ".xyz"))
    e = energies(os.path.join(base, d, d + ".energies.res"))
    out["h2bp86_Z"] = np.array([{"H": 1}[s] for s in sym], dtype=np.int32)
    out["h2bp86_xyz_angstrom"] = pos
    out["h2bp86_energy_ids"] = np.array(sorted(e), dtype=np.int32)
    out["h2bp86_energy_values"] = np.array([e[k] for k in sorted(e)])
    # open shell: triplet O2 / 6-31G / PBE (LibXC), unrestricted density and Fock matrices
    d = "TestSystem_O2_TRIP_6_31G_PBE_TDA"
    for spin in ("alpha", "beta"):
        out["o2_P" + spin[0]] = h5_dataset(os.path.join(base, d, d + ".dmat.unres.h5"), "densityMatrix_" + spin)
        out["o2_F" + spin[0]] = h5_dataset(os.path.join(base, d, d + ".FockMatrix.unres.h5"), "FockMatrix_" + spin)
    sym, pos = xyz(os.path.join(base, d, d + ".xyz"))
    e = energies(os.path.join(base, d, d + ".energies.unres"))
    out["o2_Z"] = np.array([{"O": 8}[s] for s in sym], dtype=np.int32)
    out["o2_xyz_angstrom"] = pos
    out["o2_energy_ids"] = np.array(sorted(e), dtype=np.int32)
    out["o2_energy_values"] = np.array([e[k] for k in sorted(e)])
    np.savez(os.path.join(HERE, "fixtures.npz"), **out)


def grid_factory_cases():
    """C60 geometry (bohr) and the volume-integration windows of GridFactory_test.cpp."""
    sup = read("src/testsupply/SystemController__TEST_SUPPLY.cpp")
    block = sup.split("case TEST_SYSTEM_CONTROLLERS::C60_MINBAS: {\n      std::vector<std::shared_ptr<Atom>> atoms;")[1]
    block = block.split("_testGeometries[TEST_SYSTEM_CONTROLLERS::C60_MINBAS]")[0]
    xyz = [[float(a), float(b), float(c)] for a, b, c in
           re.findall(r'getAtomType\("C"\),\s*([-0-9.eE]+),\s*([-0-9.eE]+),\s*([-0-9.eE]+)\)', block)]
    text = read("src/grid/construction/GridFactory_test.cpp")
    cases = []
    for m in re.finditer(r"(?m)^TEST_F\(GridFactoryTest, (\w+)\) \{(.*?)\n\}", text, re.S):
        body = m.group(2)
        g = re.search(r"GridFactory tGridFactory\(Options::GRID_TYPES::(\w+), (\d+), Options::RADIAL_GRID_TYPES::(\w+),"
                      r"\s*Options::SPHERICAL_GRID_TYPES::LEBEDEV, (\d+), ([-0-9.eE]+)\)", body)
        near = re.search(r"\n\s*EXPECT_NEAR\(sum, 1.0, ([-0-9.eE]+)\)", body)
        gt = re.search(r"\n\s*EXPECT_GT\(sum, 1.0 \+ ([-0-9.eE]+)\)", body)
        lt = re.search(r"\n\s*EXPECT_LT\(sum, 1.0 - ([-0-9.eE]+)\)", body)
        if not g or not near:
            continue
        cases.append(dict(name=m.group(1), grid_type=g.group(1), smoothing=int(g.group(2)), radial=g.group(3),
                          accuracy=int(g.group(4)), weight_threshold=float(g.group(5)), near=float(near.group(1)),
                          greater=float(gt.group(1)) if gt else None, less=float(lt.group(1)) if lt else None))
    return dict(c60_bohr=xyz, sphere_radius=7.0, cases=cases)


def hilbert_case():
    """2x2x2 points and the expected order (HilbertRTreeSorting_test.cpp:32-54)."""
    text = read("src/grid/HilbertRTreeSorting_test.cpp").split("TEST_F(HilbertRTreeSortingTest, 2x2x2)")[1]
    text = text.split("TEST_F(")[0]
    def row_major(name, rows, cols):
        vals = re.search(name + r" <<(.+?);", text, re.S).group(1)
        return np.array([float(v) for v in vals.replace("\n", " ").split(",")]).reshape(rows, cols)
    return dict(weights=row_major("weights", 1, 8).ravel().tolist(), coords=row_major("coords", 3, 8).T.tolist(),
                wref=row_major("wref", 1, 8).ravel().tolist(), pref=row_major("pref", 3, 8).T.tolist())


def basis_entry(src, element, out):
    """One element's entry of a Serenity basis file, kept in Turbomole format."""
    text = read("data/basis/" + src)
    m = re.search(r"\n%s\s+%s\s*\n\*\n(.*?)\n\*" % (element, re.escape(src)), text, re.S)
    with open(os.path.join(HERE, out), "w") as f:
        f.write("# %s entry of Serenity data/basis/%s (Turbomole format)\n$basis\n*\n%s   %s\n*\n"
                % (element.capitalize(), src, element, src))
        f.write(m.group(1).strip("\n") + "\n*\n$end\n")


def sc_basis():
    basis_entry("CC-PV5Z", "sc", "SC-CC-PV5Z")
    basis_entry("6-31GS", "h", "H-6-31GS")
    basis_entry("6-31G", "o", "O-6-31G")


if __name__ == "__main__":
    for name, fn in (("kat_ao.json", kat_ao), ("kat_density.json", kat_density), ("kat_fock.json", kat_fock),
                     ("kat_funcpot_h2.json", kat_funcpot), ("grid_factory.json", grid_factory_cases),
                     ("hilbert.json", hilbert_case), ("kat_nadd_h2.json", kat_nadd)):
        data = fn()
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(data, f, indent=1)
        print(name, len(data.get("entries", data.get("F", data.get("cases", [])))))
    fixtures()
    sc_basis()
    print("fixtures.npz, SC-CC-PV5Z, H-6-31GS, O-6-31G written")
