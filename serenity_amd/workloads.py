"""Seeded synthetic systems of the XC Fock-build benchmark configurations.

All geometries are built in Angstrom and returned in bohr; the seed below is
the one every benchmark and test uses unless it passes its own.

* ``water_cluster``  (H2O)_n on an nx x ny x nz lattice (O-H 0.9572 A,
  H-O-H 104.52 deg, 2.9 A spacing, +-0.25 A uniform jitter of every O, uniform
  random SO(3) orientation from a normalised 4-vector of N(0,1) draws).
* ``alkane``  all-anti CkH2k+2 (C-C 1.53 A, C-H 1.09 A, tetrahedral angles,
  +-0.02 A uniform jitter of every atom).
* ``benzene_water``  C6H6 (C-C 1.39 A, C-H 1.09 A) plus waters whose O lies
  4-7 A from the ring centre, each with a random orientation; a water is
  rejected if its O comes closer than 2.5 A to any heavy atom already placed
  or any of its atoms closer than 1.5 A to any atom already placed.
* ``synthetic_density``  P = 2 C C^T, C = S^(-1/2) Q, Q the QR factor of a
  seeded N(0,1) n x nocc matrix.
"""
import numpy as np

from .xc import ANGSTROM_TO_BOHR, BasisController, molecular_grid

SEED = 20261202

Z_OF = {"H": 1, "He": 2, "C": 6, "N": 7, "O": 8}


def _rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def water_monomer():
    """O, H, H in Angstrom, O at the origin."""
    r, th = 0.9572, np.deg2rad(104.52)
    return np.array([[0.0, 0.0, 0.0],
                     [r * np.sin(th / 2), 0.0, r * np.cos(th / 2)],
                     [-r * np.sin(th / 2), 0.0, r * np.cos(th / 2)]])


def water_cluster(nx, ny, nz, spacing=2.9, jitter=0.25, seed=SEED):
    rng = np.random.default_rng(seed)
    mono = water_monomer()
    Z, xyz = [], []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                o = np.array([i, j, k], dtype=float) * spacing + rng.uniform(-jitter, jitter, 3)
                R = _rotation(rng)
                for a, pos in zip((8, 1, 1), mono):
                    Z.append(a)
                    xyz.append(o + R @ pos)
    return np.array(Z, dtype=np.int32), np.array(xyz) * ANGSTROM_TO_BOHR


def water_chain(n, seed=SEED):
    """(H2O)_n for n = 1..48: the first n molecules of a 4 x 4 x 3 lattice."""
    if not 1 <= n <= 48:
        raise ValueError("water count must be in 1..48")
    Z, xyz = water_cluster(4, 4, 3, seed=seed)
    return Z[:3 * n], xyz[:3 * n]


def _tetra_pair(u1, u2):
    """The two remaining tetrahedral directions at an atom bonded along u1, u2."""
    b = -(u1 + u2)
    b /= np.linalg.norm(b)
    nrm = np.cross(u1, u2)
    nrm /= np.linalg.norm(nrm)
    ca = np.sqrt(1.0 / 3.0)
    sa = np.sqrt(2.0 / 3.0)
    return b * ca + nrm * sa, b * ca - nrm * sa


def alkane(k, jitter=0.02, seed=SEED):
    """All-anti CkH2k+2 (k >= 1); carbons first, then hydrogens."""
    if k < 1:
        raise ValueError("alkane needs k >= 1")
    rng = np.random.default_rng(seed)
    cc, ch = 1.53, 1.09
    half = np.deg2rad(109.4712206) / 2
    dx, dy = cc * np.sin(half), cc * np.cos(half)
    C = np.array([[i * dx, (i % 2) * dy, 0.0] for i in range(k)])
    H = []
    if k == 1:
        for v in ([1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]):
            H.append(C[0] + ch * np.array(v, float) / np.sqrt(3.0))
    else:
        def virtual(i):  # continue the zig-zag one carbon beyond either end
            return np.array([i * dx, (i % 2) * dy, 0.0])
        for i in range(k):
            prev = C[i - 1] if i > 0 else virtual(-1)
            nxt = C[i + 1] if i < k - 1 else virtual(k)
            u1 = (prev - C[i]) / np.linalg.norm(prev - C[i])
            u2 = (nxt - C[i]) / np.linalg.norm(nxt - C[i])
            h1, h2 = _tetra_pair(u1, u2)
            H.append(C[i] + ch * h1)
            H.append(C[i] + ch * h2)
            if i == 0:
                H.append(C[i] + ch * u1)
            if i == k - 1:
                H.append(C[i] + ch * u2)
    xyz = np.vstack([C, np.array(H)])
    xyz = xyz + rng.uniform(-jitter, jitter, xyz.shape)
    Z = np.array([6] * k + [1] * len(H), dtype=np.int32)
    return Z, xyz * ANGSTROM_TO_BOHR


def benzene():
    ang = np.deg2rad(60.0) * np.arange(6)
    C = np.stack([1.39 * np.cos(ang), 1.39 * np.sin(ang), np.zeros(6)], 1)
    H = np.stack([2.48 * np.cos(ang), 2.48 * np.sin(ang), np.zeros(6)], 1)
    return np.array([6] * 6 + [1] * 6, dtype=np.int32), np.vstack([C, H])


def benzene_water(nwater=24, rmin=4.0, rmax=7.0, seed=SEED):
    """Benzene plus waters; returns (Z, xyz_bohr, subsystem index per atom)."""
    rng = np.random.default_rng(seed)
    Zb, Xb = benzene()
    Z, X, sub = list(Zb), list(Xb), [0] * len(Zb)
    mono = water_monomer()
    placed = 0
    tries = 0
    while placed < nwater:
        tries += 1
        if tries > 100000:
            raise RuntimeError("could not place the waters")
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        o = d * rng.uniform(rmin, rmax)
        R = _rotation(rng)
        atoms = [o + R @ p for p in mono]
        Xa = np.array(X)
        heavy = Xa[np.array(Z) > 1]
        if np.min(np.linalg.norm(heavy - o, axis=1)) < 2.5:
            continue
        if min(np.min(np.linalg.norm(Xa - a, axis=1)) for a in atoms) < 1.5:
            continue
        placed += 1
        for a, pos in zip((8, 1, 1), atoms):
            Z.append(a)
            X.append(pos)
            sub.append(placed)
    return np.array(Z, dtype=np.int32), np.array(X) * ANGSTROM_TO_BOHR, np.array(sub, dtype=np.int32)


def n_electrons(Z):
    return int(np.sum(Z))


def synthetic_density(S, nocc, seed=SEED):
    """P = 2 C C^T with C = S^(-1/2) Q (Q: QR of a seeded N(0,1) n x nocc matrix)."""
    n = S.shape[0]
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, nocc)))
    e, U = np.linalg.eigh(S)
    Sm12 = (U / np.sqrt(e)) @ U.T
    Cocc = Sm12 @ Q
    return np.asfortranarray(2.0 * Cocc @ Cocc.T)


CONFIGS = {
    1: dict(name="(H2O)8/def2-SVP/LDA", system=("water", (2, 2, 2)), basis="DEF2-SVP", functional="LDA"),
    2: dict(name="C20H42/def2-SVP/PBE", system=("alkane", 20), basis="DEF2-SVP", functional="PBE"),
    3: dict(name="(H2O)64/def2-TZVP/PBE", system=("water", (4, 4, 4)), basis="DEF2-TZVP", functional="PBE"),
    4: dict(name="C6H6+24H2O/def2-SVP/PBE+TF", system=("benzene_water", 24), basis="DEF2-SVP", functional="PBE"),
}


def geometry(system, seed=SEED):
    kind, arg = system
    if kind == "water":
        return water_cluster(*arg, seed=seed)
    if kind == "alkane":
        return alkane(arg, seed=seed)
    if kind == "benzene_water":
        Z, X, _ = benzene_water(arg, seed=seed)
        return Z, X
    raise ValueError(kind)


def build_system(config, accuracy=4, seed=SEED, spherical=True):
    """Geometry, basis, grid and synthetic density of a benchmark configuration."""
    cfg = CONFIGS[config] if isinstance(config, int) else config
    Z, X = geometry(cfg["system"], seed=seed)
    basis = BasisController.from_file(cfg["basis"], Z, X, spherical=spherical)
    grid = molecular_grid(Z, X, accuracy=accuracy)
    S = basis.overlap()
    P = synthetic_density(S, n_electrons(Z) // 2, seed=seed)
    return dict(cfg=cfg, Z=Z, X=X, basis=basis, grid=grid, S=S, P=P)


def build_embedding(nwater=24, basis="DEF2-SVP", accuracy=4, seed=SEED, spherical=True):
    """Benzene + waters as 1 + nwater subsystems on the supersystem grid.

    Each subsystem has its own basis (its atoms only) and a synthetic closed-shell
    density P_i = 2 C C^T (C = S_i^-1/2 Q_i, seed + i); the grid is the molecular
    grid of all atoms, shared by every subsystem.  Subsystem 0 is benzene."""
    Z, X, sub = benzene_water(nwater, seed=seed)
    grid = molecular_grid(Z, X, accuracy=accuracy)
    subsystems = []
    for i in range(int(sub.max()) + 1):
        m = sub == i
        b = BasisController.from_file(basis, Z[m], X[m], spherical=spherical)
        P = synthetic_density(b.overlap(), n_electrons(Z[m]) // 2, seed=seed + i)
        subsystems.append(dict(Z=Z[m], X=X[m], basis=b, P=P))
    return dict(Z=Z, X=X, grid=grid, subsystems=subsystems)
