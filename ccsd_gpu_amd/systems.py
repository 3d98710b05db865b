"""Seeded synthetic molecular systems of the benchmark configurations.

  * water(seed, sigma)      H2O near equilibrium (O-H 0.9572 A, HOH 104.52 deg),
                            plus N(0, sigma) Bohr on every coordinate;
  * water_cluster(n, seed)  (H2O)_n of rigid monomers with uniformly random
                            orientations (unit quaternions), each new oxygen
                            2.8 +- 0.2 A from an existing one and no closer
                            than 2.6 A to any other, intermolecular atom pairs
                            kept >= 1.5 A apart;
  * random_cnoh(seed, nheavy) closed-shell C/N/O/H molecule: a random tree of
                            heavy atoms with bond lengths U(1.0, 1.5) A, free
                            valences (C4, N3, O2) filled with H at 1.0 A.
All randomness comes from numpy.random.default_rng(seed).
"""
import math

import numpy as np

from .gto.mole import BOHR, M

R_OH = 0.9572
THETA_HOH = math.radians(104.52)


def _monomer():
    """O at the origin, H atoms in the xz-plane (Angstrom)."""
    h = THETA_HOH / 2
    return np.array([[0.0, 0.0, 0.0],
                     [R_OH * math.sin(h), 0.0, R_OH * math.cos(h)],
                     [-R_OH * math.sin(h), 0.0, R_OH * math.cos(h)]])


def _rand_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _atoms_string(symbols, xyz):
    return "; ".join(f"{s} {x:.10f} {y:.10f} {z:.10f}" for s, (x, y, z) in zip(symbols, xyz))


def water(seed=0, sigma_bohr=0.02, basis="cc-pvdz", cart=False):
    rng = np.random.default_rng(seed)
    xyz = _monomer() + rng.normal(scale=sigma_bohr * BOHR, size=(3, 3))
    return M(_atoms_string(["O", "H", "H"], xyz), basis=basis, cart=cart)


def water_cluster_xyz(n, seed=0):
    rng = np.random.default_rng(seed)
    mono = _monomer()
    placed = []
    oxy = []
    for w in range(n):
        for _ in range(10000):
            if not oxy:
                o = np.zeros(3)
            else:
                anchor = oxy[rng.integers(len(oxy))]
                v = rng.normal(size=3)
                v /= np.linalg.norm(v)
                o = anchor + v * rng.uniform(2.6, 3.0)
                if any(np.linalg.norm(o - p) < 2.6 for p in oxy):
                    continue
            atoms = mono @ _rand_rotation(rng).T + o
            if placed:
                others = np.concatenate(placed)
                dmin = np.min(np.linalg.norm(atoms[:, None, :] - others[None, :, :], axis=-1))
                if dmin < 1.5:
                    continue
            placed.append(atoms)
            oxy.append(o)
            break
        else:
            raise RuntimeError("could not place water monomer")
    return np.concatenate(placed)


def water_cluster(n, seed=0, basis="cc-pvdz", cart=False):
    xyz = water_cluster_xyz(n, seed)
    return M(_atoms_string(["O", "H", "H"] * n, xyz), basis=basis, cart=cart)


_VALENCE = {"C": 4, "N": 3, "O": 2}


def random_cnoh(seed, nheavy, basis="cc-pvdz", cart=False):
    rng = np.random.default_rng(seed)
    for _attempt in range(1000):
        sym = list(rng.choice(["C", "N", "O"], size=nheavy, p=[0.6, 0.2, 0.2]))
        pos = [np.zeros(3)]
        used = [0] * nheavy
        ok = True
        for a in range(1, nheavy):
            cands = [b for b in range(a) if used[b] < _VALENCE[sym[b]] - (0 if b == 0 else 1)]
            if not cands:
                ok = False
                break
            b = cands[rng.integers(len(cands))]
            for _ in range(200):
                v = rng.normal(size=3)
                v /= np.linalg.norm(v)
                p = pos[b] + v * rng.uniform(1.0, 1.5)
                if all(np.linalg.norm(p - q) >= 1.2 for i, q in enumerate(pos) if i != b):
                    break
            else:
                ok = False
                break
            pos.append(p)
            used[b] += 1
            used[a] += 1
        if not ok:
            continue
        hs = []
        for a in range(nheavy):
            for _ in range(_VALENCE[sym[a]] - used[a]):
                for _try in range(200):
                    v = rng.normal(size=3)
                    v /= np.linalg.norm(v)
                    p = pos[a] + v * 1.0
                    if all(np.linalg.norm(p - q) >= 0.9 for i, q in enumerate(pos) if i != a) and \
                            all(np.linalg.norm(p - q) >= 0.9 for q in hs):
                        hs.append(p)
                        break
                else:
                    ok = False
                    break
            if not ok:
                break
        if not ok:
            continue
        symbols = sym + ["H"] * len(hs)
        xyz = np.array(pos + hs)
        return M(_atoms_string(symbols, xyz), basis=basis, cart=cart)
    raise RuntimeError("could not build molecule")


def cnoh_batch(count=64, seed=0, basis="cc-pvdz", cart=False):
    rng = np.random.default_rng(seed)
    return [random_cnoh(int(rng.integers(1 << 30)), int(rng.integers(3, 13)), basis=basis, cart=cart)
            for _ in range(count)]
