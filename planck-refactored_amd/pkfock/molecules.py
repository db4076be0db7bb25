"""Synthetic molecules for the Fock-build workloads (deterministic or seeded).

Geometries are returned as (Z list, coordinates in Angstrom, N x 3).  The unit
conversion to Bohr uses the reference's ANGSTROM_TO_BOHR = 1.8897259886
(src/base/tables.h:8), which is not the reciprocal of its BOHR_TO_ANGSTROM.
"""
import math

import numpy as np

ANGSTROM_TO_BOHR = 1.8897259886

_TET = math.acos(-1.0 / 3.0)  # 109.4712 degrees


def water(r_oh=0.9572, angle_deg=104.52):
    """H2O with O at the origin and the H atoms in the xz plane."""
    h = math.radians(angle_deg) / 2.0
    x, z = r_oh * math.sin(h), r_oh * math.cos(h)
    return [8, 1, 1], np.array([[0.0, 0.0, 0.0], [x, 0.0, z], [-x, 0.0, z]])


def alkane(n_carbon, r_cc=1.54, r_ch=1.09):
    """All-trans n-CnH(2n+2): planar carbon zig-zag in the xy plane, tetrahedral
    angles everywhere, CH2 hydrogens symmetric above/below the plane, CH3 staggered.
    Atom order: all carbons, then the hydrogens of each carbon in turn."""
    half = _TET / 2.0
    dx, dy = r_cc * math.sin(half), r_cc * math.cos(half)
    C = np.array([[i * dx, (dy / 2.0) * (1 if i % 2 else -1), 0.0] for i in range(n_carbon)])
    Z = [6] * n_carbon
    H = []
    for i in range(n_carbon):
        c = C[i]
        # bisector pointing away from the backbone neighbours, in plane
        out = np.array([0.0, 1.0 if i % 2 else -1.0, 0.0])
        # the two out-of-plane hydrogens: tetrahedral about the in-plane bisector
        a = math.pi - _TET / 2.0  # angle between H-C and the backbone bisector (-out)
        for s in (1.0, -1.0):
            d = out * math.cos(_TET / 2.0) + np.array([0.0, 0.0, s * math.sin(_TET / 2.0)])
            H.append(c + r_ch * d / np.linalg.norm(d))
        if i == 0 or i == n_carbon - 1:
            # terminal methyl: third H along the backbone direction, staggered
            nb = C[1] if i == 0 else C[n_carbon - 2]
            bond = (c - nb) / np.linalg.norm(c - nb)
            # rotate the bond direction so the H-C-C angle is tetrahedral and
            # the H lies in the backbone plane (anti to the next-nearest carbon)
            perp = np.cross(np.array([0.0, 0.0, 1.0]), bond)
            perp /= np.linalg.norm(perp)
            side = -1.0 if np.dot(perp, out) > 0 else 1.0
            ang = math.pi - _TET
            d = bond * math.cos(ang) + side * perp * math.sin(ang)
            # keep the in-plane terminal H on the side opposite the out-of-plane pair
            if np.dot(d, out) > 0:
                d = bond * math.cos(ang) - side * perp * math.sin(ang)
            H.append(c + r_ch * d)
        del a
    Z += [1] * len(H)
    return Z, np.vstack([C, np.array(H)])


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def water_cluster(seed, n_min=1, n_max=16, n=None, min_oo=2.6):
    """Seeded water cluster: n ~ U{n_min..n_max} (unless n is given), O atoms
    uniform in a cube of side 3.1 n^(1/3) Angstrom with rejection of O-O < 2.6,
    rigid H2O (0.9572 / 104.52) with a uniform random orientation (unit
    quaternion from four normals).  RNG: numpy PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if n is None:
        n = int(rng.integers(n_min, n_max + 1))
    side = 3.1 * n ** (1.0 / 3.0)
    Zw, Xw = water()
    Xw = Xw - Xw[0]
    O = []
    tries = 0
    while len(O) < n:
        tries += 1
        if tries > 100000:
            raise RuntimeError("water_cluster: could not place %d molecules" % n)
        p = rng.uniform(0.0, side, size=3)
        if all(np.linalg.norm(p - q) >= min_oo for q in O):
            O.append(p)
    Z, X = [], []
    for p in O:
        R = _random_rotation(rng)
        Z += Zw
        X.append(p + Xw @ R.T)
    return Z, np.vstack(X)
