"""Deterministic synthetic solvated systems for the benchmark configurations.

Builds OpenMM ``System``/``State`` objects (the engine's input contract) with
no network and no force-field files:

* an AMBER-typed linear chain: random-walk heavy-atom backbone (C-C 0.153 nm,
  ~111 degree angles, random torsions with n in {1,2,3} and k in [0,10]
  kJ/mol), 0-3 hydrogens per heavy atom bonded at 0.109 nm and constrained
  (H-bond clusters of 1-3 constraints, 2-4 atoms), sigma in [0.30,0.40] nm,
  epsilon in [0.2,0.9] kJ/mol, charges drawn from N(0, 0.3) -- hydrogens
  positive as in AMBER (the draw's magnitude, at most 0.45 e) -- and made
  neutral over the heavy atoms;
  1-2 and 1-3 pairs excluded, 1-4 pairs scaled 1/1.2 (Coulomb) and 1/2 (LJ);
* an optional 24-atom ligand built the same way (GAFF-like random parameters);
* TIP3P water (O-H 0.09572 nm, H-H 0.15139 nm, q -0.834/+0.417,
  sigma_O 0.315075 nm, eps_O 0.635968 kJ/mol) at 1.0 g/cm^3 on a jittered
  lattice that avoids the solute, rigid (SETTLE);
* Na+/Cl- ions;
* cubic, truncated-octahedral or rhombic-dodecahedral cells in OpenMM's
  reduced form; PME, 1.0 nm cutoff, Ewald tolerance 5e-4, dispersion
  correction, CMMotionRemover(1).

Bonded equilibrium values are the built geometry, so the only strain at the
start is between molecules; the engine's minimiser removes it.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from .state import OpenMMState
from .units import ewaldAlpha, pmeMinimumGridSize
from .system import (CMMotionRemover, HarmonicAngleForce, HarmonicBondForce, MonteCarloBarostat,
                     NonbondedForce, OpenMMSystem, PeriodicTorsionForce, PositionRestraintForce,
                     boxVectorsForShape)

WATER_DENSITY = 33.43  # molecules / nm^3 at 1.0 g/cm^3
TIP3P = dict(r_oh=0.09572, r_hh=0.15139, q_o=-0.834, q_h=0.417, sig_o=0.315075, eps_o=0.635968)
NA = dict(mass=22.99, q=1.0, sig=0.243928, eps=0.365846)
CL = dict(mass=35.45, q=-1.0, sig=0.440104, eps=0.4184)
SHAPE_VOLUME = {"cube": 1.0, "octahedron": 4.0 * math.sqrt(3.0) / 9.0, "dodecahedron": math.sqrt(0.5)}


@dataclass
class Built:
    system: OpenMMSystem
    state: OpenMMState
    description: dict


class _Topology:
    def __init__(self):
        self.pos: List[np.ndarray] = []
        self.mass: List[float] = []
        self.q: List[float] = []
        self.sig: List[float] = []
        self.eps: List[float] = []
        self.bonds: List[Tuple[int, int, float, float]] = []
        self.cons: List[Tuple[int, int, float]] = []
        self.angles: List[Tuple[int, int, int, float, float]] = []
        self.tors: List[Tuple[int, int, int, int, int, float, float]] = []
        self.exc: List[Tuple[int, int, float, float, float]] = []
        self.heavy: List[int] = []
        self.grid = {}
        self.tpos: List[Tuple[float, float, float]] = []  # pos as Python floats, for the distance scans

    def add(self, p, m, q, s, e):
        self.pos.append(np.asarray(p, dtype=np.float64))
        self.tpos.append((float(self.pos[-1][0]), float(self.pos[-1][1]), float(self.pos[-1][2])))
        self.mass.append(m)
        self.q.append(q)
        self.sig.append(s)
        self.eps.append(e)
        return len(self.pos) - 1

    @property
    def n(self):
        return len(self.pos)

    def truncate(self, n: int, n_cons: int, key):
        """Drop the atoms from index n on and the constraints from n_cons on."""
        for i in range(n, len(self.pos)):
            self.grid[key(self.pos[i])].remove(i)
        for lst in (self.pos, self.tpos, self.mass, self.q, self.sig, self.eps):
            del lst[n:]
        del self.cons[n_cons:]


def _angle(a, b, c):
    u, v = a - b, c - b
    cs = np.dot(u, v) / (np.linalg.norm(u) * np.linalg.norm(v))
    return float(math.acos(max(-1.0, min(1.0, cs))))


def _dihedral(p0, p1, p2, p3):
    b1, b2, b3 = p1 - p0, p2 - p1, p3 - p2
    n1, n2 = np.cross(b1, b2), np.cross(b2, b3)
    x = np.dot(n1, n2)
    y = np.linalg.norm(b2) * np.dot(b1, n2)
    return float(math.atan2(y, x))


def _chain(top: _Topology, rng, n_atoms: int, center, radius: float, wall: float):
    """A self-avoiding random-walk chain of exactly n_atoms (heavy + H).

    Heavy atoms are kept 0.30 nm from every atom they are not bonded or
    angled to where the walk can manage it (else the roomiest of 600 tries);
    an atom's hydrogens are placed once both its heavy neighbours are known,
    in the sp3 slots those bonds leave (two on an inner atom, three on an
    end, turned to the roomiest of 12 orientations), so every H-C-X angle
    starts near 109.5 degrees.  When no try leaves 0.27 nm around the new
    atom and 0.21 nm around the hydrogen slots it fixes, the walk steps back
    (one atom, then more if the same step keeps failing) and grows again from
    there: looser limits leave hydrogens inside the 0.3 nm LJ cores of other
    chain atoms, which minimisation turns into strained, unstable geometry."""
    start = top.n
    heavy: List[int] = []
    hs_of = {}
    cell = 0.3

    def key(p):
        return (int(math.floor(p[0] / cell)), int(math.floor(p[1] / cell)), int(math.floor(p[2] / cell)))

    def nearest(p, skip):
        k = key(p)
        best = 1.0
        p0, p1, p2 = float(p[0]), float(p[1]), float(p[2])
        tpos, grid = top.tpos, top.grid
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get((k[0] + dx, k[1] + dy, k[2] + dz), ()):
                        if j not in skip:
                            # (d0^2 + d1^2) + d2^2: the order numpy's sum of three takes
                            a = tpos[j]
                            d0, d1, d2 = a[0] - p0, a[1] - p1, a[2] - p2
                            d = d0 * d0 + d1 * d1 + d2 * d2
                            if d < best:
                                best = d
        return math.sqrt(best)

    def clash(p, skip, limit=0.25):
        return nearest(p, skip) < limit

    def put(p, m, s, e):
        i = top.add(p, m, 0.0, s, e)
        top.grid.setdefault(key(p), []).append(i)
        return i

    counts = []
    total = 0
    while total < n_atoms:
        h = int(rng.choice([0, 1, 1, 2, 2]))
        h = min(h, n_atoms - total - 1)
        counts.append(h)
        total += 1 + h
    center = np.asarray(center, dtype=np.float64)
    prev_dir = rng.normal(size=3)
    prev_dir /= np.linalg.norm(prev_dir)
    cos_t, sin_t = math.cos(math.radians(109.5)), math.sin(math.radians(109.5))
    half = math.radians(0.5 * 109.5)

    def unit(v):
        return v / np.linalg.norm(v)

    def any_perpendicular(b):
        e = np.cross(b, rng.normal(size=3))
        return unit(e)

    def inner_slots(c, pa, pb):
        """The two sp3 directions an atom at c keeps free between bonds to
        pa and pb: off the bisector, in the plane normal to the bonds."""
        b0, b1 = unit(pa - c), unit(pb - c)
        u = -(b0 + b1)
        if np.linalg.norm(u) < 1e-3:
            u = any_perpendicular(b0)
        u = unit(u)
        w = np.cross(b0, b1)
        w = unit(w) if np.linalg.norm(w) > 1e-3 else unit(np.cross(u, b0))
        return [math.cos(half) * u + math.sin(half) * w, math.cos(half) * u - math.sin(half) * w]

    def slot_room(i, count, pa, pb):
        """Clearance of the `count` roomiest hydrogen slots atom i would get
        with heavy neighbours at pa and pb."""
        c = top.pos[i]
        room = sorted((nearest(c + 0.109 * s, {i, heavy[-2]}) for s in inner_slots(c, pa, pb)), reverse=True)
        return room[count - 1]

    def place_hydrogens(i, count, nbrs):
        """`count` hydrogens on heavy atom i in sp3 slots: the two slots
        off the bisector of an inner atom's bonds, or the three slots of
        the 109.5-degree cone around an end atom's bond; the roomiest
        slots win."""
        c = top.pos[i]
        bonds = [unit(top.pos[j] - c) for j in nbrs]
        if len(bonds) == 2:
            slots = inner_slots(c, top.pos[nbrs[0]], top.pos[nbrs[1]])
        else:
            b = bonds[0] if bonds else unit(rng.normal(size=3))
            e1 = any_perpendicular(b)
            e2 = np.cross(b, e1)
            phi0 = rng.uniform(0, 2 * math.pi)
            skip0 = {i} | set(nbrs)
            best_slots, best_room = None, -1.0
            for turn in range(12):
                phi = phi0 + turn * math.pi / 18  # 10-degree turns over one third of a circle
                cand = [cos_t * b + sin_t * (math.cos(phi + t) * e1 + math.sin(phi + t) * e2)
                        for t in (0.0, 2 * math.pi / 3, 4 * math.pi / 3)]
                rooms = sorted((nearest(c + 0.109 * s, skip0) for s in cand), reverse=True)
                if rooms[count - 1] > best_room:
                    best_slots, best_room = cand, rooms[count - 1]
            slots = best_slots
        skip = {i} | set(nbrs)
        room = [nearest(c + 0.109 * s, skip) for s in slots]
        hs = []
        for s in sorted(range(len(slots)), key=lambda t: -room[t])[:count]:
            hs.append(put(c + 0.109 * slots[s], 1.008, 0.2649, 0.0657))
            top.cons.append((i, hs[-1], 0.109))
        hs_of[i] = hs

    marks = []  # (atoms, constraints, walk direction) before each step
    fails = {}
    backtracks = 0
    k = 0
    while k < len(counts):
        mark = (top.n, len(top.cons), prev_dir.copy())
        if k < len(marks):
            marks[k] = mark
        else:
            marks.append(mark)
        if k == 0:
            # the first atom: near the centre, clear of anything already built
            p, room = None, -1.0
            for attempt in range(400):
                q = center + rng.normal(0, 0.05 + 0.005 * attempt, 3)
                r = nearest(q, set())
                if r > room:
                    p, room = q, r
                if r >= 0.3:
                    break
        else:
            last = heavy[-1]
            skip = {last}
            if len(heavy) > 1:
                skip |= {heavy[-2]} | set(hs_of[heavy[-2]])
            # ~111 degree bonds inside the sphere first; then outside it;
            # then any bend from 90 to 150 degrees; else the roomiest try.
            best, p, room = None, None, -1.0
            inmost, inmost_d = None, float("inf")
            for attempt in range(600):
                perp = rng.normal(size=3)
                perp -= np.dot(perp, prev_dir) * prev_dir
                perp /= np.linalg.norm(perp)
                bend = 111.0 + rng.normal(0, 3) if attempt < 400 else rng.uniform(90, 150)
                theta = math.radians(180.0 - bend)
                d = math.cos(theta) * prev_dir + math.sin(theta) * perp
                q = top.pos[last] + 0.153 * d
                outward = np.linalg.norm(q - center)
                if outward < inmost_d:
                    inmost, inmost_d = d, outward
                # a hard wall keeps every periodic image of the chain apart
                if outward > wall or (attempt < 200 and outward > radius):
                    continue
                r = nearest(q, skip)
                # the bond also fixes where the last atom's hydrogens go
                if counts[k - 1] and len(heavy) > 1 and r > room:
                    r = min(r, slot_room(last, counts[k - 1], top.pos[heavy[-2]], q) + 0.06)
                if r > room:
                    p, best, room = q, d, r
                if r >= 0.30:
                    break
            if room < 0.27 and k >= 2 and backtracks < len(counts) // 2:
                # no room here: undo the last `depth` steps and grow again
                backtracks += 1
                fails[k] = fails.get(k, 0) + 1
                depth = min(k - 1, 1 + fails[k] // 8)
                n0, c0, d0 = marks[k - depth]
                top.truncate(n0, c0, key)
                for j in heavy[k - depth - 1:k - 1]:
                    hs_of.pop(j, None)  # their hydrogens went in with the undone steps
                del heavy[k - depth:]
                prev_dir = d0
                k -= depth
                continue
            if best is None:  # every try crossed the wall: turn inwards
                best = inmost
                p = top.pos[last] + 0.153 * best
            prev_dir = best / np.linalg.norm(best)
        i = put(p, 12.011, float(rng.uniform(0.30, 0.40)), float(rng.uniform(0.2, 0.9)))
        heavy.append(i)
        # the previous atom now has both neighbours: its hydrogens go in
        if k >= 1:
            place_hydrogens(heavy[-2], counts[k - 1], heavy[max(0, k - 2):k - 1] + [i])
        k += 1
    place_hydrogens(heavy[-1], counts[-1], heavy[-2:-1])
    # bonded terms from the built geometry
    for a, b in zip(heavy[:-1], heavy[1:]):
        top.bonds.append((a, b, float(np.linalg.norm(top.pos[b] - top.pos[a])), 259408.0))
    for a, b, c in zip(heavy[:-2], heavy[1:-1], heavy[2:]):
        top.angles.append((a, b, c, _angle(top.pos[a], top.pos[b], top.pos[c]), 418.4))
    for k, i in enumerate(heavy):
        nbh = [heavy[j] for j in (k - 1, k + 1) if 0 <= j < len(heavy)]
        hs = hs_of[i]
        for h in hs:
            for j in nbh:
                top.angles.append((h, i, j, _angle(top.pos[h], top.pos[i], top.pos[j]), 292.88))
        for x in range(len(hs)):
            for y in range(x + 1, len(hs)):
                top.angles.append((hs[x], i, hs[y], _angle(top.pos[hs[x]], top.pos[i], top.pos[hs[y]]), 276.14))
    for a, b, c, d in zip(heavy[:-3], heavy[1:-2], heavy[2:-1], heavy[3:]):
        n = int(rng.integers(1, 4))
        top.tors.append((a, b, c, d, n, float(rng.uniform(0, math.pi)), float(rng.uniform(0, 10))))
    # charges: N(0, 0.3) draws, with AMBER's sign convention for hydrogens --
    # a hydrogen takes the magnitude of its draw, capped at 0.45 e (AMBER's
    # hydrogens are all positive, at most ~0.45 e) -- neutralised over the
    # heavy atoms.  A negatively charged hydrogen has nothing to keep a water
    # hydrogen (no LJ) away but the water oxygen's repulsion; such pairs sat
    # 0.11 nm apart in stiff wells that 2 fs steps could not integrate.
    idx = np.arange(start, top.n)
    q = rng.normal(0, 0.3, idx.shape[0])
    is_h = np.array([top.mass[i] < 1.5 for i in idx])
    q[is_h] = np.minimum(np.abs(q[is_h]), 0.45)
    heavy_k = ~is_h
    q[heavy_k] -= q.sum() / max(1, int(heavy_k.sum()))
    for k, i in enumerate(idx):
        top.q[i] = float(q[k])
    top.heavy.extend(heavy)
    # exclusions (1-2, 1-3) and scaled 1-4 pairs from the bond graph
    adj = {int(i): set() for i in idx}
    for a, b, *_ in top.bonds:
        if a >= start:
            adj[a].add(b)
            adj[b].add(a)
    for a, b, _ in top.cons:
        if a >= start:
            adj[a].add(b)
            adj[b].add(a)
    for i in idx:
        i = int(i)
        one2 = adj[i]
        one3 = set().union(*[adj[j] for j in one2]) - {i} - one2
        one4 = set().union(*[adj[j] for j in one3]) - {i} - one2 - one3
        for j in sorted(one2 | one3):
            if j > i:
                top.exc.append((i, j, 0.0, 1.0, 0.0))
        for j in sorted(one4):
            if j > i:
                sig = 0.5 * (top.sig[i] + top.sig[j])
                eps = 0.5 * math.sqrt(top.eps[i] * top.eps[j])
                top.exc.append((i, j, top.q[i] * top.q[j] / 1.2, sig, eps))
    return heavy


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def build(n_chain: int, n_waters: int, n_ions: int = 0, shape: str = "cube", seed: int = 0,
          n_ligand: int = 0, restraint_k: float = 0.0, barostat: bool = False,
          grid: Optional[Tuple[int, int, int]] = None) -> Built:
    """A solvated chain at 1.0 g/cm^3 in a cell of the given shape."""
    rng = np.random.default_rng(seed)
    solute_volume = 0.0125 * (n_chain + n_ligand)
    volume = n_waters / WATER_DENSITY + solute_volume + n_ions * 0.03
    width = (volume / SHAPE_VOLUME[shape]) ** (1.0 / 3.0)
    box = np.asarray(boxVectorsForShape(shape, width))
    center = 0.5 * (box[0] + box[1] + box[2])
    # half the narrowest perpendicular width: volume over the largest face
    inradius = 0.5 * volume / max(np.linalg.norm(np.cross(box[1], box[2])), np.linalg.norm(np.cross(box[2], box[0])),
                                  np.linalg.norm(np.cross(box[0], box[1])))
    top = _Topology()
    if n_chain:
        wall = inradius - 0.35
        r = max(0.8, min(0.85 * wall, (3 * solute_volume * 2.5 / (4 * math.pi)) ** (1 / 3)))
        _chain(top, rng, n_chain, center, r, wall)
    if n_ligand:
        _chain(top, rng, n_ligand, center + np.array([0.8, 0.0, 0.0]), 0.4, inradius - 0.35)
    solute = np.array(top.pos) if top.n else np.zeros((0, 3))
    # lattice sites in fractional coordinates, jittered, away from the solute
    need = n_waters + n_ions
    inv = np.linalg.inv(box)
    from scipy.spatial import cKDTree
    tree = cKDTree(solute) if solute.shape[0] else None
    factor = 1.15
    while True:
        per = (need * factor / (1.0 if tree is None else 1.0)) ** (1 / 3)
        widths = [volume / np.linalg.norm(np.cross(box[(a + 1) % 3], box[(a + 2) % 3])) for a in range(3)]
        mean_w = float(np.mean(widths))
        counts = [max(2, int(math.ceil(per * w / mean_w))) for w in widths]
        g = np.stack(np.meshgrid(*[(np.arange(c) + 0.5) / c for c in counts], indexing="ij"), -1).reshape(-1, 3)
        g = g + rng.uniform(-0.04, 0.04, g.shape) / np.asarray(counts)
        sites = g @ box
        if tree is not None:
            # distance to the nearest solute atom, including periodic images
            keep = np.ones(sites.shape[0], dtype=bool)
            for ia in (-1, 0, 1):
                for ib in (-1, 0, 1):
                    for ic in (-1, 0, 1):
                        shift = ia * box[0] + ib * box[1] + ic * box[2]
                        d, _ = tree.query(sites + shift, distance_upper_bound=0.32)
                        keep &= ~np.isfinite(d)
            sites = sites[keep]
        if sites.shape[0] >= need:
            break
        factor *= 1.1
    order = rng.permutation(sites.shape[0])
    sites = sites[order[:need]]
    # ions first (alternating Na/Cl), then water
    for k in range(n_ions):
        ion = NA if k % 2 == 0 else CL
        top.add(sites[k], ion["mass"], ion["q"], ion["sig"], ion["eps"])
    half_hh = 0.5 * TIP3P["r_hh"]
    h_height = math.sqrt(TIP3P["r_oh"] ** 2 - half_hh ** 2)
    local = np.array([[0.0, 0.0, 0.0], [half_hh, h_height, 0.0], [-half_hh, h_height, 0.0]])
    local -= (15.999 * local[0] + 1.008 * local[1] + 1.008 * local[2]) / 18.015
    water0 = top.n
    for k in range(n_waters):
        R = _random_rotation(rng)
        p = sites[n_ions + k] + local @ R.T
        o = top.add(p[0], 15.999, TIP3P["q_o"], TIP3P["sig_o"], TIP3P["eps_o"])
        h1 = top.add(p[1], 1.008, TIP3P["q_h"], 0.0, 0.0)
        h2 = top.add(p[2], 1.008, TIP3P["q_h"], 0.0, 0.0)
        top.cons += [(o, h1, TIP3P["r_oh"]), (o, h2, TIP3P["r_oh"]), (h1, h2, TIP3P["r_hh"])]
        top.exc += [(o, h1, 0.0, 1.0, 0.0), (o, h2, 0.0, 1.0, 0.0), (h1, h2, 0.0, 1.0, 0.0)]
    n = top.n
    pos = np.array(top.pos).reshape(-1)
    ex = np.array(top.exc, dtype=object) if top.exc else np.zeros((0, 5))

    def col(rows, k, dtype):
        return np.array([r[k] for r in rows], dtype=dtype)

    nb = NonbondedForce(
        charge=np.array(top.q), sigma=np.array(top.sig), epsilon=np.array(top.eps),
        exceptionP1=col(top.exc, 0, np.int32), exceptionP2=col(top.exc, 1, np.int32),
        exceptionChargeProd=col(top.exc, 2, np.float64), exceptionSigma=col(top.exc, 3, np.float64),
        exceptionEpsilon=col(top.exc, 4, np.float64), cutoff=1.0, ewaldTolerance=5e-4,
        grid=tuple(grid) if grid else (0, 0, 0), useDispersionCorrection=True, recipForceGroup=1)
    forces = [
        HarmonicBondForce(col(top.bonds, 0, np.int32), col(top.bonds, 1, np.int32),
                          col(top.bonds, 2, np.float64), col(top.bonds, 3, np.float64)),
        HarmonicAngleForce(col(top.angles, 0, np.int32), col(top.angles, 1, np.int32),
                           col(top.angles, 2, np.int32), col(top.angles, 3, np.float64),
                           col(top.angles, 4, np.float64)),
        PeriodicTorsionForce(col(top.tors, 0, np.int32), col(top.tors, 1, np.int32), col(top.tors, 2, np.int32),
                             col(top.tors, 3, np.int32), col(top.tors, 4, np.int32),
                             col(top.tors, 5, np.float64), col(top.tors, 6, np.float64)),
        nb,
        CMMotionRemover(1),
    ]
    if restraint_k > 0 and top.heavy:
        hv = np.array(sorted(top.heavy), dtype=np.int32)
        p3 = pos.reshape(-1, 3)
        forces.append(PositionRestraintForce(hv, np.full(hv.shape[0], restraint_k), p3[hv, 0].copy(),
                                             p3[hv, 1].copy(), p3[hv, 2].copy()))
    if barostat:
        forces.append(MonteCarloBarostat(1.0, 300.0, 25, seed))
    bv = tuple(tuple(float(x) for x in r) for r in box)
    system = OpenMMSystem(boxVectors=bv, masses=np.array(top.mass),
                          constraintP1=col(top.cons, 0, np.int32), constraintP2=col(top.cons, 1, np.int32),
                          constraintDistance=col(top.cons, 2, np.float64), forces=forces)
    state = OpenMMState(boxVectors=bv, positions=pos, velocities=np.zeros_like(pos))
    desc = dict(atoms=n, chain=n_chain, ligand=n_ligand, waters=n_waters, ions=n_ions, shape=shape,
                width_nm=width, volume_nm3=float(abs(np.linalg.det(box))), seed=seed)
    return Built(system, state, desc)


def config(index: int, seed: Optional[int] = None) -> Built:
    """The benchmark configurations, sized to their stated atom counts.

    0: 2,373 atoms: a 304-atom chain + 689 TIP3P + 2 ions, cube (~2.9 nm),
       PME mesh 32^3
    1: 26,749 atoms: a 2,600-atom chain + 24-atom ligand + 8 ions + 8,039
       TIP3P, truncated octahedron at 1.0 g/cm^3 (d ~ 7.0 nm; the stated
       6.5 nm does not hold the stated water at that density), mesh 64^3
    2: ~100k atoms: a 1,000-atom solute + 33,000 TIP3P, rhombic dodecahedron,
       barostat and heavy-atom restraints (k = 1000), mesh 96^3
    4: ~250k atoms: a 20,000-atom chain + 76,667 TIP3P, cube, mesh 128^3
    """
    s = 0x20261202 + index if seed is None else seed
    if index == 0:
        return build(304, 689, 2, "cube", s, grid=(32, 32, 32))
    if index == 1:
        return build(2600, 8039, 8, "octahedron", s, n_ligand=24, grid=(64, 64, 64))
    if index == 2:
        return build(1000, 33000, 0, "dodecahedron", s, restraint_k=1000.0, barostat=True, grid=(96, 96, 96))
    if index == 4:
        return build(20000, 76667, 0, "cube", s, grid=(128, 128, 128))
    raise ValueError(f"no configuration {index} (the 64-replica ensemble is `ensemble()`)")


@dataclass
class Replica:
    built: Built
    seed: int
    temperature: float


def ensemble_plan(count: int = 64, seed: Optional[int] = None, low: int = 2000, high: int = 30000):
    """Sizes, shapes, seeds and temperatures of the batched ensemble: atom
    counts log-uniform on [low, high], cubic and octahedral cells
    alternating, per-replica seeds, temperatures uniform on [290, 320] K."""
    s = 0x20261202 + 3 if seed is None else seed
    rng = np.random.default_rng(s)
    sizes = np.exp(rng.uniform(np.log(low), np.log(high), count))
    temps = rng.uniform(290.0, 320.0, count)
    seeds = rng.integers(1, 2 ** 31 - 1, count)
    plan = []
    for k in range(count):
        n = int(sizes[k])
        chain = max(40, int(round(0.1 * n)))
        ions = 2 * max(1, n // 10000)
        waters = max(10, (n - chain - ions) // 3)
        shape = "cube" if k % 2 == 0 else "octahedron"
        plan.append(dict(n_chain=chain, n_waters=waters, n_ions=ions, shape=shape,
                         seed=int(seeds[k]), temperature=float(temps[k]),
                         grid=ensemble_mesh(chain, waters, ions, shape)))
    return plan


def ensemble_mesh(n_chain: int, n_waters: int, n_ions: int, shape: str) -> Tuple[int, int, int]:
    """A replica's PME mesh: n^3 with n the smallest multiple of 8 at or above
    the Ewald minimum for the cell's longest axis at tolerance 5e-4 (1.0 nm
    cutoff), kept within 32-64 -- 32^3 for the smallest replicas, 56^3 for the
    largest (finer than the minimum, never coarser)."""
    volume = n_waters / WATER_DENSITY + 0.0125 * n_chain + n_ions * 0.03
    width = (volume / SHAPE_VOLUME[shape]) ** (1.0 / 3.0)
    alpha = ewaldAlpha(1.0, 5e-4)
    n = pmeMinimumGridSize(alpha, width, 5e-4)
    n = min(64, max(32, 8 * math.ceil(n / 8)))
    return (n, n, n)


def build_planned(p: dict) -> Built:
    """One replica of ``ensemble_plan`` (module level, so worker processes
    can build replicas in parallel)."""
    return build(p["n_chain"], p["n_waters"], p["n_ions"], p["shape"], p["seed"], grid=p["grid"])


def ensemble(count: int = 64, seed: Optional[int] = None, low: int = 2000, high: int = 30000):
    """The batched ensemble: ``count`` independent solvated chains (see
    ``ensemble_plan``), each with its own seed and temperature."""
    out = []
    for p in ensemble_plan(count, seed, low, high):
        b = build(p["n_chain"], p["n_waters"], p["n_ions"], p["shape"], p["seed"], grid=p["grid"])
        out.append(Replica(b, p["seed"], p["temperature"]))
    return out
