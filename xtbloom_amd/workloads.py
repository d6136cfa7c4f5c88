"""Seeded synthetic ragged batches for the five benchmark workloads.

All lengths are in bohr.  Every generator is deterministic for a given seed.

  water_batch          B x H2O: O at the origin, r(OH) = 1.81 + U(+-0.05),
                       angle 104.5 deg + U(+-3 deg), random rigid rotation.
  alkane               CnH2n+2 all-anti chain, C-C 2.89, C-H 2.06, tetrahedral
                       angles; backbone torsions 180 deg + N(0, sigma); then
                       N(0, 0.02) jitter on every coordinate.
  substituted alkane   with probability 0.5 one terminal CH3 hydrogen replaced
                       by OH (O-H 1.81) or NH2 (N-H 1.91).
  point charges        uniform in a spherical shell around the centroid,
                       values N(0, 0.3) shifted to sum to zero, fixed gamma.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

TETRAHEDRAL = np.arccos(-1.0 / 3.0)  # 109.47 deg
R_CC, R_CH, R_OH, R_NH = 2.89, 2.06, 1.81, 1.91
R_CO, R_CN = 2.70, 2.78  # C-O 1.43 A, C-N 1.47 A (not fixed by the workload text)
POINT_CHARGE_GAMMA = 999.0  # bare point charge (the reference's gamma999 convention)


@dataclass
class RaggedBatch:
    """Host-side ragged batch in the xtbloom_batch_t layout."""
    atom_offsets: np.ndarray            # int64 [B+1]
    atomic_numbers: np.ndarray          # int32 [sum N]
    positions: np.ndarray               # float64 [sum N, 3]
    molecular_charges: np.ndarray       # float64 [B]
    unpaired_electrons: np.ndarray      # int32 [B]
    point_offsets: np.ndarray | None = None   # int64 [B+1]
    point_positions: np.ndarray | None = None  # float64 [P, 3]
    point_charges: np.ndarray | None = None    # float64 [P]
    point_gammas: np.ndarray | None = None     # float64 [P]
    names: list = field(default_factory=list)

    @property
    def batch_size(self) -> int:
        return len(self.atom_offsets) - 1

    def system(self, i):
        a0, a1 = self.atom_offsets[i], self.atom_offsets[i + 1]
        out = dict(Z=self.atomic_numbers[a0:a1], xyz=self.positions[a0:a1],
                   charge=float(self.molecular_charges[i]), uhf=int(self.unpaired_electrons[i]))
        if self.point_offsets is not None:
            p0, p1 = self.point_offsets[i], self.point_offsets[i + 1]
            out.update(pc_xyz=self.point_positions[p0:p1], pc_q=self.point_charges[p0:p1],
                       pc_gamma=self.point_gammas[p0:p1])
        return out

    def subset(self, indices):
        systems = [self.system(i) for i in indices]
        return from_systems(systems, [self.names[i] for i in indices] if self.names else None)


def from_systems(systems, names=None) -> RaggedBatch:
    offs = np.zeros(len(systems) + 1, dtype=np.int64)
    offs[1:] = np.cumsum([len(s["Z"]) for s in systems])
    batch = RaggedBatch(
        atom_offsets=offs,
        atomic_numbers=np.concatenate([np.asarray(s["Z"], dtype=np.int32) for s in systems]),
        positions=np.concatenate([np.asarray(s["xyz"], dtype=np.float64).reshape(-1, 3)
                                  for s in systems]),
        molecular_charges=np.array([s.get("charge", 0.0) for s in systems], dtype=np.float64),
        unpaired_electrons=np.array([s.get("uhf", 0) for s in systems], dtype=np.int32),
        names=list(names) if names else [])
    if any("pc_q" in s for s in systems):
        poff = np.zeros(len(systems) + 1, dtype=np.int64)
        poff[1:] = np.cumsum([len(s.get("pc_q", [])) for s in systems])
        batch.point_offsets = poff
        batch.point_positions = np.concatenate(
            [np.asarray(s.get("pc_xyz", np.zeros((0, 3))), dtype=np.float64).reshape(-1, 3)
             for s in systems])
        batch.point_charges = np.concatenate(
            [np.asarray(s.get("pc_q", []), dtype=np.float64) for s in systems])
        batch.point_gammas = np.concatenate(
            [np.asarray(s.get("pc_gamma", []), dtype=np.float64) for s in systems])
    return batch


def _random_rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a*a+b*b-c*c-d*d, 2*(b*c-a*d), 2*(b*d+a*c)],
                     [2*(b*c+a*d), a*a-b*b+c*c-d*d, 2*(c*d-a*b)],
                     [2*(b*d-a*c), 2*(c*d+a*b), a*a-b*b-c*c+d*d]])


def water(rng) -> dict:
    r1 = 1.81 + rng.uniform(-0.05, 0.05)
    r2 = 1.81 + rng.uniform(-0.05, 0.05)
    ang = np.deg2rad(104.5 + rng.uniform(-3.0, 3.0))
    xyz = np.array([[0.0, 0.0, 0.0], [r1, 0.0, 0.0], [r2 * np.cos(ang), r2 * np.sin(ang), 0.0]])
    return dict(Z=np.array([8, 1, 1]), xyz=xyz @ _random_rotation(rng).T, charge=0.0, uhf=0)


def water_batch(batch=8, seed=0) -> RaggedBatch:
    rng = np.random.default_rng(seed)
    return from_systems([water(rng) for _ in range(batch)], [f"h2o_{i}" for i in range(batch)])


def _place(a, b, c, bond, angle, torsion):
    """Natural extension reference frame: atom d bonded to c with angle b-c-d and torsion a-b-c-d."""
    bc = c - b
    bc /= np.linalg.norm(bc)
    n = np.cross(b - a, bc)
    n /= np.linalg.norm(n)
    m = np.cross(n, bc)
    d2 = np.array([-bond * np.cos(angle), bond * np.sin(angle) * np.cos(torsion),
                   bond * np.sin(angle) * np.sin(torsion)])
    return c + d2[0] * bc + d2[1] * m + d2[2] * n


def alkane(k, rng, torsion_sigma_deg=15.0, jitter=0.02, substituent=None) -> dict:
    """CkH(2k+2); substituent in {None, 'OH', 'NH2'} replaces one terminal CH3 hydrogen."""
    assert k >= 2
    carbons = [np.zeros(3), np.array([R_CC, 0.0, 0.0])]
    half = np.pi - TETRAHEDRAL
    carbons.append(carbons[1] + R_CC * np.array([np.cos(half), np.sin(half), 0.0]))
    ref = np.array([0.0, 1.0, 0.0])
    for i in range(3, k):
        tors = np.pi + np.deg2rad(rng.normal(0.0, torsion_sigma_deg))
        carbons.append(_place(carbons[i - 3], carbons[i - 2], carbons[i - 1], R_CC,
                              TETRAHEDRAL, tors))
    carbons = carbons[:k]
    Z = [6] * k
    xyz = list(carbons)
    hydrogens = []  # (carbon index, position)
    for i in range(k):
        c = carbons[i]
        nbrs = [carbons[j] for j in (i - 1, i + 1) if 0 <= j < k]
        if len(nbrs) == 2:
            u = (c - nbrs[0]) / np.linalg.norm(c - nbrs[0]) + (c - nbrs[1]) / np.linalg.norm(c - nbrs[1])
            u /= np.linalg.norm(u)
            nrm = np.cross(nbrs[0] - c, nbrs[1] - c)
            nrm /= np.linalg.norm(nrm)
            half_hch = 0.5 * TETRAHEDRAL
            for sgn in (1.0, -1.0):
                hydrogens.append((i, c + R_CH * (np.cos(half_hch) * u + sgn * np.sin(half_hch) * nrm)))
        else:
            axis = (c - nbrs[0]) / np.linalg.norm(c - nbrs[0])
            other = carbons[i + 2] if i == 0 and k > 2 else (carbons[i - 2] if k > 2 else ref + c)
            e1 = (other - nbrs[0]) - np.dot(other - nbrs[0], axis) * axis
            if np.linalg.norm(e1) < 1e-8:
                e1 = np.cross(axis, [0.0, 0.0, 1.0])
            e1 /= np.linalg.norm(e1)
            e2 = np.cross(axis, e1)
            cz, sz = np.cos(np.pi - TETRAHEDRAL), np.sin(np.pi - TETRAHEDRAL)
            for j in range(3):
                phi = np.pi + j * 2.0 * np.pi / 3.0  # staggered, first H anti to the chain
                hydrogens.append((i, c + R_CH * (cz * axis + sz * (np.cos(phi) * e1 + np.sin(phi) * e2))))
    if substituent is not None:
        # replace the first hydrogen of the terminal (last) carbon: heavy atom along the C-H bond
        idx = next(j for j, (ci, _) in enumerate(hydrogens) if ci == k - 1)
        ci, hpos = hydrogens.pop(idx)
        c = carbons[ci]
        u = (hpos - c) / np.linalg.norm(hpos - c)
        heavy_z, r_cx, r_xh, nh = (8, R_CO, R_OH, 1) if substituent == "OH" else (7, R_CN, R_NH, 2)
        x = c + r_cx * u
        Z_extra, xyz_extra = [heavy_z], [x]
        prev = carbons[ci - 1]
        for tors in ((np.pi,) if nh == 1 else (2.0 * np.pi / 3.0, 4.0 * np.pi / 3.0)):
            xyz_extra.append(_place(prev, c, x, r_xh, np.deg2rad(108.0), tors))
            Z_extra.append(1)
    for _, h in hydrogens:
        Z.append(1)
        xyz.append(h)
    if substituent is not None:
        Z += Z_extra
        xyz += xyz_extra
    xyz = np.array(xyz) + rng.normal(0.0, jitter, size=(len(xyz), 3))
    return dict(Z=np.array(Z, dtype=np.int32), xyz=xyz, charge=0.0, uhf=0)


def point_charges(xyz, rng, count, r_min, r_max, sigma=0.3, gamma=POINT_CHARGE_GAMMA):
    centroid = np.asarray(xyz).mean(axis=0)
    d = rng.normal(size=(count, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(r_min ** 3, r_max ** 3, size=count) ** (1.0 / 3.0)
    q = rng.normal(0.0, sigma, size=count)
    q -= q.mean()
    return dict(pc_xyz=centroid + d * r[:, None], pc_q=q, pc_gamma=np.full(count, gamma))


def uniform_alkanes(batch=128, k=20, seed=0) -> RaggedBatch:
    rng = np.random.default_rng(seed)
    return from_systems([alkane(k, rng) for _ in range(batch)],
                        [f"C{k}H{2*k+2}_{i}" for i in range(batch)])


def ragged_alkanes(batch=512, kmin=2, kmax=38, seed=0, substitute=True, torsion_sigma_deg=15.0,
                   points=0, r_min=10.0, r_max=20.0) -> RaggedBatch:
    rng = np.random.default_rng(seed)
    systems, names = [], []
    for i in range(batch):
        k = int(rng.integers(kmin, kmax + 1))
        sub = None
        if substitute and rng.random() < 0.5:
            sub = "OH" if rng.random() < 0.5 else "NH2"
        s = alkane(k, rng, torsion_sigma_deg=torsion_sigma_deg, substituent=sub)
        if points:
            s.update(point_charges(s["xyz"], rng, points, r_min, r_max))
        systems.append(s)
        names.append(f"C{k}" + (f"-{sub}" if sub else ""))
    return from_systems(systems, names)


WORKLOADS = {
    # name: (generator, kwargs); sizes are those of the benchmark configurations
    "water8": (water_batch, dict(batch=8)),
    "c20h42_128": (uniform_alkanes, dict(batch=128, k=20)),
    "ragged512": (ragged_alkanes, dict(batch=512)),
    "embedded256": (ragged_alkanes, dict(batch=256, kmin=6, kmax=26, substitute=True, points=500,
                                         r_min=10.0, r_max=20.0)),
    "long64": (ragged_alkanes, dict(batch=64, kmin=50, kmax=99, substitute=False,
                                    torsion_sigma_deg=20.0, points=1000, r_min=15.0, r_max=30.0)),
    # systems above 1024 AOs (1082-1322): the global-memory kernel instantiations
    "large4": (ragged_alkanes, dict(batch=4, kmin=180, kmax=220, substitute=False)),
}


def make(name, seed=0, **overrides) -> RaggedBatch:
    gen, kw = WORKLOADS[name]
    kw = dict(kw)
    kw.update(overrides)
    return gen(seed=seed, **kw)
