"""Seeded synthetic systems for the benchmark workloads.

This replaces the reference's setup path (BoxBuilder.cu:5-145, 248-276;
CompoundBuilder.cpp:7-49, 225-318) with a generator of chains and single-site
solvent, producing the same data the reference's step consumes:

* chains are self-avoiding random walks (bond 0.15 nm, bond angle 109.5 deg,
  uniform torsion), cut into compounds of 128 particles; a placement is
  rejected when closer than 0.30 nm to any bead outside its 1-2/1-3/1-4 set;
* compound particles are stored re-imaged around their particle 0, with
  particle 0 inside the box (as the compound kernel leaves them,
  Engine.cu:993-996); key particle and confining sphere follow
  Compound::calcParticleSphere (Bodies.cuh:407-423);
* bonded terms inside a compound use compound-local indices; terms spanning
  two compounds go to the bridge between them, with bridge-local indices and
  particles added in first-seen order (CompoundBridge::addBondParticles,
  Bodies.cuh:551-557);
* every bonded tuple adds mutual LJ exclusions (distributeLJIgnores,
  CompoundBuilder.cpp:307-318);
* solvent sits on a jittered cubic lattice at 33 nm^-3, skipping sites within
  0.23 nm of a bead; velocities are Maxwell at T0 and prev = x - v dt.
Type 0 is the solvent type (ATOMTYPE_SOL, Forcefield.cuh:15).
"""
from __future__ import annotations

import math

import numpy as np

from .abi import ANGLE_DTYPE, DIHEDRAL_DTYPE, PAIR_DTYPE, LimaSystem

GAS_R = 8.3145                   # J/(mol K), Engine.cu:356
BOND_B0 = 0.15                   # nm
BOND_KB = 2.5e8                  # J/mol/nm^2  (2.5e5 kJ/mol/nm^2)
ANGLE_THETA0 = float(np.float32(math.radians(109.5)))
ANGLE_K = 4.0e5                  # J/mol/rad^2 (400 kJ/mol/rad^2)
DIHEDRAL_K = 1000.0              # J/mol
DIHEDRAL_N = 3
DIHEDRAL_PHI0 = 0.0
SOLVENT_PARAMS = (0.018, 0.3166, 650.0)   # kg/mol, nm, J/mol
EXCLUSION_DIST = 0.30            # nm, bead-bead self-avoidance of the chain walk
# Lattice sites closer than this to a bead are skipped.  0.23 nm reproduces the
# stated solvent counts of the workloads at 33 nm^-3 (~24,000 in the 10 nm box
# with 12,800 beads, ~160,000 in the 18 nm box with 51,200 beads).
SOLVENT_EXCLUSION = 0.23         # nm
DT = 2e-6                        # ns (2 fs)
COMPOUND_SIZE = 128


def make_forcefield(rng: np.random.Generator, n_types: int) -> np.ndarray:
    ff = np.zeros((32, 3), dtype=np.float32)
    ff[:, :] = (0.012, 0.32, 100.0)          # unused slots: benign values
    ff[0] = SOLVENT_PARAMS
    ff[1:n_types + 1, 0] = rng.uniform(12.0, 16.0, n_types) * 1e-3
    ff[1:n_types + 1, 1] = rng.uniform(0.30, 0.35, n_types)
    ff[1:n_types + 1, 2] = rng.uniform(0.1, 0.5, n_types) * 1000.0
    return ff


class _Grid:
    """Periodic cell hash for the self-avoidance test."""

    def __init__(self, L: float, h: float):
        self.L = L
        self.n = max(1, int(L // h))
        self.cells: dict = {}
        self._nbr: dict = {}

    def key(self, p):
        # any consistent cell of edge >= h serves: the 3x3x3 block then holds every bead within h
        L, n = self.L, self.n
        return (int(math.floor((float(p[0]) % L) / L * n)) % n, int(math.floor((float(p[1]) % L) / L * n)) % n,
                int(math.floor((float(p[2]) % L) / L * n)) % n)

    def add(self, idx, p):
        self.cells.setdefault(self.key(p), []).append(idx)

    def remove(self, idx, p):
        self.cells[self.key(p)].remove(idx)

    def _around(self, key):
        """The distinct cells of the 3x3x3 block around a cell, in a fixed order (cached)."""
        got = self._nbr.get(key)
        if got is None:
            cx, cy, cz = key
            rng = (-1, 0, 1) if self.n >= 3 else tuple(range(self.n))
            got = []
            for dx in rng:
                for dy in rng:
                    for dz in rng:
                        k = ((cx + dx) % self.n, (cy + dy) % self.n, (cz + dz) % self.n)
                        if k not in got:
                            got.append(k)
            self._nbr[key] = got
        return got

    def near(self, p):
        out = []
        cells = self.cells
        for k in self._around(self.key(p)):
            c = cells.get(k)
            if c:
                out.extend(c)
        return out


def _cross(a, b):
    """a x b of two 3-vectors, each component a_i*b_j - a_j*b_i as np.cross rounds it."""
    a0, a1, a2 = float(a[0]), float(a[1]), float(a[2])
    b0, b1, b2 = float(b[0]), float(b[1]), float(b[2])
    return np.array([a1 * b2 - a2 * b1, a2 * b0 - a0 * b2, a0 * b1 - a1 * b0])


def _frame(a, b, c):
    """The NeRF frame (bc, m, n) at c for the chain a-b-c."""
    bc = c - b
    bc /= np.linalg.norm(bc)
    n = _cross(b - a, bc)
    nn = np.linalg.norm(n)
    if nn < 1e-9:
        n = _cross(bc, np.array([1.0, 0.0, 0.0]))
        if np.linalg.norm(n) < 1e-9:
            n = _cross(bc, np.array([0.0, 1.0, 0.0]))
        nn = np.linalg.norm(n)
    n /= nn
    return bc, _cross(n, bc), n


def _place_in(c, frame, bond, angle, torsion):
    """_place with the frame of (a, b, c) given."""
    bc, m, n = frame
    d2 = np.array([-bond * math.cos(angle), bond * math.sin(angle) * math.cos(torsion),
                   bond * math.sin(angle) * math.sin(torsion)])
    return c + d2[0] * bc + d2[1] * m + d2[2] * n


def _place(a, b, c, bond, angle, torsion):
    """NeRF: position of d given a, b, c, |cd|, angle bcd and torsion abcd."""
    bc = c - b
    bc /= np.linalg.norm(bc)
    n = _cross(b - a, bc)
    nn = np.linalg.norm(n)
    if nn < 1e-9:
        n = _cross(bc, np.array([1.0, 0.0, 0.0]))
        if np.linalg.norm(n) < 1e-9:
            n = _cross(bc, np.array([0.0, 1.0, 0.0]))
        nn = np.linalg.norm(n)
    n /= nn
    m = _cross(n, bc)
    d2 = np.array([-bond * math.cos(angle), bond * math.sin(angle) * math.cos(torsion),
                   bond * math.sin(angle) * math.sin(torsion)])
    return c + d2[0] * bc + d2[1] * m + d2[2] * n


def _random_unit(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def _walk_chains(rng, n_chains, n_beads_chain, L, confine, compound_size):
    """Self-avoiding walks (unwrapped coordinates), one per chain."""
    pos = np.zeros((n_chains * n_beads_chain, 3))
    grid = _Grid(L, EXCLUSION_DIST)
    excl2 = EXCLUSION_DIST ** 2

    conf2_lo = conf2_hi = 0.0
    if confine is not None:
        conf2_lo, conf2_hi = confine * confine * (1 - 1e-9), confine * confine * (1 + 1e-9)

    def ok(p, gidx, chain_start, comp_first):
        if confine is not None and comp_first is not None:
            # |p - first| > confine, decided in plain floats away from the bound (np.linalg.norm at it)
            dx, dy, dz = p[0] - comp_first[0], p[1] - comp_first[1], p[2] - comp_first[2]
            r2 = dx * dx + dy * dy + dz * dz
            if r2 > conf2_hi or (r2 >= conf2_lo and np.linalg.norm(p - comp_first) > confine):
                return False
        cand = [j for j in grid.near(p) if not (chain_start <= j and gidx - 3 <= j < gidx)]
        if not cand:
            return True
        d = pos[cand] - p
        d -= L * np.round(d / L)
        return bool(np.min(np.einsum("ij,ij->i", d, d)) >= excl2)

    for ch in range(n_chains):
        start = ch * n_beads_chain
        restarts = 0
        while True:
            placed = 0
            fails = 0
            while placed < n_beads_chain:
                g = start + placed
                comp_first = None
                if placed % compound_size != 0:
                    comp_first = pos[start + (placed // compound_size) * compound_size]
                success = False
                frame = _frame(pos[g - 3], pos[g - 2], pos[g - 1]) if placed >= 3 else None
                for _ in range(40):
                    if placed == 0:
                        p = rng.uniform(0, L, 3)
                    elif placed == 1:
                        p = pos[g - 1] + BOND_B0 * _random_unit(rng)
                    elif placed == 2:
                        a = pos[g - 1] + _random_unit(rng)
                        p = _place(a, pos[g - 2], pos[g - 1], BOND_B0, ANGLE_THETA0, rng.uniform(0, 2 * math.pi))
                    else:
                        p = _place_in(pos[g - 1], frame, BOND_B0, ANGLE_THETA0, rng.uniform(0, 2 * math.pi))
                    if ok(p, g, start, comp_first):
                        success = True
                        break
                if success:
                    pos[g] = p
                    grid.add(g, p)
                    placed += 1
                    continue
                fails += 1
                if fails > 200 or placed < 4:
                    break
                back = min(placed - 1, 4 + int(rng.integers(0, 6)))
                for _ in range(back):
                    placed -= 1
                    grid.remove(start + placed, pos[start + placed])
            if placed == n_beads_chain:
                break
            for k in range(placed):
                grid.remove(start + k, pos[start + k])
            restarts += 1
            if restarts > 50:
                raise RuntimeError("could not place chain; box too crowded")
    return pos


def _compound_sphere(p: np.ndarray):
    """Compound::calcParticleSphere (Bodies.cuh:407-423) in FP32."""
    p = p.astype(np.float32)
    n = len(p)
    com = np.zeros(3, dtype=np.float32)
    inv = np.float32(1.0) / np.float32(n)
    for i in range(n):
        com = (com + (p[i] * inv).astype(np.float32)).astype(np.float32)
    d = np.sqrt(((p - com).astype(np.float32) ** 2).sum(axis=1, dtype=np.float32)).astype(np.float32)
    return int(np.argmin(d)), float(d.max())


def build_system(seed: int, n_chains: int, compounds_per_chain: int, box_len: float, n_types: int,
                 n_solvents: int | None, T0: float = 300.0, dt: float = DT,
                 solvent_density: float = 33.0, compound_size: int = COMPOUND_SIZE,
                 forcefield: np.ndarray | None = None, confine: float | None = None,
                 central_first: bool = False, sphere_margin: float = 0.0,
                 auto_density: bool = False) -> LimaSystem:
    """Build one seeded system.  n_solvents=None keeps every free lattice site.

    confine: largest distance of a compound's particles from its first walk bead
    (default 0.38 L, less a margin in boxes under 5 nm).  central_first: number
    each compound from the bead nearest its centroid, so that the particle every
    other one is re-imaged against (Engine.cu:993-996) sits in the middle.
    sphere_margin: added to each compound's confining sphere (Bodies.cuh:407-423),
    which widens its neighbour-list radii for chains that swell during the run.
    auto_density: when fewer lattice sites than n_solvents are free of the chain, rebuild with a
    lattice 10% denser (same seed), until enough are."""
    rng = np.random.default_rng(seed)
    ff = make_forcefield(rng, n_types) if forcefield is None else forcefield.astype(np.float32)
    L = float(box_len)
    n_beads_chain = compounds_per_chain * compound_size
    # every compound stays well inside half a box of its particle 0, against which
    # its positions are re-imaged: 0.38 L, less a margin in boxes under 5 nm
    if confine is None:
        confine = 0.38 * L - max(0.0, 0.6 - 0.12 * L)
    walk = (_walk_chains(rng, n_chains, n_beads_chain, L, confine, compound_size) if n_chains > 0
            else np.zeros((0, 3)))
    n_chain = len(walk)
    # walk order -> particle index: identity, or each compound's central bead first
    w2g = np.arange(n_chain)
    if central_first:
        for c0 in range(0, n_chain, compound_size):
            seg = walk[c0:c0 + compound_size]
            k = int(np.argmin(((seg - seg.mean(axis=0)) ** 2).sum(axis=1)))
            order = [k] + [i for i in range(len(seg)) if i != k]
            w2g[c0 + np.array(order)] = c0 + np.arange(len(seg))
    walk_g = np.zeros_like(walk)
    walk_g[w2g] = walk
    types = rng.integers(1, n_types + 1, size=n_chain).astype(np.uint8)
    masses = ff[types, 0].astype(np.float64)

    nc = n_chains * compounds_per_chain
    coff = np.arange(nc + 1, dtype=np.int32) * compound_size
    cpos = np.zeros((n_chain, 3), dtype=np.float32)
    for c in range(nc):
        seg = walk_g[coff[c]:coff[c + 1]]
        p0 = np.mod(seg[0], L)
        cpos[coff[c]:coff[c + 1]] = (p0 + (seg - seg[0])).astype(np.float32)
    key = np.zeros(nc, dtype=np.int32)
    sphere = np.zeros(nc, dtype=np.float32)
    for c in range(nc):
        key[c], sphere[c] = _compound_sphere(cpos[coff[c]:coff[c + 1]])
        sphere[c] += np.float32(sphere_margin)

    # ---- topology: compound terms, bridges, exclusions ----
    comp_of = np.arange(n_chain) // compound_size
    local_of = np.arange(n_chain) % compound_size
    c_bonds = [[] for _ in range(nc)]
    c_angles = [[] for _ in range(nc)]
    c_dihs = [[] for _ in range(nc)]
    bridges = {}   # (left, right) -> dict(particles=[global ids], bonds, angles, dihs)
    ignores = [[] for _ in range(n_chain)]

    def add_term(kind, atoms):
        comps = sorted(set(int(comp_of[a]) for a in atoms))
        for i in atoms:
            for j in atoms:
                if i == j:
                    continue
                e = (int(comp_of[j]), int(local_of[j]))
                if e not in ignores[i]:
                    ignores[i].append(e)
        if len(comps) == 1:
            c = comps[0]
            loc = [int(local_of[a]) for a in atoms]
            {"b": c_bonds, "a": c_angles, "d": c_dihs}[kind][c].append(loc)
        else:
            br = bridges.setdefault((comps[0], comps[1]), {"particles": [], "b": [], "a": [], "d": []})
            for a in atoms:
                if a not in br["particles"]:
                    br["particles"].append(a)
            br[kind].append([br["particles"].index(a) for a in atoms])

    for ch in range(n_chains):
        s = ch * n_beads_chain
        e = s + n_beads_chain
        for i in range(s, e - 1):
            add_term("b", tuple(int(w2g[k]) for k in (i, i + 1)))
        for i in range(s, e - 2):
            add_term("a", tuple(int(w2g[k]) for k in (i, i + 1, i + 2)))
        for i in range(s, e - 3):
            add_term("d", tuple(int(w2g[k]) for k in (i, i + 1, i + 2, i + 3)))

    def pack(lists, kind):
        off = np.zeros(len(lists) + 1, dtype=np.int32)
        flat = []
        for k, lst in enumerate(lists):
            off[k + 1] = off[k] + len(lst)
            flat.extend(lst)
        if kind == "b":
            arr = np.zeros(len(flat), dtype=PAIR_DTYPE)
            arr["b0"], arr["kb"] = BOND_B0, BOND_KB
        elif kind == "a":
            arr = np.zeros(len(flat), dtype=ANGLE_DTYPE)
            arr["theta_0"], arr["k_theta"] = ANGLE_THETA0, ANGLE_K
        else:
            arr = np.zeros(len(flat), dtype=DIHEDRAL_DTYPE)
            arr["phi_0"], arr["k_phi"], arr["n"] = DIHEDRAL_PHI0, DIHEDRAL_K, DIHEDRAL_N
        if flat:
            arr["idx"] = np.array(flat, dtype=np.uint32)
        return off, arr

    bond_off, bonds = pack(c_bonds, "b")
    angle_off, angles = pack(c_angles, "a")
    dih_off, dihs = pack(c_dihs, "d")
    br_keys = sorted(bridges)
    br_list = [bridges[k] for k in br_keys]
    bpo = np.zeros(len(br_list) + 1, dtype=np.int32)
    bpc, bpl = [], []
    for k, br in enumerate(br_list):
        bpo[k + 1] = bpo[k] + len(br["particles"])
        bpc.extend(int(comp_of[a]) for a in br["particles"])
        bpl.extend(int(local_of[a]) for a in br["particles"])
    bb_off, bb = pack([br["b"] for br in br_list], "b")
    ba_off, ba = pack([br["a"] for br in br_list], "a")
    bd_off, bd = pack([br["d"] for br in br_list], "d")
    ig_off = np.zeros(n_chain + 1, dtype=np.int32)
    for i in range(n_chain):
        ig_off[i + 1] = ig_off[i] + len(ignores[i])
    ig_c = np.array([e[0] for lst in ignores for e in lst], dtype=np.int32)
    ig_l = np.array([e[1] for lst in ignores for e in lst], dtype=np.int32)

    # ---- solvent lattice ----
    density = solvent_density
    while True:
        nl = max(1, int(round(L * density ** (1.0 / 3.0))))
        a = L / nl
        g = (np.arange(nl) + 0.5) * a
        sites = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
        sites = sites + rng.uniform(-0.02, 0.02, size=sites.shape)
        sites = np.mod(sites, L)
        if n_chain > 0:
            from scipy.spatial import cKDTree
            wrapped = np.mod(cpos.astype(np.float64), L)
            wrapped[wrapped >= L] = 0.0
            tree = cKDTree(wrapped, boxsize=L)
            sites_c = sites.copy()
            sites_c[sites_c >= L] = 0.0
            d, _ = tree.query(sites_c, k=1)
            sites = sites[d >= SOLVENT_EXCLUSION]
        if n_solvents is None or n_solvents <= len(sites) or not auto_density:
            break
        density *= 1.1          # a finer lattice, until enough sites are free of the chain
    if n_solvents is not None:
        if n_solvents > len(sites):
            raise ValueError(f"only {len(sites)} free lattice sites for {n_solvents} solvents")
        pick = np.sort(rng.choice(len(sites), size=n_solvents, replace=False))
        sites = sites[pick]
    spos = sites.astype(np.float32)
    ns = len(spos)

    # ---- Maxwell velocities, prev = x - v dt ----
    vc = rng.normal(size=(n_chain, 3)) * np.sqrt(GAS_R * T0 / masses)[:, None]
    vs = rng.normal(size=(ns, 3)) * math.sqrt(GAS_R * T0 / float(ff[0, 0]))
    cprev = (cpos.astype(np.float64) - vc * dt).astype(np.float32)
    sprev = (spos.astype(np.float64) - vs * dt).astype(np.float32)

    sys_ = LimaSystem(
        box_len=L, dt=dt, thermostat_target=T0, forcefield=ff,
        compound_offset=coff, key_particle_index=key, confining_particle_sphere=sphere,
        chain_positions=cpos, chain_prev_positions=cprev, chain_atom_types=types,
        bond_offset=bond_off, bonds=bonds, angle_offset=angle_off, angles=angles,
        dihedral_offset=dih_off, dihedrals=dihs,
        ignore_offset=ig_off, ignore_compound=ig_c, ignore_local=ig_l,
        bridge_particle_offset=bpo, bridge_particle_compound=np.array(bpc, dtype=np.int32),
        bridge_particle_local=np.array(bpl, dtype=np.int32),
        bridge_bond_offset=bb_off, bridge_bonds=bb, bridge_angle_offset=ba_off, bridge_angles=ba,
        bridge_dihedral_offset=bd_off, bridge_dihedrals=bd,
        solvent_positions=spos, solvent_prev_positions=sprev,
        meta={"seed": seed, "n_chains": n_chains, "compounds_per_chain": compounds_per_chain,
              "n_types": n_types},
    )
    return sys_


# ---------------------------------------------------------------------------
# Named workloads.  seed = 0x51A7 + 1000*config + replica.
# ---------------------------------------------------------------------------
SEED_BASE = 0x51A7


def small(seed_offset: int = 0) -> LimaSystem:
    """1 chain of 256 (2 compounds x 128) + 3000 solvent, 5 nm, 8 types, 300 K."""
    return build_system(SEED_BASE + seed_offset, 1, 2, 5.0, 8, 3000, T0=300.0)


# Compounds of the production and large systems are built within 1.6 nm of their first walk
# bead, numbered from their central bead and given a 1.25 nm sphere margin: with these the
# systems run their configured 10,000 and 5,000 steps (the compounds swell; the reference
# re-images them against particle 0 and lists neighbours with spheres fixed at build time).
LONG_RUN = dict(confine=1.6, central_first=True, sphere_margin=1.25)


def production() -> LimaSystem:
    """10 chains x 10 compounds x 128 (12,800 chain) + lattice solvent at 33/nm^3, 10 nm, 16 types."""
    return build_system(SEED_BASE + 1000, 10, 10, 10.0, 16, None, T0=300.0, **LONG_RUN)


def large() -> LimaSystem:
    """10 chains x 40 compounds x 128 (51,200 chain) + lattice solvent at 33/nm^3, 18 nm, 16 types."""
    return build_system(SEED_BASE + 2000, 10, 40, 18.0, 16, None, T0=300.0, **LONG_RUN)


# Compound size by box edge: 16 particles in boxes under 4.6 nm, else 32.  A compound of n
# particles straightened to all-trans torsions reaches 0.122 n / 2 nm from its central bead
# (1.0 nm at 16, 2.0 nm at 32), inside half of the smallest box it is given.
REPLICA_COMPOUND = (16, 4.6, 32)
REPLICA_CONFINE = 0.8      # nm, compound particles from their first walk bead
REPLICA_MARGIN = 1.0       # nm, added to each compound's confining sphere
REPLICA_BOX = (3.0, 6.0)   # nm, box edge range


def _replica(args) -> LimaSystem:
    r, seed_config, compound, confine, margin = args
    seed = SEED_BASE + 1000 * seed_config + r
    rng = np.random.default_rng(seed ^ 0x5EED)
    n_chain = COMPOUND_SIZE * int(rng.integers(1, 9))      # 128 .. 1024
    nsol = int(rng.integers(500, 6001))
    T0 = float(rng.uniform(280.0, 320.0))
    L = ((n_chain + nsol) / 33.0) ** (1.0 / 3.0)
    L = float(min(REPLICA_BOX[1], max(REPLICA_BOX[0], L)))
    if isinstance(compound, tuple):      # (size in boxes under edge, edge, size otherwise)
        compound = compound[0] if L < compound[1] else compound[2]
    return build_system(seed, 1, n_chain // compound, L, 16, nsol, T0=T0, compound_size=compound,
                        confine=confine, central_first=True, sphere_margin=margin, auto_density=True)


def replicas(n: int = 256, seed_config: int = 3, workers: int | None = None, compound=REPLICA_COMPOUND,
             confine: float = REPLICA_CONFINE, margin: float = REPLICA_MARGIN) -> list[LimaSystem]:
    """Uneven replicas: chain 128-1024 particles, solvent 500-6000, box edge scaled to 33 nm^-3 and
    held within 3-6 nm, each replica with its own seed and thermostat target U(280, 320) K.

    The chains of this force field straighten (only 1-2, 1-3 and 1-4 pairs are excluded, and the
    1-5 LJ repulsion at 0.15 nm bonds drives every torsion to trans over the weak dihedral), and
    the reference re-images every compound against its particle 0 and lists neighbours with
    spheres fixed at build time.  So that every replica stays valid for the run (4,000 steps
    measured, twice the configured 2,000), compounds hold 16 particles in boxes under 4.6 nm and
    32 otherwise (`compound`: an int, or (size under edge, edge, size otherwise)), are numbered
    from their central bead, are built within 0.8 nm of their first bead, and their spheres
    carry a 1.0 nm margin.  Where fewer lattice sites than the drawn solvent count are free of
    the chain, the solvent lattice is made finer.  Built in `workers` processes (default: up to
    16)."""
    import os
    args = [(r, seed_config, compound, confine, margin) for r in range(n)]
    if workers is None:
        workers = min(16, os.cpu_count() or 1)
    if workers > 1 and n >= 8:
        import multiprocessing
        from concurrent.futures import ProcessPoolExecutor
        with ProcessPoolExecutor(workers, mp_context=multiprocessing.get_context("spawn")) as ex:
            return list(ex.map(_replica, args, chunksize=2))
    return [_replica(a) for a in args]


def adjoint_replicas(n: int = 64) -> list[LimaSystem]:
    """64 replicas of 1 chain of 512 + 1,488 solvent in a 4.2 nm box, 16 types (shared force field)."""
    ff = make_forcefield(np.random.default_rng(SEED_BASE + 4000 - 1), 16)
    return [build_system(SEED_BASE + 4000 + r, 1, 4, 4.2, 16, 1488, T0=300.0, forcefield=ff) for r in range(n)]


# ---------------------------------------------------------------------------
# BoxBuilder's initial velocities (BoxBuilder.cu:248-276, BoxBuilder.cuh:56-102)
# ---------------------------------------------------------------------------
BOXBUILDER_T = 313.0                       # BoxBuilder::T [K] (BoxBuilder.cuh:63)
BOXBUILDER_R = 8.3144                      # BoxBuilder::R (BoxBuilder.cuh:64)
SOLVENT_MASS = float(np.float32(18.01528)) * 1e-3   # constexpr double from a float literal (Simulation.cuh:30)


def reference_initial_velocities(system: LimaSystem, seed: int = 290128309) -> LimaSystem:
    """The velocities BoxBuilder gives a fresh box, as previous positions.

    Solvents (createSolvent, BoxBuilder.cu:273-276): direction = norm of three
    draws of random() = (r % 10000) / 10000.f * 2 - 1 (uniform in the cube, so
    the direction is not isotropic), speed v_rms = sqrt(3 R T / M) with T = 313 K
    and M = SOLVENT_MASS; prev = pos - vel * dt, the FP64 products narrowed to
    FP32 as Float3 * double does (QuantomTypes.cuh:39).  Compounds
    (integrateCompound, BoxBuilder.cu:248-269): zero velocity, prev = pos.
    The draws come from numpy's generator seeded with BoxBuilder's srand seed
    (BoxBuilder.cuh:10); they are not the C library's rand() sequence."""
    out = system.copy()
    rng = np.random.default_rng(seed)
    n = system.n_solvents
    r = rng.integers(0, 10000, size=(n, 3))
    comp = (r.astype(np.float32) / np.float32(10000.0) * np.float32(2) - np.float32(1)).astype(np.float32)
    norm = np.sqrt(((comp * comp).sum(axis=1, dtype=np.float32))).astype(np.float32)
    norm[norm == 0] = np.float32(1)
    unit = (comp / norm[:, None]).astype(np.float32)
    v_rms = math.sqrt(3 * BOXBUILDER_R * BOXBUILDER_T / SOLVENT_MASS)
    vel = (unit.astype(np.float64) * v_rms).astype(np.float32)
    step = (vel.astype(np.float64) * system.dt).astype(np.float32)
    out.solvent_prev_positions = (system.solvent_positions.astype(np.float32) - step).astype(np.float32)
    out.chain_prev_positions = system.chain_positions.astype(np.float32).copy()
    return out
