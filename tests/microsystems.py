"""Hand-built micro-systems whose forces, energies and thermostat scalars are derived
by hand from the reference's call sites, independently of the oracle's code.

The expected values follow the reference line by line:

* the pair force on `self` from `other` is calcLJForce (Engine.cu:475-522), whose FP64
  value is -d * VirtualPathMaker::calcForce(|d|) with d = other - self
  (VirtualPathMaker.cpp:160-170; pinned separately in test_oracle_closed_forms.py);
* sigma is (sigma_i + sigma_j) * 0.5 at every site; epsilon per call site:
  solvent <- compound particle: (eps_SOL + eps_j) * 0.5          (computeLJForces, Engine.cu:643-646,
                                                                   called at Engine.cu:1112)
  compound <- neighbour compound: (eps_i + eps_j) * 0.5          (Engine.cu:668-670)
  compound <- own compound:       sqrt(eps_i * eps_j)             (Engine.cu:689-691)
  compound <- solvent:            sqrt(eps_i * eps_SOL)           (Engine.cu:724-727, called at 971)
  solvent <- solvent:             eps_SOL, sigma_SOL              (Engine.cu:711-713)
* a pair is skipped when the receiving particle's LJ_Ignores holds (other local id, other
  compound id) (Engine.cu:660 inter, 683 intra; Bodies.cuh:191-197);
* the compound particle's force starts from its bridge carry-in, which the bridge kernel
  stores (Engine.cu:1211) and the compound kernel zeroes when it loads it
  (Bodies.cuh:437-438), so it holds exactly one step's bridge terms (Engine.cu:914);
* a pair bond adds F_a = -kb (|a-b| - b0) (a-b)/|a-b| to a and -F_a to b (Engine.cu:525-548).

Forces are FP64 here; the tests compare the oracle's FP64 path to 1e-10 and the FP32
paths within rounding.
"""
from __future__ import annotations

import math

import numpy as np

from quantom_amd.abi import ANGLE_DTYPE, DIHEDRAL_DTYPE, PAIR_DTYPE, LimaSystem

# Types chosen so that every mixing rule gives a different epsilon:
#   arithmetic(1, 2) = 500, geometric(1, 2) = 300
#   solvent <- type 1: (650 + 100)/2 = 375; type 1 <- solvent: sqrt(100 * 650) = 254.95
EPS_SOL, SIG_SOL, MASS_SOL = 650.0, 0.3166, 0.018
TYPES = {0: (MASS_SOL, SIG_SOL, EPS_SOL), 1: (0.012, 0.30, 100.0), 2: (0.016, 0.34, 900.0)}


def forcefield() -> np.ndarray:
    ff = np.zeros((32, 3), dtype=np.float32)
    ff[:, :] = (0.012, 0.32, 100.0)
    for t, v in TYPES.items():
        ff[t] = v
    return ff


def f32(x):
    return float(np.float32(x))


# The reference mixes in FP32 (sigma as (s_i + s_j) * 0.5f); the FP64 oracle mixes the same
# FP32 parameters in FP64.  MIX32 selects which of the two the expectations use.
MIX32 = True


def sigma_mix(ti, tj):
    """(sigma_i + sigma_j) * 0.5f, rounded to FP32 as the reference does when MIX32."""
    if not MIX32:
        return (f32(TYPES[ti][1]) + f32(TYPES[tj][1])) * 0.5
    s = np.float32(TYPES[ti][1]) + np.float32(TYPES[tj][1])
    return float(np.float32(s * np.float32(0.5)))


def eps_arith(ti, tj):
    return (f32(TYPES[ti][2]) + f32(TYPES[tj][2])) * 0.5


def eps_geo(ti, tj):
    return math.sqrt(f32(TYPES[ti][2]) * f32(TYPES[tj][2]))


def lj(self_pos, other_pos, sigma, eps):
    """FP64 force on self from other: -d * calcForce(|d|) of VirtualPathMaker.cpp:160-170."""
    d = np.asarray(other_pos, float) - np.asarray(self_pos, float)
    r = float(np.linalg.norm(d))
    fr = sigma / r
    f6 = fr ** 6
    calc = -24.0 * (eps / (r * r)) * f6 * (1.0 - 2.0 * f6)
    return -d * calc


def lj_energy(r, sigma, eps):
    """4 eps (s^2 - s), s = (sigma/r)^6: the energy of one pair."""
    s = (sigma / r) ** 6
    return 4.0 * eps * (s * s - s)


def pair_bond(a, b, b0, kb):
    """calcPairbondForces (Engine.cu:525-548) in FP64: force on a and on b."""
    d = np.asarray(a, float) - np.asarray(b, float)
    r = float(np.linalg.norm(d))
    fa = -kb * (r - b0) * d / r
    return fa, -fa


def build(L, compounds, solvents, ignores=None, bridges=None, bonds=None, dt=2e-6, T0=300.0, prev=None):
    """compounds: [(positions [n,3], types [n])]; solvents: [m,3];
    ignores: {(compound, local): [(compound, local), ...]} (each particle's LJ_Ignores);
    bonds: {compound: [(i, j, b0, kb)]} with compound-local indices;
    bridges: [([(compound, local), ...], [(i, j, b0, kb)])] with bridge-local bond indices;
    prev: optional (chain_prev [n,3], solvent_prev [m,3]); default at rest."""
    ignores = ignores or {}
    bonds = bonds or {}
    bridges = bridges or []
    nc = len(compounds)
    coff = np.zeros(nc + 1, np.int32)
    for c, (p, _) in enumerate(compounds):
        coff[c + 1] = coff[c] + len(p)
    n_chain = int(coff[-1])
    cpos = np.zeros((n_chain, 3), np.float32)
    types = np.zeros(n_chain, np.uint8)
    for c, (p, t) in enumerate(compounds):
        cpos[coff[c]:coff[c + 1]] = np.asarray(p, np.float32)
        types[coff[c]:coff[c + 1]] = t
    key = np.zeros(nc, np.int32)
    sphere = np.zeros(nc, np.float32)
    for c in range(nc):
        p = cpos[coff[c]:coff[c + 1]].astype(np.float64)
        sphere[c] = np.float32(np.max(np.linalg.norm(p - p[0], axis=1)) if len(p) else 0.0)
    ig_off = np.zeros(n_chain + 1, np.int32)
    ig_c, ig_l = [], []
    for c in range(nc):
        for l in range(coff[c + 1] - coff[c]):
            g = coff[c] + l
            ent = ignores.get((c, l), [])
            ig_off[g + 1] = ig_off[g] + len(ent)
            ig_c += [e[0] for e in ent]
            ig_l += [e[1] for e in ent]

    def pack_bonds(lists):
        off = np.zeros(len(lists) + 1, np.int32)
        flat = []
        for k, lst in enumerate(lists):
            off[k + 1] = off[k] + len(lst)
            flat += lst
        arr = np.zeros(len(flat), PAIR_DTYPE)
        for k, (i, j, b0, kb) in enumerate(flat):
            arr[k]["b0"], arr[k]["kb"], arr[k]["idx"] = b0, kb, (i, j)
        return off, arr

    bond_off, barr = pack_bonds([bonds.get(c, []) for c in range(nc)])
    zero_off = np.zeros(nc + 1, np.int32)
    nb = len(bridges)
    bpo = np.zeros(nb + 1, np.int32)
    bpc, bpl = [], []
    for k, (parts, _) in enumerate(bridges):
        bpo[k + 1] = bpo[k] + len(parts)
        bpc += [p[0] for p in parts]
        bpl += [p[1] for p in parts]
    bb_off, bb = pack_bonds([b for _, b in bridges])
    spos = np.asarray(solvents, np.float32).reshape(-1, 3)
    if prev is None:
        cprev, sprev = cpos.copy(), spos.copy()
    else:
        cprev = np.asarray(prev[0], np.float32).reshape(-1, 3)
        sprev = np.asarray(prev[1], np.float32).reshape(-1, 3)
    return LimaSystem(
        box_len=float(L), dt=dt, thermostat_target=T0, forcefield=forcefield(),
        compound_offset=coff, key_particle_index=key, confining_particle_sphere=sphere,
        chain_positions=cpos, chain_prev_positions=cprev, chain_atom_types=types,
        bond_offset=bond_off, bonds=barr, angle_offset=zero_off, angles=np.zeros(0, ANGLE_DTYPE),
        dihedral_offset=zero_off.copy(), dihedrals=np.zeros(0, DIHEDRAL_DTYPE),
        ignore_offset=ig_off, ignore_compound=np.array(ig_c, np.int32), ignore_local=np.array(ig_l, np.int32),
        bridge_particle_offset=bpo, bridge_particle_compound=np.array(bpc, np.int32),
        bridge_particle_local=np.array(bpl, np.int32),
        bridge_bond_offset=bb_off, bridge_bonds=bb,
        bridge_angle_offset=np.zeros(nb + 1, np.int32), bridge_angles=np.zeros(0, ANGLE_DTYPE),
        bridge_dihedral_offset=np.zeros(nb + 1, np.int32), bridge_dihedrals=np.zeros(0, DIHEDRAL_DTYPE),
        solvent_positions=spos, solvent_prev_positions=sprev, meta={"micro": True})


# ---------------------------------------------------------------------------------------------
# The micro-systems.  All particles sit well inside the box, within one list radius of each
# other (2 x 0.6 nm plus the spheres), so every listed pair is listed and no image is involved.
# ---------------------------------------------------------------------------------------------
def sites_system():
    """Every call site once: compound 0 = {type 1, type 2} (an intra pair), compound 1 = {type 2}
    (inter pairs), solvents s0, s1 (solvent <- chain, chain <- solvent, solvent <- solvent)."""
    c0 = ([(2.00, 2.00, 2.00), (2.00, 2.00, 2.38)], [1, 2])
    c1 = ([(2.36, 2.00, 2.00)], [2])
    sol = [(2.00, 2.41, 2.00), (2.00, 2.41, 2.37)]
    return build(6.0, [c0, c1], sol)


def exclusions_system():
    """Intra exclusion a0 <-> a1 in compound 0; cross exclusion a2 <-> b0 between compounds 0
    and 1 (each particle's LJ_Ignores holds the other); no bonded terms, no solvent."""
    c0 = ([(2.00, 2.00, 2.00), (2.36, 2.00, 2.00), (2.18, 2.33, 2.00)], [1, 2, 1])
    c1 = ([(2.18, 2.33, 2.37), (2.50, 2.20, 2.30)], [2, 1])
    ign = {(0, 0): [(0, 1)], (0, 1): [(0, 0)], (0, 2): [(1, 0)], (1, 0): [(0, 2)]}
    return build(6.0, [c0, c1], [], ignores=ign)


BRIDGE_B0, BRIDGE_KB = 0.15, 2.5e8


def bridge_system():
    """Compounds {p0, p1} and {q0, q1}; a bridge holds one pair bond p1 - q0, stretched to 0.17 nm
    from b0 = 0.15 nm; p1 and q0 exclude each other (as every bonded tuple does); no solvent."""
    c0 = ([(2.00, 2.00, 2.00), (2.30, 2.05, 2.00)], [1, 2])
    c1 = ([(2.47, 2.05, 2.00), (2.80, 2.10, 2.10)], [1, 2])
    ign = {(0, 1): [(1, 0)], (1, 0): [(0, 1)]}
    br = [([(0, 1), (1, 0)], [(0, 1, BRIDGE_B0, BRIDGE_KB)])]
    return build(6.0, [c0, c1], [], ignores=ign, bridges=br)


def sites_expected(sysm):
    """Forces of sites_system by the reference's call sites (FP64)."""
    cp = sysm.chain_positions.astype(float)
    sp = sysm.solvent_positions.astype(float)
    t = sysm.chain_atom_types
    comp = [0, 0, 1]
    fc = np.zeros((3, 3))
    fs = np.zeros((2, 3))
    for i in range(3):
        for j in range(3):
            if i == j:
                continue
            sig = sigma_mix(t[i], t[j])
            e = eps_geo(t[i], t[j]) if comp[i] == comp[j] else eps_arith(t[i], t[j])
            fc[i] += lj(cp[i], cp[j], sig, e)
        for s in range(2):
            fc[i] += lj(cp[i], sp[s], sigma_mix(t[i], 0), eps_geo(t[i], 0))
    for s in range(2):
        for j in range(3):
            fs[s] += lj(sp[s], cp[j], sigma_mix(0, t[j]), eps_arith(0, t[j]))
        o = 1 - s
        fs[s] += lj(sp[s], sp[o], f32(SIG_SOL), f32(EPS_SOL))
    return fc, fs


def sites_energies(sysm):
    """The FP64 LJ energy terms of sites_system: half the sum over directed pairs of
    4 eps (s^2 - s) per site (lima_energies)."""
    cp = sysm.chain_positions.astype(float)
    sp = sysm.solvent_positions.astype(float)
    t = sysm.chain_atom_types
    comp = [0, 0, 1]
    out = dict(lj_intra=0.0, lj_inter=0.0, lj_compound_solvent=0.0, lj_solvent_compound=0.0, lj_solvent_solvent=0.0)
    for i in range(3):
        for j in range(3):
            if i == j:
                continue
            r = np.linalg.norm(cp[j] - cp[i])
            if comp[i] == comp[j]:
                out["lj_intra"] += 0.5 * lj_energy(r, sigma_mix(t[i], t[j]), eps_geo(t[i], t[j]))
            else:
                out["lj_inter"] += 0.5 * lj_energy(r, sigma_mix(t[i], t[j]), eps_arith(t[i], t[j]))
        for s in range(2):
            r = np.linalg.norm(sp[s] - cp[i])
            out["lj_compound_solvent"] += 0.5 * lj_energy(r, sigma_mix(t[i], 0), eps_geo(t[i], 0))
            out["lj_solvent_compound"] += 0.5 * lj_energy(r, sigma_mix(0, t[i]), eps_arith(0, t[i]))
    r = np.linalg.norm(sp[1] - sp[0])
    out["lj_solvent_solvent"] = 2 * 0.5 * lj_energy(r, f32(SIG_SOL), f32(EPS_SOL))
    return out


def exclusions_expected(sysm):
    cp = sysm.chain_positions.astype(float)
    t = sysm.chain_atom_types
    comp = [0, 0, 0, 1, 1]
    local = [0, 1, 2, 0, 1]
    excluded = {(0, 1), (1, 0), (2, 3), (3, 2)}
    f = np.zeros((5, 3))
    for i in range(5):
        for j in range(5):
            if i == j or (i, j) in excluded:
                continue
            e = eps_geo(t[i], t[j]) if comp[i] == comp[j] else eps_arith(t[i], t[j])
            f[i] += lj(cp[i], cp[j], sigma_mix(t[i], t[j]), e)
    del local
    return f


def bridge_expected(sysm):
    cp = sysm.chain_positions.astype(float)
    t = sysm.chain_atom_types
    comp = [0, 0, 1, 1]
    excluded = {(1, 2), (2, 1)}
    f = np.zeros((4, 3))
    for i in range(4):
        for j in range(4):
            if i == j or (i, j) in excluded:
                continue
            e = eps_geo(t[i], t[j]) if comp[i] == comp[j] else eps_arith(t[i], t[j])
            f[i] += lj(cp[i], cp[j], sigma_mix(t[i], t[j]), e)
    fa, fb = pair_bond(cp[1], cp[2], f32(BRIDGE_B0), f32(BRIDGE_KB))
    f[1] += fa
    f[2] += fb
    return f


def free_solvents(T_init, n=4):
    """n solvents farther apart than any list radius (no forces), each moving at the speed
    sqrt(3 R T_init / m) along a fixed direction, so that the temperature of their first
    displacement is T_init.  Returns the system."""
    dirs = np.array([(1, 2, 2), (-2, 1, 2), (2, -2, 1), (1, 1, -1)], float)[:n]
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    pos = np.array([(1.0 + 2.0 * (k % 2), 1.0 + 2.0 * ((k // 2) % 2), 1.0) for k in range(n)])
    v = dirs * math.sqrt(3 * 8.3145 * T_init / MASS_SOL)                              # nm/ns
    p32 = pos.astype(np.float32)
    prev = (pos - v * 2e-6).astype(np.float32)
    return build(6.0, [], p32, prev=(np.zeros((0, 3)), prev))
