"""Pins the oracle's neighbour-list update (Engine.cu:146-287) and its invariants.

* hand-built hysteresis cases for the compound-compound rule, including the
  reference's use of the lower id's radius in the add threshold (Engine.cu:243);
* brute-force O(N^2) set equality for the solvent and compound-solvent lists
  after the incremental cull/add with swap-remove and hash dedupe;
* the squared-distance bound the GPU build compares against is exact;
* momentum, translation and box-shift invariants of the force pass.
"""
import numpy as np
import pytest

from oracle.oracle import KIND_CC, KIND_CS, KIND_SC, KIND_SS, Oracle
from quantom_amd.abi import ANGLE_DTYPE, DIHEDRAL_DTYPE, PAIR_DTYPE, LimaSystem, RunParams
from tests.conftest import as_sets, rel_rms

F32 = np.float32


def manual_system(keys, spheres, solvents=(), L=10.0, n_per=4):
    """Compounds of n_per particles clustered at their key (index 0); no bonded terms."""
    nc = len(keys)
    pos = []
    for k in keys:
        for i in range(n_per):
            pos.append(np.asarray(k, float) + [0.0, 0.05 * i, 0.0])
    pos = np.array(pos, F32).reshape(-1, 3)
    ns = len(solvents)
    sp = np.array(solvents, F32).reshape(-1, 3)
    ff = np.zeros((32, 3), F32)
    ff[:] = (0.012, 0.32, 100.0)
    ff[0] = (0.018, 0.3166, 650.0)
    z = lambda n: np.zeros(n, np.int32)
    return LimaSystem(
        box_len=L, dt=2e-6, thermostat_target=300.0, forcefield=ff,
        compound_offset=np.arange(nc + 1, dtype=np.int32) * n_per, key_particle_index=z(nc),
        confining_particle_sphere=np.array(spheres, F32), chain_positions=pos, chain_prev_positions=pos.copy(),
        chain_atom_types=np.ones(nc * n_per, np.uint8), bond_offset=z(nc + 1), bonds=np.zeros(0, PAIR_DTYPE),
        angle_offset=z(nc + 1), angles=np.zeros(0, ANGLE_DTYPE), dihedral_offset=z(nc + 1),
        dihedrals=np.zeros(0, DIHEDRAL_DTYPE), ignore_offset=z(nc * n_per + 1), ignore_compound=z(0),
        ignore_local=z(0), bridge_particle_offset=z(1), bridge_particle_compound=z(0), bridge_particle_local=z(0),
        bridge_bond_offset=z(1), bridge_bonds=np.zeros(0, PAIR_DTYPE), bridge_angle_offset=z(1),
        bridge_angles=np.zeros(0, ANGLE_DTYPE), bridge_dihedral_offset=z(1),
        bridge_dihedrals=np.zeros(0, DIHEDRAL_DTYPE), solvent_positions=sp, solvent_prev_positions=sp.copy())


def move_compound(o, system, c, dx):
    cx, cxm, sx, sxm = o.state()
    b0, b1 = system.compound_offset[c], system.compound_offset[c + 1]
    cx[b0:b1, 0] += dx
    o.set_state(cx, cxm, sx, sxm)


def cc_pairs(o):
    return {(i, int(j)) for i, l in enumerate(o.neighbors(KIND_CC)) for j in l}


def test_compound_hysteresis_known_answer():
    """C=0.6, r0=0.5, r1=1.0: add if d < 0.6+((0.5+0.5)+0.6)=2.2, keep if d < 0.6+((0.5+1.0)+0.6)=2.7."""
    s = manual_system([[2, 5, 5], [4.5, 5, 5]], [0.5, 1.0])
    o = Oracle(s, RunParams(nlist_interval=0, enable_solvents=False))
    assert cc_pairs(o) == set()                  # d = 2.5: not added from scratch
    move_compound(o, s, 1, -0.5)                 # d = 2.0 < 2.2: added
    o.rebuild_neighbors()
    assert cc_pairs(o) == {(0, 1), (1, 0)}
    move_compound(o, s, 1, +0.5)                 # d = 2.5: kept (was member and < 2.7)
    o.rebuild_neighbors()
    assert cc_pairs(o) == {(0, 1), (1, 0)}
    move_compound(o, s, 1, +0.3)                 # d = 2.8 >= 2.7: culled
    o.rebuild_neighbors()
    assert cc_pairs(o) == set()


def test_compound_add_threshold_reads_lower_id_radius():
    """Radii swapped (r0=1.0, r1=0.5): add threshold 0.6+((1.0+1.0)+0.6)=3.2 exceeds the
    cull threshold 2.7, so d=3.0 is added from scratch (Engine.cu:243 uses the self radius)."""
    s = manual_system([[2, 5, 5], [5.0, 5, 5]], [1.0, 0.5])
    o = Oracle(s, RunParams(nlist_interval=0, enable_solvents=False))
    assert cc_pairs(o) == {(0, 1), (1, 0)}
    s2 = manual_system([[2, 5, 5], [5.0, 5, 5]], [0.5, 1.0])   # same geometry, radii the other way
    o2 = Oracle(s2, RunParams(nlist_interval=0, enable_solvents=False))
    assert cc_pairs(o2) == set()


def test_compound_pair_across_the_boundary():
    s = manual_system([[0.3, 5, 5], [9.5, 5, 5]], [0.2, 0.2])   # minimum-image distance 0.8
    o = Oracle(s, RunParams(nlist_interval=0, enable_solvents=False))
    assert cc_pairs(o) == {(0, 1), (1, 0)}


def within32(a, b, offset, C, L):
    """Engine::neighborWithinCutoff in float32 (numpy float32 ops are IEEE single)."""
    a = a.astype(F32)
    b = b.astype(F32).copy()
    L, half = F32(L), F32(L) * F32(0.5)
    b = b + L * ((a - b) > half).astype(F32)
    b = b - L * ((a - b) < -half).astype(F32)
    d = a - b
    d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    return np.sqrt(d2) < F32(C) + F32(offset)


def test_solvent_lists_equal_brute_force_after_motion(small_system):
    C = F32(0.6)
    o = Oracle(small_system, RunParams(nlist_interval=0))
    for rounds in range(3):
        _, _, sx, _ = o.state()
        sx = sx.astype(F32)
        n = len(sx)
        want = []
        for i in range(n):
            j = np.arange(n)
            lo = np.minimum(i, j)
            hi = np.maximum(i, j)
            m = within32(sx[lo], sx[hi], C, C, small_system.box_len)
            m[i] = False
            want.append(set(np.nonzero(m)[0].tolist()))
        assert as_sets(o.neighbors(KIND_SS)) == want
        o.step(7)
        o.rebuild_neighbors()


def test_compound_solvent_lists_equal_brute_force(small_system):
    C = F32(0.6)
    o = Oracle(small_system, RunParams(nlist_interval=0))
    o.step(5)
    o.rebuild_neighbors()
    cx, _, sx, _ = o.state()
    cs = as_sets(o.neighbors(KIND_CS))
    sc = as_sets(o.neighbors(KIND_SC))
    for c in range(small_system.n_compounds):
        key = cx[small_system.compound_offset[c] + small_system.key_particle_index[c]].astype(F32)
        r = F32(small_system.confining_particle_sphere[c])
        m = within32(np.broadcast_to(key, sx.shape), sx.astype(F32), r + C, C, small_system.box_len)
        assert cs[c] == set(np.nonzero(m)[0].tolist())
    # the solvent side of the same relation is its transpose
    for s_id, comps in enumerate(sc):
        assert comps == {c for c in range(small_system.n_compounds) if s_id in cs[c]}


def sqrt_bound(t):
    """Smallest float32 x >= 0 with sqrt(x) >= t (bisection over the bit patterns)."""
    lo, hi = 0, 0x7F800000
    t = F32(t)
    while lo < hi:
        mid = (lo + hi) // 2
        x = np.array([mid], np.uint32).view(F32)[0]
        if np.sqrt(x) >= t:
            hi = mid
        else:
            lo = mid + 1
    return np.array([lo], np.uint32).view(F32)[0]


@pytest.mark.parametrize("t", [1.2, 0.6 + (1.337 + 0.6), 2.7000003, 4.1])
def test_squared_distance_bound_is_exact(t):
    t = F32(t)
    b = sqrt_bound(t)
    bits = np.array([b], F32).view(np.uint32)[0]
    near = np.arange(bits - 3000, bits + 3000, dtype=np.uint32).view(F32)
    rng = np.random.default_rng(1)
    wide = (rng.uniform(0, 4 * float(t) ** 2, 20000)).astype(F32)
    for x in (near, wide):
        np.testing.assert_array_equal(np.sqrt(x) < t, x < b)


def test_momentum_conserved_for_pure_solvent():
    rng = np.random.default_rng(5)
    L = 3.0
    g = (np.arange(8) + 0.5) * (L / 8)
    sites = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    sites = np.mod(sites + rng.uniform(-0.03, 0.03, sites.shape), L)
    s = manual_system([], [], solvents=sites, L=L)
    o = Oracle(s, RunParams(nlist_interval=0), precision=64)
    _, sf, _ = o.compute_forces()
    assert np.abs(sf.sum(0)).max() < 1e-9 * np.abs(sf).sum()


def test_translation_and_box_shift_invariance(small_system):
    o = Oracle(small_system, RunParams(nlist_interval=0), precision=64)
    cf0, sf0, _ = o.compute_forces()
    cx, cxm, sx, sxm = o.state()
    L = small_system.box_len
    # shift every coordinate by one box length: hyperpos/PBC make the forces identical
    o.set_state(cx + L, cxm + L, sx + L, sxm + L)
    cf1, sf1, _ = o.compute_forces()
    assert rel_rms(cf1, cf0) < 1e-9 and rel_rms(sf1, sf0) < 1e-9
    # rigid translation by a small vector (no particle crosses a wrap boundary of its own images)
    d = np.array([0.013, -0.021, 0.007])
    o.set_state(cx + d, cxm + d, sx + d, sxm + d)
    cf2, sf2, _ = o.compute_forces()
    assert rel_rms(cf2, cf0) < 1e-6 and rel_rms(sf2, sf0) < 1e-6


def test_fp32_oracle_close_to_fp64(small_system):
    o32 = Oracle(small_system, RunParams(nlist_interval=0), precision=32)
    o64 = Oracle(small_system, RunParams(nlist_interval=0), precision=64)
    for k in (KIND_CC, KIND_CS, KIND_SC, KIND_SS):
        o64.set_neighbors(k, o32.neighbors(k))
    a, b, _ = o32.compute_forces()
    c, d, _ = o64.compute_forces()
    assert rel_rms(a, c) < 1e-5 and rel_rms(b, d) < 1e-5
