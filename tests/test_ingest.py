"""System ingest (include/lima_ingest.h): CHARMM .prm + GROMACS .gro/.top -> lima_system.

The native ingest (quantom_amd/liblima_ingest.so) is compared byte for byte with the
oracle's restatement (oracle/ingest_oracle.py): the four text files between the stages
and every array of the built system.  Both are pinned by facts derived here without
either: unit conversions in closed form, exclusion and bridge sets as unions over the
topology's terms, the atom-name -> type resolution worked out by hand for the
hand-made fixture (tests/golden/ingest/), solvent acceptance from the stated radii, the
centring and sphere geometry in FP64, and the C library's own rand() sequence.

The T4-lysozyme demo the reference ships (Demo/6lzm, Demo/CHARMM_FF) is used where the
reference tree is mounted; the built demo system is committed as
tests/golden/t4lysozyme.npz for the GPU tests, and checked here against a fresh build.
"""
import ctypes as C
import math
import os
import re
import subprocess

import numpy as np
import pytest

from tests.conftest import GOLDEN, ROOT

FIX = os.path.join(GOLDEN, "ingest")
PRM = [os.path.join(FIX, "mini_a.prm"), os.path.join(FIX, "mini_b.prm")]
GRO = os.path.join(FIX, "mini.gro")
TOP = os.path.join(FIX, "mini.top")
SMALL = dict(max_compound_particles=8, max_atoms_in_residue=6)
DEMO = "/root/reference/Demo"
FILES = ("LIMA_ffnonbonded.txt", "LIMA_ffbonded.txt", "LIMA_ffnonbonded_filtered.txt", "LIMA_ffbonded_filtered.txt")


def _fixture_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_fixture", os.path.join(FIX, "make_fixture.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _oracle_build(gro, top, prm, out_dir, **kw):
    """The oracle's three stages, writing the same four files as the product."""
    from oracle import ingest_oracle as O
    os.makedirs(out_dir, exist_ok=True)
    m = O.merge_charmm(prm)
    with open(os.path.join(out_dir, FILES[0]), "w") as f:
        f.write(O.lima_ffnonbonded_text(m))
    with open(os.path.join(out_dir, FILES[1]), "w") as f:
        f.write(O.lima_ffbonded_text(m))
    rd = lambda n: O.read_param_file(os.path.join(out_dir, n), ignores=("/",), internal=True)
    flt = O.filter_forcefield(gro, top, rd(FILES[0]), rd(FILES[1]))
    with open(os.path.join(out_dir, FILES[2]), "w") as f:
        f.write(O.filtered_nonbonded_text(flt))
    with open(os.path.join(out_dir, FILES[3]), "w") as f:
        f.write(O.filtered_bonded_text(flt))
    eff = O.load_engine_forcefield(O.read_engine_file(os.path.join(out_dir, FILES[2])),
                                   O.read_engine_file(os.path.join(out_dir, FILES[3])))
    box = {k: kw.pop(k) for k in ("box_len", "dt", "seed") if k in kw}
    ing = O.build_box(gro, top, eff, **box, **kw)
    return O.to_lima_system(ing, box.get("box_len", 7.0), box.get("dt", 1e-7), 310.0), ing


def _same_system(a, b):
    diff = [k for k in a._ARRAYS
            if getattr(a, k).shape != getattr(b, k).shape or getattr(a, k).tobytes() != getattr(b, k).tobytes()]
    return diff


def test_library_exports_every_declared_symbol():
    from quantom_amd import ingest
    with open(os.path.join(ROOT, "include", "lima_ingest.h")) as f:
        text = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
    declared = set(re.findall(r"\b(lima_\w+)\s*\(", text))
    assert declared == set(ingest.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", ingest.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    assert declared <= set(re.findall(r"\sT\s(lima_\w+)", out))


def test_random_sequence_is_the_c_library_rand():
    """BoxBuilder seeds rand() with srand(290128309) (BoxBuilder.cuh:10); the oracle's private
    generator must give the C library's own sequence."""
    from oracle.ingest_oracle import _RandR
    libc = C.CDLL(None)
    libc.srand(290128309)
    want = [libc.rand() for _ in range(1000)]
    r = _RandR(290128309)
    assert [r() for _ in range(1000)] == want


@pytest.mark.parametrize("kw", [{}, SMALL], ids=["one_compound", "four_compounds"])
def test_product_matches_oracle_on_fixture(tmp_path, kw):
    from quantom_amd import ingest
    s, info = ingest.ingest(GRO, TOP, PRM, str(tmp_path / "p"), **kw)
    o, ing = _oracle_build(GRO, TOP, PRM, str(tmp_path / "o"), **dict(kw))
    for n in FILES:
        assert (tmp_path / "p" / n).read_bytes() == (tmp_path / "o" / n).read_bytes(), n
    assert _same_system(s, o) == []
    assert info.n_random_draws == ing.n_rand


def _rows(path, section):
    """Rows of one '# section' block of a text file written by the ingest."""
    out, cur = [], None
    for line in open(path).read().split("\n"):
        if line.startswith("# "):
            cur = line[2:]
            continue
        if cur == section and line and not line.startswith("/"):
            out.append(line.split(";"))
    return out


def test_merge_rules_and_units(tmp_path):
    """Merge rules on the hand-made files: first MASS of a type wins, later NONBONDED rows
    (another file's, or an NBFIX row) overwrite, first row of a sorted type tuple wins;
    units are kcal->J (4184 for epsilon, 4183 for force constants), A->nm, Rmin/2 -> 2 Rmin/2."""
    from quantom_amd import ingest
    ingest.merge_forcefield(PRM, str(tmp_path))
    nb = {r[0]: [float(x) for x in r[1:]] for r in _rows(tmp_path / FILES[0], "ff_nonbonded")}
    assert list(nb) == ["H", "HC", "HA", "HB", "C", "CC", "CT1", "CT2", "CT3", "N", "NH1", "NH3", "O", "OC",
                        "OH1", "OT", "S"]
    assert nb["C"][0] == pytest.approx(12.011)                    # file A's mass, not file B's 13
    assert nb["C"][1:] == pytest.approx([2 * 2.1 / 10, 0.090 * 4184], rel=1e-6)   # file B's NONBONDED row
    assert nb["O"][1:] == pytest.approx([2 * 3.1 / 10, 0.200 * 4184], rel=1e-6)   # the NBFIX row
    assert nb["CT1"][1:] == pytest.approx([2 * 2.275 / 10, 0.0200 * 4184], rel=1e-6)
    assert nb["CT2"][1:] == [-1.0, -1.0]                            # no NONBONDED row
    assert nb["S"] == pytest.approx([32.06, 0.4, 0.45 * 4184], rel=1e-6)
    bonds = {(r[0], r[1]): (float(r[2]), float(r[3])) for r in _rows(tmp_path / FILES[1], "ff_bondtypes")}
    assert len(bonds) == 15                                         # 17 rows: a reversed duplicate, a file-B repeat
    assert bonds[("CT1", "CT2")] == pytest.approx((0.1538, 222.5 * 4183 * 100), rel=1e-6)  # not the reversed row
    assert bonds[("C", "CT1")] == pytest.approx((0.149, 250.0 * 4183 * 100), rel=1e-6)     # not file B's
    assert bonds[("CT1", "CT1")] == pytest.approx((0.15, 222.5 * 4183 * 100), rel=1e-6)
    angles = {tuple(r[:3]): (float(r[3]), float(r[4])) for r in _rows(tmp_path / FILES[1], "ff_angletypes")}
    th, k = angles[("CT1", "C", "NH1")]
    assert th == pytest.approx(116.5 * math.pi / 180, rel=5e-5) and k == pytest.approx(80 * 4183, rel=1e-6)
    assert ("CT1", "NH3", "HC") in angles                            # HC NH3 CT1 stored outer-sorted
    dih = {tuple(r[:4]): (float(r[4]), float(r[5]), int(r[6])) for r in _rows(tmp_path / FILES[1], "ff_dihedraltypes")}
    assert dih[("NH1", "C", "CT1", "NH1")] == pytest.approx((0.0, 0.6 * 4183, 1), rel=1e-6)  # first multiplicity
    assert dih[("X", "C", "CT1", "X")][2] == 1                       # X CT1 C X flipped: inner names sorted
    assert dih[("C", "CT2", "NH1", "C")][0] == pytest.approx(math.pi, rel=5e-5)


def test_filter_type_resolution(tmp_path):
    """Atom names resolve to the longest prefix that is a type (hand-derived for the fixture)."""
    from quantom_amd import ingest
    ingest.merge_forcefield(PRM, str(tmp_path))
    ingest.prepare_forcefield(GRO, TOP, str(tmp_path))
    types = _rows(tmp_path / FILES[2], "ff_nonbonded")
    assert [(r[0], int(r[1])) for r in types] == [("WATER", 0), ("N", 1), ("H", 2), ("C", 3), ("HA", 4), ("O", 5),
                                                  ("OT", 6)]
    # the solvent type: one O and two H of the water model (ForcefieldTypes.h:17-19)
    water = [15.999 + 2 * 1.008, 0.302905564168 + 2 * 0.040001352445, (0.50208 + 2 * 0.19246) * 1000]
    assert [float(x) for x in types[0][2:]] == pytest.approx(water, rel=3e-6)
    maps = _rows(tmp_path / FILES[2], "mappings")
    assert [tuple(r) for r in maps] == [("H1", "H"), ("CA", "C"), ("CB", "C"), ("HN", "H"), ("HA1", "HA"), ("OG", "O"),
                                        ("HG1", "H"), ("OT1", "OT"), ("OT2", "OT"), ("OW", "O"), ("HW1", "H"),
                                        ("HW2", "H")]
    fx = _fixture_module()
    slot = {"N": 1, "H": 2, "C": 3, "HA": 4, "O": 5, "OT": 6}
    name_type = {"N": "N", "H1": "H", "CA": "C", "HA": "HA", "CB": "C", "C": "C", "O": "O", "HN": "H", "HA1": "HA",
                 "OG": "O", "HG1": "H", "OT1": "OT", "OT2": "OT"}
    atoms = _rows(tmp_path / FILES[3], "atoms")
    assert [(int(a), int(t)) for a, t in atoms] == [(i + 1, slot[name_type[a[2]]]) for i, a in enumerate(fx.ATOMS)]
    bonds = {(int(r[0]), int(r[1])): (r[2], r[3], float(r[4]), float(r[5])) for r in _rows(tmp_path / FILES[3], "bonds")}
    assert len(bonds) == len(fx.BONDS)
    # CT1-CC has no bond type: the best name likeness is C-CT1 (1/2 * 1)
    assert bonds[(24, 26)] == ("CC", "CT1", pytest.approx(0.149), pytest.approx(250.0 * 4183 * 100, rel=1e-6))
    # HC-NH3 has none either: H-NH1 scores 1/2 * 2/3
    assert bonds[(1, 2)][2:] == (pytest.approx(0.0997), pytest.approx(440.0 * 4183 * 100, rel=1e-6))
    dih = _rows(tmp_path / FILES[3], "dihedrals")
    assert len(dih) == len(fx.DIHEDRALS)                            # the second [ dihedrals ] section is not read


def _system(**kw):
    from quantom_amd import ingest
    import tempfile
    with tempfile.TemporaryDirectory(prefix="ingest_") as d:   # the force-field files written on the way
        return ingest.ingest(GRO, TOP, PRM, d, **kw)


def test_compounds_bridges_and_exclusions_from_topology():
    """Four compounds of whole residues; each term inside one compound or in the bridge of
    two consecutive ones; exclusions are the union of term partners (heavy atoms only)."""
    fx = _fixture_module()
    s, info = _system(**SMALL)
    heavy = [i + 1 for i, a in enumerate(fx.ATOMS) if a[2][0] != "H"]
    assert list(info.global_ids) == heavy
    assert list(np.diff(s.compound_offset)) == [5, 4, 6, 6]
    comp_of = {}
    for c in range(s.n_compounds):
        for k in range(s.compound_offset[c], s.compound_offset[c + 1]):
            comp_of[int(info.global_ids[k])] = (c, k - s.compound_offset[c])
    terms = [t for t in fx.BONDS + fx.ANGLES + fx.DIHEDRALS if all(i in comp_of for i in t)]
    want = {g: set() for g in heavy}
    for t in terms:
        for a in t:
            want[a] |= set(t) - {a}
    for k, g in enumerate(info.global_ids):
        got = set()
        for j in range(s.ignore_offset[k], s.ignore_offset[k + 1]):
            c, l = int(s.ignore_compound[j]), int(s.ignore_local[j])
            got.add(int(info.global_ids[s.compound_offset[c] + l]))
        assert got == want[int(g)], g
    # bridges: consecutive compound pairs; particles in order of first appearance in the spanning terms
    assert s.n_bridges == 3
    for b in range(3):
        span = [t for t in terms if {comp_of[i][0] for i in t} == {b, b + 1}]
        order = []
        for t in span:
            order += [i for i in t if i not in order]
        lo, hi = s.bridge_particle_offset[b], s.bridge_particle_offset[b + 1]
        got = [int(info.global_ids[s.compound_offset[c] + l])
               for c, l in zip(s.bridge_particle_compound[lo:hi], s.bridge_particle_local[lo:hi])]
        assert got == order
        nb = [len([t for t in span if len(t) == n]) for n in (2, 3, 4)]
        assert nb == [s.bridge_bond_offset[b + 1] - s.bridge_bond_offset[b],
                      s.bridge_angle_offset[b + 1] - s.bridge_angle_offset[b],
                      s.bridge_dihedral_offset[b + 1] - s.bridge_dihedral_offset[b]]
    inside = [t for t in terms if len({comp_of[i][0] for i in t}) == 1]
    assert len(s.bonds) + len(s.angles) + len(s.dihedrals) == len(inside)


def test_box_geometry_solvents_and_velocities():
    """Centring on the mean of compound centres, key/sphere per compound, the waters kept
    (none within 0.2 nm of the chain or 0.25 nm of a kept water) and their speed v_rms."""
    fx = _fixture_module()
    s, info = _system(**SMALL)
    pos = s.chain_positions.astype(np.float64)
    centres = [pos[s.compound_offset[c]:s.compound_offset[c + 1]].mean(0) for c in range(s.n_compounds)]
    assert np.allclose(np.mean(centres, 0), 3.5, atol=2e-6)
    for c in range(s.n_compounds):
        p = pos[s.compound_offset[c]:s.compound_offset[c + 1]]
        d = np.linalg.norm(p - p.mean(0), axis=1)
        assert s.key_particle_index[c] == int(np.argmin(d))
        assert s.confining_particle_sphere[c] == pytest.approx(d.max(), abs=2e-6)
    assert np.array_equal(s.chain_prev_positions, s.chain_positions)          # compounds at rest
    off = np.array(info.offset, np.float64)
    kept = np.array([[1.0, 1.0, 1.0], [5.5, 5.0, 4.0], [6.0, 1.5, 2.0]]) + off   # waters 5, 8 and the fused 9
    assert s.n_solvents == 3 and info.n_solvent_candidates == 5
    assert np.allclose(s.solvent_positions, kept, atol=1e-6)
    v = np.linalg.norm((s.solvent_positions.astype(np.float64) - s.solvent_prev_positions) / 1e-7, axis=1)
    v_rms = math.sqrt(3 * 8.3144 * 313 / (float(np.float32(18.01528)) * 1e-3))
    assert np.allclose(v, v_rms, rtol=2e-2)
    assert info.n_random_draws == 3 * s.n_compounds + 3 * s.n_solvents
    # the first direction after the compounds' draws: BoxBuilder::random from the C library's rand()
    libc = C.CDLL(None)
    libc.srand(290128309)
    for _ in range(3 * s.n_compounds):
        libc.rand()
    z, y, x = [(libc.rand() % 10000) / 10000.0 * 2 - 1 for _ in range(3)]
    u = np.array([x, y, z]) / np.linalg.norm([x, y, z])
    d0 = (s.solvent_positions[0].astype(np.float64) - s.solvent_prev_positions[0]) / 1e-7
    assert np.allclose(d0 / np.linalg.norm(d0), u, atol=2e-2)


def test_errors_are_status_codes(tmp_path):
    from quantom_amd import ingest
    from quantom_amd.abi import LIMA_E_ARG, LIMA_E_CAPACITY
    with pytest.raises(ingest.IngestError) as e:
        ingest.merge_forcefield([str(tmp_path / "absent.prm")], str(tmp_path))
    assert e.value.status == LIMA_E_ARG and "absent.prm" in str(e.value)
    d = str(tmp_path / "ff")
    ingest.ingest(GRO, TOP, PRM, d)
    with pytest.raises(ingest.IngestError) as e:
        ingest.build_system(GRO, TOP, d, max_exclusions=2)
    assert e.value.status == LIMA_E_CAPACITY
    with pytest.raises(ingest.IngestError) as e:
        ingest.build_system(GRO, TOP, d, max_compound_particles=4, max_atoms_in_residue=2)
    assert e.value.status == LIMA_E_CAPACITY
    with pytest.raises(ingest.IngestError) as e:
        ingest.build_system(GRO, TOP, d, max_compound_particles=200)
    assert e.value.status == LIMA_E_ARG


def test_ingested_system_is_accepted_by_the_oracle_engine():
    """The built system passes the boundary's checks and steps on the CPU oracle."""
    from oracle.oracle import Oracle
    from quantom_amd.abi import RunParams
    s, _ = _system(**SMALL)
    o = Oracle(s, RunParams(cutoff=0.7, nlist_interval=10), precision=32)
    assert o.step(20) == 0


# ------------------------------------------------------------------ the shipped demo

needs_demo = pytest.mark.skipif(not os.path.isdir(DEMO), reason="the reference's Demo folder is not mounted")


def _demo_paths():
    from quantom_amd.ingest import CHARMM_FILES
    return (os.path.join(DEMO, "6lzm", "conf.gro"), os.path.join(DEMO, "6lzm", "topol.top"),
            [os.path.join(DEMO, "CHARMM_FF", f) for f in CHARMM_FILES])


@pytest.fixture(scope="module")
def demo_build(tmp_path_factory):
    from quantom_amd import ingest
    gro, top, prm = _demo_paths()
    d = tmp_path_factory.mktemp("demo")
    s, info = ingest.ingest(gro, top, prm, str(d / "p"))
    return s, info, d


@needs_demo
def test_demo_product_matches_oracle(demo_build):
    s, info, d = demo_build
    gro, top, prm = _demo_paths()
    o, ing = _oracle_build(gro, top, prm, str(d / "o"))
    for n in FILES:
        assert (d / "p" / n).read_bytes() == (d / "o" / n).read_bytes(), n
    assert _same_system(s, o) == []
    assert info.n_random_draws == ing.n_rand


def _gro_columns(path):
    """GROMACS fixed columns: residue id 0-5, residue name 5-10, atom name 10-15, serial 15-20."""
    out = []
    lines = open(path).read().split("\n")
    for line in lines[2:2 + int(lines[1])]:
        out.append((int(line[0:5]), line[5:10].strip(), line[10:15].strip(), int(line[15:20]),
                    [float(line[20 + 8 * k:28 + 8 * k]) for k in range(3)]))
    return out


def _top_sections(path):
    sec, out = None, {}
    for line in open(path).read().split("\n"):
        t = line.split(";")[0].split()
        if not t or t[0].startswith("#"):
            continue
        if t[0] == "[":
            sec = t[1]
            out.setdefault(sec, []).append([])
            continue
        if sec:
            out[sec][-1].append(t)
    return out


@needs_demo
def test_demo_against_its_input_files(demo_build):
    """T4 lysozyme: 1,293 heavy atoms in residue order, one compound per run of whole
    residues, every topology term over heavy atoms present once, the waters' oxygens."""
    s, info, _ = demo_build
    gro, top, _ = _demo_paths()
    atoms = _gro_columns(gro)
    heavy = [a for a in atoms if a[1] != "SOL" and a[2][0] != "H" and a[0] <= 200]
    assert s.n_chain == len(heavy) == 1293
    assert list(info.global_ids) == [a[3] for a in heavy]
    assert np.allclose(s.chain_positions - np.array(info.offset), np.array([a[4] for a in heavy]), atol=2e-6)
    res = {a[3]: a[0] for a in heavy}
    for c in range(s.n_compounds):                               # whole residues, up to 64 particles
        ids = info.global_ids[s.compound_offset[c]:s.compound_offset[c + 1]]
        assert len(ids) <= 64
        if c + 1 < s.n_compounds:
            assert res[int(ids[-1])] != res[int(info.global_ids[s.compound_offset[c + 1]])]
    sec = _top_sections(top)
    hs = {a[3] for a in heavy}
    place = {int(g): k for k, g in enumerate(info.global_ids)}
    got = {2: set(), 3: set(), 4: set()}
    for c in range(s.n_compounds):
        base = s.compound_offset[c]
        for arr, offs, n in ((s.bonds, s.bond_offset, 2), (s.angles, s.angle_offset, 3), (s.dihedrals, s.dihedral_offset, 4)):
            for t in arr[offs[c]:offs[c + 1]]:
                got[n].add(tuple(int(info.global_ids[base + i]) for i in t["idx"][:n]))
    for b in range(s.n_bridges):
        lo = s.bridge_particle_offset[b]
        gid = [int(info.global_ids[s.compound_offset[c] + l]) for c, l in
               zip(s.bridge_particle_compound[lo:s.bridge_particle_offset[b + 1]],
                   s.bridge_particle_local[lo:s.bridge_particle_offset[b + 1]])]
        for arr, offs, n in ((s.bridge_bonds, s.bridge_bond_offset, 2), (s.bridge_angles, s.bridge_angle_offset, 3),
                             (s.bridge_dihedrals, s.bridge_dihedral_offset, 4)):
            for t in arr[offs[b]:offs[b + 1]]:
                got[n].add(tuple(gid[i] for i in t["idx"][:n]))
    want = {n: {tuple(int(x) for x in r[:n]) for r in rows if all(int(x) in hs for x in r[:n])}
            for n, rows in ((2, sec["bonds"][0]), (3, sec["angles"][0]), (4, sec["dihedrals"][0]))}
    for n in (2, 3, 4):
        assert got[n] == want[n], n
    assert len(place) == s.n_chain
    oxygens = [a for a in atoms if a[1] == "SOL" and a[2][0] == "O"]
    assert info.n_solvent_candidates == len(oxygens) and s.n_solvents == 5000


@needs_demo
def test_committed_demo_fixture_is_current(demo_build):
    from quantom_amd.abi import LimaSystem
    s, _, _ = demo_build
    ref = LimaSystem.load(os.path.join(GOLDEN, "t4lysozyme.npz"))
    assert _same_system(s, ref) == []


CPP_CALLER = r"""
#include "lima_ingest.h"
#include <cstdio>
#include <cstdlib>
// Environment::Environment's build (Environment.cpp:16-45) through the C ABI
int main(int argc, char** argv) {
    const char* prm[] = {argv[1], argv[2]};
    if (lima_ff_merge(prm, 2, argv[5]) != LIMA_OK) { std::printf("merge %s\n", lima_ingest_last_error()); return 1; }
    if (lima_ff_prepare(argv[3], argv[4], argv[5]) != LIMA_OK) { std::printf("prepare %s\n", lima_ingest_last_error()); return 1; }
    lima_ingest_params ip;
    lima_ingest_default_params(&ip);
    ip.max_compound_particles = 8;
    ip.max_atoms_in_residue = 6;
    lima_ingest* in = nullptr;
    if (lima_ingest_create(argv[3], argv[4], argv[5], &ip, &in) != LIMA_OK) { std::printf("create %s\n", lima_ingest_last_error()); return 1; }
    lima_system s;
    lima_ingest_system(in, &s);
    lima_ingest_info info;
    lima_ingest_get_info(in, &info);
    std::printf("%d %d %d %d %lld", s.n_compounds, s.n_bridges, s.compound_offset[s.n_compounds], s.n_solvents,
                (long long)info.n_random_draws);
    for (int i = 0; i < 3 * s.n_solvents; i++) std::printf(" %a", s.solvent_prev_positions[i]);
    std::printf("\n");
    lima_ingest_destroy(in);
    lima_ingest* bad = nullptr;
    int st = lima_ingest_create("/nonexistent.gro", argv[4], argv[5], &ip, &bad);
    std::printf("%d %d\n", st, bad == nullptr);
    return 0;
}
"""


def test_cpp_caller_builds_the_fixture_through_the_c_abi(tmp_path):
    """A C++ program written against include/lima_ingest.h runs all three stages and gets
    what the Python binding gets; a missing file is a status code, not exit()."""
    from quantom_amd import ingest
    from quantom_amd.abi import LIMA_E_ARG
    src = tmp_path / "caller.cpp"
    src.write_text(CPP_CALLER)
    exe = tmp_path / "caller"
    libdir = os.path.dirname(ingest.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe), "-L", libdir, "-llima_ingest", f"-Wl,-rpath,{libdir}"], check=True, capture_output=True)
    d = tmp_path / "ff"
    d.mkdir()
    out = subprocess.run([str(exe), PRM[0], PRM[1], GRO, TOP, str(d)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    first, second = out.stdout.strip().split("\n")
    s, info = ingest.build_system(GRO, TOP, str(d), **SMALL)
    f = first.split()
    assert [int(x) for x in f[:5]] == [s.n_compounds, s.n_bridges, s.n_chain, s.n_solvents, info.n_random_draws]
    assert [float.fromhex(x) for x in f[5:]] == [float(v) for v in s.solvent_prev_positions.ravel()]
    assert second.split() == [str(LIMA_E_ARG), "1"]


@needs_demo
@pytest.mark.parametrize("kw", [dict(ignore_hydrogens=0), dict(min_residue_id=10, max_residue_id=40),
                                dict(add_solvents=0, box_len=9.0)], ids=["with_hydrogens", "residues_10_40", "no_solvent"])
def test_demo_product_matches_oracle_other_options(tmp_path, kw):
    """The builder's options (Environment.cpp:21-23; ENABLE_SOLVENTS; BOX_LEN) on the demo."""
    from quantom_amd import ingest
    gro, top, prm = _demo_paths()
    s, info = ingest.ingest(gro, top, prm, str(tmp_path / "p"), **kw)
    okw = dict(kw)
    if "ignore_hydrogens" in okw:
        okw["ignore_hydrogens"] = bool(okw["ignore_hydrogens"])
    if "add_solvents" in okw:
        okw["solvents"] = bool(okw.pop("add_solvents"))
    o, ing = _oracle_build(gro, top, prm, str(tmp_path / "o"), **okw)
    assert _same_system(s, o) == []
    assert info.n_random_draws == ing.n_rand
    if kw.get("ignore_hydrogens") == 0:
        assert s.n_chain == 2610 and s.n_compounds == 66             # every protein atom of residues <= 200
