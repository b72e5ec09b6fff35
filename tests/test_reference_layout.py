"""The reference-layout converter (include/lima_reference_layout.hpp).

The reference's arrays (Compound, CompoundState, Solvent, CompoundBridgeCompact,
ForceField; Bodies.cuh, Forcefield.cuh) are written here byte for byte from numpy
dtypes with explicit offsets, for MAX_COMPOUND_PARTICLES = 64 (the reference's
default, Constants.cuh:58) and 128.  The sizes and offsets they must have are the
ones a layout probe of the reference headers reported: Compound 24,036 B at 64 and
30,116 B at 128 (singlebonds at 12,180, anglebonds 14,232, dihedrals 19,356, key
particle 30,108, sphere 30,112), CompoundState 772 / 1,540 B, CompoundBridgeCompact
2,352 B.  A C++ caller built against the header converts them to lima_system and
writes a stepped state back; every array must come back exactly.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT

F3 = ("<f4", (3,))
LJ_IGN = np.dtype([("local", "u1", (32,)), ("comp", "u1", (32,)), ("n", "u1")])
PAIR = np.dtype({"names": ["b0", "kb", "idx"], "formats": ["<f4", "<f4", ("<u4", (2,))], "offsets": [0, 4, 8],
                 "itemsize": 16})
ANGLE = np.dtype({"names": ["t0", "k", "idx"], "formats": ["<f4", "<f4", ("<u4", (3,))], "offsets": [0, 4, 8],
                  "itemsize": 20})
DIH = np.dtype({"names": ["p0", "k", "n", "idx"], "formats": ["<f4", "<f4", "u1", ("<u4", (4,))],
                "offsets": [0, 4, 8, 12], "itemsize": 28})


def compound_dtype(M):
    o_lj = 4 + 26 * M
    o_gid = o_lj + 65 * M
    o_com = o_gid + 4 * M
    o_sb = o_com + 12 + 4
    o_ab = o_sb + 128 * 16 + 4
    o_d = o_ab + 256 * 20 + 4
    o_key = o_d + 384 * 28
    names = ["n", "prev", "forces", "types", "colors", "lj", "gid", "com", "nsb", "sb", "nab", "ab", "nd", "d", "key",
             "sphere"]
    formats = ["u1", ("<f4", (M, 3)), ("<f4", (M, 3)), ("u1", (M,)), ("u1", (M,)), (LJ_IGN, (M,)), ("<u4", (M,)), F3,
               "<u2", (PAIR, (128,)), "<u2", (ANGLE, (256,)), "<u2", (DIH, (384,)), "<i4", "<f4"]
    offsets = [0, 4, 4 + 12 * M, 4 + 24 * M, 4 + 25 * M, o_lj, o_gid, o_com, o_com + 12, o_sb, o_sb + 2048, o_ab,
               o_ab + 5120, o_d, o_key, o_key + 4]
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": o_key + 8})


def state_dtype(M):
    return np.dtype({"names": ["pos", "n"], "formats": [("<f4", (M, 3)), "u1"], "offsets": [0, 12 * M],
                     "itemsize": 12 * M + 4})


BRIDGE = np.dtype({"names": ["refs", "types", "n", "nsb", "sb", "nab", "ab", "nd", "d"],
                   "formats": [("<i4", (32, 2)), ("u1", (32,)), "<i4", "<u2", (PAIR, (32,)), "<u2", (ANGLE, (32,)),
                               "<u2", (DIH, (32,))],
                   "offsets": [0, 256, 288, 292, 296, 808, 812, 1452, 1456], "itemsize": 2352})
SOLVENT = np.dtype([("pos", "<f4", (3,)), ("prev", "<f4", (3,))])


def test_dtypes_match_the_probed_reference_layout():
    c64, c128 = compound_dtype(64), compound_dtype(128)
    assert c64.itemsize == 24036 and c128.itemsize == 30116
    off = {k: c128.fields[k][1] for k in ("prev", "forces", "types", "lj", "sb", "ab", "d", "key", "sphere")}
    assert off == {"prev": 4, "forces": 1540, "types": 3076, "lj": 3332, "sb": 12180, "ab": 14232, "d": 19356,
                   "key": 30108, "sphere": 30112}
    assert state_dtype(64).itemsize == 772 and state_dtype(128).itemsize == 1540
    assert BRIDGE.itemsize == 2352 and SOLVENT.itemsize == 24 and LJ_IGN.itemsize == 65


def to_reference_bytes(s, M):
    """The reference's arrays for one LimaSystem, filled field by field."""
    nc = s.n_compounds
    off = s.compound_offset
    comps = np.zeros(nc, compound_dtype(M))
    states = np.zeros(nc, state_dtype(M))
    for c in range(nc):
        a, b = off[c], off[c + 1]
        n = b - a
        assert n <= M
        comps[c]["n"] = n
        states[c]["n"] = n
        comps[c]["prev"][:n] = s.chain_prev_positions[a:b]
        states[c]["pos"][:n] = s.chain_positions[a:b]
        comps[c]["types"][:n] = s.chain_atom_types[a:b]
        comps[c]["gid"][:n] = np.arange(a, b)
        for p in range(n):
            e0, e1 = s.ignore_offset[a + p], s.ignore_offset[a + p + 1]
            lj = comps[c]["lj"][p]
            lj["n"] = e1 - e0
            lj["local"][:e1 - e0] = s.ignore_local[e0:e1]
            lj["comp"][:e1 - e0] = s.ignore_compound[e0:e1]
        for name, arr, offs, dst, cnt in (("b", s.bonds, s.bond_offset, "sb", "nsb"),
                                          ("a", s.angles, s.angle_offset, "ab", "nab"),
                                          ("d", s.dihedrals, s.dihedral_offset, "d", "nd")):
            t = arr[offs[c]:offs[c + 1]]
            comps[c][cnt] = len(t)
            tgt = comps[c][dst][:len(t)]
            tgt["idx"] = t["idx"]
            if name == "b":
                tgt["b0"], tgt["kb"] = t["b0"], t["kb"]
            elif name == "a":
                tgt["t0"], tgt["k"] = t["theta_0"], t["k_theta"]
            else:
                tgt["p0"], tgt["k"], tgt["n"] = t["phi_0"], t["k_phi"], t["n"]
        comps[c]["key"] = s.key_particle_index[c]
        comps[c]["sphere"] = s.confining_particle_sphere[c]
    nb = len(s.bridge_particle_offset) - 1
    bridges = np.zeros(nb, BRIDGE)
    for k in range(nb):
        p0, p1 = s.bridge_particle_offset[k], s.bridge_particle_offset[k + 1]
        br = bridges[k]
        br["n"] = p1 - p0
        br["refs"][:p1 - p0, 0] = s.bridge_particle_compound[p0:p1]
        br["refs"][:p1 - p0, 1] = s.bridge_particle_local[p0:p1]
        for arr, offs, dst, cnt in ((s.bridge_bonds, s.bridge_bond_offset, "sb", "nsb"),
                                    (s.bridge_angles, s.bridge_angle_offset, "ab", "nab"),
                                    (s.bridge_dihedrals, s.bridge_dihedral_offset, "d", "nd")):
            t = arr[offs[k]:offs[k + 1]]
            br[cnt] = len(t)
            tgt = br[dst][:len(t)]
            tgt["idx"] = t["idx"]
            names = list(t.dtype.names)
            tgt[tgt.dtype.names[0]], tgt[tgt.dtype.names[1]] = t[names[0]], t[names[1]]
            if dst == "d":
                tgt["n"] = t["n"]
    solv = np.zeros(s.n_solvents, SOLVENT)
    solv["pos"] = s.solvent_positions
    solv["prev"] = s.solvent_prev_positions
    ff = np.ascontiguousarray(s.forcefield, dtype="<f4")
    return comps, states, bridges, solv, ff


CALLER = r"""
#include "lima_reference_layout.hpp"
#include <cstdio>
#include <vector>
using namespace lima::ref;
template <class T> std::vector<T> rd(const char* path, size_t n) {
    std::vector<T> v(n ? n : 1);
    FILE* f = std::fopen(path, "rb");
    if (n && std::fread(v.data(), sizeof(T), n, f) != n) std::exit(3);
    std::fclose(f);
    return v;
}
template <class T> void wr(FILE* f, const T* p, size_t n) { if (n) std::fwrite(p, sizeof(T), n, f); }
template <int M> int run(const char* dir, int nc, int ns, int nb) {
    std::string d(dir);
    auto comps = rd<Compound<M>>((d + "/comp.bin").c_str(), nc);
    auto states = rd<CompoundState<M>>((d + "/state.bin").c_str(), nc);
    auto bridges = rd<CompoundBridgeCompact>((d + "/bridge.bin").c_str(), nb);
    auto solv = rd<Solvent>((d + "/solv.bin").c_str(), ns);
    auto ff = rd<ForceField>((d + "/ff.bin").c_str(), 1);
    System sys;
    from_reference<M>(sys, comps.data(), states.data(), nc, solv.data(), ns, bridges.data(), nb, ff[0], 7.5f, 2e-6, 300.f);
    const lima_system& s = sys.system;
    const int nch = s.compound_offset[s.n_compounds];
    FILE* f = std::fopen((d + "/out.bin").c_str(), "wb");
    int32_t head[4] = {s.n_compounds, nch, s.n_solvents, s.n_bridges};
    wr(f, head, 4);
    wr(f, s.compound_offset, s.n_compounds + 1);
    wr(f, s.key_particle_index, s.n_compounds);
    wr(f, s.confining_particle_sphere, s.n_compounds);
    wr(f, s.chain_positions, 3 * nch); wr(f, s.chain_prev_positions, 3 * nch); wr(f, s.chain_atom_types, nch);
    wr(f, s.bond_offset, s.n_compounds + 1); wr(f, s.bonds, s.bond_offset[s.n_compounds]);
    wr(f, s.angle_offset, s.n_compounds + 1); wr(f, s.angles, s.angle_offset[s.n_compounds]);
    wr(f, s.dihedral_offset, s.n_compounds + 1); wr(f, s.dihedrals, s.dihedral_offset[s.n_compounds]);
    wr(f, s.ignore_offset, nch + 1); wr(f, s.ignore_compound, s.ignore_offset[nch]); wr(f, s.ignore_local, s.ignore_offset[nch]);
    wr(f, s.bridge_particle_offset, s.n_bridges + 1);
    const int nbp = s.bridge_particle_offset[s.n_bridges];
    wr(f, s.bridge_particle_compound, nbp); wr(f, s.bridge_particle_local, nbp);
    wr(f, s.bridge_bond_offset, s.n_bridges + 1); wr(f, s.bridge_bonds, s.bridge_bond_offset[s.n_bridges]);
    wr(f, s.bridge_angle_offset, s.n_bridges + 1); wr(f, s.bridge_angles, s.bridge_angle_offset[s.n_bridges]);
    wr(f, s.bridge_dihedral_offset, s.n_bridges + 1); wr(f, s.bridge_dihedrals, s.bridge_dihedral_offset[s.n_bridges]);
    wr(f, s.solvent_positions, 3 * s.n_solvents); wr(f, s.solvent_prev_positions, 3 * s.n_solvents);
    wr(f, &s.box_len, 1); wr(f, &s.dt, 1); wr(f, &s.thermostat_target, 1);
    wr(f, &s.forcefield, 1);
    std::fclose(f);
    // a stepped state (positions + 1 nm, prev - 1 nm) written back into the reference arrays
    std::vector<float> cx(s.chain_positions, s.chain_positions + 3 * nch), cxm(s.chain_prev_positions, s.chain_prev_positions + 3 * nch);
    std::vector<float> sx(s.solvent_positions, s.solvent_positions + 3 * ns), sxm(s.solvent_prev_positions, s.solvent_prev_positions + 3 * ns);
    for (float& v : cx) v += 1.f;
    for (float& v : cxm) v -= 1.f;
    for (float& v : sx) v += 1.f;
    for (float& v : sxm) v -= 1.f;
    write_back<M>(cx.data(), cxm.data(), sx.data(), sxm.data(), comps.data(), states.data(), nc, solv.data(), ns);
    f = std::fopen((d + "/back.bin").c_str(), "wb");
    wr(f, comps.data(), nc); wr(f, states.data(), nc); wr(f, solv.data(), ns);
    std::fclose(f);
    return 0;
}
int main(int argc, char** argv) {
    const int M = std::atoi(argv[2]), nc = std::atoi(argv[3]), ns = std::atoi(argv[4]), nb = std::atoi(argv[5]);
    try {
        return M == 64 ? run<64>(argv[1], nc, ns, nb) : run<128>(argv[1], nc, ns, nb);
    } catch (const std::exception& e) {
        std::printf("%s\n", e.what());
        return 2;
    }
}
"""


@pytest.fixture(scope="module")
def caller(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ missing")
    d = tmp_path_factory.mktemp("reflayout")
    src = d / "caller.cpp"
    src.write_text(CALLER)
    exe = d / "caller"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True, capture_output=True)
    return str(exe)


def _read(buf, pos, dtype, n):
    dt = np.dtype(dtype)
    a = np.frombuffer(buf, dt, n, pos)
    return a, pos + dt.itemsize * n


@pytest.mark.parametrize("M,fixture", [(64, "tiny_system"), (128, "small_system")])
def test_reference_arrays_convert_and_write_back_exactly(caller, tmp_path, request, M, fixture):
    from quantom_amd.abi import ANGLE_DTYPE, DIHEDRAL_DTYPE, PAIR_DTYPE
    s = request.getfixturevalue(fixture)
    comps, states, bridges, solv, ff = to_reference_bytes(s, M)
    for name, a in (("comp", comps), ("state", states), ("bridge", bridges), ("solv", solv), ("ff", ff)):
        a.tofile(tmp_path / f"{name}.bin")
    nb = len(bridges)
    r = subprocess.run([caller, str(tmp_path), str(M), str(s.n_compounds), str(s.n_solvents), str(nb)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    buf = (tmp_path / "out.bin").read_bytes()
    head, p = _read(buf, 0, "<i4", 4)
    nc, nch, ns, nbr = (int(v) for v in head)
    assert (nc, nch, ns, nbr) == (s.n_compounds, s.n_chain, s.n_solvents, nb)
    checks = [("compound_offset", "<i4", nc + 1), ("key_particle_index", "<i4", nc),
              ("confining_particle_sphere", "<f4", nc), ("chain_positions", "<f4", 3 * nch),
              ("chain_prev_positions", "<f4", 3 * nch), ("chain_atom_types", "u1", nch),
              ("bond_offset", "<i4", nc + 1), ("bonds", PAIR_DTYPE, len(s.bonds)),
              ("angle_offset", "<i4", nc + 1), ("angles", ANGLE_DTYPE, len(s.angles)),
              ("dihedral_offset", "<i4", nc + 1), ("dihedrals", DIHEDRAL_DTYPE, len(s.dihedrals)),
              ("ignore_offset", "<i4", nch + 1), ("ignore_compound", "<i4", len(s.ignore_compound)),
              ("ignore_local", "<i4", len(s.ignore_local)), ("bridge_particle_offset", "<i4", nb + 1),
              ("bridge_particle_compound", "<i4", len(s.bridge_particle_compound)),
              ("bridge_particle_local", "<i4", len(s.bridge_particle_local)),
              ("bridge_bond_offset", "<i4", nb + 1), ("bridge_bonds", PAIR_DTYPE, len(s.bridge_bonds)),
              ("bridge_angle_offset", "<i4", nb + 1), ("bridge_angles", ANGLE_DTYPE, len(s.bridge_angles)),
              ("bridge_dihedral_offset", "<i4", nb + 1), ("bridge_dihedrals", DIHEDRAL_DTYPE, len(s.bridge_dihedrals)),
              ("solvent_positions", "<f4", 3 * ns), ("solvent_prev_positions", "<f4", 3 * ns)]
    for name, dt, n in checks:
        got, p = _read(buf, p, dt, n)
        want = np.asarray(getattr(s, name)).astype(dt).reshape(-1)
        assert got.tobytes() == want.tobytes(), name
    (box,), p = _read(buf, p, "<f4", 1)
    (dt_,), p = _read(buf, p, "<f8", 1)
    (t0,), p = _read(buf, p, "<f4", 1)
    assert (box, dt_, t0) == (np.float32(7.5), 2e-6, np.float32(300.0))
    got, p = _read(buf, p, "<f4", 96)
    assert np.array_equal(got, ff.reshape(-1)) and p == len(buf)
    # the written-back state, read through the reference layout
    back = (tmp_path / "back.bin").read_bytes()
    c2, q = _read(back, 0, compound_dtype(M), nc)
    st2, q = _read(back, q, state_dtype(M), nc)
    sv2, q = _read(back, q, SOLVENT, ns)
    off = s.compound_offset
    for c in range(nc):
        n = off[c + 1] - off[c]
        assert np.array_equal(st2[c]["pos"][:n], s.chain_positions[off[c]:off[c + 1]] + np.float32(1))
        assert np.array_equal(c2[c]["prev"][:n], s.chain_prev_positions[off[c]:off[c + 1]] - np.float32(1))
        assert c2[c].tobytes()[4 + 12 * M:] == comps[c].tobytes()[4 + 12 * M:]     # nothing else touched
    assert np.array_equal(sv2["pos"], s.solvent_positions + np.float32(1))
    assert np.array_equal(sv2["prev"], s.solvent_prev_positions - np.float32(1))


def test_oversized_compound_is_rejected(caller, tmp_path, tiny_system):
    comps, states, bridges, solv, ff = to_reference_bytes(tiny_system, 64)
    comps[0]["n"] = 65                                     # more particles than MAX_COMPOUND_PARTICLES
    states[0]["n"] = 65
    for name, a in (("comp", comps), ("state", states), ("bridge", bridges), ("solv", solv), ("ff", ff)):
        a.tofile(tmp_path / f"{name}.bin")
    r = subprocess.run([caller, str(tmp_path), "64", str(len(comps)), str(len(solv)), str(len(bridges))],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "compound particle count" in r.stdout
