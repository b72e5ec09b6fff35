"""CPU restatement of the reference's system ingest (TEST INFRASTRUCTURE ONLY).

Imported by tests/ only, as the checker of quantom_amd's native ingest
(quantom_amd/csrc/ingest.cpp).  It follows three reference programs step by
step, with FP32 arithmetic done in numpy float32 in the reference's operation
order, so that its output can be compared byte for byte:

1. The CHARMM parameter merge (LIMA_ForcefieldMaker.cpp:116-141,
   ForcefieldMerger.h:10-240) and its text output (Filehandling.h:225-286).
2. The per-system filter (LIMA_ForcefieldMaker.cpp:19-114,
   ForcefieldTypes.h:17-838) and its text output (Filehandling.h:139-223).
3. The molecule and box build (Forcefield.cu:10-218, CompoundBuilder.cpp:7-479,
   Bodies.cuh:115-198, 389-423, 500-640, 644-700, BoxBuilder.cu:37-63,
   194-276, 431-461, BoxBuilder.cuh:9-11, 60-102).

Text readers follow each program's own tokeniser: the parameter programs skip
lines whose first non-blank character is listed (Filehandling.h:29-47) and
split on blanks and tabs (external files) or blanks and ';' (their own files,
Filehandling.h:50-79); the engine-side reader skips by the first character
only (Filehandling.cpp:7-15); the molecule builder splits .gro/.top lines on
single blanks (CompoundBuilder.cpp:324-463).

Where the reference's behaviour is undefined the restatement picks the
defined reading and says so: FTHelpers::isSorted with one name a prefix of the
other (ForcefieldTypes.h:42-56, no return) is taken as "sorted" when the left
name is the shorter; Float3(random(), random(), random()) (BoxBuilder.cu:274)
evaluates its arguments right to left, as GCC and MSVC do for this call.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

F32 = np.float32

_libc = C.CDLL(None)
_libc.strtof.restype = C.c_float
_libc.strtof.argtypes = [C.c_char_p, C.c_void_p]
_libc.strtod.restype = C.c_double
_libc.strtod.argtypes = [C.c_char_p, C.c_void_p]


class IngestError(Exception):
    """Where the reference prints and exits."""


def stof(s: str) -> np.float32:
    """std::stof: strtof on the token (correctly rounded to float)."""
    return F32(_libc.strtof(s.encode(), None))


def stod_f(s: str) -> np.float32:
    """A double literal narrowed to float (Float3(stod(..)), CompoundBuilder.cpp:454)."""
    return F32(_libc.strtod(s.encode(), None))


def stoi(s: str) -> int:
    """std::stoi: optional blanks and sign, then at least one digit."""
    t = s.lstrip(" \t\n\r\f\v")
    i = 1 if t[:1] in "+-" else 0
    j = i
    while j < len(t) and t[j].isdigit():
        j += 1
    if j == i:
        raise IngestError(f"stoi: no digits in {s!r}")
    return int(t[:j])


def to_string(x) -> str:
    """std::to_string(float): "%f" of the value."""
    return "%f" % float(x)


def _lines(path: str) -> list[str]:
    with open(path, "rb") as f:
        text = f.read().decode("latin-1")
    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    return lines


def _split(line: str, seps: str) -> list[str]:
    out, cur = [], []
    for ch in line:
        if ch in seps:
            if cur:
                out.append("".join(cur))
            cur = []
        else:
            cur.append(ch)
    if cur:
        out.append("".join(cur))
    return out


def read_param_file(path: str, ignores=(";", "#", "!"), internal=False) -> list[list[str]]:
    """Reader::readFile (Filehandling.h:81-110)."""
    rows = []
    for line in _lines(path):
        if len(line) == 0:
            continue
        first = " "
        for ch in line:
            if ch not in " \t":
                first = ch
                break
        if first in ignores:
            continue
        rows.append(_split(line, " ;" if internal else " \t"))
    return rows


def read_engine_file(path: str) -> list[list[str]]:
    """Filehandler::readFile (Filehandling.cpp:18-66)."""
    rows = []
    for line in _lines(path):
        if len(line) == 0 or line[0] in ";/":
            continue
        rows.append(_split(line, " ;"))
    return rows


# ----------------------------------------------------------------------------------------------
# 1. CHARMM merge (ForcefieldMerger.h)
# ----------------------------------------------------------------------------------------------

def _string_sorted(a: str, b: str) -> bool:
    """FTHelpers::isSorted (ForcefieldTypes.h:42-56); the no-return case reads as True."""
    p = 0
    while len(a) > p and len(b) > p:
        if ord(a[p]) < ord(b[p]):
            return True
        if ord(a[p]) > ord(b[p]):
            return False
        p += 1
    if len(a) > len(b):
        return False
    return True


def _outer_sort(a: str, b: str) -> tuple[str, str]:
    """Bondtype::sort / Angletype::sort(l, r) (ForcefieldTypes.h:399-423, 547-562)."""
    p = 0
    while len(a) > p and len(b) > p:
        if ord(a[p]) < ord(b[p]):
            return a, b
        if ord(a[p]) > ord(b[p]):
            return b, a
        p += 1
    if len(a) > len(b):
        return b, a
    return a, b


def sort_bond(t):
    return _outer_sort(t[0], t[1])


def sort_angle(t):
    a, c = _outer_sort(t[0], t[2])
    return a, t[1], c


def sort_dihedral(t):
    """Dihedraltype::sort (ForcefieldTypes.h:706-721)."""
    t1, t2, t3, t4 = t
    if t1 != t4:
        if not _string_sorted(t1, t4):
            return t4, t3, t2, t1
    else:
        if not _string_sorted(t2, t3):
            return t4, t3, t2, t1
    return t1, t2, t3, t4


@dataclass
class NBType:
    name: str
    mass: np.float32
    sigma: np.float32 = F32(-1)
    epsilon: np.float32 = F32(-1)
    atnum_local: int = 0


@dataclass
class Merged:
    nonbonded: list = field(default_factory=list)    # NBType
    bonds: list = field(default_factory=list)        # ((t1, t2), b0, kb)
    angles: list = field(default_factory=list)       # ((t1, t2, t3), theta0, ktheta)
    dihedrals: list = field(default_factory=list)    # ((t1..t4), phi0, kphi, n)


_PRM_STATES = {"ATOMS", "BONDS", "ANGLES", "DIHEDRALS", "NONBONDED", "IMPROPER", "CMAP"}


def _deg_to_rad(x):
    return F32(F32(F32(x * F32(2)) * stof("3.1415")) / F32(360))


def merge_charmm(paths: list[str]) -> Merged:
    """ForcefieldMerger::mergeForcefields (ForcefieldMerger.h:10-176): per file a fresh
    section state; the first ATOMS mass of a type wins, later NONBONDED rows overwrite
    sigma/epsilon (including rows after NBFIX, which keeps the NONBONDED state), the first
    bonded row of a sorted type tuple wins."""
    m = Merged()
    nb_index: dict[str, int] = {}
    bond_seen, angle_seen, dih_seen = set(), set(), set()
    for path in paths:
        state = None
        for row in read_param_file(path):
            if not row:
                continue
            if row[0] in _PRM_STATES:
                state = row[0]
            if state == "ATOMS":
                if len(row) < 4 or row[2] in nb_index:
                    continue
                nb_index[row[2]] = len(m.nonbonded)
                m.nonbonded.append(NBType(row[2], stof(row[3])))
            elif state == "NONBONDED":
                if len(row) < 4 or row[0] not in nb_index:
                    continue
                rec = m.nonbonded[nb_index[row[0]]]
                rec.epsilon = F32(abs(F32(stof(row[2]) * F32(4184))))
                rec.sigma = F32(F32(stof(row[3]) * F32(2)) / F32(10))
            elif state == "BONDS":
                if len(row) < 4:
                    continue
                key = sort_bond((row[0], row[1]))
                if key in bond_seen:
                    continue
                bond_seen.add(key)
                b0 = F32(stof(row[3]) * stof("0.1"))
                kb = F32(F32(stof(row[2]) * F32(4183)) * F32(100))
                m.bonds.append((key, b0, kb))
            elif state == "ANGLES":
                if len(row) < 5:
                    continue
                key = sort_angle((row[0], row[1], row[2]))
                if key in angle_seen:
                    continue
                angle_seen.add(key)
                m.angles.append((key, _deg_to_rad(stof(row[4])), F32(stof(row[3]) * F32(4183))))
            elif state == "DIHEDRALS":
                if len(row) < 7:
                    continue
                key = sort_dihedral((row[0], row[1], row[2], row[3]))
                if key in dih_seen:
                    continue
                dih_seen.add(key)
                m.dihedrals.append((key, _deg_to_rad(stof(row[6])), F32(stof(row[4]) * F32(4183)), stoi(row[5])))
    return m


def lima_ffnonbonded_text(m: Merged) -> str:
    """Printer::printFFNonbonded (Filehandling.h:225-242)."""
    out = ["// ----------------################ Forcefield Non-bonded ################---------------- //\n\n",
           "// ---------------- Non-bonded parameters ---------------- //\n",
           "// {atom_type \t mass [g/mol] \t sigma [nm] \t epsilon [J/mol]} //\n",
           "# ff_nonbonded\n"]
    for t in m.nonbonded:
        out.append(f"{t.name};{to_string(t.mass)};{to_string(t.sigma)};{to_string(t.epsilon)}\n")
    out.append("\n\n\n\n")
    return "".join(out)


def lima_ffbonded_text(m: Merged) -> str:
    """Printer::printFFBonded (Filehandling.h:244-286)."""
    out = ["// ----------------################ Forcefield Bonded ################---------------- //\n\n",
           "// ---------------- Bondtype parameters ---------------- //\n",
           "// {atom_types \t b_0 [nm] \t k_b [J/(mol*nm^2)] \t } //\n",
           "# ff_bondtypes\n"]
    for (t, b0, kb) in m.bonds:
        out.append(f"{t[0]};{t[1]};{to_string(b0)};{to_string(kb)}\n")
    out += ["\n\n\n\n", "// ---------------- Angletype parameters ---------------- //\n",
            "// {atom_types \t theta_0 [rad] \t k_theta [J/(mol*rad^2)] \t } //\n", "# ff_angletypes\n"]
    for (t, th, k) in m.angles:
        out.append(f"{t[0]};{t[1]};{t[2]};{to_string(th)};{to_string(k)}\n")
    out += ["\n\n\n\n", "// ---------------- Dihedraltype parameters ---------------- //\n",
            "// {atom_types \t phi_0 [rad] \t k_phi [J/(mol)] \t n} //\n", "# ff_dihedraltypes\n"]
    for (t, phi, k, n) in m.dihedrals:
        out.append(f"{t[0]};{t[1]};{t[2]};{t[3]};{to_string(phi)};{to_string(k)};{n}\n")
    out.append("\n\n\n\n")
    return "".join(out)


# ----------------------------------------------------------------------------------------------
# 2. Per-system filter (LIMA_ForcefieldMaker.cpp:19-114, ForcefieldTypes.h)
# ----------------------------------------------------------------------------------------------

WATER_MASS = F32(15.999000 + 2 * 1.008000)                       # ForcefieldTypes.h:17
WATER_SIGMA = F32(0.302905564168 + 2 * 0.040001352445)           # :18
WATER_EPSILON = F32(F32(stof("0.50208") + F32(F32(2) * stof("0.19246"))) * F32(1000))   # :19


_LIKENESS: dict = {}


def _likeness(a: str, b: str) -> np.float32:
    """FTHelpers::calcLikeness (ForcefieldTypes.h:29-40), cached per name pair."""
    key = (a, b)
    if key not in _LIKENESS:
        _LIKENESS[key] = _likeness_of(a, b)
    return _LIKENESS[key]


def _likeness_of(a: str, b: str) -> np.float32:
    scale = F32(F32(1) / F32(max(len(a), len(b))))
    like = F32(0)
    for i in range(min(len(a), len(b))):
        if a[i] == b[i]:
            like = F32(like + scale)
        else:
            break
    return like


def _best_match(query, table, threshold, early_break):
    """getBondFromTypes / getAngleFromTypes / getDihedralFromTypes
    (ForcefieldTypes.h:492-510, 646-666, 802-822): the first entry of the highest
    product of per-position likeness; angles and dihedrals stop at a product of 1."""
    best, best_like = None, F32(0)
    for entry in table:
        types = entry[0]
        like = _likeness(query[0], types[0])
        for q, t in zip(query[1:], types[1:]):
            like = F32(like * _likeness(q, t))
        if like > best_like:
            best_like, best = like, entry
        if early_break and like == F32(1):
            break
    if best is not None and best_like > F32(threshold):
        return best
    raise IngestError(f"no parameters for types {query}")


@dataclass
class Filtered:
    mappings: list                 # (conf name, type) in insertion order
    types: list                    # NBType with atnum_local, water first
    atoms: list                    # (topology id, type id)
    bonds: list                    # (ids, types, b0, kb)
    angles: list
    dihedrals: list


def _parse_lima_nonbonded(rows):
    """NB_Atomtype::parseNonbonded (ForcefieldTypes.h:172-204)."""
    recs = [NBType("WATER", WATER_MASS, WATER_SIGMA, WATER_EPSILON)]
    state = None
    for row in rows:
        if len(row) == 2:
            state = _FT_STATE.get(row[1], state)
            continue
        if state == "nb":
            recs.append(NBType(row[0], stof(row[1]), stof(row[2]), stof(row[3])))
    return recs


_FT_STATE = {"ff_nonbonded": "nb", "ff_bondtypes": "bond", "bonds": "bond", "ff_angletypes": "angle",
             "angles": "angle", "ff_dihedraltypes": "dih", "dihedrals": "dih", "pairs": "pair"}


def _parse_lima_bonded(rows):
    """parseFFBondtypes / parseFFAngletypes / parseFFDihedraltypes (ForcefieldTypes.h:426-449,
    572-599, 725-750); the angle parser stops at the dihedral section, the others read on."""
    bonds, angles, dihs = [], [], []
    state = None
    angle_done = False
    for row in rows:
        if len(row) == 2:
            state = _FT_STATE.get(row[1], state)
            if state == "dih":
                angle_done = True
            continue
        if state == "bond":
            bonds.append((sort_bond((row[0], row[1])), stof(row[2]), stof(row[3])))
        elif state == "angle" and not angle_done:
            angles.append((sort_angle((row[0], row[1], row[2])), stof(row[3]), stof(row[4])))
        elif state == "dih":
            dihs.append((sort_dihedral((row[0], row[1], row[2], row[3])), stof(row[4]), stof(row[5]), stoi(row[6])))
    return bonds, angles, dihs


def _topology_atoms(rows):
    """Atom::parseTopolAtoms (ForcefieldTypes.h:310-340): (id, bonded type, atom name)."""
    state = None
    atoms = []
    for row in rows:
        if len(row) == 3:
            if row[1] == "atoms":
                state = "atoms"
            elif row[1] == "bonds":
                state = "done"
            continue
        if state == "atoms":
            atoms.append((stoi(row[0]), row[1], row[4]))
        if state == "done":
            break
    return atoms


def _topology_terms(rows):
    """parseTopolBondtypes / parseTopolAngletypes / parseTopolDihedraltypes
    (ForcefieldTypes.h:451-475, 601-630, 752-783)."""
    bonds, angles, dihs = [], [], []
    state = None
    angles_open, dih_entered, dih_closed = True, False, False
    for row in rows:
        if len(row) == 3 and row[0][0] == "[":
            state = _FT_STATE.get(row[1], state)
            if state == "dih":
                angles_open = False
            if dih_entered:
                dih_closed = True
            continue
        if state == "bond":
            bonds.append((stoi(row[0]), stoi(row[1])))
        elif state == "angle" and angles_open:
            angles.append((stoi(row[0]), stoi(row[1]), stoi(row[2])))
        elif state == "dih" and not dih_closed:
            if len(row) != 5:
                continue
            dihs.append(tuple(stoi(x) for x in row[:4]))
            dih_entered = True
    return bonds, angles, dihs


def filter_forcefield(conf_path: str, top_path: str, nonbonded_text_rows, bonded_text_rows) -> Filtered:
    """prepareForcefieldForSimulation (LIMA_ForcefieldMaker.cpp:70-114)."""
    conf_names = []
    for row in read_param_file(conf_path):
        if len(row) == 6 and row[1] not in conf_names:
            conf_names.append(row[1])
    ff = _parse_lima_nonbonded(nonbonded_text_rows)
    first = {}
    for i, t in enumerate(ff):
        first.setdefault(t.name, i)
    filtered = [NBType(ff[0].name, ff[0].mass, ff[0].sigma, ff[0].epsilon, 0)]
    in_filtered = {ff[0].name}
    mappings, mapped = [], set()
    for name in conf_names:
        alias = name
        while alias not in first:
            if len(alias) > 1:
                alias = alias[:-1]
            else:
                raise IngestError(f"no non-bonded type for conf atom name {name!r}")
        rec = ff[first[alias]]
        if alias not in in_filtered:
            filtered.append(NBType(rec.name, rec.mass, rec.sigma, rec.epsilon, len(filtered)))
            in_filtered.add(alias)
        if name != alias and name not in mapped:
            mappings.append((name, alias))
            mapped.add(name)
    to_alias = dict(mappings)
    top_rows = read_param_file(top_path)
    atoms = []
    type_id = {t.name: t.atnum_local for t in filtered}
    for (aid, btype, name) in _topology_atoms(top_rows):
        alias = to_alias.get(name, name)
        if alias not in type_id:
            raise IngestError(f"topology atom {aid} ({name}) has no type")
        atoms.append((aid, btype, type_id[alias]))
    ffb, ffa, ffd = _parse_lima_bonded(bonded_text_rows)
    tb, ta, td = _topology_terms(top_rows)
    btype = lambda i: atoms[i - 1][1]
    memo: dict = {}

    def match(q, table, threshold, early):      # the search depends only on the type tuple
        if q not in memo:
            memo[q] = _best_match(q, table, threshold, early)
        return memo[q]

    bonds = []
    for ids in tb:
        q = sort_bond((btype(ids[0]), btype(ids[1])))
        e = match(q, ffb, 0.01, False)
        bonds.append((ids, q, e[1], e[2]))
    angles = []
    for ids in ta:
        q = sort_angle(tuple(btype(i) for i in ids))
        e = match(q, ffa, 0.01, True)
        angles.append((ids, q, e[1], e[2]))
    dihs = []
    for ids in td:
        q = sort_dihedral(tuple(btype(i) for i in ids))
        e = match(q, ffd, 0.001, True)
        dihs.append((ids, q, e[1], e[2], e[3]))
    return Filtered(mappings, filtered, [(a[0], a[2]) for a in atoms], bonds, angles, dihs)


def filtered_nonbonded_text(f: Filtered) -> str:
    """Printer::printForcefieldSummary (Filehandling.h:139-165)."""
    out = ["#pragma once\n",
           "// ----------------################ Forcefield Non-bonded ################---------------- //\n\n",
           "// ---------------- Mappings ---------------- //\n", "# mappings\n"]
    for (l, r) in f.mappings:
        out.append(f"{l};{r}\n")
    out += ["\n\n\n\n", "// ---------------- Non-bonded parameters ---------------- //\n",
            "// {atom_type \t type_id \t mass [g/mol] \t sigma [nm] \t epsilon [J/mol]} //\n", "# ff_nonbonded\n"]
    for t in f.types:
        out.append(f"{t.name};{t.atnum_local};{to_string(t.mass)};{to_string(t.sigma)};{to_string(t.epsilon)}\n")
    out.append("\n\n\n\n")
    return "".join(out)


def filtered_bonded_text(f: Filtered) -> str:
    """Printer::printForcefield (Filehandling.h:167-223)."""
    out = ["// ----------------################ Forcefield Non-bonded ################---------------- //\n\n",
           "// Atoms {particle id [simulation specific]\tatomtype_id [simulation specific]} //\n", "# atoms\n"]
    for (i, t) in f.atoms:
        out.append(f"{i};{t}\n")
    out += ["\n\n\n\n", "// ---------------- Bonds ---------------- //\n",
            "// {ID_p1 \t ID_p2 \t Atomtype \t Atomtype \t b_0 [nm] \t k_b [J/(mol * nm^2)]} //\n", "# bonds\n"]
    for (ids, t, b0, kb) in f.bonds:
        out.append(f"{ids[0]};{ids[1]};{t[0]};{t[1]};{to_string(b0)};{to_string(kb)}\n")
    out += ["\n\n\n\n", "// ---------------- Angles ---------------- //\n",
            "// {Atom-IDs \t Atomtypes \t theta_0 [rad] \t k_theta [J/(mol * rad^2)} //\n", "# angles\n"]
    for (ids, t, th, k) in f.angles:
        out.append(f"{ids[0]};{ids[1]};{ids[2]};{t[0]};{t[1]};{t[2]};{to_string(th)};{to_string(k)}\n")
    out += ["\n\n\n\n", "// ---------------- Dihedrals ---------------- //\n",
            "// {Atom IDs \t Atomtypes \t phi_0 [rad] \t k_phi [J/(mol * rad^2)] \t n} //\n", "# dihedrals\n"]
    for (ids, t, phi, k, n) in f.dihedrals:
        out.append(f"{ids[0]};{ids[1]};{ids[2]};{ids[3]};{t[0]};{t[1]};{t[2]};{t[3]};"
                   f"{to_string(phi)};{to_string(k)};{n}\n")
    out.append("\n\n\n\n")
    return "".join(out)


# ----------------------------------------------------------------------------------------------
# 3. Molecule and box (Forcefield.cu, CompoundBuilder.cpp, BoxBuilder.cu)
# ----------------------------------------------------------------------------------------------

@dataclass
class EngineForcefield:
    params: np.ndarray              # (32, 3) mass kg/mol, sigma, epsilon; unused slots -1
    atom_type: dict                 # topology id -> type id
    bonds: list                     # (id1, id2, b0, kb)
    angles: list
    dihedrals: list                 # (ids..., phi0, |kphi|, n)


def load_engine_forcefield(nonbonded_rows, bonded_rows, max_types: int = 32) -> EngineForcefield:
    """ForceFieldMaker::buildForcefield (Forcefield.cu:10-218)."""
    types = []
    state = None
    for row in nonbonded_rows:
        if row[0][0] == "#":
            state = row[1] if row[1] in ("ff_nonbonded", "atoms", "bonds", "angles", "dihedrals") else state
            continue
        if state == "ff_nonbonded":
            types.append((stof(row[2]), stof(row[3]), stof(row[4])))
    if len(types) > max_types:
        raise IngestError(f"{len(types)} atom types exceed the {max_types} slots")
    params = np.full((32, 3), -1, np.float32)
    for i, (m, s, e) in enumerate(types):
        params[i] = (F32(float(m) * 1e-3), s, e)
    atom_type, bonds, angles, dihs = {}, [], [], []
    state = None
    for row in bonded_rows:
        if row[0][0] == "#":
            state = row[1] if row[1] in ("ff_nonbonded", "atoms", "bonds", "angles", "dihedrals") else state
            continue
        if state == "atoms":
            atom_type[stoi(row[0])] = stoi(row[1])
        elif state == "bonds":
            bonds.append((stoi(row[0]), stoi(row[1]), stof(row[4]), stof(row[5])))
        elif state == "angles":
            angles.append((stoi(row[0]), stoi(row[1]), stoi(row[2]), stof(row[6]), stof(row[7])))
        elif state == "dihedrals":
            dihs.append((stoi(row[0]), stoi(row[1]), stoi(row[2]), stoi(row[3]), stof(row[8]),
                         F32(abs(stof(row[9]))), stoi(row[10])))
    return EngineForcefield(params, atom_type, bonds, angles, dihs)


def parse_gro(path: str):
    """CompoundBuilder::parseGRO (CompoundBuilder.cpp:415-479): (serial, name, residue name,
    residue id, position) for every line the reference accepts."""
    out = []
    for line in _lines(path):
        words = [w for w in line.split(" ") if w not in ("", "[", "]")]
        if not words:
            raise IngestError(f"{path}: blank line")
        digits = "".join(c for c in words[0] if c.isdigit() and c.isascii())
        name = "".join(c for c in words[0] if not (c.isdigit() and c.isascii()))
        valid = digits != "" and name != ""
        if len(words) == 5:       # an atom name fused with a 5-digit serial (CompoundBuilder.h:90-106)
            if words[1][0] == "O":
                words = [words[0], "OW", words[1][2:]] + words[2:5]
            else:
                words = [words[0], "HW", words[1][3:]] + words[2:5]
        if not valid or len(words) != 6:
            continue
        pos = np.array([stod_f(words[3]), stod_f(words[4]), stod_f(words[5])], np.float32)
        out.append((stoi(words[2]), words[1], name, int(digits), pos))
    return out


def parse_top(path: str):
    """CompoundBuilder::parseTOP (CompoundBuilder.cpp:324-360)."""
    recs = []
    for line in _lines(path):
        rec = [w for w in line.split(" ") if w not in ("", "[", "]")]
        if rec and rec[0] == ";":
            continue
        recs.append(rec)
    return recs


def _len3(v) -> np.float32:
    """Float3::len (QuantomTypes.cuh:60): sqrtf((x*x + y*y) + z*z)."""
    x, y, z = (F32(c) for c in v)
    return F32(np.sqrt(F32(F32(F32(x * x) + F32(y * y)) + F32(z * z))))


def _scale(v, a: float) -> np.ndarray:
    """Float3 * double (QuantomTypes.cuh:39): FP64 product, narrowed."""
    return np.array([F32(float(c) * float(a)) for c in v], np.float32)


def _com(points: np.ndarray) -> np.ndarray:
    """Compound::calcCOM (Bodies.cuh:362-369)."""
    inv = F32(F32(1) / F32(len(points)))
    com = np.zeros(3, np.float32)
    for p in points:
        com = (com + _scale(p, inv)).astype(np.float32)
    return com


@dataclass
class Compound:
    types: list = field(default_factory=list)
    pos: list = field(default_factory=list)
    gid: list = field(default_factory=list)
    bonds: list = field(default_factory=list)       # (local a, local b, b0, kb)
    angles: list = field(default_factory=list)
    dihedrals: list = field(default_factory=list)
    ignores: list = field(default_factory=list)     # per particle [(local, compound)]


@dataclass
class Bridge:
    left: int
    right: int
    refs: list = field(default_factory=list)        # (gid, compound, local)
    bonds: list = field(default_factory=list)       # bridge-local indices
    angles: list = field(default_factory=list)
    dihedrals: list = field(default_factory=list)


class _RandR:
    """The C library's rand() after srand(seed), from glibc's own random_r on a private
    128-byte state (the default generator of rand())."""

    class _Data(C.Structure):
        _fields_ = [("fptr", C.c_void_p), ("rptr", C.c_void_p), ("state", C.c_void_p), ("rand_type", C.c_int),
                    ("rand_deg", C.c_int), ("rand_sep", C.c_int), ("end_ptr", C.c_void_p)]

    def __init__(self, seed: int):
        self.buf = C.create_string_buffer(128)
        self.data = self._Data()
        if _libc.initstate_r(C.c_uint(seed), self.buf, C.c_size_t(128), C.byref(self.data)) != 0:
            raise IngestError("initstate_r failed")

    def __call__(self) -> int:
        out = C.c_int32()
        _libc.random_r(C.byref(self.data), C.byref(out))
        return out.value


def _random_unit(rng) -> float:
    """BoxBuilder::random (BoxBuilder.cuh:100-102)."""
    return float(F32(F32(F32(F32(rng() % 10000) / F32(10000)) * F32(2)) - F32(1)))


def _random_direction(rng) -> np.ndarray:
    z = _random_unit(rng)
    y = _random_unit(rng)
    x = _random_unit(rng)
    v = np.array([x, y, z], np.float32)
    ln = _len3(v)
    if ln == 0:
        return np.zeros(3, np.float32)
    return _scale(v, F32(F32(1) / ln))


SOLVENT_MASS = float(F32(18.01528)) * 1e-3                   # Simulation.cuh:30
BOXBUILDER_VRMS = math.sqrt(3 * 8.3144 * 313 / SOLVENT_MASS)  # BoxBuilder.cuh:63-71


@dataclass
class Ingested:
    compounds: list
    bridges: list
    solvents: np.ndarray            # (ns, 3)
    solvent_prev: np.ndarray
    forcefield: np.ndarray          # (32, 3)
    key: list
    sphere: list
    n_rand: int


def build_molecule(conf_path: str, top_path: str, eff: EngineForcefield, max_residue_id: int = 200,
                   min_residue_id: int = 0, ignore_hydrogens: bool = True, max_compound_particles: int = 64,
                   max_atoms_in_residue: int = 32, max_ignores: int = 32):
    """CompoundBuilder::buildMolecule (CompoundBuilder.cpp:7-48)."""
    records = parse_gro(conf_path)
    comps: list[Compound] = []
    bridges: list[Bridge] = []
    ref = {}
    cur_res = -1
    for (serial, name, resname, resid, pos) in records:              # loadParticles (83-150)
        if resid < min_residue_id:
            continue
        if resid > max_residue_id:
            break
        if ignore_hydrogens and name[0] == "H":
            continue
        if resname == "SOL":
            continue
        if resid != cur_res:
            if not comps or len(comps[-1].types) + max_atoms_in_residue > max_compound_particles:
                if comps:
                    bridges.append(Bridge(len(comps) - 1, len(comps)))
                comps.append(Compound())
            cur_res = resid
        c = comps[-1]
        if len(c.types) == max_compound_particles:
            raise IngestError("compound full")
        if serial > len(eff.atom_type) or serial == 0:
            raise IngestError(f"atom {serial} has no type")
        ref[serial] = (len(comps) - 1, len(c.types))
        c.types.append(eff.atom_type[serial])
        c.pos.append(pos)
        c.gid.append(serial)
        c.ignores.append([])

    def lookup(table, ids):
        for e in table:
            if tuple(e[:len(ids)]) == ids:
                return e
        raise IngestError(f"no parameters for term {ids}")

    mode, dih_sections = None, 0
    for rec in parse_top(top_path):                                   # loadTopology (152-175)
        if not rec:
            mode = None
            continue
        if mode is None:
            mode = {"bonds": 2, "angles": 3, "dihedrals": 4}.get(rec[0])
            if mode == 4:
                dih_sections += 1
            continue
        if mode == 4 and dih_sections > 1:
            continue
        ids = tuple(stoi(x) for x in rec[:mode])
        if not all(i in ref for i in ids):
            continue
        refs = [ref[i] for i in ids]
        e = lookup({2: eff.bonds, 3: eff.angles, 4: eff.dihedrals}[mode], ids)
        for a in range(mode):                                         # distributeLJIgnores (307-318)
            for b in range(mode):
                if a == b:
                    continue
                lst = comps[refs[a][0]].ignores[refs[a][1]]
                tgt = (refs[b][1], refs[b][0])
                if tgt not in lst:
                    if len(lst) == max_ignores:
                        raise IngestError("exclusion list full")
                    lst.append(tgt)
        cids = sorted(set(r[0] for r in refs))
        if len(cids) == 1:
            c = comps[cids[0]]
            loc = tuple(r[1] for r in refs)
            {2: c.bonds, 3: c.angles, 4: c.dihedrals}[mode].append(loc + tuple(e[mode:]))
        else:
            # GenericBond (Bodies.cuh:118-131): the first compound, and the last other one seen
            other = [r[0] for r in refs if r[0] != refs[0][0]][-1]
            key = tuple(sorted((refs[0][0], other)))
            br = next((b for b in bridges if (b.left, b.right) == key), None)
            if br is None:
                raise IngestError(f"term {ids} spans compounds {key} without a bridge")
            for g, r in zip(ids, refs):
                if not any(x[0] == g for x in br.refs):
                    if len(br.refs) == 32:
                        raise IngestError("bridge full")
                    br.refs.append((g, r[0], r[1]))
            loc = tuple(next(k for k, x in enumerate(br.refs) if x[0] == g) for g in ids)
            {2: br.bonds, 3: br.angles, 4: br.dihedrals}[mode].append(loc + tuple(e[mode:]))
    keys, spheres = [], []
    for c in comps:                                                   # calcParticleSphere (407-423)
        pts = np.array(c.pos, np.float32)
        com = _com(pts)
        closest, far, ci = F32(np.inf), F32(-np.inf), 0
        for i, p in enumerate(pts):
            d = _len3((p - com).astype(np.float32))
            ci = i if d < closest else ci
            closest = min(closest, d)
            far = max(far, d)
        keys.append(ci)
        spheres.append(far)
    return comps, bridges, keys, spheres, records


def build_box(conf_path: str, top_path: str, eff: EngineForcefield, box_len: float = 7.0, dt: float = 1e-7,
              seed: int = 290128309, solvents: bool = True, **molecule_kw) -> Ingested:
    """Environment::Environment's build (Environment.cpp:16-45): molecule, BoxBuilder
    centring (BoxBuilder.cu:37-63), compounds at rest (248-269) and the .gro waters'
    oxygens placed where there is room (194-218, 431-461) with v_rms velocities (273-276)."""
    comps, bridges, keys, spheres, records = build_molecule(conf_path, top_path, eff, **molecule_kw)
    rng = _RandR(seed)
    n_rand = 0
    inv = F32(F32(1) / F32(len(comps)))
    mcom = np.zeros(3, np.float32)
    for c in comps:
        mcom = (mcom + _scale(_com(np.array(c.pos, np.float32)), inv)).astype(np.float32)
    offset = (np.full(3, F32(box_len) / F32(2), np.float32) - mcom).astype(np.float32)
    for c in comps:
        c.pos = [(p + offset).astype(np.float32) for p in c.pos]
        _random_direction(rng)
        n_rand += 3
    chain = np.array([p for c in comps for p in c.pos], np.float32).reshape(-1, 3)
    sol, sol_prev = [], []
    if solvents:
        for (serial, name, resname, resid, pos) in records:          # getSolventPositions (51-62)
            if resname != "SOL" or name[0] != "O":
                continue
            p = (pos + offset).astype(np.float32)
            d = chain - p
            dc = np.sqrt(((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]).astype(np.float32))
            if len(dc) and float(dc.min()) < 0.2:
                continue
            if sol:
                s = np.array(sol, np.float32) - p
                ds = np.sqrt(((s[:, 0] * s[:, 0] + s[:, 1] * s[:, 1]) + s[:, 2] * s[:, 2]).astype(np.float32))
                if float(ds.min()) < 0.25:
                    continue
            vel = _scale(_random_direction(rng), BOXBUILDER_VRMS)
            n_rand += 3
            sol.append(p)
            sol_prev.append((p - _scale(vel, dt)).astype(np.float32))
    return Ingested(comps, bridges, np.array(sol, np.float32).reshape(-1, 3),
                    np.array(sol_prev, np.float32).reshape(-1, 3), eff.params.copy(), keys, spheres, n_rand)


def to_lima_system(ing: Ingested, box_len: float, dt: float, thermostat_target: float = 310.0):
    """The ingested box in the engine's boundary layout (quantom_amd.abi.LimaSystem)."""
    from quantom_amd.abi import ANGLE_DTYPE, DIHEDRAL_DTYPE, PAIR_DTYPE, LimaSystem

    def offsets(counts):
        return np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)

    def pack(rows, dtype, n):
        a = np.zeros(len(rows), dtype)
        for k, r in enumerate(rows):
            a[k]["idx"] = r[:n]
            vals = r[n:]
            if dtype is PAIR_DTYPE:
                a[k]["b0"], a[k]["kb"] = vals
            elif dtype is ANGLE_DTYPE:
                a[k]["theta_0"], a[k]["k_theta"] = vals
            else:
                a[k]["phi_0"], a[k]["k_phi"], a[k]["n"] = vals
        return a

    cs = ing.compounds
    pos = np.array([p for c in cs for p in c.pos], np.float32).reshape(-1, 3)
    ign = [t for c in cs for lst in c.ignores for t in [lst]]
    return LimaSystem(
        box_len=box_len, dt=dt, thermostat_target=thermostat_target, forcefield=ing.forcefield,
        compound_offset=offsets([len(c.types) for c in cs]),
        key_particle_index=np.array(ing.key, np.int32), confining_particle_sphere=np.array(ing.sphere, np.float32),
        chain_positions=pos, chain_prev_positions=pos.copy(),
        chain_atom_types=np.array([t for c in cs for t in c.types], np.uint8),
        bond_offset=offsets([len(c.bonds) for c in cs]), bonds=pack([b for c in cs for b in c.bonds], PAIR_DTYPE, 2),
        angle_offset=offsets([len(c.angles) for c in cs]),
        angles=pack([b for c in cs for b in c.angles], ANGLE_DTYPE, 3),
        dihedral_offset=offsets([len(c.dihedrals) for c in cs]),
        dihedrals=pack([b for c in cs for b in c.dihedrals], DIHEDRAL_DTYPE, 4),
        ignore_offset=offsets([len(l) for l in ign]),
        ignore_compound=np.array([t[1] for l in ign for t in l], np.int32),
        ignore_local=np.array([t[0] for l in ign for t in l], np.int32),
        bridge_particle_offset=offsets([len(b.refs) for b in ing.bridges]),
        bridge_particle_compound=np.array([r[1] for b in ing.bridges for r in b.refs], np.int32),
        bridge_particle_local=np.array([r[2] for b in ing.bridges for r in b.refs], np.int32),
        bridge_bond_offset=offsets([len(b.bonds) for b in ing.bridges]),
        bridge_bonds=pack([x for b in ing.bridges for x in b.bonds], PAIR_DTYPE, 2),
        bridge_angle_offset=offsets([len(b.angles) for b in ing.bridges]),
        bridge_angles=pack([x for b in ing.bridges for x in b.angles], ANGLE_DTYPE, 3),
        bridge_dihedral_offset=offsets([len(b.dihedrals) for b in ing.bridges]),
        bridge_dihedrals=pack([x for b in ing.bridges for x in b.dihedrals], DIHEDRAL_DTYPE, 4),
        solvent_positions=ing.solvents, solvent_prev_positions=ing.solvent_prev)
