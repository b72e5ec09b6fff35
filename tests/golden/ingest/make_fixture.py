"""Writes the small ingest fixture (mini_a.prm, mini_b.prm, mini.gro, mini.top).

Hand-made inputs in the CHARMM parameter and GROMACS formats, sized so that every
rule of the ingest shows up once: duplicate and later-file parameter rows, an NBFIX
row, a bond type found only by name likeness, hydrogens and terms that name them,
two [ dihedrals ] sections, water lines with a fused five-digit serial, and waters
that are rejected for lying too close to the chain or to another water.
Run from this directory: `python make_fixture.py`.
"""
import math

PRM_A = """* hand-made parameter file A
*
read param card flex append

ATOMS
MASS  -1  H        1.00800 ! polar H
MASS  -1  HC       1.00800 ! N-ter H
MASS  -1  HA       1.00800 ! aliphatic H
MASS  -1  HB       1.00800
MASS  -1  C       12.01100 ! carbonyl C
MASS  -1  CC      12.01100
MASS  -1  CT1     12.01100
MASS  -1  CT2     12.01100
MASS  -1  CT3     12.01100
MASS  -1  N       14.00700
MASS  -1  NH1     14.00700
MASS  -1  NH3     14.00700
MASS  -1  O       15.99900
MASS  -1  OC      15.99900
MASS  -1  OH1     15.99900
MASS  -1  OT      15.99940 ! water O

BONDS
!atom types  Kb  b0
CT1  NH3   200.000     1.4800
CT1  CT3   222.500     1.5380
C    CT1   250.000     1.4900
C    O     620.000     1.2300
C    NH1   370.000     1.3450
H    NH1   440.000     0.9970
CT2  NH1   320.000     1.4300
C    CT2   250.000     1.4900
CT1  NH1   320.000     1.4300
CT1  CT2   222.500     1.5380
CT2  OH1   428.000     1.4200
CT2  CT1   999.000     9.9999 ! reversed duplicate, dropped
CC   OC    525.000     1.2600
CT1  HB    309.000     1.1110
H    OH1   545.000     0.9600

ANGLES
!atom types     Ktheta    Theta0
NH3  CT1  CT3    67.700    110.00
NH3  CT1  C      43.700    110.00
CT3  CT1  C      52.000    108.00
CT1  C    O      80.000    121.00
CT1  C    NH1    80.000    116.50
O    C    NH1    80.000    122.50
C    NH1  CT2    50.000    120.00
NH1  CT2  C      50.000    107.00
CT2  C    O      80.000    121.00
CT2  C    NH1    80.000    116.50
C    NH1  CT1    50.000    120.00
NH1  CT1  CT2    70.000    113.50
NH1  CT1  C      50.000    107.00
CT1  CT2  OH1    75.700    110.10
CT1  C    NH1    99.000    199.00 ! duplicate, dropped
NH1  CT1  CT3    70.000    113.50
CT3  CT1  CC     52.000    108.00
NH1  CT1  CC     50.000    107.00
CT1  CC   OC     40.000    118.00
OC   CC   OC    100.000    124.00
HC   NH3  CT1    30.000    109.50
CT2  OH1  H      57.500    106.00

DIHEDRALS
!atom types             Kchi    n   delta
X    CT1  C    X        0.0000  1     0.00
X    C    NH1  X        2.5000  2   180.00
X    CT2  C    X        0.0000  1     0.00
X    CT1  CT2  X        0.2000  3     0.00
X    CT1  NH1  X        0.0000  1     0.00
NH1  C    CT1  NH1      0.6000  1     0.00
NH1  C    CT1  NH1      0.2000  2   180.00 ! second multiplicity of the same types, dropped
CT3  CT1  C    O        1.4000  1     0.00
CT1  C    NH1  CT2      1.6000  1     0.00
C    CT2  NH1  C        0.2000  1   180.00
NH1  C    CT2  NH1      0.6000  1     0.00
CT1  NH1  C    CT2      1.6000  1     0.00
C    CT1  NH1  C        0.2000  1   180.00
NH1  CT1  CT2  OH1      0.2000  3     0.00
CT1  C    NH1  CT1      1.6000  1     0.00
C    NH1  CT1  CC       0.3000  2     0.00
NH1  CT1  CC   OC       0.0500  6   180.00
HC   NH3  CT1  CT3      0.1000  3     0.00

IMPROPER
O    X    X    C      120.0000         0      0.0000

NONBONDED nbxmod  5 atom cdiel fshift vatom vdistance vfswitch -
cutnb 14.0 ctofnb 12.0 ctonnb 10.0 eps 1.0 e14fac 1.0 wmin 1.5
H        0.0       -0.046      0.2245
HA       0.0       -0.022      1.3200
C        0.0       -0.110      2.0000
CT1      0.0       -0.0200     2.2750   0.0 -0.01 1.9
N        0.0       -0.200      1.8500
O        0.0       -0.120      1.7000
OT       0.0       -0.1521     1.7682

NBFIX
!              Emin         Rmin
O        NH1      -0.200       3.100 ! overrides O, NONBONDED state continues

HBOND CUTHB 0.5

END
"""

PRM_B = """* hand-made parameter file B
*
ATOMS
MASS  -1  C       13.00000 ! duplicate type, first mass kept
MASS  -1  S       32.06000

BONDS
C    CT1   111.000     1.1100 ! already defined by file A, dropped
CT1  CT1   222.500     1.5000

NONBONDED nbxmod  5 atom cdiel fshift vatom vdistance vfswitch -
C        0.0       -0.090      2.1000 ! later file overrides
S        0.0       -0.450      2.0000

END
"""

# residue id, residue name, atom name, topology type, x, y, z
ATOMS = [
    (1, "ALA", "N", "NH3"), (1, "ALA", "H1", "HC"), (1, "ALA", "CA", "CT1"), (1, "ALA", "HA", "HB"),
    (1, "ALA", "CB", "CT3"), (1, "ALA", "C", "C"), (1, "ALA", "O", "O"),
    (2, "GLY", "N", "NH1"), (2, "GLY", "HN", "H"), (2, "GLY", "CA", "CT2"), (2, "GLY", "HA1", "HB"),
    (2, "GLY", "C", "C"), (2, "GLY", "O", "O"),
    (3, "SER", "N", "NH1"), (3, "SER", "HN", "H"), (3, "SER", "CA", "CT1"), (3, "SER", "CB", "CT2"),
    (3, "SER", "OG", "OH1"), (3, "SER", "HG1", "H"), (3, "SER", "C", "C"), (3, "SER", "O", "O"),
    (4, "ALA", "N", "NH1"), (4, "ALA", "HN", "H"), (4, "ALA", "CA", "CT1"), (4, "ALA", "CB", "CT3"),
    (4, "ALA", "C", "CC"), (4, "ALA", "OT1", "OC"), (4, "ALA", "OT2", "OC"),
]

BONDS = [(1, 2), (1, 3), (3, 4), (3, 5), (3, 6), (6, 7), (6, 8), (8, 9), (8, 10), (10, 11), (10, 12), (12, 13),
         (12, 14), (14, 15), (14, 16), (16, 17), (16, 20), (17, 18), (18, 19), (20, 21), (20, 22), (22, 23),
         (22, 24), (24, 25), (24, 26), (26, 27), (26, 28)]
ANGLES = [(2, 1, 3), (1, 3, 5), (1, 3, 6), (5, 3, 6), (3, 6, 7), (3, 6, 8), (7, 6, 8), (6, 8, 10), (8, 10, 12),
          (10, 12, 13), (10, 12, 14), (12, 14, 16), (14, 16, 17), (14, 16, 20), (16, 17, 18), (17, 18, 19),
          (16, 20, 21), (16, 20, 22), (20, 22, 24), (22, 24, 25), (22, 24, 26), (24, 26, 27), (24, 26, 28),
          (27, 26, 28)]
DIHEDRALS = [(2, 1, 3, 5), (1, 3, 6, 8), (5, 3, 6, 7), (3, 6, 8, 10), (6, 8, 10, 12), (8, 10, 12, 14),
             (10, 12, 14, 16), (12, 14, 16, 20), (14, 16, 17, 18), (14, 16, 20, 22), (16, 20, 22, 24),
             (20, 22, 24, 26), (22, 24, 26, 27)]
IMPROPERS = [(6, 3, 8, 7), (20, 16, 22, 21)]


def positions():
    """Heavy atoms on a helix along x (no three consecutive atoms on a line, no four in a
    plane); each hydrogen 0.1 nm beside the heavy atom before it."""
    out = []
    heavy = 0
    last = (2.0, 3.0, 3.0)
    for (res, rn, name, t) in ATOMS:
        if name[0] == "H":
            x, y, z = last[0], last[1] + 0.1, last[2]
        else:
            a = 2.1 * heavy
            x, y, z = 2.0 + 0.13 * heavy, 3.0 + 0.09 * math.cos(a), 3.0 + 0.09 * math.sin(a)
            last = (x, y, z)
            heavy += 1
        out.append((x, y, z))
    return out


def gro():
    lines = ["hand-made tetrapeptide in water", " 40"]
    pos = positions()
    for i, ((res, rn, name, t), (x, y, z)) in enumerate(zip(ATOMS, pos), start=1):
        lines.append("%5d%-5s%5s%5d%8.3f%8.3f%8.3f" % (res, rn, name, i, x, y, z))
    near = pos[4]                              # the third heavy atom (CB of residue 1)
    waters = [(5, (1.0, 1.0, 1.0)),            # kept
              (6, (near[0], near[1] + 0.12, near[2])),   # within 0.2 nm of a chain atom: dropped
              (7, (1.1, 1.05, 1.0)),           # within 0.25 nm of water 5: dropped
              (8, (5.5, 5.0, 4.0))]            # kept
    serial = len(ATOMS)
    for res, (x, y, z) in waters:
        for name, d in (("OW", 0.0), ("HW1", 0.08), ("HW2", -0.08)):
            serial += 1
            lines.append("%5d%-5s%5s%5d%8.3f%8.3f%8.3f" % (res, "SOL", name, serial, x + d, y, z))
    # a water whose serials have five digits: the atom name runs into the serial
    for name, d, s in (("OW", 0.0, 10001), ("HW1", 0.08, 10002), ("HW2", -0.08, 10003)):
        lines.append("%5d%-5s%5s%5d%8.3f%8.3f%8.3f" % (9, "SOL", name, s, 6.0 + d, 1.5, 2.0))
    lines.append("   7.00000   7.00000   7.00000")
    return "\n".join(lines) + "\n"


def top():
    L = [";", ";\tFile 'mini.top' was written by hand", ";", "", "; Include forcefield parameters",
         '#include "mini.ff/forcefield.itp"', "", "[ moleculetype ]", "; Name            nrexcl",
         "Peptide_chain_A     3", "", "[ atoms ]",
         ";   nr       type  resnr residue  atom   cgnr     charge       mass"]
    for i, (res, rn, name, t) in enumerate(ATOMS, start=1):
        L.append("%6d %10s %6d %6s %6s %6d %10.3f %10.3f" % (i, t, res, rn, name, i, 0.0, 1.0))
    L += ["", "[ bonds ]", ";  ai    aj funct            c0            c1            c2            c3"]
    L += ["%5d %5d %5d " % (a, b, 1) for a, b in BONDS]
    L += ["", "[ pairs ]", ";  ai    aj funct"]
    L += ["%5d %5d %5d " % (a, d, 1) for a, b, c, d in DIHEDRALS]
    L += ["", "[ angles ]", ";  ai    aj    ak funct"]
    L += ["%5d %5d %5d %5d " % (a, b, c, 5) for a, b, c in ANGLES]
    L += ["", "[ dihedrals ]", ";  ai    aj    ak    al funct"]
    L += ["%5d %5d %5d %5d %5d " % (a, b, c, d, 9) for a, b, c, d in DIHEDRALS]
    L += ["", "[ dihedrals ]", ";  ai    aj    ak    al funct"]
    L += ["%5d %5d %5d %5d %5d " % (a, b, c, d, 2) for a, b, c, d in IMPROPERS]
    L += ["", "; Include water topology", '#include "mini.ff/tip3p.itp"', "", "[ system ]", "; Name",
          "hand-made tetrapeptide in water", "", "[ molecules ]", "; Compound        #mols",
          "Peptide_chain_A     1", "SOL                 5"]
    return "\n".join(L) + "\n"


if __name__ == "__main__":
    for name, text in (("mini_a.prm", PRM_A), ("mini_b.prm", PRM_B), ("mini.gro", gro()), ("mini.top", top())):
        with open(name, "w") as f:
            f.write(text)
