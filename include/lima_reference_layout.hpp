/* lima_reference_layout.hpp — the reference's in-memory system (Box/Simulation
 * arrays of Compound, CompoundState, Solvent, CompoundBridgeBundleCompact and
 * ForceField) converted to lima_system, and the stepped state written back.
 *
 * The structs below restate the byte layout of Bodies.cuh:19-25, 61-113,
 * 169-198, 203-292, 303-472, 644-732 and Forcefield.cuh:12-27, for a reference
 * build with MAX_COMPOUND_PARTICLES = MAXP (Constants.cuh:58: 64 by default)
 * and LIMA_DEBUGMODE defined (Compound::particle_global_ids present).  A caller
 * that holds the reference's arrays passes them with a reinterpret_cast; the
 * static_asserts pin every size and offset the two layouts must share.
 *
 *   Compound<64>  24,036 B    Compound<128>  30,116 B
 *   CompoundState<64> 772 B   CompoundState<128> 1,540 B
 *   CompoundBridgeCompact 2,352 B, NeighborList 2,180 B, Solvent 24 B, ForceField 384 B
 *
 * Header-only C++17; needs only lima.h.
 */
#pragma once
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "lima.h"

namespace lima {
namespace ref {

struct Float3 {            /* QuantomTypes.cuh:33: three floats, no padding */
    float x, y, z;
};
struct Solvent {           /* Bodies.cuh:19-25 */
    Float3 pos, pos_tsub1;
};
struct PairBond {          /* Bodies.cuh:61-80 */
    float b0, kb;
    uint32_t atom_indexes[2];
};
struct AngleBond {         /* Bodies.cuh:82-93 */
    float theta_0, k_theta;
    uint32_t atom_indexes[3];
};
struct DihedralBond {      /* Bodies.cuh:95-108 */
    float phi_0, k_phi;
    uint8_t n;
    uint32_t atom_indexes[4];
};
struct LJ_Ignores {        /* Bodies.cuh:169-198 */
    static const int max_ids = 32;
    uint8_t local_ids[max_ids];
    uint8_t compound_ids[max_ids];
    uint8_t n_ignores;
};
struct NeighborList {      /* Bodies.cuh:203-292, NEIGHBORLIST_MAX_* of Constants.cuh:41-42 */
    uint16_t neighborcompound_ids[64];
    uint16_t n_compound_neighbors;
    uint16_t neighborsolvent_ids[1024];
    uint16_t n_solvent_neighbors;
};

constexpr int kMaxPairBonds = 128, kMaxAngleBonds = 256, kMaxDihedrals = 384;   /* Bodies.cuh:300-302 */

template <int MAXP> struct CompoundState {   /* Bodies.cuh:303-315 */
    Float3 positions[MAXP];
    uint8_t n_particles;
};

template <int MAXP> struct Compound {        /* Bodies.cuh:318-347 */
    uint8_t n_particles;
    Float3 prev_positions[MAXP];
    Float3 forces[MAXP];
    uint8_t atom_types[MAXP];
    uint8_t atom_color_types[MAXP];
    LJ_Ignores lj_ignore_list[MAXP];
    uint32_t particle_global_ids[MAXP];
    Float3 center_of_mass;
    uint16_t n_singlebonds;
    PairBond singlebonds[kMaxPairBonds];
    uint16_t n_anglebonds;
    AngleBond anglebonds[kMaxAngleBonds];
    uint16_t n_dihedrals;
    DihedralBond dihedrals[kMaxDihedrals];
    int32_t key_particle_index;
    float confining_particle_sphere;
};

constexpr int kBridgeParticles = 32, kBridgeTerms = 32, kBridgesInBundle = 64;   /* Constants.cuh:109-113 */

struct ParticleRefCompact {                   /* Bodies.cuh:644-650 */
    int32_t compound_id, local_id_compound;
};
struct CompoundBridgeCompact {                /* Bodies.cuh:652-725 */
    ParticleRefCompact particle_refs[kBridgeParticles];
    uint8_t atom_types[kBridgeParticles];
    int32_t n_particles;
    uint16_t n_singlebonds;
    PairBond singlebonds[kBridgeTerms];
    uint16_t n_anglebonds;
    AngleBond anglebonds[kBridgeTerms];
    uint16_t n_dihedrals;
    DihedralBond dihedrals[kBridgeTerms];
};
struct CompoundBridgeBundleCompact {          /* Bodies.cuh:727-732 */
    CompoundBridgeCompact compound_bridges[kBridgesInBundle];
    int32_t n_bridges;
};
struct ParticleParameters {                   /* Forcefield.cuh:17-21 */
    float mass, sigma, epsilon;
};
struct ForceField {                           /* Forcefield.cuh:23-27 */
    ParticleParameters particle_parameters[32];
};

static_assert(sizeof(Float3) == 12 && sizeof(Solvent) == 24, "Float3/Solvent layout");
static_assert(sizeof(PairBond) == 16 && sizeof(AngleBond) == 20 && sizeof(DihedralBond) == 28, "bond layout");
static_assert(offsetof(DihedralBond, atom_indexes) == 12, "DihedralBond layout");
static_assert(sizeof(LJ_Ignores) == 65 && sizeof(NeighborList) == 2180, "LJ_Ignores/NeighborList layout");
static_assert(offsetof(NeighborList, neighborsolvent_ids) == 130 && offsetof(NeighborList, n_solvent_neighbors) == 2178,
              "NeighborList layout");
static_assert(sizeof(CompoundState<64>) == 772 && sizeof(CompoundState<128>) == 1540, "CompoundState layout");
static_assert(sizeof(Compound<64>) == 24036 && sizeof(Compound<128>) == 30116, "Compound layout");
static_assert(offsetof(Compound<128>, prev_positions) == 4 && offsetof(Compound<128>, forces) == 1540 &&
              offsetof(Compound<128>, atom_types) == 3076 && offsetof(Compound<128>, lj_ignore_list) == 3332 &&
              offsetof(Compound<128>, singlebonds) == 12180 && offsetof(Compound<128>, anglebonds) == 14232 &&
              offsetof(Compound<128>, dihedrals) == 19356 && offsetof(Compound<128>, key_particle_index) == 30108 &&
              offsetof(Compound<128>, confining_particle_sphere) == 30112, "Compound<128> offsets");
static_assert(sizeof(CompoundBridgeCompact) == 2352 && offsetof(CompoundBridgeCompact, singlebonds) == 296 &&
              offsetof(CompoundBridgeCompact, dihedrals) == 1456, "CompoundBridgeCompact layout");
static_assert(sizeof(ForceField) == 384 && sizeof(lima_forcefield) == 384, "ForceField layout");

/* The flattened arrays of one system and the lima_system that points into them.
 * Non-copyable: `system` holds pointers into this object's vectors. */
class System {
public:
    System() = default;
    System(const System&) = delete;
    System& operator=(const System&) = delete;

    lima_system system{};
    std::vector<int32_t> compound_offset, key_particle_index, bond_offset, angle_offset, dihedral_offset;
    std::vector<float> sphere, chain_positions, chain_prev_positions, solvent_positions, solvent_prev_positions;
    std::vector<uint8_t> chain_atom_types;
    std::vector<lima_pair_bond> bonds, bridge_bonds;
    std::vector<lima_angle_bond> angles, bridge_angles;
    std::vector<lima_dihedral_bond> dihedrals, bridge_dihedrals;
    std::vector<int32_t> ignore_offset, ignore_compound, ignore_local;
    std::vector<int32_t> bridge_particle_offset, bridge_particle_compound, bridge_particle_local;
    std::vector<int32_t> bridge_bond_offset, bridge_angle_offset, bridge_dihedral_offset;

    void point() {
        lima_system& s = system;
        s.n_compounds = (int32_t)key_particle_index.size();
        s.compound_offset = compound_offset.data();
        s.key_particle_index = key_particle_index.data();
        s.confining_particle_sphere = sphere.data();
        s.chain_positions = chain_positions.data();
        s.chain_prev_positions = chain_prev_positions.data();
        s.chain_atom_types = chain_atom_types.data();
        s.bond_offset = bond_offset.data(); s.bonds = bonds.data();
        s.angle_offset = angle_offset.data(); s.angles = angles.data();
        s.dihedral_offset = dihedral_offset.data(); s.dihedrals = dihedrals.data();
        s.ignore_offset = ignore_offset.data();
        s.ignore_compound = ignore_compound.data();
        s.ignore_local = ignore_local.data();
        s.n_bridges = (int32_t)bridge_particle_offset.size() - 1;
        s.bridge_particle_offset = bridge_particle_offset.data();
        s.bridge_particle_compound = bridge_particle_compound.data();
        s.bridge_particle_local = bridge_particle_local.data();
        s.bridge_bond_offset = bridge_bond_offset.data(); s.bridge_bonds = bridge_bonds.data();
        s.bridge_angle_offset = bridge_angle_offset.data(); s.bridge_angles = bridge_angles.data();
        s.bridge_dihedral_offset = bridge_dihedral_offset.data(); s.bridge_dihedrals = bridge_dihedrals.data();
        s.n_solvents = (int32_t)(solvent_positions.size() / 3);
        s.solvent_positions = solvent_positions.data();
        s.solvent_prev_positions = solvent_prev_positions.data();
    }
};

namespace detail {
inline void fail(const std::string& what) { throw std::invalid_argument("reference layout: " + what); }
inline void put3(std::vector<float>& v, const Float3& p) {
    v.push_back(p.x); v.push_back(p.y); v.push_back(p.z);
}
inline lima_pair_bond conv(const PairBond& b) { return {b.b0, b.kb, {b.atom_indexes[0], b.atom_indexes[1]}}; }
inline lima_angle_bond conv(const AngleBond& b) {
    return {b.theta_0, b.k_theta, {b.atom_indexes[0], b.atom_indexes[1], b.atom_indexes[2]}};
}
inline lima_dihedral_bond conv(const DihedralBond& b) {
    return {b.phi_0, b.k_phi, (uint32_t)b.n, {b.atom_indexes[0], b.atom_indexes[1], b.atom_indexes[2], b.atom_indexes[3]}};
}
}  // namespace detail

/* Box contents -> lima_system (Engine::Engine reads the same fields, Engine.cu:9-38).
 * box_len is BOX_LEN (Constants.cuh:35), dt is Box::dt (Simulation.cuh:60), the
 * target is the thermostat's (the reference's handleBoxtemp uses 310 K). */
template <int MAXP>
void from_reference(System& out, const Compound<MAXP>* compounds, const CompoundState<MAXP>* states, int n_compounds,
                    const Solvent* solvents, int n_solvents, const CompoundBridgeCompact* bridges, int n_bridges,
                    const ForceField& ff, float box_len, double dt, float thermostat_target = 310.0f) {
    if (n_compounds < 0 || n_solvents < 0 || n_bridges < 0) detail::fail("negative count");
    if ((n_compounds && (!compounds || !states)) || (n_solvents && !solvents) || (n_bridges && !bridges))
        detail::fail("null array");
    out.compound_offset.assign(1, 0);
    out.bond_offset.assign(1, 0);
    out.angle_offset.assign(1, 0);
    out.dihedral_offset.assign(1, 0);
    out.ignore_offset.assign(1, 0);
    for (int c = 0; c < n_compounds; c++) {
        const Compound<MAXP>& cp = compounds[c];
        const int n = cp.n_particles;
        if (n > MAXP || states[c].n_particles != cp.n_particles) detail::fail("compound particle count");
        if (cp.n_singlebonds > kMaxPairBonds || cp.n_anglebonds > kMaxAngleBonds || cp.n_dihedrals > kMaxDihedrals)
            detail::fail("compound term count");
        for (int p = 0; p < n; p++) {
            detail::put3(out.chain_positions, states[c].positions[p]);
            detail::put3(out.chain_prev_positions, cp.prev_positions[p]);
            out.chain_atom_types.push_back(cp.atom_types[p]);
            const LJ_Ignores& ig = cp.lj_ignore_list[p];
            if (ig.n_ignores > LJ_Ignores::max_ids) detail::fail("LJ_Ignores count");
            for (int e = 0; e < ig.n_ignores; e++) {
                out.ignore_compound.push_back(ig.compound_ids[e]);
                out.ignore_local.push_back(ig.local_ids[e]);
            }
            out.ignore_offset.push_back((int32_t)out.ignore_local.size());
        }
        out.compound_offset.push_back(out.compound_offset.back() + n);
        out.key_particle_index.push_back(cp.key_particle_index);
        out.sphere.push_back(cp.confining_particle_sphere);
        for (int k = 0; k < cp.n_singlebonds; k++) out.bonds.push_back(detail::conv(cp.singlebonds[k]));
        for (int k = 0; k < cp.n_anglebonds; k++) out.angles.push_back(detail::conv(cp.anglebonds[k]));
        for (int k = 0; k < cp.n_dihedrals; k++) out.dihedrals.push_back(detail::conv(cp.dihedrals[k]));
        out.bond_offset.push_back((int32_t)out.bonds.size());
        out.angle_offset.push_back((int32_t)out.angles.size());
        out.dihedral_offset.push_back((int32_t)out.dihedrals.size());
    }
    out.bridge_particle_offset.assign(1, 0);
    out.bridge_bond_offset.assign(1, 0);
    out.bridge_angle_offset.assign(1, 0);
    out.bridge_dihedral_offset.assign(1, 0);
    for (int b = 0; b < n_bridges; b++) {
        const CompoundBridgeCompact& br = bridges[b];
        if (br.n_particles < 0 || br.n_particles > kBridgeParticles || br.n_singlebonds > kBridgeTerms ||
            br.n_anglebonds > kBridgeTerms || br.n_dihedrals > kBridgeTerms)
            detail::fail("bridge count");
        for (int p = 0; p < br.n_particles; p++) {
            out.bridge_particle_compound.push_back(br.particle_refs[p].compound_id);
            out.bridge_particle_local.push_back(br.particle_refs[p].local_id_compound);
        }
        for (int k = 0; k < br.n_singlebonds; k++) out.bridge_bonds.push_back(detail::conv(br.singlebonds[k]));
        for (int k = 0; k < br.n_anglebonds; k++) out.bridge_angles.push_back(detail::conv(br.anglebonds[k]));
        for (int k = 0; k < br.n_dihedrals; k++) out.bridge_dihedrals.push_back(detail::conv(br.dihedrals[k]));
        out.bridge_particle_offset.push_back((int32_t)out.bridge_particle_local.size());
        out.bridge_bond_offset.push_back((int32_t)out.bridge_bonds.size());
        out.bridge_angle_offset.push_back((int32_t)out.bridge_angles.size());
        out.bridge_dihedral_offset.push_back((int32_t)out.bridge_dihedrals.size());
    }
    out.solvent_positions.clear();
    out.solvent_prev_positions.clear();
    for (int s = 0; s < n_solvents; s++) {
        detail::put3(out.solvent_positions, solvents[s].pos);
        detail::put3(out.solvent_prev_positions, solvents[s].pos_tsub1);
    }
    out.system = lima_system{};
    out.system.box_len = box_len;
    out.system.dt = dt;
    out.system.thermostat_target = thermostat_target;
    for (int t = 0; t < 32; t++)
        out.system.forcefield.particle_parameters[t] = {ff.particle_parameters[t].mass, ff.particle_parameters[t].sigma,
                                                        ff.particle_parameters[t].epsilon};
    out.point();
}

/* The bundle the reference keeps in the Box (Simulation.cuh, CompoundBridgeBundleCompact). */
template <int MAXP>
void from_reference(System& out, const Compound<MAXP>* compounds, const CompoundState<MAXP>* states, int n_compounds,
                    const Solvent* solvents, int n_solvents, const CompoundBridgeBundleCompact* bundle,
                    const ForceField& ff, float box_len, double dt, float thermostat_target = 310.0f) {
    const int nb = bundle ? bundle->n_bridges : 0;
    if (nb < 0 || nb > kBridgesInBundle) detail::fail("bridge bundle count");
    from_reference<MAXP>(out, compounds, states, n_compounds, solvents, n_solvents, bundle ? bundle->compound_bridges : nullptr,
                         nb, ff, box_len, dt, thermostat_target);
}

/* Stepped state (lima_get_state's arrays, [n*3] floats) back into the
 * reference's arrays: CompoundState::positions, Compound::prev_positions and
 * Solvent{pos, pos_tsub1}.  Compound sizes are taken from `compounds`. */
template <int MAXP>
void write_back(const float* chain_positions, const float* chain_prev_positions, const float* solvent_positions,
                const float* solvent_prev_positions, Compound<MAXP>* compounds, CompoundState<MAXP>* states,
                int n_compounds, Solvent* solvents, int n_solvents) {
    size_t g = 0;
    for (int c = 0; c < n_compounds; c++) {
        const int n = compounds[c].n_particles;
        if (n > MAXP) detail::fail("compound particle count");
        for (int p = 0; p < n; p++, g++) {
            states[c].positions[p] = {chain_positions[3 * g], chain_positions[3 * g + 1], chain_positions[3 * g + 2]};
            compounds[c].prev_positions[p] = {chain_prev_positions[3 * g], chain_prev_positions[3 * g + 1],
                                              chain_prev_positions[3 * g + 2]};
        }
    }
    for (int s = 0; s < n_solvents; s++) {
        solvents[s].pos = {solvent_positions[3 * s], solvent_positions[3 * s + 1], solvent_positions[3 * s + 2]};
        solvents[s].pos_tsub1 = {solvent_prev_positions[3 * s], solvent_prev_positions[3 * s + 1],
                                 solvent_prev_positions[3 * s + 2]};
    }
}

}  // namespace ref
}  // namespace lima
