/*
 * lima_ingest.h — C ABI of the system ingest: CHARMM parameter files plus a GROMACS
 * .gro/.top pair in, a lima_system (lima.h) out.
 *
 * The reference builds a simulation from files in two programs:
 *
 *  1. LIMA_ForcefieldMaker (LIMA_ForcefieldMaker.cpp:145-155) merges CHARMM .prm files
 *     into LIMA_ffnonbonded.txt / LIMA_ffbonded.txt (mergeForcefieldFiles, :116-141;
 *     ForcefieldMerger.h:10-176; Printer, Filehandling.h:225-286), then filters them for
 *     one system into LIMA_ffnonbonded_filtered.txt / LIMA_ffbonded_filtered.txt
 *     (prepareForcefieldForSimulation, :70-114; ForcefieldTypes.h; Printer,
 *     Filehandling.h:139-223).
 *  2. Quantom's Environment (Environment.cpp:16-45) reads the filtered files
 *     (ForceFieldMaker::buildForcefield, Forcefield.cu:10-218), groups the molecule's
 *     residues into compounds and bridges (CompoundBuilder::buildMolecule,
 *     CompoundBuilder.cpp:7-318) and fills the box (BoxBuilder.cu:37-63, 194-276).
 *
 * The three entry points below are those stages, with the reference's file formats
 * between them, so that files written by either side are read by the other.  Every
 * call returns a lima.h status code instead of calling exit(); lima_ingest_last_error()
 * gives the message of the calling thread's most recent failure.  The library is host
 * code only: it never touches a GPU.
 */
#ifndef LIMA_INGEST_H
#define LIMA_INGEST_H

#include <stdint.h>

#include "lima.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Stage 1.  Merges CHARMM parameter files in the given order and writes
 * out_dir/LIMA_ffnonbonded.txt and out_dir/LIMA_ffbonded.txt.
 * Replaces mergeForcefieldFiles (LIMA_ForcefieldMaker.cpp:116-141), whose file list is
 * par_all36_lipid.prm, par_all36_na.prm, par_all36m_prot.prm (:124-126). */
int lima_ff_merge(const char* const* prm_paths, int32_t n_paths, const char* out_dir);

/* Stage 2.  Reads ff_dir/LIMA_ffnonbonded.txt and ff_dir/LIMA_ffbonded.txt, matches them
 * to the atoms and bonded terms of conf_gro / topol_top, and writes
 * ff_dir/LIMA_ffnonbonded_filtered.txt and ff_dir/LIMA_ffbonded_filtered.txt.
 * Replaces prepareForcefieldForSimulation (LIMA_ForcefieldMaker.cpp:70-114). */
int lima_ff_prepare(const char* conf_gro, const char* topol_top, const char* ff_dir);

/* Options of stage 3; lima_ingest_default_params() gives the reference's values. */
typedef struct lima_ingest_params {
    int32_t min_residue_id;         /* Environment.cpp:21 (0) */
    int32_t max_residue_id;         /* Environment.cpp:22 (200) */
    int32_t ignore_hydrogens;       /* Environment.cpp:23 (1) */
    int32_t max_compound_particles; /* MAX_COMPOUND_PARTICLES (Constants.cuh:58: 64); at most 128 */
    int32_t max_atoms_in_residue;   /* MAX_ATOMS_IN_RESIDUE (Constants.cuh:90: 32) */
    int32_t max_bridge_particles;   /* MAX_PARTICLES_IN_BRIDGE (Constants.cuh:110: 32) */
    int32_t max_exclusions;         /* LJ_Ignores::max_ids (Bodies.cuh:170: 32) */
    int32_t add_solvents;           /* ENABLE_SOLVENTS (Constants.cuh:13: on) */
    float box_len;                  /* BOX_LEN (Constants.cuh:35: 7 nm) */
    double dt;                      /* Simulation::dt (Simulation.cuh:152: 1e-7 ns) */
    float thermostat_target;        /* handleBoxtemp's target (Engine.cu:363: 310 K) */
    uint32_t seed;                  /* BoxBuilder's srand seed (BoxBuilder.cuh:10: 290128309) */
} lima_ingest_params;

void lima_ingest_default_params(lima_ingest_params* out);

typedef struct lima_ingest lima_ingest;

/* Stage 3.  Reads ff_dir/LIMA_ff*_filtered.txt and conf_gro / topol_top, builds the
 * molecule's compounds, bridges and exclusions, centres it in the box, places the
 * .gro file's water oxygens where there is room and gives them BoxBuilder's start
 * velocities.  Replaces Environment::Environment's build (Environment.cpp:16-45):
 * ForceFieldMaker::buildForcefield, CompoundBuilder::buildMolecule,
 * BoxBuilder::addSingleMolecule and BoxBuilder::solvateBox(Simulation*, vector<Float3>*). */
int lima_ingest_create(const char* conf_gro, const char* topol_top, const char* ff_dir,
                       const lima_ingest_params* params, lima_ingest** out);

/* Fills *out with the built system (lima.h layout).  The arrays belong to the handle and
 * stay valid until lima_ingest_destroy. */
int lima_ingest_system(const lima_ingest* ingest, lima_system* out);

typedef struct lima_ingest_info {
    int32_t n_compounds, n_bridges, n_chain, n_solvents;
    int32_t n_atom_types;           /* filled force-field slots; the rest hold -1 */
    int32_t n_solvent_candidates;   /* water oxygens in the .gro file */
    int64_t n_random_draws;         /* rand() calls BoxBuilder made */
    float offset[3];                /* the centring shift (BoxBuilder.cu:38-39) */
} lima_ingest_info;

int lima_ingest_get_info(const lima_ingest* ingest, lima_ingest_info* out);

/* The .gro serial number of every chain particle, in lima_system order
 * (Compound::particle_global_ids, Bodies.cuh:331). */
int lima_ingest_global_ids(const lima_ingest* ingest, int32_t* out);

void lima_ingest_destroy(lima_ingest* ingest);

const char* lima_ingest_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* LIMA_INGEST_H */
