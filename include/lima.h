/*
 * lima.h — C ABI of the MI355X (gfx950) implementation of LIMA_MD's per-step path.
 *
 * The reference drives its step through the C++ class `Engine`
 * (Quantom/Quantom/Engine.cuh:109-208): `Engine(Simulation*, ForceField)`
 * (Engine.cu:9-38), `deviceMaster()` -> `step()` (Engine.cu:44-47, 400-451) and
 * `hostMaster()` (Engine.cu:49-75).  State lives in `Simulation`/`Box`
 * (Simulation.cuh:35-178) as AoS structs with compile-time capacities
 * (Bodies.cuh:19-25, 169-472, 644-732; Constants.cuh:35-113).
 *
 * This header replaces that boundary with plain C: the system is described by
 * flat arrays that restate the reference fields (capacities become runtime
 * sizes), the engine owns all device memory, and every call returns a status
 * code instead of calling exit().  One engine may hold several independent
 * systems ("replicas"); a single system is a batch of one.
 *
 * Units are the reference's: nm, ns, kg/mol, J/mol; force is J/(mol*nm).
 * An engine handle is single-threaded: calls on one handle must not overlap.
 * Distinct handles are independent.
 */
#ifndef LIMA_H
#define LIMA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LIMA_ABI_VERSION 1

/* ---- status codes (replace printf+exit in Engine.cuh:47-55, Bodies.cuh:215-245,
 *      Engine.cu:383-386 and the device flag critical_error_encountered,
 *      Simulation.cuh:61) ---------------------------------------------------- */
enum {
    LIMA_OK = 0,
    LIMA_E_HIP = 1,            /* a HIP runtime call failed */
    LIMA_E_CAPACITY = 2,       /* a neighbour list or staging buffer overflowed */
    LIMA_E_CRITICAL_FORCE = 3, /* |F| > 2e12 on a chain particle (Engine.cu:1050-1053) */
    LIMA_E_NAN = 4,            /* NaN solvent position (Engine.cu:1151-1154) */
    LIMA_E_ARG = 5,            /* invalid argument or system description */
    LIMA_E_UNSUPPORTED = 6     /* valid request this build does not implement */
};

/* ---- force field: ForceField{ParticleParameters[32]} (Forcefield.cuh:12-27).
 *      Type 0 is the solvent type (ATOMTYPE_SOL, Forcefield.cuh:15). -------- */
#define LIMA_MAX_ATOM_TYPES 32
typedef struct lima_particle_params {
    float mass;    /* kg/mol */
    float sigma;   /* nm */
    float epsilon; /* J/mol */
} lima_particle_params;

typedef struct lima_forcefield {
    lima_particle_params particle_parameters[LIMA_MAX_ATOM_TYPES];
} lima_forcefield;

/* ---- bonded terms, compound-local (or bridge-local) particle indices
 *      PairBond / AngleBond / DihedralBond (Bodies.cuh:61-113) -------------- */
typedef struct lima_pair_bond {
    float b0, kb;
    uint32_t atom_indexes[2];
} lima_pair_bond;

typedef struct lima_angle_bond {
    float theta_0, k_theta;
    uint32_t atom_indexes[3];
} lima_angle_bond;

typedef struct lima_dihedral_bond {
    float phi_0, k_phi;
    uint32_t n;               /* periodicity (uint8_t in the reference) */
    uint32_t atom_indexes[4];
} lima_dihedral_bond;

/* ---- one system: restates Box/Simulation/Compound/CompoundState/Solvent/
 *      LJ_Ignores/CompoundBridgeCompact with runtime sizes.
 *      All arrays are read during lima_create and never retained. ---------- */
typedef struct lima_system {
    float box_len;              /* BOX_LEN (Constants.cuh:35), per system */
    double dt;                  /* Box::dt [ns] (Simulation.cuh:60, 152) */
    float thermostat_target;    /* K; target temperature of this system */
    lima_forcefield forcefield; /* per system (replicas may differ) */

    /* compounds: compound c owns chain particles [compound_offset[c], compound_offset[c+1]) */
    int32_t n_compounds;
    const int32_t* compound_offset;     /* [n_compounds+1] */
    const int32_t* key_particle_index;  /* [n_compounds] Compound::key_particle_index (Bodies.cuh:347) */
    const float* confining_particle_sphere; /* [n_compounds] (Bodies.cuh:348) */
    const float* chain_positions;       /* [n_chain*3] CompoundState::positions (Bodies.cuh:313) */
    const float* chain_prev_positions;  /* [n_chain*3] Compound::prev_positions (Bodies.cuh:323) */
    const uint8_t* chain_atom_types;    /* [n_chain]   Compound::atom_types (Bodies.cuh:325) */

    /* compound bonded terms, concatenated per compound with offsets [n_compounds+1]
       (Compound::singlebonds/anglebonds/dihedrals, Bodies.cuh:338-345) */
    const int32_t* bond_offset;      const lima_pair_bond* bonds;
    const int32_t* angle_offset;     const lima_angle_bond* angles;
    const int32_t* dihedral_offset;  const lima_dihedral_bond* dihedrals;

    /* LJ exclusions per chain particle: entries [ignore_offset[i], ignore_offset[i+1])
       hold (compound id, compound-local id) pairs (LJ_Ignores, Bodies.cuh:169-198) */
    const int32_t* ignore_offset;    /* [n_chain+1] */
    const int32_t* ignore_compound;
    const int32_t* ignore_local;

    /* bridges: bonded terms spanning compounds (CompoundBridgeCompact, Bodies.cuh:652-725).
       Bridge b references particles [bridge_particle_offset[b], ...+1) by
       (compound id, compound-local id); its terms use bridge-local indices. */
    int32_t n_bridges;
    const int32_t* bridge_particle_offset;   /* [n_bridges+1] */
    const int32_t* bridge_particle_compound;
    const int32_t* bridge_particle_local;
    const int32_t* bridge_bond_offset;      const lima_pair_bond* bridge_bonds;
    const int32_t* bridge_angle_offset;     const lima_angle_bond* bridge_angles;
    const int32_t* bridge_dihedral_offset;  const lima_dihedral_bond* bridge_dihedrals;

    /* solvents: Solvent{pos, pos_tsub1} (Bodies.cuh:19-25) */
    int32_t n_solvents;
    const float* solvent_positions;      /* [n_solvents*3] */
    const float* solvent_prev_positions; /* [n_solvents*3] */
} lima_system;

/* ---- run parameters (compile-time constants of Constants.cuh made runtime) */
/* NONE: lambda = 1.  BERENDSEN: every step, lambda = sqrt(1 + dt/tau (T0/T - 1)).
 * REFERENCE: Engine::hostMaster's cadence (Engine.cu:49-75, 314-393): every 40th
 * step, from the two-step kinetic energy, lambda = T0/T for one step; both
 * clamp lambda to [0.9, 1.1]. */
enum { LIMA_THERMOSTAT_NONE = 0, LIMA_THERMOSTAT_BERENDSEN = 1, LIMA_THERMOSTAT_REFERENCE = 2 };
enum { LIMA_ACCUM_FP32 = 0, LIMA_ACCUM_FIXED64 = 1 };

typedef struct lima_run_params {
    float cutoff;               /* CUTOFF (Constants.cuh:52); list radii are 2*cutoff (+spheres) */
    int32_t nlist_interval;     /* STEPS_PER_NLIST_UPDATE (Constants.cuh:71) */
    int32_t thermostat;         /* LIMA_THERMOSTAT_* */
    double thermostat_tau;      /* ns, Berendsen coupling time */
    int32_t accumulation;       /* LIMA_ACCUM_* */
    int32_t enable_solvents;    /* ENABLE_SOLVENTS (Constants.cuh:13) */
    int32_t enable_dihedrals;   /* ENABLE_DIHEDRALS (Constants.cuh:25) */
    int32_t max_compound_neighbors;  /* NEIGHBORLIST_MAX_COMPOUNDS (Constants.cuh:41) */
    int32_t max_compound_solvents;   /* NEIGHBORLIST_MAX_SOLVENTS for compound lists (Constants.cuh:42) */
    int32_t max_solvent_neighbors;   /* NEIGHBORLIST_MAX_SOLVENTS for solvent lists (Constants.cuh:42) */
    int32_t device;             /* HIP device ordinal */
} lima_run_params;

/* FP64 energy report (Analyzer::analyzeEnergy, Analyzer.cu:137-207, made in-step) */
typedef struct lima_energies {
    double bond, angle, dihedral;       /* reference bonded energies, accumulated */
    double lj_intra, lj_inter, lj_compound_solvent, lj_solvent_compound, lj_solvent_solvent;
                                         /* 1/2 * sum over directed pairs of 4*eps*(s^2-s) */
    double potential;                    /* sum of all of the above */
    double kinetic;                      /* sum 1/2 m |x - x_prev|^2 / dt^2 */
    double temperature;                  /* 2 KE / (3 N R), R = 8.3145 */
    double virial;                       /* sum over directed pairs of (r_j - r_i) . F_i / 2 */
    double thermostat_scalar;            /* lambda applied on the most recent step */
} lima_energies;

typedef struct lima_engine lima_engine;

/* library */
const char* lima_version(void);
int lima_device_count(int* count);

/* Engine(Simulation*, ForceField) (Engine.cu:9-38): copies the systems to
   the device and builds the neighbour lists (handleNLISTS(force), Engine.cu:35). */
int lima_create(const lima_system* systems, int32_t n_systems,
                const lima_run_params* params, lima_engine** out);
void lima_destroy(lima_engine* engine);
const char* lima_last_error(const lima_engine* engine);

/* deviceMaster()+hostMaster() repeated n_steps times (Engine.cu:44-75, 400-451):
   bridge, compound and solvent forces, integration, PBC, thermostat and the
   scheduled neighbour-list rebuilds, all on the device.  Synchronises once at
   the end and reports device error flags. */
int lima_step(lima_engine* engine, int32_t n_steps);
/* Queue n_steps without synchronising (errors surface at the next lima_sync). */
int lima_step_async(lima_engine* engine, int32_t n_steps);
int lima_sync(lima_engine* engine);

/* Engine::updateNeighborLists (Engine.cu:146-287), synchronous and on the device */
int lima_rebuild_neighbors(lima_engine* engine);

/* state in lima_system layout, all systems concatenated (chain then solvent per system order) */
int lima_get_state(lima_engine* engine, float* chain_positions, float* chain_prev_positions,
                   float* solvent_positions, float* solvent_prev_positions);
int lima_set_state(lima_engine* engine, const float* chain_positions, const float* chain_prev_positions,
                   const float* solvent_positions, const float* solvent_prev_positions);
/* total force used to integrate each particle on the most recent step */
int lima_get_forces(lima_engine* engine, float* chain_forces, float* solvent_forces);
/* forces on the current state without advancing it (the step's force pass only) */
int lima_compute_forces(lima_engine* engine, float* chain_forces, float* solvent_forces);

/* FP64 energies of the current state, one record per system */
int lima_get_energies(lima_engine* engine, lima_energies* out);
/* lima_step with an energy series: the first step and every `every`-th after it
   also sum the FP64 energies on the device (Analyzer::analyzeEnergy's per-step
   series, Analyzer.cu:137-207, made in-step).  A record holds the potential
   terms at the step's positions and the kinetic energy / temperature of the
   step's displacement (what the thermostat uses).  out: [ceil(n_steps/every)][n_systems].
   One host synchronisation at the end. */
int lima_step_observed(lima_engine* engine, int32_t n_steps, int32_t every, lima_energies* out);
/* No counterpart in the reference: for callers that cannot allocate device memory
   (tight serving loops, code that forbids hipMalloc).  lima_step_observed with its
   energy series staged in a fixed device array of the library (4 MiB per device,
   84 bytes per record and system: ceil(n_steps/every) x n_systems records at most
   49,000 or so) instead of a per-call allocation, and with the plain steps between
   records replayed from launch graphs already built (or launched directly), never
   captured.  Calls no hipMalloc/hipFree.  Larger requests return LIMA_E_ARG.  The
   array is shared by every engine on a device: calls that overlap on one device
   (from several host threads) race.  Results equal lima_step_observed's. */
int lima_step_observed_noalloc(lima_engine* engine, int32_t n_steps, int32_t every, lima_energies* out);
/* lima_step with the per-step logs the reference keeps for one system
   (Engine::offloadLoggingData / offloadTrainData, Engine.cu:288-310; the
   compoundKernel's data log, Engine.cu:1019-1046), copied out every 100 steps.
   logdata:   [n_steps][10] floats (logdata.bin): for particle 0 of the system's
              first compound, the speed |x' - x|/dt, the kinetic energy
              calcKineticEnergy gives for those one-step-apart positions,
              thread 0's potE slot (its first dihedral's energy, which the
              reference assigns, else its first bond and angle), |F|; then six 0.
   traindata: [n_steps][n_compounds][slots_per_compound][3][3] floats
              (sim_traindata.bin, data_GAN): position after the step (-1 in slots
              past the compound), the compound<-solvent LJ force (Float3(-1) with
              solvents off), and a third Float3 written 0.
   Either output may be NULL (not both); slots_per_compound is the reference
   build's MAX_COMPOUND_PARTICLES and must cover the system's largest compound. */
int lima_step_logged(lima_engine* engine, int32_t n_steps, int32_t system, int32_t slots_per_compound,
                     float* logdata, float* traindata);

/* neighbour lists (NeighborList, Bodies.cuh:203-292) as global ids within a system.
   counts: [n_compounds] / [n_solvents]; ids: stride = the matching capacity. */
int lima_get_neighbor_counts(lima_engine* engine, int32_t* compound_compound_count,
                             int32_t* compound_solvent_count, int32_t* solvent_solvent_count);
int lima_get_neighbors(lima_engine* engine, int32_t* compound_compound_ids,
                       int32_t* compound_solvent_ids, int32_t* solvent_solvent_ids);
int lima_set_neighbors(lima_engine* engine,
                       const int32_t* compound_compound_count, const int32_t* compound_compound_ids,
                       const int32_t* compound_solvent_count, const int32_t* compound_solvent_ids,
                       const int32_t* solvent_solvent_count, const int32_t* solvent_solvent_ids);

/* Engine::timings (Engine.cuh:123): [0] device step time us, [1] rebuild time us,
   [2] host time us, [3] steps run, [4] rebuilds run */
int lima_get_timings(lima_engine* engine, double* out5);
/* Average device time in microseconds of each kernel of the step, each launched
   `reps` times alone on the engine's stream between HIP events, on the current
   state: [0] solvent<-chain all-pairs, [1] compound listed LJ, [2] compound
   sums + integrate (forces only), [3] solvent integrate (forces only),
   [4] bridge, [5] full neighbour-list rebuild, [6] solvent listed LJ,
   [7] compound local forces (bridge carry-in, bonded, intra-compound LJ). */
int lima_time_kernels(lima_engine* engine, int32_t reps, double* out8);

/* In-step timing of the all-pairs kernel (solvent_chain_kernel, the step's dominant
 * kernel; solventForceKernel's compound loop, Engine.cu:1099-1117).  With on != 0 one step
 * per neighbour-list period (the second after each rebuild; every 8th step without
 * rebuilds) adds, on the device's constant-rate clock, the span from the kernel's first
 * block start to its last block end, as it runs among the step's other kernels, to a
 * running total (cleared when the clock is switched on; the launch graph is re-captured).
 * lima_get_kernel_clock returns the total in microseconds and the launches counted; the
 * difference of two reads is the mean over the steps between them.  No counterpart in the
 * reference (its Engine::timings, Engine.cuh:123, are host timers per kernel group). */
int lima_set_kernel_clock(lima_engine* engine, int32_t on);
int lima_get_kernel_clock(lima_engine* engine, double* total_us, int64_t* launches);
/* Launch shape chosen for this engine (see lima_autotune), up to twelve values
   (the first n_out are written): solvents per lane of the all-pairs kernel, its
   target chunk size, its work items, the most chunks of one solvent tile, graph
   replay on/off, solvent cells, solvent tiles, systems, the grid of the listed
   compound kernel, lanes per solvent of the listed solvent kernel, the form of
   the listed compound kernel for compounds of at most 64 particles (1 packed,
   2 (FP32) plain with each item's sources split over both wavefronts, else 0),
   and where the thermostat runs: 0 a kernel of its own after the integrators,
   1 the last of both integrators' blocks of each system, 2 the last compound
   integrator block, run after the solvent integrator (same sums in all). */
int lima_get_config(lima_engine* engine, int32_t* out, int32_t n_out);
/* Picks the launch shape (solvents per lane and chunk size of the all-pairs
   kernel, grid of the listed compound kernel, and, in FP32, one or four lanes
   per solvent of the listed solvent kernel) by timing whole steps of each shape,
   trial_steps at a time (more when that is under 8 ms, in whole list cycles),
   from the current state; the state, lists and
   thermostat scalars are restored (the last-step force buffers hold a trial
   step's).  Choices are cached per system shape; LIMA_AUTOTUNE_CACHE names a
   file that keeps them across processes. */
int lima_autotune(lima_engine* engine, int32_t trial_steps);

/* ---- parameters and reverse-mode gradients of trajectory losses.
 * The reference has no differentiable path; these calls take the same force
 * field (ForceField, Forcefield.cuh:12-27) and bonded constants (PairBond /
 * AngleBond, Bodies.cuh:61-88) as lima_create and expose them as a parameter
 * vector per system:
 *   [0..31]  sigma of types 0..31      [32..63] epsilon of types 0..31
 *   [64] kb  [65] b0  [66] k_theta  [67] theta_0
 * The four bonded entries act as constants shared by every pair bond / angle of
 * the system (bridges included): setting them writes every term, and their
 * gradient is the sum of the per-term derivatives (the derivative for a common
 * shift of every term's constant).  get returns the first term's constants.  On
 * a system whose terms carry different constants (an ingested molecule), set
 * keeps the per-term constants when entries 64-67 come back unchanged and
 * returns LIMA_E_UNSUPPORTED for a change of the differing kind. */
#define LIMA_N_PARAMETERS 68

int lima_get_parameters(lima_engine* engine, int32_t system, double* params);
int lima_set_parameters(lima_engine* engine, int32_t system, const double* params);

/* Loss over the frames x_0 .. x_{K-1} of the trajectory that starts at the
 * current state (x_0 = current positions, integrated with the engine's step;
 * the thermostat must be LIMA_THERMOSTAT_NONE):
 *   L_r = sum_k alpha_r (E_k,r - E*_k,r)^2 + beta_r sum_i |F_k,i - F*_k,i|^2
 * with E the potential energy of lima_energies and F the force of each particle
 * on frame k.  NULL weights or targets read as zeros. */
typedef struct lima_adjoint_spec {
    int32_t n_frames;                  /* K >= 1 */
    int32_t checkpoint_interval;       /* frames per stored state; segments start on list rebuilds.  0: one
                                          segment when its frames, residuals and lists fit in half the free
                                          device memory, else 20.  Negative values are an error. */
    const double* energy_weight;       /* [n_systems] alpha */
    const double* force_weight;        /* [n_systems] beta */
    const double* energy_target;       /* [K][n_systems] E* (J/mol) */
    const float* chain_force_target;   /* [K][n_chain][3] F*, systems concatenated */
    const float* solvent_force_target; /* [K][n_solvents][3] */
} lima_adjoint_spec;

/* Forward pass over the K frames, then the adjoint recurrence backwards with
 * checkpoints: the last segment's frames are kept from the forward pass, each
 * earlier segment is recomputed from its checkpoint (neighbour lists restored
 * exactly).  loss: [n_systems]; grad: [n_systems][LIMA_N_PARAMETERS] or NULL
 * (loss only).  The engine's state, lists and step count are left as they were. */
int lima_adjoint_run(lima_engine* engine, const lima_adjoint_spec* spec, double* loss, double* grad);

/* The same K frames, recorded: forces [K][n_chain][3], [K][n_solvents][3] and
 * potential energies [K][n_systems] (any output may be NULL).  The engine's
 * state is left as it was. */
int lima_record_frames(lima_engine* engine, int32_t n_frames, float* chain_forces, float* solvent_forces,
                       double* potential);

#ifdef __cplusplus
}
#endif

#endif /* LIMA_H */
