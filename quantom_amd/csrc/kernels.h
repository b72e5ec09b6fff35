// kernels.h — gfx950 kernels of one LIMA_MD step and of the neighbour-list build.
//
// Mapping to the reference (Quantom/Quantom/Engine.cu):
//   bridge_kernel              compoundBridgeKernel          1168-1223
//   solvent_chain_kernel       solventForceKernel 1099-1117 (all-pairs solvent <- every chain particle)
//   compound_listed_kernel     compoundKernel 939-974 (inter-compound and solvent-list LJ)
//   compound_integrate_kernel  compoundKernel 918-931, 981-1056 (bonded, intra-LJ, integrate, PBC)
//   solvent_integrate_kernel   solventForceKernel 1121-1155 (solvent-solvent LJ, integrate, PBC)
//   thermostat_kernel          Engine::handleBoxtemp 362-393, as an FP64 device reduction
//   nl_*                       Engine::updateNeighborLists 146-287, on the device
//
// Every kernel is replica-aware: entities carry their replica id and read the
// box length, time step and type tables of that replica, so one launch covers
// any number of independent systems.
#pragma once
#include <type_traits>
#include "bonded.h"
#include "device_common.h"

namespace lima {

struct I3 {
    long long x, y, z;
};

struct StepArgs {
    const ReplicaInfo* reps;
    int n_replicas;
    // chain particles
    const float4* cpos;      // current (x, y, z, type bits)
    float4* cpos_next;
    float4* cprev;
    float4* cforce;
    float4* fbridge;
    const uint32_t* intra_mask;   // [n_chain*4]
    const int* xoff;              // cross-compound exclusions [n_chain+1]
    const int2* xent;             // (compound, local)
    // compounds
    int n_comp;
    const int* comp_off;          // [n_comp+1]
    const int* comp_rep;
    const int* comp_key;          // compound-local key index
    const float* comp_sphere;
    const int* bond_off; const BondD* bonds;
    const int* angle_off; const AngleD* angles;
    const int* dih_off; const DihD* dihs;
    const int* inc_off; const uint32_t* inc;   // per chain particle: term incidences
    // bridges
    int n_bridges;
    const int* br_off; const int* br_glob; const int* br_rep;
    const int* brb_off; const BondD* brb;
    const int* bra_off; const AngleD* bra;
    const int* brd_off; const DihD* brd;
    // solvents
    int n_solv;
    const float4* spos;
    float4* spos_next;
    float4* sprev;
    float4* sforce;
    const int* solv_rep;
    const int4* stiles;           // (replica, begin, end, tile index) per solvent-integrate tile
    // type tables per replica
    const float* mass;            // [R*32]
    // pair parameters (sigma^2, 24 eps, -48 eps, sigma) with each call site's mixing rule
    const float4* tab_sc;         // [R*32] solvent <- chain type, arithmetic eps
    const float4* tab_cs;         // [R*32] chain type <- solvent, geometric eps
    const float4* tab_inter;      // [R*1024] arithmetic
    const float4* tab_intra;      // [R*1024] geometric
    const float4* tab_ss;         // [R]
    // lists
    int cap_cc, cap_cs, cap_ss;
    const int* cc_cnt; const int* cc_ids;
    const int* cs_cnt; const int* cs_ids;
    const int* ss_cnt; const int* ss_ids;   // rows: ss_ids[s*ss_ld + k], ss_ld = cap_ss rounded up to 4
    int ss_ld;
    float4* ss_part;              // solvent-solvent local sums
    float4* ss_img;               // the solvent's position after the reference's re-imaging sequence
    const int* cw_off; const int* cw_total;
    const int* item_comp;         // [work item] its compound (written with cw_off at each rebuild)
    const int* run_off; const int* run_cs_off; const int* run_total; const int* run_comp;   // packed listed kernel's runs
    const float2* tab_inter_ab;   // [R*1024] lj_ab of tab_inter, [source type * 32 + target type]
    // partial forces
    float4* partC;  I3* partC_fx;   // [slot*128 + p]
    float4* cloc;   I3* cloc_fx;    // [n_chain] bridge + bonded + intra-compound LJ (compound_local_kernel)
    float4* partS;  I3* partS_fx;   // [chunk*n_solv + s]
    const SCWork* scwork; int n_scwork;
    int chunk_cap;                // chain particles of the largest solvent<-chain chunk (LDS tile size)
    // thermostat
    const float* lambda;          // [R] applied this step
    double* ke_comp;              // [n_comp]
    double* ke_stile;             // [n_stiles]
    int ke_ref;                   // 1: the thermostat sums read the reference's two-step kinetic energy
    // per-step logs of one replica (the compoundKernel's data log, Engine.cu:1019-1046)
    int log_rep;                  // logged replica, -1: none
    int log_slots;                // particle slots per compound (the reference's MAX_COMPOUND_PARTICLES)
    int log_ring;                 // steps the device rings hold; step k writes slot k % log_ring
    const long long* step_count;  // [R] steps done before this one
    // the step's thermostat, run by the last integrator block of each replica where thermo_cnt is
    // set (otherwise by thermostat_kernel after the integrators)
    int* thermo_cnt;              // [R] integrator blocks of the replica done this step (back to 0 after)
    const int* thermo_total;      // [R] its integrator blocks
    float* lambda_next;           // [R] the lambda array itself (the thermostat writes the next step's)
    float* lambda_applied;        // [R]
    double* temperature;          // [R]
    long long* step_count_w;      // [R] the step counters (the thermostat advances them)
    int thermo_mode;
    int thermo_in_solv;           // 1: the solvent integrator's blocks count too (else it runs first)
    unsigned long long* kclock;   // [4] span of the step's solvent_chain_kernel and its running sum, or null
    float* log_data;              // [log_ring][10]
    float* train_data;            // [log_ring][compounds of log_rep][log_slots][3 x Float3], or null
    // energies: [R*kEnergyTerms]
    double* energy;
    uint32_t* err;                // [4]
    int enable_solvents, enable_dihedrals;
    int loc_nb, loc_na, loc_nd;   // most bonds, angles, dihedrals of any compound (compound_local LDS strides)
    int listed_split;             // FP32: compound_listed_kernel splits the sources of compounds of <= 64 particles over both waves
};

enum { kE_bond = 0, kE_angle, kE_dih, kE_intra, kE_inter, kE_cs, kE_sc, kE_ss, kE_virial, kE_kin, kEnergyTerms };

constexpr int kChunkParticles = 2048;   // chain particles staged per solvent<-chain work item
constexpr int kChunkCompounds = 64;

LIMA_DEV int ftype(float4 p) { return __float_as_int(p.w); }

LIMA_DEV void block_atomic_energy(double v, double* dst) {
    __shared__ double red[16];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < (int)(blockDim.x + 63) / 64; w++) s += red[w];
        if (s != 0.0) atomicAdd(dst, s);
    }
}

// ---------------------------------------------------------------------------
// compoundBridgeKernel (Engine.cu:1168-1223): one thread per bridge, terms in
// index order, forces stored (not added) into the carry-in buffer.
// ---------------------------------------------------------------------------
// Bridges: bonded terms spanning two compounds (CompoundBridgeCompact, Bodies.cuh:644-700),
// at most 32 particles each (MAX_PARTICLES_IN_BRIDGE, Constants.cuh:110).
constexpr int kBridgeParticles = 32;

// compoundBridgeKernel (Engine.cu:1168-1223), one wavefront per bridge.  The bridge's
// particles are gathered into LDS and re-imaged against its first particle; the lanes
// evaluate up to kBridgeChunk terms of one kind at a time, and lane k then adds the
// results for particle k in term order (bonds, then angles, then dihedrals, each kind
// summed separately), the order of a serial pass over the terms, so the carry-in forces
// are the serial ones bit for bit.  They are stored (not added) into fbridge.
constexpr int kBridgeChunk = 64;
static_assert(kMaxTypes * kMaxTypes % kCompoundBlock == 0 && kSolventChunk % kCompoundBlock == 0,
              "staging loops stride by the block size");

template <bool ENERGY>
__global__ __launch_bounds__(64) void bridge_kernel(StepArgs a) {
    __shared__ F3 pos[kBridgeParticles];
    __shared__ F3 res[kBridgeChunk * 4];
    __shared__ int who[kBridgeChunk * 4];
    __shared__ double en[kBridgeChunk];
    const int b = blockIdx.x, lane = threadIdx.x;
    const ReplicaInfo rp = a.reps[a.br_rep[b]];
    const int p0 = a.br_off[b], np = min(a.br_off[b + 1] - p0, kBridgeParticles);
    const int tb = a.brb_off[b], nb = a.brb_off[b + 1] - tb;
    const int ta = a.bra_off[b], na = a.bra_off[b + 1] - ta;
    const int td = a.brd_off[b], nd = a.enable_dihedrals ? a.brd_off[b + 1] - td : 0;
    // each kind's first chunk of term records is loaded together with the positions
    const BondD bd0 = lane < nb ? a.brb[tb + lane] : BondD{};
    const AngleD ad0 = lane < na ? a.bra[ta + lane] : AngleD{};
    const DihD dd0 = lane < nd ? a.brd[td + lane] : DihD{};
    F3 me = f3(0.f, 0.f, 0.f);
    if (lane < np) {
        const float4 q = a.cpos[a.br_glob[p0 + lane]];
        me = f3(q.x, q.y, q.z);
        pos[lane] = me;
    }
    __syncthreads();
    if (lane > 0 && lane < np) {
        const F3 ref = pos[0];
        hyperpos(ref.x, ref.y, ref.z, me.x, me.y, me.z, rp.L, rp.half);
        pos[lane] = me;
    }
    __syncthreads();
    F3 ps = f3(0.f, 0.f, 0.f), as = ps, ds = ps;
    double eb = 0, ea = 0, ed = 0;
    // one kind of term: compute a chunk into LDS, then each particle's lane adds its slots in order
    auto sweep = [&](int nterms, int slots, auto&& eval, F3& acc, double& esum) {
        for (int c0 = 0; c0 < nterms; c0 += kBridgeChunk) {
            const int nt = min(kBridgeChunk, nterms - c0);
            if (lane < nt) {
                double e = 0;
                eval(c0, &res[lane * 4], &who[lane * 4], &e);
                en[lane] = e;
            }
            __syncthreads();
            if (lane < np)
                for (int t = 0; t < nt; t++)
                    for (int k = 0; k < slots; k++)
                        if (who[t * 4 + k] == lane) acc = f3_add(acc, res[t * 4 + k]);
            if (ENERGY && lane == 0)
                for (int t = 0; t < nt; t++) esum += en[t];
            __syncthreads();
        }
    };
    sweep(nb, 2, [&](int c0, F3* r, int* w, double* e) {
        const BondD bd = c0 == 0 ? bd0 : a.brb[tb + c0 + lane];
        pairbond(pos[bd.a0], pos[bd.a1], bd.b0, bd.kb, r, e);
        w[0] = bd.a0; w[1] = bd.a1;
    }, ps, eb);
    sweep(na, 3, [&](int c0, F3* r, int* w, double* e) {
        const AngleD ad = c0 == 0 ? ad0 : a.bra[ta + c0 + lane];
        anglebond(pos[ad.a0], pos[ad.a1], pos[ad.a2], ad.t0, ad.k, r, e);
        w[0] = ad.a0; w[1] = ad.a1; w[2] = ad.a2;
    }, as, ea);
    sweep(nd, 4, [&](int c0, F3* r, int* w, double* e) {
        const DihD dd = c0 == 0 ? dd0 : a.brd[td + c0 + lane];
        dihedral(pos[dd.a0], pos[dd.a1], pos[dd.a2], pos[dd.a3], dd.p0, dd.k, dd.n, r, e);
        w[0] = dd.a0; w[1] = dd.a1; w[2] = dd.a2; w[3] = dd.a3;
    }, ds, ed);
    if (lane < np) {
        const F3 f = f3_add(f3_add(f3_add(f3(0.f, 0.f, 0.f), ps), as), ds);
        a.fbridge[a.br_glob[p0 + lane]] = make_float4(f.x, f.y, f.z, 0.0f);
    }
    if (ENERGY && lane == 0) {
        double* e = a.energy + a.br_rep[b] * kEnergyTerms;
        atomicAdd(e + kE_bond, eb);
        atomicAdd(e + kE_angle, ea);
        atomicAdd(e + kE_dih, ed);
    }
}

// ---------------------------------------------------------------------------
// The solvent's own position is re-imaged against each compound's particle 0 in
// compound order before its LJ sums (Engine.cu:1111).  replay_images continues
// that sequence over compounds [cb, ce) for the SPT solvents of each lane, from
// particle-0 positions staged in LDS; every thread of the block calls it.
// ---------------------------------------------------------------------------
constexpr int kReplayStage = 256;

template <int SPT>
LIMA_DEV void replay_images(const StepArgs& a, int cb, int ce, float4* p0s, float* sx, float* sy, float* sz, float L,
                            float half) {
    for (int c0 = cb; c0 < ce; c0 += kReplayStage) {
        const int nc = min(kReplayStage, ce - c0);
        __syncthreads();
        for (int i = threadIdx.x; i < nc; i += blockDim.x) p0s[i] = a.cpos[a.comp_off[c0 + i]];
        __syncthreads();
        for (int i = 0; i < nc; i++) {
            const float4 q = p0s[i];
#pragma unroll
            for (int k = 0; k < SPT; k++) hyperpos(q.x, q.y, q.z, sx[k], sy[k], sz[k], L, half);
        }
    }
}

// ---------------------------------------------------------------------------
// solventForceKernel's compound loop (Engine.cu:1099-1117): every solvent
// against every chain particle of its replica, no cutoff.  Work item =
// (solvent tile, compound chunk).  The chunk's particles are staged in LDS as
// (x, y, z, A) + B (the pair constants of lj_ab); every lane holds SPT solvents.  The solvent's own
// position is re-imaged against each compound's particle 0 in compound order,
// exactly as the reference mutates it (Engine.cu:1111), by replaying the
// compounds before the chunk (replay_images).  Per-compound partial sums match
// computeLJForces' local sum (Engine.cu:640-651).
// ---------------------------------------------------------------------------
// In-step timing of the all-pairs kernel, on the device's constant-rate clock (wall_clock64):
// kclock[0] = the earliest start of the first kClockHead blocks (dispatched first, one per
// XCD), kclock[1] = the latest end of any block (work items differ in length, so any block
// may be the last); the step's thermostat kernel, which runs after it on the same stream,
// adds the span to kclock[2] and counts it in kclock[3], then clears kclock[0..1].  A clock
// read stalls its wave, so the start is read by the first blocks only.
constexpr int kClockHead = 8;

template <int SPT, bool FIXED, bool ENERGY>
__global__ __launch_bounds__(256) void solvent_chain_kernel(StepArgs a) {
    // dynamic LDS sized to the largest chunk of this engine: [cap] float4 + [cap] float2
    extern __shared__ float4 dyn_lds[];
    float4* tile = dyn_lds;
    float* tile_b = reinterpret_cast<float*>(dyn_lds + a.chunk_cap);
    __shared__ int tile_off[kChunkCompounds + 1];
    __shared__ float2 tab[kMaxTypes];
    __shared__ float4 p0s[kReplayStage];
    const int tid = threadIdx.x;
    if (a.kclock && tid == 0 && blockIdx.x < kClockHead) atomicMin(&a.kclock[0], (unsigned long long)wall_clock64());
    const SCWork wk = a.scwork[blockIdx.x];
    const ReplicaInfo rp = a.reps[wk.replica];
    const float L = rp.L, half = rp.half;
    if (tid < kMaxTypes) {
        const float4 v = a.tab_sc[wk.replica * kMaxTypes + tid];
        tab[tid] = lj_ab(v.x, v.y);
    }
    const int c0 = wk.comp_begin, ncomp = wk.comp_end - wk.comp_begin;
    const int base = a.comp_off[c0], total = a.comp_off[wk.comp_end] - base;
    for (int k = tid; k <= ncomp; k += blockDim.x) tile_off[k] = a.comp_off[c0 + k] - base;
    __syncthreads();
    for (int k = tid; k < total; k += blockDim.x) {
        float4 p = a.cpos[base + k];
        const float2 ab = tab[ftype(p)];
        tile[k] = make_float4(p.x, p.y, p.z, ab.x);
        tile_b[k] = ab.y;
    }
    float sx[SPT], sy[SPT], sz[SPT];
    float fx[SPT], fy[SPT], fz[SPT];
    long long ix[SPT], iy[SPT], iz[SPT];
    double en[SPT], vir[SPT];
#pragma unroll
    for (int k = 0; k < SPT; k++) {
        int s = wk.solv_begin + tid + k * blockDim.x;
        float4 p = a.spos[s < wk.solv_end ? s : wk.solv_begin];
        sx[k] = p.x; sy[k] = p.y; sz[k] = p.z;
        fx[k] = fy[k] = fz[k] = 0.f;
        ix[k] = iy[k] = iz[k] = 0;
        en[k] = vir[k] = 0.0;
    }
    // the reference's sequential re-imaging over the compounds before the chunk
    replay_images<SPT>(a, rp.comp_begin, c0, p0s, sx, sy, sz, L, half);
    __syncthreads();
    for (int ci = 0; ci < ncomp; ci++) {
        const int j0 = tile_off[ci], j1 = tile_off[ci + 1];
        const float4 q0 = tile[j0];
#pragma unroll
        for (int k = 0; k < SPT; k++) hyperpos(q0.x, q0.y, q0.z, sx[k], sy[k], sz[k], L, half);
        float lx[SPT], ly[SPT], lz[SPT];
#pragma unroll
        for (int k = 0; k < SPT; k++) lx[k] = ly[k] = lz[k] = 0.f;
#pragma unroll 4
        for (int j = j0; j < j1; j++) {
            const float4 q = tile[j];
            const float b = tile_b[j];
#pragma unroll
            for (int k = 0; k < SPT; k++) {
                if (ENERGY) lj_acc_energy(sx[k], sy[k], sz[k], q.x, q.y, q.z, q.w, b, lx[k], ly[k], lz[k], en[k], vir[k]);
                else lj_acc(sx[k], sy[k], sz[k], q.x, q.y, q.z, q.w, b, lx[k], ly[k], lz[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < SPT; k++) {
            if (FIXED) {
                uint32_t bad = 0;
                ix[k] += to_fixed(lx[k], bad); iy[k] += to_fixed(ly[k], bad); iz[k] += to_fixed(lz[k], bad);
                if (bad) atomicOr(a.err, bad);
            } else { fx[k] += lx[k]; fy[k] += ly[k]; fz[k] += lz[k]; }
        }
    }
    double esum = 0, vsum = 0;
#pragma unroll
    for (int k = 0; k < SPT; k++) {
        int s = wk.solv_begin + tid + k * blockDim.x;
        if (s < wk.solv_end) {
            size_t slot = (size_t)wk.chunk * a.n_solv + s;
            if (FIXED) a.partS_fx[slot] = I3{ix[k], iy[k], iz[k]};
            else a.partS[slot] = make_float4(fx[k], fy[k], fz[k], 0.f);
            esum += en[k];
            vsum += vir[k];
        }
    }
    if (ENERGY) {
        block_atomic_energy(esum, a.energy + wk.replica * kEnergyTerms + kE_sc);
        block_atomic_energy(vsum, a.energy + wk.replica * kEnergyTerms + kE_virial);
    }
    if (a.kclock) {
        __syncthreads();   // every wave of the block is done
        if (tid == 0) atomicMax(&a.kclock[1], (unsigned long long)wall_clock64());
    }
}

// ---------------------------------------------------------------------------
// compoundKernel's listed LJ (Engine.cu:939-974): work item = (compound,
// source chunk); sources are one neighbour compound (arithmetic eps, LJ
// exclusions) or 256 entries of the compound's solvent list (geometric eps).
// Sources are re-imaged against the compound's particle 0 (Engine.cu:949, 967).
// One thread per target particle (two, each over half the sources, when the
// compound fits one wavefront); partial forces go to the compound's slots.
// ---------------------------------------------------------------------------
LIMA_DEV int find_compound(const int* cw_off, int n, int w) {
    int lo = 0, hi = n;   // largest c with cw_off[c] <= w
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (cw_off[mid] <= w) lo = mid; else hi = mid;
    }
    return lo;
}

template <bool FIXED, bool ENERGY>
__global__ __launch_bounds__(128) void compound_listed_kernel(StepArgs a) {
    __shared__ float4 src[kSolventChunk];
    __shared__ float2 tab[kMaxTypes * kMaxTypes];   // (A, B) of lj_ab: 8-byte entries keep the per-lane reads conflict-light
    __shared__ float4 xch[kCompoundBlock / 2];      // the second wave's partial sums of a split item
    __shared__ int cached_rep;
    const int p = threadIdx.x;
    if (p == 0) cached_rep = -1;
    const int total = *a.cw_total;
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        const int c = a.item_comp[w];
        const int k = w - a.cw_off[c];
        const int r = a.comp_rep[c];
        const ReplicaInfo rp = a.reps[r];
        const float L = rp.L, half = rp.half;
        const int base = a.comp_off[c], n = a.comp_off[c + 1] - base;
        // one thread per particle sums the item's sources whole, in list order, as the packed form
        // does.  With a.listed_split (FP32 only, an autotuner choice for small systems) a compound
        // of at most one wavefront's particles has both waves take particle pi, each one contiguous
        // half of the item's sources, the halves added in order at the end.
        const bool split = !FIXED && a.listed_split && n <= kCompoundBlock / 2;
        const int pi = split ? (p & (kCompoundBlock / 2 - 1)) : p;
        const int hs = split ? (p / (kCompoundBlock / 2)) : 0;
        const bool active = pi < n;
        float4 me = a.cpos[base + (active ? pi : 0)];
        const float4 p0 = a.cpos[base];
        hyperpos(p0.x, p0.y, p0.z, me.x, me.y, me.z, L, half);
        const int tp = ftype(me);
        const int ncc = a.cc_cnt[c];
        __syncthreads();
        float lx = 0.f, ly = 0.f, lz = 0.f;
        double en = 0.0, vir = 0.0;
        if (k < ncc) {
            if (cached_rep != r) {
                // a compile-time stride, so that the eight loads per lane are in flight together
                float4 v[kMaxTypes * kMaxTypes / kCompoundBlock];
#pragma unroll
                for (int u = 0; u < kMaxTypes * kMaxTypes / kCompoundBlock; u++)
                    v[u] = a.tab_inter[(size_t)r * kMaxTypes * kMaxTypes + p + u * kCompoundBlock];
#pragma unroll
                for (int u = 0; u < kMaxTypes * kMaxTypes / kCompoundBlock; u++) tab[p + u * kCompoundBlock] = lj_ab(v[u].x, v[u].y);
            }
            const int nb = a.cc_ids[(size_t)c * a.cap_cc + k];
            const int nbase = a.comp_off[nb], nn = a.comp_off[nb + 1] - nbase;
            if (p < nn) {
                float4 q = a.cpos[nbase + p];
                hyperpos(p0.x, p0.y, p0.z, q.x, q.y, q.z, L, half);
                src[p] = q;
            }
            // LJ_Ignores of this particle against the neighbour (Engine.cu:660)
            uint32_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
            if (active) {
                for (int e = a.xoff[base + pi]; e < a.xoff[base + pi + 1]; e++) {
                    int2 x = a.xent[e];
                    if (x.x == nb) {
                        uint32_t bit = 1u << (x.y & 31);
                        int wsel = x.y >> 5;
                        m0 |= wsel == 0 ? bit : 0u; m1 |= wsel == 1 ? bit : 0u;
                        m2 |= wsel == 2 ? bit : 0u; m3 |= wsel == 3 ? bit : 0u;
                    }
                }
            }
            __syncthreads();
            if (p == 0) cached_rep = r;
            // exclusions across compounds exist only between bridged compounds: a wave
            // with none against this neighbour takes the loop without the mask test
            const bool wave_clean = __ballot((m0 | m1 | m2 | m3) != 0u) == 0ull;
            const int hn = split ? (nn + 1) / 2 : nn;
            const int jlo = hs * hn, jhi = min(nn, jlo + hn);
            if (active && wave_clean) {
                const float2* trow = tab + tp;
#pragma unroll 4
                for (int j = jlo; j < jhi; j++) {
                    const float4 q = src[j];
                    const float2 pr = trow[ftype(q) * kMaxTypes];
                    const float dx = q.x - me.x, dy = q.y - me.y, dz = q.z - me.z;
                    const float r2 = pair_r2(dx, dy, dz);
                    float inv;
                    const float fs = lj_scalar(r2, pr.x, pr.y, inv);
                    pair_add(dx, dy, dz, fs, lx, ly, lz);
                    if (ENERGY) {
                        en += lj_energy(inv, pr.x, pr.y);
                        pair_virial(r2, fs, vir);
                    }
                }
            } else if (active) {
                for (int g = 0; g * 32 < nn; g++) {
                    const uint32_t mw = g == 0 ? m0 : g == 1 ? m1 : g == 2 ? m2 : m3;
                    const int jend = min(jhi, g * 32 + 32);
                    for (int j = max(jlo, g * 32); j < jend; j++) {
                        const float4 q = src[j];
                        const float2 pr = tab[ftype(q) * kMaxTypes + tp];
                        float dx = q.x - me.x, dy = q.y - me.y, dz = q.z - me.z;
                        float r2 = pair_r2(dx, dy, dz);
                        float inv;
                        float fs = lj_scalar(r2, pr.x, pr.y, inv);
                        const bool ex = (mw >> (j & 31)) & 1u;
                        fs = ex ? 0.0f : fs;
                        pair_add(dx, dy, dz, fs, lx, ly, lz);
                        if (ENERGY && !ex) {
                            en += lj_energy(inv, pr.x, pr.y);
                            pair_virial(r2, fs, vir);
                        }
                    }
                }
            }
            if (ENERGY) {
                block_atomic_energy(en, a.energy + r * kEnergyTerms + kE_inter);
                block_atomic_energy(vir, a.energy + r * kEnergyTerms + kE_virial);
            }
        } else {
            const int kk = k - ncc;
            const int j0 = kk * kSolventChunk;
            const int nsrc = min(kSolventChunk, a.cs_cnt[c] - j0);
#pragma unroll
            for (int u = 0; u < kSolventChunk / kCompoundBlock; u++) {
                const int jj = p + u * kCompoundBlock;
                if (jj < nsrc) {
                    float4 q = a.spos[a.cs_ids[(size_t)c * a.cap_cs + j0 + jj]];
                    hyperpos(p0.x, p0.y, p0.z, q.x, q.y, q.z, L, half);
                    src[jj] = q;
                }
            }
            __syncthreads();
            const float4 pv = a.tab_cs[r * kMaxTypes + tp];
            const float2 pr = lj_ab(pv.x, pv.y);
            const int hn = split ? (nsrc + 1) / 2 : nsrc;
            const int jlo = hs * hn, jhi = min(nsrc, jlo + hn);
            if (active) {
                for (int j = jlo; j < jhi; j++) {
                    const float4 q = src[j];
                    if (ENERGY) lj_acc_energy(me.x, me.y, me.z, q.x, q.y, q.z, pr.x, pr.y, lx, ly, lz, en, vir);
                    else lj_acc(me.x, me.y, me.z, q.x, q.y, q.z, pr.x, pr.y, lx, ly, lz);
                }
            }
            if (ENERGY) {
                block_atomic_energy(en, a.energy + r * kEnergyTerms + kE_cs);
                block_atomic_energy(vir, a.energy + r * kEnergyTerms + kE_virial);
            }
        }
        if (split) {
            if (hs == 1) xch[pi] = make_float4(lx, ly, lz, 0.f);
            __syncthreads();
            if (hs == 0 && active) a.partC[(size_t)w * kCompoundBlock + pi] = make_float4(lx + xch[pi].x, ly + xch[pi].y,
                                                                                     lz + xch[pi].z, 0.f);
        } else if (active) {
            const size_t slot = (size_t)w * kCompoundBlock + p;
            if (FIXED) {
                uint32_t bad = 0;
                a.partC_fx[slot] = I3{to_fixed(lx, bad), to_fixed(ly, bad), to_fixed(lz, bad)};
                if (bad) atomicOr(a.err, bad);
            } else a.partC[slot] = make_float4(lx, ly, lz, 0.f);
        }
    }
}

// ---------------------------------------------------------------------------
// compound_listed_kernel for engines whose compounds all hold at most 16 * PPL particles
// (PPL 1, 2 or 4 particles per lane).  A block is one wavefront of four groups of 16 lanes and
// takes four runs at once, one per group.  A run is either one solvent-chunk work item (its
// 256 sources staged and summed in two halves of 128), or up to cc_run(PPL) = 8 / PPL
// consecutive neighbour-compound items of one compound (at most 128 sources, staged together).
// Lane pi of a group holds particles pi, pi + 16, ... of the run's compound and sums each
// item's sources whole in list order, so every work item is still one sequential sum per
// particle written to its own partial slot, as in compound_listed_kernel.  A group reads only
// its own LDS row, so wavefront-level ordering is the only synchronisation.  Groups may belong
// to different replicas: the inter-compound pair constants come from the global table
// tab_inter_ab.
// ---------------------------------------------------------------------------
constexpr int kPackLanes = 16;
constexpr int kPackGroups = 64 / kPackLanes;
constexpr int kPackRow = 128;                 // staged sources per group
__host__ __device__ constexpr int cc_run(int ppl) { return kPackRow / (kPackLanes * ppl); }

template <int W> LIMA_DEV double group_sum(double v) {
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

LIMA_DEV void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool FIXED, bool ENERGY, int PPL>
__global__ __launch_bounds__(64) void compound_listed_packed_kernel(StepArgs a) {
    static_assert(PPL == 1 || PPL == 2 || PPL == 4, "16, 32 or 64 particles per compound");
    constexpr int RUN = cc_run(PPL);
    // rows padded by one entry: the four groups' rows start in different LDS banks
    __shared__ float4 src[kPackGroups][kPackRow + 1];
    __shared__ int run_nb[kPackGroups][RUN], run_off[kPackGroups][RUN + 1];
    const int p = threadIdx.x, g = p / kPackLanes, pi = p % kPackLanes;
    float4* sg = src[g];
    const int n_ccr = a.run_total[0], total = n_ccr + a.run_total[1];
    for (int s0 = blockIdx.x * kPackGroups; s0 < total; s0 += gridDim.x * kPackGroups) {
        const int sr = s0 + g;
        const bool valid = sr < total;
        const int si = valid ? sr : s0;
        const int c = a.run_comp[si];
        const bool is_cc = si < n_ccr;
        const int ncc = a.cc_cnt[c];
        const int r = a.comp_rep[c];
        const ReplicaInfo rp = a.reps[r];
        const float L = rp.L, half = rp.half;
        const int base = a.comp_off[c], n = a.comp_off[c + 1] - base;
        const float4 p0 = a.cpos[base];
        float mx[PPL], my[PPL], mz[PPL];
        int tp[PPL], xb[PPL], xe[PPL];
        bool act[PPL];
#pragma unroll
        for (int q = 0; q < PPL; q++) {
            const int pidx = pi + kPackLanes * q;
            act[q] = valid && pidx < n;
            const int gi = base + (act[q] ? pidx : 0);
            float4 me = a.cpos[gi];
            hyperpos(p0.x, p0.y, p0.z, me.x, me.y, me.z, L, half);                 // Engine.cu:923
            mx[q] = me.x; my[q] = me.y; mz[q] = me.z;
            tp[q] = ftype(me);
            xb[q] = act[q] ? a.xoff[gi] : 0;
            xe[q] = act[q] ? a.xoff[gi + 1] : 0;
        }
        // the run's items: neighbour items w_first .. w_first + nitems - 1, or one solvent chunk
        const int k0 = is_cc ? (si - a.run_off[c]) * RUN : ncc + (si - n_ccr - a.run_cs_off[c]);
        const int nitems = is_cc ? min(RUN, ncc - k0) : 1;
        const int w_first = a.cw_off[c] + k0;
        double en = 0.0, vir = 0.0;
        auto store = [&](int w, int q, float lx, float ly, float lz) {
            const size_t slot = (size_t)w * kCompoundBlock + pi + kPackLanes * q;
            if (FIXED) {
                uint32_t bad = 0;
                a.partC_fx[slot] = I3{to_fixed(lx, bad), to_fixed(ly, bad), to_fixed(lz, bad)};
                if (bad) atomicOr(a.err, bad);
            } else a.partC[slot] = make_float4(lx, ly, lz, 0.f);
        };
        // The pair loops, for the first NQ particle slots of each lane: when every group of the
        // wavefront holds a compound of at most 16 particles the second slot is idle, and the
        // loops skip it (a wavefront-uniform choice; the active slots' sums are unchanged)
        auto pairs = [&](auto nq_tag) {
        constexpr int NQ = decltype(nq_tag)::value;
        wave_sync();   // every group of the wavefront is done with the previous round's rows
        if (valid && is_cc) {
            // the run's neighbour compounds, re-imaged against this compound's particle 0 (Engine.cu:949)
            if (pi < nitems) {
                const int nb = a.cc_ids[(size_t)c * a.cap_cc + k0 + pi];
                run_nb[g][pi] = nb;
                run_off[g][pi + 1] = a.comp_off[nb + 1] - a.comp_off[nb];
            }
            wave_sync();
            if (pi == 0) {
                run_off[g][0] = 0;
                for (int t = 0; t < nitems; t++) run_off[g][t + 1] += run_off[g][t];
            }
            wave_sync();
            for (int t = 0; t < nitems; t++) {
                const int nbase = a.comp_off[run_nb[g][t]], o = run_off[g][t], nn = run_off[g][t + 1] - o;
                for (int j = pi; j < nn; j += kPackLanes) {
                    float4 qv = a.cpos[nbase + j];
                    hyperpos(p0.x, p0.y, p0.z, qv.x, qv.y, qv.z, L, half);
                    sg[o + j] = qv;
                }
            }
            wave_sync();
            const float2* tab = a.tab_inter_ab + (size_t)r * kMaxTypes * kMaxTypes;
            for (int t = 0; t < nitems; t++) {
                const int nb = run_nb[g][t], j0 = run_off[g][t], j1 = run_off[g][t + 1];
                // LJ_Ignores of each particle against this neighbour (Engine.cu:660); only
                // particles at a bridged junction have entries
                uint32_t m0[PPL], m1[PPL];
                bool clean = true;
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    m0[q] = m1[q] = 0u;
                    for (int e = xb[q]; e < xe[q]; e++) {
                        const int2 x = a.xent[e];
                        if (x.x == nb) {
                            const uint32_t bit = 1u << (x.y & 31);
                            m0[q] |= x.y < 32 ? bit : 0u;
                            m1[q] |= x.y >= 32 ? bit : 0u;
                        }
                    }
                    clean = clean && (m0[q] | m1[q]) == 0u;
                }
                float lx[PPL], ly[PPL], lz[PPL];
#pragma unroll
                for (int q = 0; q < NQ; q++) lx[q] = ly[q] = lz[q] = 0.f;
                if (clean) {
#pragma unroll 2
                    for (int j = j0; j < j1; j++) {
                        const float4 qv = sg[j];
                        const int ts = ftype(qv) * kMaxTypes;
#pragma unroll
                        for (int q = 0; q < NQ; q++) {
                            const float2 pr = tab[ts + tp[q]];
                            const float dx = qv.x - mx[q], dy = qv.y - my[q], dz = qv.z - mz[q];
                            const float r2 = pair_r2(dx, dy, dz);
                            float inv;
                            const float fs = lj_scalar(r2, pr.x, pr.y, inv);
                            pair_add(dx, dy, dz, fs, lx[q], ly[q], lz[q]);
                            if (ENERGY && act[q]) {
                                en += lj_energy(inv, pr.x, pr.y);
                                pair_virial(r2, fs, vir);
                            }
                        }
                    }
                } else {
                    for (int j = j0; j < j1; j++) {
                        const float4 qv = sg[j];
                        const int ts = ftype(qv) * kMaxTypes, jl = j - j0;
#pragma unroll
                        for (int q = 0; q < NQ; q++) {
                            const float2 pr = tab[ts + tp[q]];
                            const float dx = qv.x - mx[q], dy = qv.y - my[q], dz = qv.z - mz[q];
                            const float r2 = pair_r2(dx, dy, dz);
                            float inv;
                            float fs = lj_scalar(r2, pr.x, pr.y, inv);
                            const bool ex = (((jl < 32 ? m0[q] : m1[q]) >> (jl & 31)) & 1u) != 0u;
                            fs = ex ? 0.0f : fs;
                            pair_add(dx, dy, dz, fs, lx[q], ly[q], lz[q]);
                            if (ENERGY && act[q] && !ex) {
                                en += lj_energy(inv, pr.x, pr.y);
                                pair_virial(r2, fs, vir);
                            }
                        }
                    }
                }
#pragma unroll
                for (int q = 0; q < NQ; q++)
                    if (act[q]) store(w_first + t, q, lx[q], ly[q], lz[q]);
            }
        } else if (valid) {
            // 256 entries of the compound's solvent list (Engine.cu:967), staged in two halves
            const int j0 = (k0 - ncc) * kSolventChunk;
            const int nsrc = min(kSolventChunk, a.cs_cnt[c] - j0);
            const int* ids = a.cs_ids + (size_t)c * a.cap_cs + j0;
            float2 pr_cs[PPL];
#pragma unroll
            for (int q = 0; q < PPL; q++) {
                const float4 pv = a.tab_cs[r * kMaxTypes + tp[q]];
                pr_cs[q] = lj_ab(pv.x, pv.y);
            }
            float lx[PPL], ly[PPL], lz[PPL];
#pragma unroll
            for (int q = 0; q < NQ; q++) lx[q] = ly[q] = lz[q] = 0.f;
            for (int h0 = 0; h0 < nsrc; h0 += kPackRow) {
                const int hn = min(kPackRow, nsrc - h0);
                if (h0 > 0) wave_sync();   // the first half is summed before the row is rewritten
#pragma unroll
                for (int u = 0; u < kPackRow / kPackLanes; u++) {
                    const int jj = pi + u * kPackLanes;
                    if (jj < hn) {
                        float4 qv = a.spos[ids[h0 + jj]];
                        hyperpos(p0.x, p0.y, p0.z, qv.x, qv.y, qv.z, L, half);
                        sg[jj] = qv;
                    }
                }
                wave_sync();
#pragma unroll 2
                for (int j = 0; j < hn; j++) {
                    const float4 qv = sg[j];
#pragma unroll
                    for (int q = 0; q < NQ; q++) {
                        if (ENERGY) {
                            double e0 = 0.0, w0 = 0.0;
                            lj_acc_energy(mx[q], my[q], mz[q], qv.x, qv.y, qv.z, pr_cs[q].x, pr_cs[q].y, lx[q], ly[q],
                                          lz[q], e0, w0);
                            if (act[q]) { en += e0; vir += w0; }
                        } else lj_acc(mx[q], my[q], mz[q], qv.x, qv.y, qv.z, pr_cs[q].x, pr_cs[q].y, lx[q], ly[q], lz[q]);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < NQ; q++)
                if (act[q]) store(w_first, q, lx[q], ly[q], lz[q]);
        }
        };
        if (PPL > 1 && __ballot(valid && n > kPackLanes) == 0ull) pairs(std::integral_constant<int, 1>{});
        else pairs(std::integral_constant<int, PPL>{});
        if (ENERGY) {
            en = group_sum<kPackLanes>(en);
            vir = group_sum<kPackLanes>(vir);
            if (valid && pi == 0) {
                if (en != 0.0) atomicAdd(a.energy + r * kEnergyTerms + (is_cc ? kE_inter : kE_cs), en);
                if (vir != 0.0) atomicAdd(a.energy + r * kEnergyTerms + kE_virial, vir);
            }
        }
    }
}

// lj_ab of every inter-compound table entry: the global copy the packed listed kernel reads
__global__ void ab_table_kernel(const float4* t, float2* ab, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ab[i] = lj_ab(t[i].x, t[i].y);
}

// ---------------------------------------------------------------------------
// compoundKernel's local forces (Engine.cu:918-937, 678-700): bridge carry-in,
// bonded terms (term forces computed once, then gathered per particle in term
// index order = the serialized scatter's order, Engine.cu:765-771) and the
// intra-compound LJ (geometric eps, exclusions), summed in that order.  It
// needs no neighbour list, so it runs beside the listed kernel; the sum goes to
// cloc (int64 in the fixed-point form).
// ---------------------------------------------------------------------------
// BLOCK threads (64 when every compound of the engine fits one wavefront, else 128); the term
// results are staged in dynamic LDS sized to the engine's largest compound (local_lds_bytes):
// [BLOCK] float4 positions, [1024] float2 pair constants, then the bond, angle and dihedral
// results at strides a.loc_nb, a.loc_na, a.loc_nd terms.
template <int BLOCK>
inline size_t local_lds_bytes(int nb, int na, int nd) {
    return (size_t)BLOCK * sizeof(float4) + (size_t)kMaxTypes * kMaxTypes * sizeof(float2) +
           (size_t)(2 * nb + 3 * na + 4 * nd) * sizeof(F3);
}

template <bool FIXED, bool ENERGY, int BLOCK>
__global__ __launch_bounds__(BLOCK) void compound_local_kernel(StepArgs a) {
    static_assert(BLOCK == 64 || BLOCK == 128, "one or two wavefronts");
    extern __shared__ float4 dyn_lds[];
    float4* pos = dyn_lds;
    float2* tab = reinterpret_cast<float2*>(dyn_lds + BLOCK);
    F3* rb = reinterpret_cast<F3*>(tab + kMaxTypes * kMaxTypes);
    F3* ra = rb + 2 * a.loc_nb;
    F3* rd = ra + 3 * a.loc_na;
    const int c = blockIdx.x;
    const int p = threadIdx.x;
    const int r = a.comp_rep[c];
    const ReplicaInfo rp = a.reps[r];
    const float L = rp.L, half = rp.half;
    const int base = a.comp_off[c], n = a.comp_off[c + 1] - base;
    const bool active = p < n;
    const int gi = base + (active ? p : 0);
    float4 me = a.cpos[gi];
    const float4 p0 = a.cpos[base];
    hyperpos(p0.x, p0.y, p0.z, me.x, me.y, me.z, L, half);                 // Engine.cu:923
    pos[p] = me;
    {
        // a compile-time stride, so that the eight loads per lane are in flight together
        float4 v[kMaxTypes * kMaxTypes / BLOCK];
#pragma unroll
        for (int u = 0; u < kMaxTypes * kMaxTypes / BLOCK; u++)
            v[u] = a.tab_intra[(size_t)r * kMaxTypes * kMaxTypes + p + u * BLOCK];
#pragma unroll
        for (int u = 0; u < kMaxTypes * kMaxTypes / BLOCK; u++) tab[p + u * BLOCK] = lj_ab(v[u].x, v[u].y);
    }
    __syncthreads();
    double eb = 0, ea = 0, ed = 0;
    const int b0 = a.bond_off[c], nb = a.bond_off[c + 1] - b0;
    const int a0 = a.angle_off[c], na = a.angle_off[c + 1] - a0;
    const int d0 = a.dih_off[c], nd = a.dih_off[c + 1] - d0;
    auto P = [&](int i) { float4 q = pos[i]; return f3(q.x, q.y, q.z); };
    for (int t = p; t < nb; t += BLOCK) {
        BondD bd = a.bonds[b0 + t];
        pairbond(P(bd.a0), P(bd.a1), bd.b0, bd.kb, &rb[t * 2], &eb);
    }
    for (int t = p; t < na; t += BLOCK) {
        AngleD ad = a.angles[a0 + t];
        anglebond(P(ad.a0), P(ad.a1), P(ad.a2), ad.t0, ad.k, &ra[t * 3], &ea);
    }
    if (a.enable_dihedrals) {
        for (int t = p; t < nd; t += BLOCK) {
            DihD dd = a.dihs[d0 + t];
            dihedral(P(dd.a0), P(dd.a1), P(dd.a2), P(dd.a3), dd.p0, dd.k, dd.n, &rd[t * 4], &ed);
        }
    }
    __syncthreads();
    float4 fb = a.fbridge[gi];
    F3 force = f3(fb.x, fb.y, fb.z);                                        // Engine.cu:918
    F3 bs = f3(0.f, 0.f, 0.f), as = bs, ds = bs;
    if (active) {
        for (int e = a.inc_off[gi]; e < a.inc_off[gi + 1]; e++) {
            const uint32_t v = a.inc[e];
            const uint32_t kind = v >> 30, slot = v & 3u, term = (v >> 2) & 0xFFFFFFFu;
            if (kind == 0) bs = f3_add(bs, rb[term * 2 + slot]);
            else if (kind == 1) as = f3_add(as, ra[term * 3 + slot]);
            else ds = f3_add(ds, rd[term * 4 + slot]);
        }
    }
    long long fx_i = 0, fy_i = 0, fz_i = 0;
    uint32_t bad = 0;
    if (FIXED) {
        fx_i = to_fixed(force.x, bad) + to_fixed(bs.x, bad) + to_fixed(as.x, bad) + (a.enable_dihedrals ? to_fixed(ds.x, bad) : 0);
        fy_i = to_fixed(force.y, bad) + to_fixed(bs.y, bad) + to_fixed(as.y, bad) + (a.enable_dihedrals ? to_fixed(ds.y, bad) : 0);
        fz_i = to_fixed(force.z, bad) + to_fixed(bs.z, bad) + to_fixed(as.z, bad) + (a.enable_dihedrals ? to_fixed(ds.z, bad) : 0);
    } else {
        force = f3_add(force, bs);
        force = f3_add(force, as);
        if (a.enable_dihedrals) force = f3_add(force, ds);
    }
    // computeIntramolecularLJForces (Engine.cu:678-700)
    float lx = 0.f, ly = 0.f, lz = 0.f;
    double e_intra = 0, vir = 0;
    if (active) {
        const int tp = ftype(me);
        const uint32_t* mk = a.intra_mask + (size_t)gi * 4;
        const uint32_t m[4] = {mk[0], mk[1], mk[2], mk[3]};
        for (int g = 0; g * 32 < n; g++) {
            const uint32_t mw = g == 0 ? m[0] : g == 1 ? m[1] : g == 2 ? m[2] : m[3];
            const int jend = min(n, g * 32 + 32);
            for (int j = g * 32; j < jend; j++) {
                const float4 q = pos[j];
                const float2 pr = tab[ftype(q) * kMaxTypes + tp];
                float dx = q.x - me.x, dy = q.y - me.y, dz = q.z - me.z;
                float r2 = pair_r2(dx, dy, dz);
                float inv;
                float fs = lj_scalar(r2, pr.x, pr.y, inv);
                const bool ex = (j == p) || ((mw >> (j & 31)) & 1u);
                fs = ex ? 0.0f : fs;
                pair_add(dx, dy, dz, fs, lx, ly, lz);
                if (ENERGY && !ex) {
                    e_intra += lj_energy(inv, pr.x, pr.y);
                    pair_virial(r2, fs, vir);
                }
            }
        }
    }
    if (FIXED) {
        fx_i += to_fixed(lx, bad); fy_i += to_fixed(ly, bad); fz_i += to_fixed(lz, bad);
        if (bad) atomicOr(a.err, bad);
    } else force = f3_add(force, f3(lx, ly, lz));
    // the sum so far, in the order above; compound_integrate_kernel adds the listed partials to it
    if (active) {
        if (FIXED) a.cloc_fx[gi] = I3{fx_i, fy_i, fz_i};
        else a.cloc[gi] = make_float4(force.x, force.y, force.z, 0.f);
    }
    // thread 0's potE slot of the logged replica's first compound (Engine.cu:1027): the dihedral
    // assigns it (Engine.cu:612), else its bond and angle terms
    if (a.log_rep == r && c == rp.comp_begin && p == 0) {
        float* o = a.log_data + (size_t)(a.step_count[r] % a.log_ring) * 10;
        o[2] = (float)((a.enable_dihedrals && nd > 0) ? ed : eb + ea);
    }
    if (ENERGY) {
        block_atomic_energy(eb, a.energy + r * kEnergyTerms + kE_bond);
        block_atomic_energy(ea, a.energy + r * kEnergyTerms + kE_angle);
        block_atomic_energy(ed, a.energy + r * kEnergyTerms + kE_dih);
        block_atomic_energy(e_intra, a.energy + r * kEnergyTerms + kE_intra);
        block_atomic_energy(vir, a.energy + r * kEnergyTerms + kE_virial);
    }
}

// ---------------------------------------------------------------------------
// Thermostat (Engine::handleBoxtemp, Engine.cu:362-393, made a device FP64
// reduction).  One wavefront per replica, fixed summation order.
//   mode 1 (Berendsen): T = sum m|dx|^2 / dt^2 / (3 N R) every step;
//     lambda = sqrt(1 + dt/tau (T0/T - 1)), clamped to [0.9, 1.1], T=0 guard.
//   mode 2 (the reference's cadence, Engine.cu:49-75, 314-393): after step k,
//     with s = k + 1 steps done, s % 40 == 0 evaluates T from the per-particle
//     FP32 kinetic energies of the positions two steps apart,
//     T = (float)(sum / N * 2 / (3 * 8.3145)), and, when s > 20, sets
//     lambda = clamp((float)T0 / T, 0.9, 1.1) for step s only; every other
//     step runs with lambda = 1.  step_count[r] counts the steps done.
// ---------------------------------------------------------------------------
constexpr int kThermostatEvery = 40;   // STEPS_PER_THERMOSTAT (Constants.cuh:82)

// One replica's thermostat by one wavefront (lane = its lane index).
LIMA_DEV void thermostat_replica(int r, const ReplicaInfo& rp, int lane, const double* ke_comp, const double* ke_stile,
                                 float* lambda, float* lambda_applied, double* temperature, long long* step_count,
                                 int mode, unsigned long long* kclock) {
    double s = 0.0;
    for (int c = rp.comp_begin + lane; c < rp.comp_end; c += 64) s += ke_comp[c];
    for (int t = rp.stile_begin + lane; t < rp.stile_end; t += 64) s += ke_stile[t];
    s = wave_sum(s);
    if (lane == 0) {
        lambda_applied[r] = lambda[r];
        const long long done = step_count[r] + 1;
        step_count[r] = done;
        if (kclock && r == 0) {   // this step's span of the all-pairs kernel
            const unsigned long long b = kclock[0], e = kclock[1];
            if (b != ~0ull && e > b) {
                kclock[2] += e - b;
                kclock[3] += 1;
            }
            kclock[0] = ~0ull;
            kclock[1] = 0ull;
        }
        const double ntot = (double)((rp.chain_end - rp.chain_begin) + (rp.solv_end - rp.solv_begin));
        float lam = 1.0f;
        if (mode == 2) {
            if (done % kThermostatEvery == 0) {
                const double avg = ntot > 0 ? s / ntot : 0.0;
                float temp = (float)(avg * 2.0f / (3.0f * 8.3145));
                temp = temp == 0.f ? 1.f : temp;
                temperature[r] = temp;
                float sc = rp.T0 / temp;
                sc = fminf(sc, 1.1f);
                sc = fmaxf(sc, 0.9f);
                if (done > 20) lam = sc;
            }
        } else {
            const double temp = ntot > 0 ? s / (rp.dt * rp.dt) / (3.0 * ntot * 8.3145) : 0.0;
            temperature[r] = temp;
            if (mode == 1) {
                double t = temp == 0.0 ? 1.0 : temp;
                double l = sqrt(1.0 + rp.dt / rp.tau * ((double)rp.T0 / t - 1.0));
                l = l < 1.1 ? l : 1.1;
                l = l > 0.9 ? l : 0.9;
                lam = (float)l;
            }
        }
        lambda[r] = lam;
    }
}

__global__ __launch_bounds__(64) void thermostat_kernel(const ReplicaInfo* reps, int n_rep, const double* ke_comp,
                                                      const double* ke_stile, float* lambda, float* lambda_applied,
                                                      double* temperature, long long* step_count, int mode,
                                                      unsigned long long* kclock) {
    const int r = blockIdx.x;
    if (r >= n_rep) return;
    thermostat_replica(r, reps[r], threadIdx.x, ke_comp, ke_stile, lambda, lambda_applied, temperature, step_count, mode,
                       kclock);
}

// The thermostat at the end of an integrator block: the replica's last block to finish runs
// it (a count per replica, no waiting), so no separate launch follows the integrators.  The
// release fence orders this block's kinetic-energy partial before its count; the last block's
// acquire fence makes every partial of the replica visible.  The sums are thermostat_kernel's.
LIMA_DEV void thermostat_if_last(const StepArgs& a, int r) {
    if (!a.thermo_cnt) return;
    __shared__ int last;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(a.thermo_cnt + r, 1) == a.thermo_total[r] - 1;
    __syncthreads();
    if (last) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (threadIdx.x < 64)
            thermostat_replica(r, a.reps[r], threadIdx.x, a.ke_comp, a.ke_stile, a.lambda_next, a.lambda_applied,
                               a.temperature, a.step_count_w, a.thermo_mode, a.kclock);
        if (threadIdx.x == 0) a.thermo_cnt[r] = 0;
    }
}

// ---------------------------------------------------------------------------
// compoundKernel remainder (Engine.cu:939-1056): the local sum plus the listed
// partials in list order, integration, PBC.  ADVANCE=false computes forces
// (and the kinetic energy) without changing the state.
// ---------------------------------------------------------------------------
template <bool ADVANCE, bool FIXED, bool ENERGY>
__global__ __launch_bounds__(128) void compound_integrate_kernel(StepArgs a) {
    __shared__ float4 pos[kCompoundBlock];
    __shared__ double red[2];
    const int c = blockIdx.x;
    const int p = threadIdx.x;
    const int r = a.comp_rep[c];
    const ReplicaInfo rp = a.reps[r];
    const float L = rp.L, half = rp.half;
    const int base = a.comp_off[c], n = a.comp_off[c + 1] - base;
    const bool active = p < n;
    const int gi = base + (active ? p : 0);
    float4 me = a.cpos[gi];
    const float4 p0 = a.cpos[base];
    hyperpos(p0.x, p0.y, p0.z, me.x, me.y, me.z, L, half);                 // Engine.cu:923
    F3 force = f3(0.f, 0.f, 0.f);
    long long fx_i = 0, fy_i = 0, fz_i = 0;
    if (active) {
        if (FIXED) { const I3 v = a.cloc_fx[gi]; fx_i = v.x; fy_i = v.y; fz_i = v.z; }
        else { const float4 v = a.cloc[gi]; force = f3(v.x, v.y, v.z); }
    }
    // listed partials: neighbour compounds (list order), then solvent chunks
    const bool logged = ADVANCE && a.log_rep == r;
    F3 f_sol = f3(-1.f, -1.f, -1.f);   // computeSolventToCompoundLJForces' result, Float3(-1) when unset (Engine.cu:919)
    if (active) {
        const int s0 = a.cw_off[c], s1 = a.cw_off[c + 1];
        // eight slots' loads in flight at once, then added in slot order
        for (int s = s0; s < s1; s += 8) {
            if (FIXED) {
                I3 v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) v[u] = s + u < s1 ? a.partC_fx[(size_t)(s + u) * kCompoundBlock + p] : I3{0, 0, 0};
#pragma unroll
                for (int u = 0; u < 8; u++) { fx_i += v[u].x; fy_i += v[u].y; fz_i += v[u].z; }
            } else {
                float4 v[8];
#pragma unroll
                for (int u = 0; u < 8; u++)
                    v[u] = s + u < s1 ? a.partC[(size_t)(s + u) * kCompoundBlock + p] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                for (int u = 0; u < 8; u++)
                    if (s + u < s1) { force.x += v[u].x; force.y += v[u].y; force.z += v[u].z; }
            }
        }
        if (logged && a.enable_solvents) {
            long long sx_i = 0, sy_i = 0, sz_i = 0;
            float gx = 0.f, gy = 0.f, gz = 0.f;
            for (int s = s0 + a.cc_cnt[c]; s < s1; s++) {
                const size_t slot = (size_t)s * kCompoundBlock + p;
                if (FIXED) { I3 v = a.partC_fx[slot]; sx_i += v.x; sy_i += v.y; sz_i += v.z; }
                else { float4 v = a.partC[slot]; gx += v.x; gy += v.y; gz += v.z; }
            }
            f_sol = FIXED ? f3(from_fixed(sx_i), from_fixed(sy_i), from_fixed(sz_i)) : f3(gx, gy, gz);
        }
    }
    if (FIXED) force = f3(from_fixed(fx_i), from_fixed(fy_i), from_fixed(fz_i));
    if (active) a.cforce[gi] = make_float4(force.x, force.y, force.z, 0.f);
    if (active && exact_len(force.x, force.y, force.z) > 2e12f) atomicOr(a.err, kErrCriticalForce);   // Engine.cu:1050
    double ke = 0.0, kt = 0.0;   // one-step m|dx|^2; what the thermostat sums
    const double m = (double)a.mass[r * kMaxTypes + ftype(me)];
    float4 pv = a.cprev[gi];
    if (ADVANCE) {
        float px = me.x, py = me.y, pz = me.z, qx = pv.x, qy = pv.y, qz = pv.z, dx, dy, dz;
        const double mts = rp.dt / m * rp.dt;
        integrate(px, py, pz, qx, qy, qz, force.x, force.y, force.z, mts, a.lambda[r], L, half, dx, dy, dz);
        if (active) {
            a.cprev[gi] = make_float4(qx, qy, qz, 0.f);
            ke = m * ((double)dx * dx + (double)dy * dy + (double)dz * dz);
        }
        __syncthreads();
        pos[p] = make_float4(px, py, pz, me.w);
        __syncthreads();
        // PBC on particle 0, then re-image every particle against it (Engine.cu:993-996)
        float4 z = pos[0];
        float zx = pbc1(z.x, L), zy = pbc1(z.y, L), zz = pbc1(z.z, L);
        hyperpos(zx, zy, zz, px, py, pz, L, half);
        if (active) a.cpos_next[gi] = make_float4(px, py, pz, me.w);
        if (logged) {
            const long long k = a.step_count[r];
            const int ring = (int)(k % a.log_ring);
            const int cl = c - rp.comp_begin;
            if (a.train_data && p < a.log_slots) {
                // data_GAN (Engine.cu:1036-1043): position, solvent LJ force, and a third Float3 left zero
                float* t = a.train_data + (((size_t)ring * (rp.comp_end - rp.comp_begin) + cl) * a.log_slots + p) * 9;
                const bool in = p < n;
                t[0] = in ? px : -1.f; t[1] = in ? py : -1.f; t[2] = in ? pz : -1.f;
                t[3] = f_sol.x; t[4] = f_sol.y; t[5] = f_sol.z;
                t[6] = 0.f; t[7] = 0.f; t[8] = 0.f;
            }
            if (cl == 0 && p == 0) {
                // outdata of compound 0, particle 0 (Engine.cu:1019-1028): speed, kinetic energy of the
                // one-step-apart positions, |F| (thread 0's potE slot comes from compound_local_kernel)
                float bx = qx, by = qy, bz = qz;
                hyperpos(px, py, pz, bx, by, bz, L, half);
                float* o = a.log_data + (size_t)ring * 10;
                o[0] = (float)((double)exact_len(fsub(px, bx), fsub(py, by), fsub(pz, bz)) / rp.dt);
                o[1] = ref_kinetic(px, py, pz, bx, by, bz, m, rp.dt, L, half);
                o[3] = exact_len(force.x, force.y, force.z);
                for (int j = 4; j < 10; j++) o[j] = 0.f;
            }
        }
        // the logged position of this step against the one two steps back (Engine.cu:333-337)
        kt = !active ? 0.0 : a.ke_ref ? (double)ref_kinetic(px, py, pz, pv.x, pv.y, pv.z, m, rp.dt, L, half) : ke;
    } else if (active) {
        float qx = pv.x, qy = pv.y, qz = pv.z;
        hyperpos(me.x, me.y, me.z, qx, qy, qz, L, half);
        float dx = me.x - qx, dy = me.y - qy, dz = me.z - qz;
        ke = m * ((double)dx * dx + (double)dy * dy + (double)dz * dz);
    }
    // kinetic-energy partials of this compound (FP64, fixed order)
    if (ADVANCE) {
        double ks = wave_sum(kt);
        __syncthreads();
        if ((p & 63) == 0) red[p >> 6] = ks;
        __syncthreads();
        if (p == 0) a.ke_comp[c] = red[0] + (blockDim.x > 64 ? red[1] : 0.0);
        thermostat_if_last(a, r);
    }
    if (ENERGY) {
        double ks = wave_sum(ke);
        __syncthreads();
        if ((p & 63) == 0) red[p >> 6] = ks;
        __syncthreads();
        if (p == 0) atomicAdd(a.energy + r * kEnergyTerms + kE_kin, red[0] + (blockDim.x > 64 ? red[1] : 0.0));
    }
}

// ---------------------------------------------------------------------------
// solventForceKernel's listed part (Engine.cu:1111, 1121-1123, 706-720): the
// solvent's own position is re-imaged against every compound's particle 0 in
// compound order, then against every listed neighbour in list order, as the
// reference mutates it; the solvent-solvent LJ local sum and the final self
// image are written for the integrator.  One thread per solvent; it runs
// concurrently with the all-pairs kernel.
//
// One wavefront per block, 64 solvents: work item = (solvent tile, quarter).  The
// lists are rows of ss_ld ids (a multiple of kListChunk).  The wavefront reads
// the rows of its 64 solvents kListedChunk entries at a time, four rows per load
// instruction (one 64-byte segment each), and transposes them through LDS; each
// lane then walks its row from LDS.  Every list byte is read from HBM once, in
// whole segments, which leaves L2 to the position gathers.  Items are taken in
// XCD-contiguous order (xcd_item), so the tiles of one replica share one L2.
// ---------------------------------------------------------------------------
constexpr int kListChunk = 16;                 // row length granule of the solvent lists
constexpr int kListedChunk = 16;               // list entries per lane per segment of the listed kernel
constexpr int kXcds = 8;

// Workgroups are dispatched round-robin over the XCDs (block b runs on XCD
// b % 8).  With the grid padded to a multiple of 8, XCD x takes the contiguous
// item range [x * per, (x + 1) * per).
LIMA_DEV int xcd_item(int b, int nblocks) {
    const int per = nblocks / kXcds;
    return (b % kXcds) * per + b / kXcds;
}
inline int xcd_grid(int n) { return (n + kXcds - 1) / kXcds * kXcds; }

constexpr int kListedItems = 4;   // wavefront items per 256-solvent tile

template <bool ENERGY>
__global__ __launch_bounds__(64) void solvent_listed_kernel(StepArgs a, int n_items) {
    constexpr int CH = kListedChunk;          // list entries per lane per segment
    constexpr int RPL = 64 / CH;              // rows per load instruction
    constexpr int STR = CH + 4;               // LDS row stride in ints (16-byte aligned rows)
    static_assert(CH % 4 == 0 && 64 % CH == 0 && kListChunk % CH == 0, "segments tile the rows");
    // the wavefront's list stage; the replay's particle-0 stage aliases it (used before the lists)
    constexpr int kStageInts = 64 * STR > kReplayStage * 4 ? 64 * STR : kReplayStage * 4;
    __shared__ __align__(16) int stage[kStageInts];
    const int item = xcd_item(blockIdx.x, gridDim.x);
    if (item >= n_items) return;
    const int4 tile = a.stiles[item / kListedItems];
    const int wrow0 = tile.y + (item % kListedItems) * 64;   // the wavefront's first solvent
    if (wrow0 >= tile.z) return;
    const int r = tile.x;
    const ReplicaInfo rp = a.reps[r];
    const float L = rp.L, half = rp.half;
    const int s = wrow0 + threadIdx.x;
    const bool active = s < tile.z;
    const int si = active ? s : tile.y;
    const float4 me = a.spos[si];
    float sx = me.x, sy = me.y, sz = me.z;
    replay_images<1>(a, rp.comp_begin, rp.comp_end, reinterpret_cast<float4*>(stage), &sx, &sy, &sz, L, half);   // Engine.cu:1111
    __syncthreads();   // the stage is reused for the lists below
    const float4 pv = a.tab_ss[r];
    const float2 pss = lj_ab(pv.x, pv.y);
    float lx = 0.f, ly = 0.f, lz = 0.f;
    double en = 0, vir = 0;
    const int cnt = active ? a.ss_cnt[si] : 0;
    const int lane = threadIdx.x;
    int* st = stage;                            // rows beyond tile.z have count 0
    int wmax = cnt;
    for (int off = 32; off > 0; off >>= 1) wmax = max(wmax, __shfl_xor(wmax, off, 64));
    const int sub = lane / CH, col = lane % CH;
    // this lane loads entry k0 + col of rows RPL i + sub (i < CH); the list buffer has 256
    // spare rows, so rows past the tile (count 0) stay inside it
    int rcs[CH];
#pragma unroll
    for (int i = 0; i < CH; i++) rcs[i] = __shfl(cnt, RPL * i + sub, 64);
    // 32-bit element offsets (the buffer is under 2^32 bytes, checked at creation) keep one
    // offset register per load instead of a 64-bit address each
    const unsigned rbase = (unsigned)(wrow0 + sub) * (unsigned)a.ss_ld;
    const unsigned rstep = (unsigned)RPL * (unsigned)a.ss_ld;
    // cooperative, coalesced read of 64 row segments, all loads in flight at once; a row
    // already finished re-reads its entry 0 (no new line) instead of its unused tail.  The
    // next segment's ids are loaded while the current one is summed.
    int v[CH];
    auto load_ids = [&](int k0) {
#pragma unroll
        for (int i = 0; i < CH; i++)
            v[i] = a.ss_ids[rbase + (unsigned)i * rstep + (unsigned)(k0 + col < rcs[i] ? k0 + col : 0)];
    };
    if (wmax > 0) load_ids(0);
    for (int k0 = 0; k0 < wmax; k0 += CH) {
#pragma unroll
        for (int i = 0; i < CH; i++) st[(RPL * i + sub) * STR + col] = v[i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int n = min(CH, cnt - k0);   // this lane's entries in the segment (may be <= 0)
        const int* mine = st + lane * STR;
        // all of the segment's neighbour positions in flight at once; slots past the row's
        // count read the solvent's own position and are not summed
        float4 q[CH];
#pragma unroll
        for (int u = 0; u < CH; u += 4) {
            const int4 i4 = *reinterpret_cast<const int4*>(mine + u);
            q[u + 0] = a.spos[u + 0 < n ? i4.x : si];
            q[u + 1] = a.spos[u + 1 < n ? i4.y : si];
            q[u + 2] = a.spos[u + 2 < n ? i4.z : si];
            q[u + 3] = a.spos[u + 3 < n ? i4.w : si];
        }
        // The next segment's ids are requested after the positions.  The wait counter
        // retires loads in issue order, so a wait for a position (an L2 hit) then does not
        // also wait for the ids (HBM); they have the whole segment's sums to arrive.  The
        // request is unconditional (past the last segment every row re-reads its entry 0):
        // behind a branch, the waits after it would be counted for the path without it.
        __builtin_amdgcn_sched_barrier(0);
        load_ids(k0 + CH);
        __builtin_amdgcn_sched_barrier(0);
        if (__ballot(!(active && n >= CH)) == 0ull) {
            // every lane has a full segment: no per-pair branches, so no load is moved
            // into one (where its wait would cover every load in flight)
#pragma unroll
            for (int u = 0; u < CH; u++) {
                hyperpos(q[u].x, q[u].y, q[u].z, sx, sy, sz, L, half);
                if (ENERGY) lj_acc_energy(sx, sy, sz, q[u].x, q[u].y, q[u].z, pss.x, pss.y, lx, ly, lz, en, vir);
                else lj_acc(sx, sy, sz, q[u].x, q[u].y, q[u].z, pss.x, pss.y, lx, ly, lz);
            }
        } else {
            // a row ends in this segment: every slot is evaluated and a slot past the row's
            // end keeps the lane's image and sums (selects, the same values as skipping it)
#pragma unroll
            for (int u = 0; u < CH; u++) {
                const bool take = u < n;
                float mx = sx, my = sy, mz = sz, fx = lx, fy = ly, fz = lz;
                double e = en, w = vir;
                hyperpos(q[u].x, q[u].y, q[u].z, mx, my, mz, L, half);
                if (ENERGY) lj_acc_energy(mx, my, mz, q[u].x, q[u].y, q[u].z, pss.x, pss.y, fx, fy, fz, e, w);
                else lj_acc(mx, my, mz, q[u].x, q[u].y, q[u].z, pss.x, pss.y, fx, fy, fz);
                sx = take ? mx : sx;
                sy = take ? my : sy;
                sz = take ? mz : sz;
                lx = take ? fx : lx;
                ly = take ? fy : ly;
                lz = take ? fz : lz;
                if (ENERGY) {
                    en = take ? e : en;
                    vir = take ? w : vir;
                }
            }
        }
        // every lane is done with the stage before the next segment overwrites it
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (active) {
        a.ss_part[s] = make_float4(lx, ly, lz, 0.f);
        a.ss_img[s] = make_float4(sx, sy, sz, 0.f);
    }
    if (ENERGY) {
        block_atomic_energy(en, a.energy + r * kEnergyTerms + kE_ss);
        block_atomic_energy(vir, a.energy + r * kEnergyTerms + kE_virial);
    }
}

// The same listed sum with four lanes per solvent, for systems too small to fill
// the device (one wavefront per 16 solvents).  Every lane of a solvent's group
// replays the whole re-imaging sequence (it is sequential by definition), and lane
// g evaluates the pairs 4b + g, each with the image after its own entry; the four
// partial sums are added in lane order.  The images are the reference's; the float
// sums are grouped differently, so this form serves FP32 accumulation only.
constexpr int kSplitLanes = 4;
constexpr int kSplitItems = 256 / (64 / kSplitLanes);   // wavefront items per 256-solvent tile

template <bool ENERGY>
__global__ __launch_bounds__(64) void solvent_listed_split_kernel(StepArgs a, int n_items) {
    __shared__ __align__(16) int stage[kReplayStage * 4];   // the replay's particle-0 stage, then ids
    const int item = xcd_item(blockIdx.x, gridDim.x);
    if (item >= n_items) return;
    const int4 tile = a.stiles[item / kSplitItems];
    const int lane = threadIdx.x, grp = lane / kSplitLanes, g = lane % kSplitLanes;
    const int wrow0 = tile.y + (item % kSplitItems) * (64 / kSplitLanes);
    if (wrow0 >= tile.z) return;
    const int r = tile.x;
    const ReplicaInfo rp = a.reps[r];
    const float L = rp.L, half = rp.half;
    const int s = wrow0 + grp;
    const bool active = s < tile.z;
    const int si = active ? s : tile.y;
    const float4 me = a.spos[si];
    float sx = me.x, sy = me.y, sz = me.z;
    replay_images<1>(a, rp.comp_begin, rp.comp_end, reinterpret_cast<float4*>(stage), &sx, &sy, &sz, L, half);   // Engine.cu:1111
    __syncthreads();
    const float4 pv = a.tab_ss[r];
    const float2 pss = lj_ab(pv.x, pv.y);
    float lx = 0.f, ly = 0.f, lz = 0.f;
    double en = 0, vir = 0;
    const int cnt = active ? a.ss_cnt[si] : 0;
    int wmax = cnt;
    for (int off = 32; off > 0; off >>= 1) wmax = max(wmax, __shfl_xor(wmax, off, 64));
    const int* row = a.ss_ids + (size_t)si * a.ss_ld;
    int4* st4 = reinterpret_cast<int4*>(stage);
    auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    // lane g of a group reads entries k0 + 4g .. k0 + 4g + 3 of the row (16-byte aligned: the rows
    // are multiples of 16 ids); past the row's count it re-reads entries 4g .. 4g + 3 (in bounds,
    // not used).  The next segment's ids are requested after the current segment's positions, so
    // the waits for those (L2) do not also wait for the ids (HBM).
    auto load_ids = [&](int k0) {
        return *reinterpret_cast<const int4*>(row + (k0 < cnt ? k0 : 0) + 4 * g);
    };
    int4 nxt = wmax > 0 ? load_ids(0) : make_int4(0, 0, 0, 0);
    for (int k0 = 0; k0 < wmax; k0 += 16) {
        st4[lane] = k0 < cnt ? nxt : make_int4(si, si, si, si);
        wave_sync();
        const int n = min(16, cnt - k0);
        float4 q[16];
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int4 i4 = st4[grp * kSplitLanes + b];
            q[4 * b + 0] = a.spos[4 * b + 0 < n ? i4.x : si];
            q[4 * b + 1] = a.spos[4 * b + 1 < n ? i4.y : si];
            q[4 * b + 2] = a.spos[4 * b + 2 < n ? i4.z : si];
            q[4 * b + 3] = a.spos[4 * b + 3 < n ? i4.w : si];
        }
        __builtin_amdgcn_sched_barrier(0);
        nxt = load_ids(k0 + 16);
        __builtin_amdgcn_sched_barrier(0);
        wave_sync();   // the stage is rewritten by the next segment
        // FULL: every lane's row has the whole segment (no per-entry conditions, so no load is
        // moved into a branch); otherwise an entry past the row's end keeps the image (select)
        auto sums = [&](auto full_tag) {
            constexpr bool FULL = decltype(full_tag)::value;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                float ex = sx, ey = sy, ez = sz, qx = 0.f, qy = 0.f, qz = 0.f;
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int u = 4 * b + t;
                    if (FULL) {
                        hyperpos(q[u].x, q[u].y, q[u].z, sx, sy, sz, L, half);
                    } else {
                        float mx = sx, my = sy, mz = sz;
                        hyperpos(q[u].x, q[u].y, q[u].z, mx, my, mz, L, half);
                        sx = u < n ? mx : sx; sy = u < n ? my : sy; sz = u < n ? mz : sz;
                    }
                    const bool mine = t == g;   // this lane's pair: the image after its own entry
                    ex = mine ? sx : ex; ey = mine ? sy : ey; ez = mine ? sz : ez;
                    qx = mine ? q[u].x : qx; qy = mine ? q[u].y : qy; qz = mine ? q[u].z : qz;
                }
                if (FULL || 4 * b + g < n) {
                    if (ENERGY) lj_acc_energy(ex, ey, ez, qx, qy, qz, pss.x, pss.y, lx, ly, lz, en, vir);
                    else lj_acc(ex, ey, ez, qx, qy, qz, pss.x, pss.y, lx, ly, lz);
                }
            }
        };
        if (__ballot(!(active && n >= 16)) == 0ull) sums(std::true_type{});
        else sums(std::false_type{});
    }
    // the group's partial sums in lane order
    const int g0 = grp * kSplitLanes;
    float tx = __shfl(lx, g0, 64), ty = __shfl(ly, g0, 64), tz = __shfl(lz, g0, 64);
#pragma unroll
    for (int j = 1; j < kSplitLanes; j++) {
        tx += __shfl(lx, g0 + j, 64); ty += __shfl(ly, g0 + j, 64); tz += __shfl(lz, g0 + j, 64);
    }
    if (active && g == 0) {
        a.ss_part[s] = make_float4(tx, ty, tz, 0.f);
        a.ss_img[s] = make_float4(sx, sy, sz, 0.f);
    }
    if (ENERGY) {
        block_atomic_energy(en, a.energy + r * kEnergyTerms + kE_ss);
        block_atomic_energy(vir, a.energy + r * kEnergyTerms + kE_virial);
    }
}

// ---------------------------------------------------------------------------
// solventForceKernel remainder (Engine.cu:1129-1155): force = compound
// contributions (chunk partials in order) + the solvent-solvent local sum;
// integrate, PBC, NaN check.  Tiles never straddle replicas.
// ---------------------------------------------------------------------------
template <bool ADVANCE, bool FIXED, bool ENERGY>
__global__ __launch_bounds__(256) void solvent_integrate_kernel(StepArgs a) {
    __shared__ double red[4];
    const int4 tile = a.stiles[blockIdx.x];
    const int r = tile.x;
    const ReplicaInfo rp = a.reps[r];
    const float L = rp.L, half = rp.half;
    const int s = tile.y + threadIdx.x;
    const bool active = s < tile.z;
    const int si = active ? s : tile.y;
    const float4 me = a.spos[si];
    const float4 img = a.ss_img[si];
    float sx = img.x, sy = img.y, sz = img.z;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    long long ix = 0, iy = 0, iz = 0;
    // eight chunks' loads in flight at once, then added in chunk order
    for (int k = 0; k < rp.chunk_count; k += 8) {
        if (FIXED) {
            I3 v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = k + u < rp.chunk_count ? a.partS_fx[(size_t)(k + u) * a.n_solv + si] : I3{0, 0, 0};
#pragma unroll
            for (int u = 0; u < 8; u++) { ix += v[u].x; iy += v[u].y; iz += v[u].z; }
        } else {
            float4 v[8];
#pragma unroll
            for (int u = 0; u < 8; u++)
                v[u] = k + u < rp.chunk_count ? a.partS[(size_t)(k + u) * a.n_solv + si] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (k + u < rp.chunk_count) { fx += v[u].x; fy += v[u].y; fz += v[u].z; }
        }
    }
    const float4 l = a.ss_part[si];
    if (FIXED) {
        uint32_t bad = 0;
        ix += to_fixed(l.x, bad); iy += to_fixed(l.y, bad); iz += to_fixed(l.z, bad);
        if (bad) atomicOr(a.err, bad);
        fx = from_fixed(ix); fy = from_fixed(iy); fz = from_fixed(iz);
    } else {
        fx += l.x; fy += l.y; fz += l.z;
    }
    if (active) a.sforce[s] = make_float4(fx, fy, fz, 0.f);
    const double m = (double)a.mass[r * kMaxTypes + 0];
    double ke = 0.0, kt = 0.0;   // one-step m|dx|^2; what the thermostat sums
    float4 pv = a.sprev[si];
    if (ADVANCE) {
        float qx = pv.x, qy = pv.y, qz = pv.z, dx, dy, dz;
        const double mts = rp.dt / m * rp.dt;
        integrate(sx, sy, sz, qx, qy, qz, fx, fy, fz, mts, a.lambda[r], L, half, dx, dy, dz);
        sx = pbc1(sx, L); sy = pbc1(sy, L); sz = pbc1(sz, L);
        if (active) {
            if (sx != sx) atomicOr(a.err, kErrNaN);                        // Engine.cu:1151
            a.spos_next[s] = make_float4(sx, sy, sz, 0.f);
            a.sprev[s] = make_float4(qx, qy, qz, 0.f);
            ke = m * ((double)dx * dx + (double)dy * dy + (double)dz * dz);
            // the logged position of this step against the one two steps back (Engine.cu:347-351)
            kt = a.ke_ref ? (double)ref_kinetic(sx, sy, sz, pv.x, pv.y, pv.z, m, rp.dt, L, half) : ke;
        }
    } else if (active) {
        float qx = pv.x, qy = pv.y, qz = pv.z;
        hyperpos(me.x, me.y, me.z, qx, qy, qz, L, half);
        float dx = me.x - qx, dy = me.y - qy, dz = me.z - qz;
        ke = m * ((double)dx * dx + (double)dy * dy + (double)dz * dz);
    }
    if (ADVANCE) {
        double ks = wave_sum(kt);
        __syncthreads();
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ks;
        __syncthreads();
        if (threadIdx.x == 0) {
            double tot = 0;
            for (int w = 0; w < (int)blockDim.x / 64; w++) tot += red[w];
            a.ke_stile[tile.w] = tot;
        }
        if (a.thermo_in_solv) thermostat_if_last(a, r);
    }
    if (ENERGY) {
        double ks = wave_sum(ke);
        __syncthreads();
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ks;
        __syncthreads();
        if (threadIdx.x == 0) {
            double tot = 0;
            for (int w = 0; w < (int)blockDim.x / 64; w++) tot += red[w];
            atomicAdd(a.energy + r * kEnergyTerms + kE_kin, tot);
        }
    }
}

// sets every replica's step counter (graph-safe: a kernel, not a memset)
__global__ void set_step_count_kernel(long long* step_count, int n_rep, long long v) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n_rep) step_count[r] = v;
}

// ===========================================================================
// Neighbour lists (Engine::updateNeighborLists, Engine.cu:208-287).
// The reference's incremental cull/add with hash dedupe yields, per pair:
//   compound-compound (lo < hi): member = (was && d < C+((r_lo+r_hi)+C)) || d < C+((r_lo+r_lo)+C)
//                                 (the add threshold reads the lower id's radius twice, Engine.cu:243)
//   compound-solvent:            member = d < C + (r_c + C)
//   solvent-solvent (lo < hi):   member = d < C + C
// with d = |a - hyperpos(a, b)| in FP32, a = the lower id (or the compound key).
// Lists are produced in ascending id order (compound lists) or cell-scan order
// (solvent lists); the sets are bit-exact, the order is this build's own.
// ===========================================================================
struct NlistArgs {
    const ReplicaInfo* reps;
    const float4* cpos;
    const float4* spos;
    int n_comp, n_solv;
    const int* comp_off; const int* comp_rep; const int* comp_key; const float* comp_sphere;
    const int* solv_rep;
    float cutoff;
    int cap_cc, cap_cs, cap_ss;
    int* cc_cnt; int* cc_ids;
    int* cs_cnt; int* cs_ids;
    int* ss_cnt; int* ss_ids;   // rows of ss_ld
    int ss_ld;
    int n_cells;
    int* cell_cnt; int* cell_start; int* cell_cursor; int* cell_ids; int* solv_cell;
    const int* cell_rep;          // replica of each cell
    const float* comp_cs_bound;   // per compound: bound(C + (r + C)) for the squared distance
    int* seg_cnt;                 // [n_comp * max_segs] compound-solvent members per solvent segment
    int max_segs;
    int* cw_off; int* cw_total;
    int* item_comp;
    int run_len;                  // neighbour items per run of the packed listed kernel (0: not used)
    int* run_off; int* run_cs_off; int* run_total; int* run_comp;   // run_total[2]: neighbour runs, solvent chunks
    uint32_t* err;
    int enable_solvents;
};

LIMA_DEV float4 key_pos(const NlistArgs& a, int c) { return a.cpos[a.comp_off[c] + a.comp_key[c]]; }

// One wavefront per compound, candidates of its replica in ascending id order.
// The compound's previous list is first turned into an LDS bitmask (it is the
// only reader of its own list), then overwritten with the new one.
constexpr int kMaxReplicaCompounds = 8192;

__global__ __launch_bounds__(64) void nl_compound_compound_kernel(NlistArgs a) {
    __shared__ uint32_t was_bits[kMaxReplicaCompounds / 32];
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const int r = a.comp_rep[i];
    const ReplicaInfo rp = a.reps[r];
    const int ncr = rp.comp_end - rp.comp_begin;
    if (ncr > kMaxReplicaCompounds) {
        if (lane == 0) atomicOr(a.err, kErrCapacity);
        return;
    }
    const float L = rp.L, half = rp.half, C = a.cutoff;
    for (int w = lane; w < (ncr + 31) / 32; w += 64) was_bits[w] = 0u;
    __syncthreads();
    const int old_cnt = a.cc_cnt[i];
    for (int e = lane; e < old_cnt; e += 64) {
        const int j = a.cc_ids[(size_t)i * a.cap_cc + e] - rp.comp_begin;
        atomicOr(&was_bits[j >> 5], 1u << (j & 31));
    }
    __syncthreads();
    const float4 ki = key_pos(a, i);
    int count = 0;
    for (int j0 = rp.comp_begin; j0 < rp.comp_end; j0 += 64) {
        const int j = j0 + lane;
        bool member = false;
        if (j < rp.comp_end && j != i) {
            const int lo = min(i, j), hi = max(i, j);
            const float4 klo = lo == i ? ki : key_pos(a, j);
            const float4 khi = hi == i ? ki : key_pos(a, j);
            const float rlo = a.comp_sphere[lo], rhi = a.comp_sphere[hi];
            const int jl = j - rp.comp_begin;
            const bool was = (was_bits[jl >> 5] >> (jl & 31)) & 1u;
            const bool keep = within(klo.x, klo.y, klo.z, khi.x, khi.y, khi.z, fadd(fadd(rlo, rhi), C), C, L, half);
            const bool add = within(klo.x, klo.y, klo.z, khi.x, khi.y, khi.z, fadd(fadd(rlo, rlo), C), C, L, half);
            member = (was && keep) || add;
        }
        const unsigned long long ball = __ballot(member);
        const int before = __popcll(ball & ((1ull << lane) - 1ull));
        if (member) {
            const int pos = count + before;
            if (pos < a.cap_cc) a.cc_ids[(size_t)i * a.cap_cc + pos] = j;
        }
        count += __popcll(ball);
    }
    if (lane == 0) {
        if (count > a.cap_cc) { atomicOr(a.err, kErrCapacity); atomicMax(a.err + 1, count); }
        a.cc_cnt[i] = min(count, a.cap_cc);
    }
}

// Compound -> solvent lists in ascending solvent order, in two passes over a
// (compound, 2048-solvent segment) grid: count members per segment, then write
// each segment's members at the prefix of the counts before it.
constexpr int kSegSolvents = 2048;

LIMA_DEV bool cs_member(const NlistArgs& a, float4 ki, float bound, int s, float L, float half) {
    const float4 q = a.spos[s];
    // |d| < C + (r + C)  <=>  d^2 < bound (correctly rounded sqrt is monotone)
    return dist2_reimaged(ki.x, ki.y, ki.z, q.x, q.y, q.z, L, half) < bound;
}

__global__ __launch_bounds__(256) void nl_cs_count_kernel(NlistArgs a) {
    __shared__ int wsum[4];
    const int i = blockIdx.x, seg = blockIdx.y;
    const ReplicaInfo rp = a.reps[a.comp_rep[i]];
    const int s0 = rp.solv_begin + seg * kSegSolvents;
    const int s1 = min(s0 + kSegSolvents, rp.solv_end);
    const float4 ki = key_pos(a, i);
    const float bound = a.comp_cs_bound[i];
    int c = 0;
    for (int s = s0 + (int)threadIdx.x; s < s1; s += blockDim.x) c += cs_member(a, ki, bound, s, rp.L, rp.half);
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) a.seg_cnt[(size_t)i * a.max_segs + seg] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ __launch_bounds__(256) void nl_cs_write_kernel(NlistArgs a) {
    __shared__ int wave_cnt[4];
    const int i = blockIdx.x, seg = blockIdx.y;
    const ReplicaInfo rp = a.reps[a.comp_rep[i]];
    const int nseg = (rp.solv_end - rp.solv_begin + kSegSolvents - 1) / kSegSolvents;
    const int* sc = a.seg_cnt + (size_t)i * a.max_segs;
    if (seg == 0 && threadIdx.x == 0) {
        int total = 0;
        for (int k = 0; k < nseg; k++) total += sc[k];
        if (total > a.cap_cs) { atomicOr(a.err, kErrCapacity); atomicMax(a.err + 2, total); }
        a.cs_cnt[i] = min(total, a.cap_cs);
    }
    if (seg >= nseg) return;
    int count = 0;
    for (int k = 0; k < seg; k++) count += sc[k];
    const int s0 = rp.solv_begin + seg * kSegSolvents;
    const int s1 = min(s0 + kSegSolvents, rp.solv_end);
    const float4 ki = key_pos(a, i);
    const float bound = a.comp_cs_bound[i];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int b = s0; b < s1; b += 256) {
        const int s = b + threadIdx.x;
        const bool member = s < s1 && cs_member(a, ki, bound, s, rp.L, rp.half);
        const unsigned long long ball = __ballot(member);
        if (lane == 0) wave_cnt[wid] = __popcll(ball);
        __syncthreads();
        int before = __popcll(ball & ((1ull << lane) - 1ull));
        for (int w = 0; w < wid; w++) before += wave_cnt[w];
        if (member && count + before < a.cap_cs) a.cs_ids[(size_t)i * a.cap_cs + count + before] = s;
        count += wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
        __syncthreads();
    }
}

LIMA_DEV int cell_coord(float v, float inv, int n) {
    int c = (int)(v * inv);
    c = c < 0 ? 0 : c;
    return c >= n ? n - 1 : c;
}

__global__ void nl_cell_zero_kernel(int* cell_cnt, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cell_cnt[i] = 0;
}

// Wavefront-aggregated counter update: the lanes of each distinct cell in the wavefront add their
// number with one atomic, made by the cell's first lane; returns each lane's slot (the counter's
// old value plus the lane's rank among the cell's lanes).  Consecutive solvents mostly share a
// few cells, so the atomics on one counter drop from one per solvent to one per wavefront.
LIMA_DEV int wave_cell_add(int* counter, int cell, bool valid) {
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long todo = __ballot(valid);
    int slot = 0;
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int lc = __shfl(cell, leader);
        const unsigned long long same = __ballot(valid && cell == lc) & todo;
        int base = 0;
        if (lane == leader) base = atomicAdd(counter + lc, __popcll(same));
        base = __shfl(base, leader);
        if ((same >> lane) & 1ull) slot = base + __popcll(same & below);
        todo &= ~same;
    }
    return slot;
}

__global__ void nl_cell_count_kernel(NlistArgs a) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = s < a.n_solv;
    int cell = 0;
    if (valid) {
        const ReplicaInfo rp = a.reps[a.solv_rep[s]];
        const float4 q = a.spos[s];
        const int n = rp.ncell;
        const int cx = cell_coord(q.x, rp.cell_inv, n), cy = cell_coord(q.y, rp.cell_inv, n), cz = cell_coord(q.z, rp.cell_inv, n);
        cell = rp.cell_begin + (cx * n + cy) * n + cz;
        a.solv_cell[s] = cell;
    }
    wave_cell_add(a.cell_cnt, cell, valid);
}

// exclusive scan of cell counts (single block of 1024)
__global__ __launch_bounds__(1024) void nl_cell_scan_kernel(NlistArgs a) {
    __shared__ int part[1024];
    const int n = a.n_cells;
    const int per = (n + 1023) / 1024;
    const int b = threadIdx.x * per;
    int s = 0;
    for (int k = 0; k < per && b + k < n; k++) s += a.cell_cnt[b + k];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = threadIdx.x ? part[threadIdx.x - 1] : 0;
    for (int k = 0; k < per && b + k < n; k++) {
        a.cell_start[b + k] = run;
        a.cell_cursor[b + k] = run;
        run += a.cell_cnt[b + k];
    }
    if (threadIdx.x == 1023) a.cell_start[n] = part[1023];
}

__global__ void nl_cell_scatter_kernel(NlistArgs a) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = s < a.n_solv;
    const int cell = valid ? a.solv_cell[s] : 0;
    const int pos = wave_cell_add(a.cell_cursor, cell, valid);
    if (valid) a.cell_ids[pos] = s;
}

// sort each cell's ids ascending (rank sort in LDS), making the lists deterministic
__global__ __launch_bounds__(256) void nl_cell_sort_kernel(NlistArgs a) {
    constexpr int CAP = 2048;
    __shared__ int ids[CAP];
    const int cell = blockIdx.x;
    const int b = a.cell_start[cell], n = a.cell_start[cell + 1] - b;
    if (n > CAP) {
        if (threadIdx.x == 0) { atomicOr(a.err, kErrCapacity); atomicMax(a.err + 3, n); }
        return;
    }
    for (int k = threadIdx.x; k < n; k += blockDim.x) ids[k] = a.cell_ids[b + k];
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int v = ids[k];
        int rank = 0;
        for (int e = 0; e < n; e++) rank += ids[e] < v;
        a.cell_ids[b + rank] = v;
    }
}

// One block per home cell: the solvents of the (deduplicated) 27 neighbouring
// cells are staged in LDS in a fixed cell order (each cell sorted by id); every
// wavefront tests one home solvent against all candidates.  The pair predicate is
// evaluated with the lower id as the static point, as the reference's add loop
// does (Engine.cu:262-264), and compared as d^2 < bound(C + C).
// 3456 staged candidates (54 KB of positions, 13.5 KB of ids, 8 KB of queues)
// let two blocks share a CU.
constexpr int kCellCandidates = 3456;
constexpr int kSsQueue = 256;   // per-wavefront ring of prefiltered candidate slots
static_assert(kCellCandidates <= 65535 && kCellCandidates % 128 == 0, "16-bit slots, whole passes of 128");

LIMA_DEV bool ss_member(int s, float4 me, int j, float4 q, float L, float half, float bound) {
    const bool lo_self = s < j;
    const float ax = lo_self ? me.x : q.x, ay = lo_self ? me.y : q.y, az = lo_self ? me.z : q.z;
    const float bx = lo_self ? q.x : me.x, by = lo_self ? q.y : me.y, bz = lo_self ? q.z : me.z;
    return (j != s) & (dist2_reimaged(ax, ay, az, bx, by, bz, L, half) < bound);
}

// 1024 threads: each block carries 16 waves to hide LDS and ballot latency, and
// two blocks fit one CU
constexpr int kSolventListThreads = 1024;

__global__ __launch_bounds__(kSolventListThreads) void nl_solvent_solvent_kernel(NlistArgs a) {
    __shared__ float4 cand_pos[kCellCandidates];
    __shared__ int cand_id[kCellCandidates];
    __shared__ unsigned short queue[kSolventListThreads / 64][kSsQueue];
    __shared__ int cell_b[27], cell_n[27], cell_off[28];
    // cells of one replica are contiguous: XCD-contiguous order keeps a replica's positions in one L2
    const int cell = xcd_item(blockIdx.x, gridDim.x);
    if (cell >= a.n_cells) return;
    const int h0 = a.cell_start[cell], h1 = a.cell_start[cell + 1];
    if (h0 == h1) return;
    const ReplicaInfo rp = a.reps[a.cell_rep[cell]];
    const float L = rp.L, half = rp.half, bound = rp.ss_bound;
    const int n = rp.ncell;
    const int local = cell - rp.cell_begin;
    const int cx = local / (n * n), cy = (local / n) % n, cz = local % n;
    const int span = n >= 3 ? 3 : n;
    const int nnb = span * span * span;
    // neighbour cell k = (ox, oy, oz) in a fixed order; grids narrower than 3 cells take every cell once
    auto nb_cell = [&](int k) {
        const int ox = k / (span * span), oy = (k / span) % span, oz = k % span;
        const int x = (n >= 3) ? (cx - 1 + ox + n) % n : ox;
        const int y = (n >= 3) ? (cy - 1 + oy + n) % n : oy;
        const int z = (n >= 3) ? (cz - 1 + oz + n) % n : oz;
        return rp.cell_begin + (x * n + y) * n + z;
    };
    // all cell ranges at once, then an exclusive scan of their sizes
    if (threadIdx.x < nnb) {
        const int c = nb_cell(threadIdx.x);
        const int b = a.cell_start[c];
        cell_b[threadIdx.x] = b;
        cell_n[threadIdx.x] = a.cell_start[c + 1] - b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int k = 0; k < nnb; k++) {
            cell_off[k] = acc;
            acc += cell_n[k];
        }
        cell_off[nnb] = acc;
    }
    __syncthreads();
    const int total = cell_off[nnb];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    // With 3 or more cells per side each neighbour cell has one image next to the home cell, and
    // its candidates are staged shifted into it, so that the prefilter below is a plain distance
    // (the exact test reads the raw positions); narrower grids stage them raw.
    const bool shifted = n >= 3;
    const unsigned long long below = (1ull << lane) - 1ull;
    const int nhome = h1 - h0;
    const float rb = sqrtf(bound) + 1e-5f * L + 1e-6f;
    const float pre = rb * rb;
    // below this the cheap distance certifies membership: per axis it differs from the exactly
    // rounded re-imaged difference by a few roundings of coordinates up to 2L, far below the margin
    const float ri = sqrtf(bound) - (1e-5f * L + 1e-6f);
    const float inner = ri * ri;
    unsigned short* q = queue[wid];
    // The candidates, in the fixed cell order, pass through LDS kCellCandidates at a time; each
    // home solvent takes every chunk in turn, so its row is in candidate order whatever the chunks.
    // Between chunks the row's running count (not yet clamped to the capacity) waits in ss_cnt.
    for (int c0 = 0; c0 < total; c0 += kCellCandidates) {
        const int c1 = min(total, c0 + kCellCandidates);
        const bool last = c1 == total;
        if (c0 > 0) {
            // every wavefront is done with the previous chunk, and the running counts it stored
            // are visible to the loads below
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __syncthreads();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        // stage: wavefront w copies the parts of cells w, w + nw, ... that fall in [c0, c1)
        for (int k = wid; k < nnb; k += nw) {
            const int b = cell_b[k], o = cell_off[k], m = cell_n[k];
            const int e0 = max(0, c0 - o), e1 = min(m, c1 - o);
            if (e0 >= e1) continue;
            float sx = 0.f, sy = 0.f, sz = 0.f;
            if (shifted) {
                const int ux = cx - 1 + k / 9, uy = cy - 1 + (k / 3) % 3, uz = cz - 1 + k % 3;
                sx = ux < 0 ? -L : ux >= n ? L : 0.f;
                sy = uy < 0 ? -L : uy >= n ? L : 0.f;
                sz = uz < 0 ? -L : uz >= n ? L : 0.f;
            }
            for (int e = e0 + lane; e < e1; e += 64) {
                const int id = a.cell_ids[b + e];
                const float4 x = a.spos[id];
                cand_id[o + e - c0] = id;
                cand_pos[o + e - c0] = make_float4(x.x + sx, x.y + sy, x.z + sz, x.w);
            }
        }
        // pad to whole passes of 128 with NaN positions, which no distance test accepts
        const int nc = c1 - c0;
        const int padded = (nc + 127) & ~127;
        if (threadIdx.x < padded - nc) {
            const float nan = __builtin_nanf("");
            cand_pos[nc + threadIdx.x] = make_float4(nan, nan, nan, nan);
        }
        __syncthreads();
        // Each wavefront takes home solvents in turn.  A cheap distance (minimum image on grids
        // narrower than 3 cells), with a margin far above its rounding error against the exact
        // re-imaged one, drops the candidates that cannot be members; the rest enter the
        // wavefront's queue in candidate order.  Whenever 64 are queued, the exact predicate
        // runs on them and the members are appended in that order (ballot compaction), so each
        // list is the one a test of every candidate in order would give.
        for (int h = wid; h < nhome; h += nw) {
            const int s = a.cell_ids[h0 + h];
            const float4 me = a.spos[s];
            int* row = a.ss_ids + (size_t)s * a.ss_ld;
            int count = c0 == 0 ? 0 : a.ss_cnt[s];
            unsigned head = 0, tail = 0;   // wavefront-uniform queue cursors
            auto dist2 = [&](float4 c) {   // the cheap distance of a staged candidate
                float dx = c.x - me.x, dy = c.y - me.y, dz = c.z - me.z;
                if (!shifted) {
                    dx = fabsf(dx); dy = fabsf(dy); dz = fabsf(dz);
                    dx = fminf(dx, fabsf(L - dx)); dy = fminf(dy, fabsf(L - dy)); dz = fminf(dz, fabsf(L - dz));
                }
                return dx * dx + dy * dy + dz * dz;
            };
            auto drain = [&](int m) {      // membership of the m <= 64 oldest queued slots
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");   // other lanes' queue writes
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const bool has = lane < m;
                const int e = has ? (int)q[(head + lane) & (kSsQueue - 1)] : 0;
                const int j = cand_id[e];
                const float d2 = dist2(cand_pos[e]);
                bool mem = has & (j != s) & (d2 < inner);
                // the exact predicate only for slots near the bound (wavefront-uniform branch)
                const bool band = has & !mem & (j != s);
                if (__ballot(band) != 0ull)
                    mem = mem | (band && ss_member(s, me, j, shifted ? a.spos[j] : cand_pos[e], L, half, bound));
                const unsigned long long bm = __ballot(mem);
                const int pos = count + __popcll(bm & below);
                if (mem & (pos < a.cap_ss)) row[pos] = j;
                count += __popcll(bm);
                head += m;
            };
            auto near = [&](int e) { return dist2(cand_pos[e]) < pre; };
            for (int e0 = 0; e0 < padded; e0 += 128) {
                const int ea = e0 + lane, eb = ea + 64;
                const bool ma = near(ea), mb = near(eb);
                const unsigned long long ba = __ballot(ma), bb = __ballot(mb);
                const unsigned ta = tail + __popcll(ba);
                if (ma) q[(tail + __popcll(ba & below)) & (kSsQueue - 1)] = (unsigned short)ea;
                if (mb) q[(ta + __popcll(bb & below)) & (kSsQueue - 1)] = (unsigned short)eb;
                tail = ta + __popcll(bb);
                while (tail - head >= 64u) drain(64);   // at most 191 queued, 63 left
            }
            while (tail != head) drain((int)min(64u, tail - head));
            if (lane == 0) {
                if (!last) {
                    a.ss_cnt[s] = count;
                } else {
                    if (count > a.cap_ss) { atomicOr(a.err, kErrCapacity); atomicMax(a.err + 3, count); }
                    a.ss_cnt[s] = min(count, a.cap_ss);
                }
            }
        }
    }
}

// per compound work slots: listed neighbours + ceil(solvent list / 256); exclusive scan.  With
// run_len > 0 (the packed listed kernel), also the compound's runs: ceil(neighbours / run_len)
// + ceil(solvent list / 256), scanned into run_off.
// run_len > 0: the compound's runs of run_len neighbours; run_len == 0: its work items (one per
// listed neighbour, then one per solvent chunk); run_len < 0: its solvent chunks alone
LIMA_DEV int work_count(const NlistArgs& a, int c, int run_len) {
    const int chunks = a.enable_solvents ? (a.cs_cnt[c] + kSolventChunk - 1) / kSolventChunk : 0;
    if (run_len > 0) return (a.cc_cnt[c] + run_len - 1) / run_len;
    if (run_len < 0) return chunks;
    return a.cc_cnt[c] + chunks;
}

LIMA_DEV void block_exclusive_scan(int* part, const NlistArgs& a, int run_len, int* off, int* total) {
    const int n = a.n_comp;
    const int per = (n + 1023) / 1024;
    const int b = threadIdx.x * per;
    int s = 0;
    for (int k = 0; k < per && b + k < n; k++) s += work_count(a, b + k, run_len);
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int v = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = threadIdx.x ? part[threadIdx.x - 1] : 0;
    for (int k = 0; k < per && b + k < n; k++) {
        off[b + k] = run;
        run += work_count(a, b + k, run_len);
    }
    if (threadIdx.x == 1023) { off[n] = part[1023]; *total = part[1023]; }
    __syncthreads();
}

// With run_len > 0 (the packed listed kernel) the run list holds every compound's neighbour
// runs first (run_off, run_total[0] of them), then every compound's solvent chunks (run_cs_off,
// run_total[1]), so that the groups of a wavefront mostly take runs of one kind.
__global__ __launch_bounds__(1024) void nl_work_kernel(NlistArgs a) {
    __shared__ int part[1024];
    block_exclusive_scan(part, a, 0, a.cw_off, a.cw_total);
    if (a.run_len > 0) {
        block_exclusive_scan(part, a, a.run_len, a.run_off, a.run_total);
        block_exclusive_scan(part, a, -1, a.run_cs_off, a.run_total + 1);
    }
}

// the compound of every work item, and of every run (one block per compound)
__global__ __launch_bounds__(64) void nl_item_kernel(NlistArgs a) {
    const int c = blockIdx.x;
    for (int w = a.cw_off[c] + threadIdx.x; w < a.cw_off[c + 1]; w += blockDim.x) a.item_comp[w] = c;
    if (a.run_len > 0) {
        for (int w = a.run_off[c] + threadIdx.x; w < a.run_off[c + 1]; w += blockDim.x) a.run_comp[w] = c;
        const int cs0 = a.run_total[0];
        for (int w = a.run_cs_off[c] + threadIdx.x; w < a.run_cs_off[c + 1]; w += blockDim.x) a.run_comp[cs0 + w] = c;
    }
}

}  // namespace lima
