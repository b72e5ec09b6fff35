// engine.hip — host engine and C ABI (include/lima.h) of the gfx950 LIMA_MD step.
//
// Replaces the reference's Engine (Engine.cuh:109-208; Engine.cu:9-132,
// 400-451): the engine owns all device memory (hipMalloc, no managed memory),
// runs the whole step — forces, integration, PBC, thermostat and the scheduled
// neighbour-list rebuild — on one HIP stream, and replays each rebuild cycle as
// a captured hipGraph.  Errors are status codes plus a device error word read
// once per lima_step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/lima.h"
#include "adjoint.h"
#include "kernels.h"

using namespace lima;

// Staging of lima_step_observed_noalloc (an energy series, [records][systems][kEnergyTerms]
// doubles then [records][systems] floats of lambda): a file-scope device array, one copy per
// device, held until the process ends, so that the entry point calls no allocator.  In host
// code `scratch` only names the symbol; its device address comes from hipGetSymbolAddress,
// looked up once per device at engine creation (scratch_address).  Every call on a device
// shares its copy: calls that overlap on one device race.
constexpr size_t kObservedScratchBytes = (size_t)4 << 20;
static __device__ __align__(256) unsigned char scratch[kObservedScratchBytes];

namespace {

#define HIP_TRY(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            throw LimaError(LIMA_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
        }                                                                                     \
    } while (0)

struct LimaError {
    int code;
    std::string msg;
    LimaError(int c, std::string m) : code(c), msg(std::move(m)) {}
};

template <typename T> struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { free(); }
    void alloc(size_t count) {
        free();
        n = count;
        if (count) HIP_TRY(hipMalloc(&p, count * sizeof(T)));
    }
    void upload(const std::vector<T>& v) {
        alloc(std::max<size_t>(v.size(), 1));
        if (!v.empty()) HIP_TRY(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    void zero() {
        if (p) HIP_TRY(hipMemset(p, 0, n * sizeof(T)));
    }
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kScratchDevices] = {};

// the device address of `scratch` on the current device `device`; null past kScratchDevices
unsigned char* scratch_address(int device) {
    if (device < 0 || device >= kScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (rc != hipSuccess || !ptr)
            throw LimaError(LIMA_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc));
        g_scratch_ptr[device] = static_cast<unsigned char*>(ptr);
    }
    return g_scratch_ptr[device];
}

struct LaunchConfig {
    int spt = 2;                  // solvents per thread in solvent_chain_kernel (1, 2, 4)
    int chunk_particles = 1024;   // target chain particles per solvent<-chain work item
    int listed_grid = 0;          // grid of compound_listed_kernel (0 = auto)
    int use_graph = 1;
    int listed_lanes = 1;         // lanes per solvent of the listed solvent kernel (1, or 4 in FP32)
    int listed_packed = 1;        // compounds of <= 64 particles: the packed listed kernel (1), the plain one (0),
                                  // or (FP32) the plain one with each small item's sources split over both waves (2)
    int fused_thermostat = 0;     // the thermostat in the integrators' last blocks (thermostat_if_last): 1 both
                                  // integrators count, 2 the compound one only, run after the solvent one
};

// Device buffers of lima_adjoint_run / lima_record_frames, grown on demand and
// kept with the engine.  N = chain + solvent particles, chain first.
struct AdjBufs {
    DBuf<float4> ck;          // [segments][2N]: chain pos, chain prev, solvent pos, solvent prev
    DBuf<int> ck_cc;          // [segments][NC + NC*cap_cc]: compound lists before the segment's rebuild
    DBuf<int> lists;          // [slots][list words]: slot 0 the entry lists, then a segment's periods
    DBuf<float4> X, Rs;       // [segment frames][N]: positions, 2 beta (F - F*)
    DBuf<float> Tc, Ts;       // [K][n][3] force targets of the call
    DBuf<double4> adj;        // [3][N] a_k, a_{k+1}, a_{k+2} in rotation (FP64)
    DBuf<float4> w;           // [N]
    DBuf<float4> gxc, gxs;    // position-gradient parts
    DBuf<float4> gpart;       // [work item][128] listed-term parts of the compound side
    DBuf<float> cfac, gE;     // [N] dt^2/m, [K][R] energy weights
    DBuf<int> prep;           // [N] system of each particle
    DBuf<double> beta, loss, gpar, efr, lacc;
    DBuf<uint32_t> mask;
    DBuf<int> mask_off;
    int n_cchunks = 1, n_sslices = 1, comp_chunk = 1, solv_slice = kSolventChunk;
    bool ready = false;
};

}  // namespace

struct lima_engine {
    // ---- configuration ----
    lima_run_params params{};
    LaunchConfig cfg;
    int R = 0;
    std::vector<ReplicaInfo> reps_h;
    int NC = 0, NCH = 0, NS = 0, NB = 0;
    int max_cn = 0, max_cb = 0, max_ca = 0, max_cd = 0;   // most particles, bonds, angles, dihedrals of a compound
    std::vector<int> sys_chain_begin, sys_solv_begin, sys_comp_begin;  // per system, into the concatenation
    int n_cells = 0, n_stiles = 0;
    int kmax = 0;                 // max solvent<-chain chunks in any replica
    int chunk_cap = 1;            // chain particles of the largest chunk (LDS tile of the all-pairs kernel)
    size_t max_slots = 0;
    std::vector<SCWork> scwork_h;
    std::vector<int> comp_off_h;
    int device = 0;
    // host copies behind lima_get/set_parameters
    std::vector<lima_forcefield> ff_h;
    std::vector<BondD> bonds_h, brb_h;
    std::vector<AngleD> angles_h, bra_h;
    std::vector<int> bond_off_h, angle_off_h, brb_off_h, bra_off_h, br_rep_h;

    // ---- device buffers ----
    DBuf<ReplicaInfo> reps;
    DBuf<float4> cpos[2], cprev, cforce, fbridge;
    DBuf<uint32_t> intra_mask;
    DBuf<int> xoff;
    DBuf<int2> xent;
    DBuf<int> comp_off, comp_rep, comp_key;
    DBuf<float> comp_sphere;
    DBuf<int> bond_off, angle_off, dih_off, inc_off;
    DBuf<BondD> bonds;
    DBuf<AngleD> angles;
    DBuf<DihD> dihs;
    DBuf<uint32_t> inc;
    DBuf<int> br_off, br_glob, br_rep, brb_off, bra_off, brd_off;
    DBuf<BondD> brb;
    DBuf<AngleD> bra;
    DBuf<DihD> brd;
    DBuf<float4> spos[2], sprev, sforce;
    DBuf<int> solv_rep;
    DBuf<int4> stiles;
    DBuf<float> mass;
    DBuf<float4> tab_sc, tab_cs, tab_inter, tab_intra, tab_ss;
    DBuf<int> cc_cnt, cc_ids, cs_cnt, cs_ids, ss_cnt, ss_ids, seg_cnt;
    int max_segs = 1;
    DBuf<int> cw_off, cw_total;
    DBuf<int> item_comp;          // [max_slots] compound of each work item
    DBuf<int> run_off, run_cs_off, run_total, run_comp;   // runs of the packed listed kernel (compounds of <= 64 particles)
    DBuf<float2> tab_inter_ab;    // [R*1024] lj_ab of tab_inter
    DBuf<float4> partC, partS, cloc;
    DBuf<I3> partC_fx, partS_fx, cloc_fx;
    DBuf<SCWork> scwork;
    DBuf<float> lambda, lambda_applied;
    DBuf<double> ke_comp, ke_stile, temperature, energy;
    DBuf<long long> step_count;   // per replica: steps done (the reference cadence thermostat reads it)
    DBuf<int> thermo_cnt, thermo_total, thermo_total_comp;   // the thermostat in the integrators' last blocks
    bool thermo_ok = false;       // every system has integrator blocks for it (form 1)
    bool thermo_ok_comp = false;  // every system has compounds (form 2)
    int thermo_form() const {
        return cfg.fused_thermostat == 1 && thermo_ok ? 1 : cfg.fused_thermostat == 2 && thermo_ok_comp ? 2 : 0;
    }
    bool thermo_fused() const { return thermo_form() != 0; }
    // per-step logs (lima_step_logged): replica, slots per compound, ring length and rings
    int log_rep = -1, log_slots = 0, log_ring = 1;
    bool log_train = false;
    DBuf<float> log_data, train_data;
    // in-step timing of the all-pairs kernel (lima_set_kernel_clock)
    bool kclock_on = false;
    bool kclock_step = false;     // the step being enqueued is one the clock samples
    DBuf<unsigned long long> kclock;
    DBuf<uint32_t> err;
    DBuf<int> cell_cnt, cell_start, cell_cursor, cell_ids, solv_cell, cell_rep;
    DBuf<float4> ss_part, ss_img;
    DBuf<float> comp_cs_bound;
    AdjBufs adjb;

    // the step is a fork/join DAG over four streams (all at default priority):
    // stream: solvent<-chain all-pairs -> solvent integrate -> thermostat [-> rebuild]
    // s1:     compound listed LJ -> (after s3) compound sums + integrate
    // s2:     solvent listed LJ
    // s3:     bridge -> compound local forces
    hipStream_t stream = nullptr, s1 = nullptr, s2 = nullptr, s3 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_bridge = nullptr, ev_ss = nullptr, ev_c = nullptr, ev_si = nullptr;
    hipEvent_t ev_t0 = nullptr, ev_t1 = nullptr;   // lima_step's device timer
    // Captured launch graphs: one whole rebuild cycle (replayed from its start), and one step
    // of each kind [parity][rebuild follows][clocked] for steps that do not start a cycle.
    // Both are captured together on the first stepping call after any change of shape.
    hipGraphExec_t graph_exec = nullptr;
    hipGraphExec_t step_graph[2][2][2] = {};
    bool graphs_ready = false;
    int graph_steps = 0;
    long long dev_step = -1;      // the device step counters' value after the queued work (-1: unknown)
    bool async_open = false;      // ev_t0 marks the start of steps queued by lima_step_async
    unsigned char* scratch_d = nullptr;   // this device's copy of `scratch` (no-alloc staging)

    // ---- run state ----
    int cur = 0;                  // ping-pong parity of the current state
    long long step = 0;
    long long since_rebuild = 0;
    double t_device_us = 0, t_rebuild_us = 0, t_host_us = 0;
    long long n_rebuilds = 0;
    std::string last_error;

    ~lima_engine() {
        drop_graphs();
        for (hipEvent_t e : {ev_fork, ev_bridge, ev_ss, ev_c, ev_si, ev_t0, ev_t1})
            if (e) (void)hipEventDestroy(e);
        for (hipStream_t s : {stream, s1, s2, s3})
            if (s) (void)hipStreamDestroy(s);
    }

    void create_streams() {
        HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIP_TRY(hipStreamCreateWithFlags(&s1, hipStreamNonBlocking));
        HIP_TRY(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        HIP_TRY(hipStreamCreateWithFlags(&s3, hipStreamNonBlocking));
        for (hipEvent_t* e : {&ev_fork, &ev_bridge, &ev_ss, &ev_c, &ev_si})
            HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
        HIP_TRY(hipEventCreate(&ev_t0));
        HIP_TRY(hipEventCreate(&ev_t1));
    }

    // every captured graph holds the launch shape, buffers and logging switches of its capture
    void drop_graphs() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        graph_exec = nullptr;
        for (auto& a : step_graph)
            for (auto& b : a)
                for (hipGraphExec_t& g : b) {
                    if (g) (void)hipGraphExecDestroy(g);
                    g = nullptr;
                }
        graphs_ready = false;
    }

    StepArgs step_args(int parity) {
        StepArgs a{};
        a.reps = reps.p;
        a.n_replicas = R;
        a.cpos = cpos[parity].p;
        a.cpos_next = cpos[parity ^ 1].p;
        a.cprev = cprev.p;
        a.cforce = cforce.p;
        a.fbridge = fbridge.p;
        a.intra_mask = intra_mask.p;
        a.xoff = xoff.p;
        a.xent = xent.p;
        a.n_comp = NC;
        a.comp_off = comp_off.p;
        a.comp_rep = comp_rep.p;
        a.comp_key = comp_key.p;
        a.comp_sphere = comp_sphere.p;
        a.bond_off = bond_off.p; a.bonds = bonds.p;
        a.angle_off = angle_off.p; a.angles = angles.p;
        a.dih_off = dih_off.p; a.dihs = dihs.p;
        a.inc_off = inc_off.p; a.inc = inc.p;
        a.n_bridges = NB;
        a.br_off = br_off.p; a.br_glob = br_glob.p; a.br_rep = br_rep.p;
        a.brb_off = brb_off.p; a.brb = brb.p;
        a.bra_off = bra_off.p; a.bra = bra.p;
        a.brd_off = brd_off.p; a.brd = brd.p;
        a.n_solv = NS;
        a.spos = spos[parity].p;
        a.spos_next = spos[parity ^ 1].p;
        a.sprev = sprev.p;
        a.sforce = sforce.p;
        a.solv_rep = solv_rep.p;
        a.stiles = stiles.p;
        a.mass = mass.p;
        a.tab_sc = tab_sc.p; a.tab_cs = tab_cs.p; a.tab_inter = tab_inter.p; a.tab_intra = tab_intra.p; a.tab_ss = tab_ss.p;
        a.tab_inter_ab = tab_inter_ab.p;
        a.cap_cc = params.max_compound_neighbors;
        a.cap_cs = params.max_compound_solvents;
        a.cap_ss = params.max_solvent_neighbors;
        a.cc_cnt = cc_cnt.p; a.cc_ids = cc_ids.p;
        a.cs_cnt = cs_cnt.p; a.cs_ids = cs_ids.p;
        a.ss_cnt = ss_cnt.p; a.ss_ids = ss_ids.p; a.ss_ld = ss_ld();
        a.ss_part = ss_part.p; a.ss_img = ss_img.p;
        a.cw_off = cw_off.p; a.cw_total = cw_total.p; a.item_comp = item_comp.p;
        a.run_off = run_off.p; a.run_cs_off = run_cs_off.p; a.run_total = run_total.p; a.run_comp = run_comp.p;
        a.partC = partC.p; a.partC_fx = partC_fx.p;
        a.cloc = cloc.p; a.cloc_fx = cloc_fx.p;
        a.partS = partS.p; a.partS_fx = partS_fx.p;
        a.scwork = scwork.p; a.n_scwork = (int)scwork_h.size();
        a.chunk_cap = chunk_cap;
        a.lambda = lambda.p;
        a.ke_comp = ke_comp.p;
        a.ke_stile = ke_stile.p;
        a.ke_ref = params.thermostat == LIMA_THERMOSTAT_REFERENCE ? 1 : 0;
        a.log_rep = log_rep;
        a.log_slots = log_slots;
        a.log_ring = log_ring;
        a.step_count = step_count.p;
        a.thermo_cnt = thermo_fused() ? thermo_cnt.p : nullptr;
        a.thermo_total = thermo_form() == 2 ? thermo_total_comp.p : thermo_total.p;
        a.thermo_in_solv = thermo_form() == 1 ? 1 : 0;
        a.lambda_next = lambda.p;
        a.lambda_applied = lambda_applied.p;
        a.temperature = temperature.p;
        a.step_count_w = step_count.p;
        a.thermo_mode = params.thermostat;
        a.kclock = kclock_on && kclock_step ? kclock.p : nullptr;
        a.log_data = log_data.p;
        a.train_data = log_rep >= 0 && log_train ? train_data.p : nullptr;
        a.energy = energy.p;
        a.err = err.p;
        a.enable_solvents = params.enable_solvents;
        a.enable_dihedrals = params.enable_dihedrals;
        a.loc_nb = max_cb; a.loc_na = max_ca; a.loc_nd = max_cd;
        a.listed_split = (cfg.listed_packed == 2 && params.accumulation != LIMA_ACCUM_FIXED64) ? 1 : 0;
        return a;
    }

    NlistArgs nlist_args(int parity) {
        NlistArgs a{};
        a.reps = reps.p;
        a.cpos = cpos[parity].p;
        a.spos = spos[parity].p;
        a.n_comp = NC;
        a.n_solv = NS;
        a.comp_off = comp_off.p; a.comp_rep = comp_rep.p; a.comp_key = comp_key.p; a.comp_sphere = comp_sphere.p;
        a.solv_rep = solv_rep.p;
        a.cutoff = params.cutoff;
        a.cap_cc = params.max_compound_neighbors;
        a.cap_cs = params.max_compound_solvents;
        a.cap_ss = params.max_solvent_neighbors;
        a.cc_cnt = cc_cnt.p; a.cc_ids = cc_ids.p;
        a.seg_cnt = seg_cnt.p; a.max_segs = max_segs;
        a.cs_cnt = cs_cnt.p; a.cs_ids = cs_ids.p;
        a.ss_cnt = ss_cnt.p; a.ss_ids = ss_ids.p; a.ss_ld = ss_ld();
        a.n_cells = n_cells;
        a.cell_cnt = cell_cnt.p; a.cell_start = cell_start.p; a.cell_cursor = cell_cursor.p;
        a.cell_ids = cell_ids.p; a.solv_cell = solv_cell.p;
        a.cell_rep = cell_rep.p;
        a.comp_cs_bound = comp_cs_bound.p;
        a.cw_off = cw_off.p; a.cw_total = cw_total.p; a.item_comp = item_comp.p;
        a.run_len = pack_ppl() ? cc_run(pack_ppl()) : 0;
        a.run_off = run_off.p; a.run_cs_off = run_cs_off.p; a.run_total = run_total.p; a.run_comp = run_comp.p;
        a.err = err.p;
        a.enable_solvents = params.enable_solvents;
        return a;
    }

    // solvent list row stride: capacity rounded up to 4 ints (16-byte aligned rows)
    int ss_ld() const { return (params.max_solvent_neighbors + kListChunk - 1) / kListChunk * kListChunk; }

    // ---- launches ----
    size_t chunk_lds_bytes() const { return (size_t)chunk_cap * (sizeof(float4) + sizeof(float)); }

    template <bool FIXED, bool ENERGY> void launch_solvent_chain(const StepArgs& a, hipStream_t st) {
        const int nw = (int)scwork_h.size();
        const size_t lds = chunk_lds_bytes();
        switch (cfg.spt) {
            case 1: solvent_chain_kernel<1, FIXED, ENERGY><<<nw, 256, lds, st>>>(a); break;
            case 4: solvent_chain_kernel<4, FIXED, ENERGY><<<nw, 256, lds, st>>>(a); break;
            default: solvent_chain_kernel<2, FIXED, ENERGY><<<nw, 256, lds, st>>>(a); break;
        }
    }

    // Work list of the all-pairs kernel: per replica, solvent tiles of 256*spt x
    // chunks of consecutive compounds holding ~chunk_particles chain particles;
    // heaviest items first.  Sizes the partial buffer and the LDS tile.
    void build_scwork() {
        const int tile_sz = 256 * cfg.spt;
        scwork_h.clear();
        kmax = 0;
        chunk_cap = 1;
        for (int r = 0; r < R; r++) {
            ReplicaInfo& ri = reps_h[r];
            std::vector<std::pair<int, int>> chunks;
            int c = ri.comp_begin;
            while (c < ri.comp_end) {
                int c1 = c, np = 0;
                while (c1 < ri.comp_end && (c1 - c) < kChunkCompounds) {
                    const int n = comp_off_h[c1 + 1] - comp_off_h[c1];
                    if (np > 0 && (np + n > cfg.chunk_particles || np + n > kChunkParticles)) break;
                    np += n;
                    c1++;
                }
                chunk_cap = std::max(chunk_cap, np);
                chunks.push_back({c, c1});
                c = c1;
            }
            if (ri.solv_end == ri.solv_begin) chunks.clear();
            ri.chunk_count = (int)chunks.size();
            kmax = std::max(kmax, ri.chunk_count);
            for (int t0 = ri.solv_begin; t0 < ri.solv_end; t0 += tile_sz)
                for (int k = 0; k < (int)chunks.size(); k++)
                    scwork_h.push_back(SCWork{r, t0, std::min(t0 + tile_sz, ri.solv_end), chunks[k].first, chunks[k].second, k, 0, 0});
        }
        std::stable_sort(scwork_h.begin(), scwork_h.end(), [&](const SCWork& x, const SCWork& y) {
            const long long wx = (long long)(x.solv_end - x.solv_begin) * (comp_off_h[x.comp_end] - comp_off_h[x.comp_begin]);
            const long long wy = (long long)(y.solv_end - y.solv_begin) * (comp_off_h[y.comp_end] - comp_off_h[y.comp_begin]);
            return wx > wy;
        });
        const size_t need = std::max<size_t>(1, (size_t)kmax * std::max(NS, 1));
        if (params.accumulation == LIMA_ACCUM_FIXED64) {
            if (partS_fx.n < need) partS_fx.alloc(need);
            if (!partS.p) partS.alloc(1);
        } else {
            if (partS.n < need) partS.alloc(need);
            if (!partS_fx.p) partS_fx.alloc(1);
        }
        scwork.upload(scwork_h);
    }

    // compound_local_kernel: one wavefront per compound when every compound fits one and there are
    // enough compounds to fill the device (512, two per SIMD pair), else two (small systems run
    // each compound's terms over twice the threads); term results in dynamic LDS sized to the
    // largest compound
    template <bool FIXED, bool ENERGY> void launch_compound_local(const StepArgs& a, hipStream_t st) {
        if (max_cn <= 64 && NC >= 512)
            compound_local_kernel<FIXED, ENERGY, 64><<<NC, 64, local_lds_bytes<64>(max_cb, max_ca, max_cd), st>>>(a);
        else
            compound_local_kernel<FIXED, ENERGY, 128><<<NC, 128, local_lds_bytes<128>(max_cb, max_ca, max_cd), st>>>(a);
    }

    // particles per lane of the packed listed kernel (compounds of at most 16, 32, 64 particles), 0: plain kernel
    int pack_ppl() const { return max_cn <= 16 ? 1 : max_cn <= 32 ? 2 : max_cn <= 64 ? 4 : 0; }

    // compound_listed_kernel, or its packed form when every compound fits 16 lanes x 1, 2 or 4 particles
    template <bool FIXED, bool ENERGY> void launch_compound_listed(const StepArgs& a, int grid, hipStream_t st) {
        // one-wavefront blocks: twice the blocks of the two-wavefront plain kernel's grid
        switch (cfg.listed_packed == 1 ? pack_ppl() : 0) {
            case 1: compound_listed_packed_kernel<FIXED, ENERGY, 1><<<2 * grid, 64, 0, st>>>(a); break;
            case 2: compound_listed_packed_kernel<FIXED, ENERGY, 2><<<2 * grid, 64, 0, st>>>(a); break;
            case 4: compound_listed_packed_kernel<FIXED, ENERGY, 4><<<2 * grid, 64, 0, st>>>(a); break;
            default: compound_listed_kernel<FIXED, ENERGY><<<grid, kCompoundBlock, 0, st>>>(a);
        }
    }

    // one force pass (and integration when ADVANCE) as a fork/join over four streams
    template <bool ADVANCE, bool FIXED, bool ENERGY> void launch_forces(int parity) {
        StepArgs a = step_args(parity);
        const bool solv = params.enable_solvents && NS > 0;
        // enqueue order, measured: the all-pairs kernel, the listed kernels, then the short chain
        // that gates the compounds' integration (a graph dispatches its nodes in capture order;
        // dispatching that chain first, or on the main stream, made the step 0.03-0.09 ms slower)
        HIP_TRY(hipEventRecord(ev_fork, stream));
        // main stream: all-pairs solvent <- chain (then, below, the solvent integrator)
        if (solv && !scwork_h.empty()) launch_solvent_chain<FIXED, ENERGY>(a, stream);
        // s1: compound listed LJ, then (after the local forces) the sums and integration of the compounds
        HIP_TRY(hipStreamWaitEvent(s1, ev_fork, 0));
        if (NC > 0) {
            const int grid = cfg.listed_grid > 0 ? cfg.listed_grid : 4096;
            launch_compound_listed<FIXED, ENERGY>(a, grid, s1);
        }
        // s2: the listed solvent-solvent LJ
        HIP_TRY(hipStreamWaitEvent(s2, ev_fork, 0));
        if (solv) {
            if (!FIXED && cfg.listed_lanes == kSplitLanes)
                solvent_listed_split_kernel<ENERGY><<<xcd_grid(n_stiles * kSplitItems), 64, 0, s2>>>(a, n_stiles * kSplitItems);
            else
                solvent_listed_kernel<ENERGY><<<xcd_grid(n_stiles * kListedItems), 64, 0, s2>>>(a, n_stiles * kListedItems);
        }
        HIP_TRY(hipEventRecord(ev_ss, s2));
        // s3: bridge carry-in forces, then the compounds' local forces (no list needed)
        HIP_TRY(hipStreamWaitEvent(s3, ev_fork, 0));
        if (NB > 0) bridge_kernel<ENERGY><<<NB, 64, 0, s3>>>(a);
        if (NC > 0) launch_compound_local<FIXED, ENERGY>(a, s3);
        HIP_TRY(hipEventRecord(ev_bridge, s3));
        HIP_TRY(hipStreamWaitEvent(s1, ev_bridge, 0));
        if (ADVANCE && thermo_form() == 2) {
            // the thermostat in the compound integrator's last blocks: the solvent integrator (and
            // its kinetic-energy partials) comes first, so the compound blocks' release fences fall
            // at the end of the step
            HIP_TRY(hipStreamWaitEvent(stream, ev_ss, 0));
            if (solv) solvent_integrate_kernel<ADVANCE, FIXED, ENERGY><<<n_stiles, 256, 0, stream>>>(a);
            HIP_TRY(hipEventRecord(ev_si, stream));
            HIP_TRY(hipStreamWaitEvent(s1, ev_si, 0));
            compound_integrate_kernel<ADVANCE, FIXED, ENERGY><<<NC, kCompoundBlock, 0, s1>>>(a);
            HIP_TRY(hipEventRecord(ev_c, s1));
            HIP_TRY(hipStreamWaitEvent(stream, ev_c, 0));
            return;
        }
        if (NC > 0) compound_integrate_kernel<ADVANCE, FIXED, ENERGY><<<NC, kCompoundBlock, 0, s1>>>(a);
        HIP_TRY(hipEventRecord(ev_c, s1));
        HIP_TRY(hipStreamWaitEvent(stream, ev_ss, 0));
        if (solv) solvent_integrate_kernel<ADVANCE, FIXED, ENERGY><<<n_stiles, 256, 0, stream>>>(a);
        HIP_TRY(hipStreamWaitEvent(stream, ev_c, 0));
    }

    // one step; with `energy` the force pass also accumulates the FP64 energy
    // terms (the caller clears the energy buffer first)
    void launch_step(int parity, bool energy = false) {
        const bool fx = params.accumulation == LIMA_ACCUM_FIXED64;
        if (energy) {
            if (fx) launch_forces<true, true, true>(parity);
            else launch_forces<true, false, true>(parity);
        } else {
            if (fx) launch_forces<true, true, false>(parity);
            else launch_forces<true, false, false>(parity);
        }
        // with solvents disabled the solvent state is carried across the ping-pong unchanged
        if (!params.enable_solvents && NS > 0)
            HIP_TRY(hipMemcpyAsync(spos[parity ^ 1].p, spos[parity].p, NS * sizeof(float4), hipMemcpyDeviceToDevice, stream));
        if (!thermo_fused())
            thermostat_kernel<<<R, 64, 0, stream>>>(reps.p, R, ke_comp.p, ke_stile.p, lambda.p, lambda_applied.p,
                                                temperature.p, step_count.p, params.thermostat,
                                                kclock_on && kclock_step ? kclock.p : nullptr);
        if (dev_step >= 0) dev_step++;   // the thermostat kernel advances the device step counters
    }

    // the rebuild forks too: compound lists on s1, the solvent cell pipeline on the main stream
    void launch_rebuild(int parity) {
        NlistArgs a = nlist_args(parity);
        const bool solv = params.enable_solvents && NS > 0;
        HIP_TRY(hipEventRecord(ev_fork, stream));
        HIP_TRY(hipStreamWaitEvent(s1, ev_fork, 0));
        if (NC > 0) {
            nl_compound_compound_kernel<<<NC, 64, 0, s1>>>(a);
            if (solv) {
                nl_cs_count_kernel<<<dim3(NC, max_segs), 256, 0, s1>>>(a);
                nl_cs_write_kernel<<<dim3(NC, max_segs), 256, 0, s1>>>(a);
            }
        }
        HIP_TRY(hipEventRecord(ev_c, s1));
        if (solv) {
            nl_cell_zero_kernel<<<(n_cells + 255) / 256, 256, 0, stream>>>(cell_cnt.p, n_cells);
            nl_cell_count_kernel<<<(NS + 255) / 256, 256, 0, stream>>>(a);
            nl_cell_scan_kernel<<<1, 1024, 0, stream>>>(a);
            nl_cell_scatter_kernel<<<(NS + 255) / 256, 256, 0, stream>>>(a);
            nl_cell_sort_kernel<<<n_cells, 256, 0, stream>>>(a);
            nl_solvent_solvent_kernel<<<xcd_grid(n_cells), kSolventListThreads, 0, stream>>>(a);
        }
        HIP_TRY(hipStreamWaitEvent(stream, ev_c, 0));
        launch_work_list(a, stream);
    }

    // the listed work items of the current lists: per-compound offsets, then each item's compound
    void launch_work_list(const NlistArgs& a, hipStream_t st) {
        if (NC <= 0) return;
        nl_work_kernel<<<1, 1024, 0, st>>>(a);
        nl_item_kernel<<<NC, 64, 0, st>>>(a);
    }

    // lj_ab of the inter-compound tables (after any change of them)
    void fill_ab_tables() {
        const int n = R * kMaxTypes * kMaxTypes;
        ab_table_kernel<<<(n + 255) / 256, 256, 0, stream>>>(tab_inter.p, tab_inter_ab.p, n);
        HIP_TRY(hipStreamSynchronize(stream));
    }

    int cycle_steps() const {
        const int iv = params.nlist_interval > 0 ? params.nlist_interval : 1;
        return (iv % 2 == 0) ? iv : 2 * iv;
    }

    // What distinguishes one step's launches from another's: the state parity, whether the
    // scheduled rebuild (Engine::hostMaster, Engine.cu:49-75) follows it, and whether the kernel
    // clock samples it (one step per list period: the second, clear of the rebuild's tail).
    struct StepKind {
        int parity;
        bool rebuild;
        bool clocked;
    };
    StepKind next_kind() const {
        const int iv = params.nlist_interval;
        const int pick = iv > 1 ? 1 : 0;
        StepKind k;
        k.parity = cur;
        k.rebuild = iv > 0 && since_rebuild + 1 >= iv;
        k.clocked = kclock_on && (iv > 0 ? since_rebuild == pick : step % 8 == 0);
        return k;
    }
    void enqueue_kind(const StepKind& k) {
        kclock_step = k.clocked;
        launch_step(k.parity);
        kclock_step = false;   // steps launched directly (observed, adjoint) are never clocked
        if (k.rebuild) launch_rebuild(k.parity ^ 1);
    }
    void advance(const StepKind& k) {
        cur ^= 1;
        step++;
        since_rebuild++;
        if (k.rebuild) {
            since_rebuild = 0;
            n_rebuilds++;
        }
    }
    // one step plus the rebuild it schedules
    void enqueue_one() {
        const StepKind k = next_kind();
        enqueue_kind(k);
        advance(k);
    }

    template <typename F> hipGraphExec_t capture(F&& enqueue) {
        hipGraph_t g;
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        enqueue();
        HIP_TRY(hipStreamEndCapture(stream, &g));
        hipGraphExec_t ex = nullptr;
        HIP_TRY(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
        HIP_TRY(hipGraphDestroy(g));
        // the one-time upload of the executable graph here, not inside its first replay
        HIP_TRY(hipGraphUpload(ex, stream));
        return ex;
    }

    // Captures the cycle graph and every step kind that can occur, so that no later stepping
    // call captures: replays of either kind can follow each other in any order.
    void ensure_graphs() {
        if (graphs_ready) return;
        drop_graphs();
        const int save_cur = cur;
        const long long save_step = step, save_since = since_rebuild, save_nr = n_rebuilds, save_dev = dev_step;
        const int iv = params.nlist_interval;
        if (iv > 0) {
            // the whole cycle, from parity 0 and a fresh rebuild cycle
            const int n = cycle_steps();
            cur = 0;
            since_rebuild = 0;
            graph_exec = capture([&] {
                for (int i = 0; i < n; i++) enqueue_one();
            });
            graph_steps = n;
        }
        for (int p = 0; p < 2; p++)
            for (int rb = 0; rb < (iv > 0 ? 2 : 1); rb++)
                for (int ck = 0; ck < (kclock_on ? 2 : 1); ck++) {
                    const StepKind k{p, rb != 0, ck != 0};
                    step_graph[p][rb][ck] = capture([&] { enqueue_kind(k); });
                }
        cur = save_cur;
        step = save_step;
        since_rebuild = save_since;
        n_rebuilds = save_nr;
        dev_step = save_dev;
        graphs_ready = true;
    }

    // the device step counters follow the host's count (set before any stepping call enqueues)
    void sync_step_count() {
        set_step_count_kernel<<<(R + 63) / 64, 64, 0, stream>>>(step_count.p, R, step);
        dev_step = step;
    }

    // allow_capture false: replays only graphs already built, else launches directly
    void enqueue_steps(int n, bool allow_capture = true) {
        if (n <= 0) return;
        if (dev_step != step) sync_step_count();
        const bool graphs = cfg.use_graph != 0 && (allow_capture || graphs_ready);
        if (graphs) ensure_graphs();
        const int cyc = cycle_steps();
        while (n > 0) {
            if (graphs && graph_exec && cur == 0 && since_rebuild == 0 && n >= cyc) {
                HIP_TRY(hipGraphLaunch(graph_exec, stream));
                step += cyc;
                dev_step += cyc;
                n_rebuilds += cyc / params.nlist_interval;
                n -= cyc;
                continue;
            }
            const StepKind k = next_kind();
            if (graphs) {
                HIP_TRY(hipGraphLaunch(step_graph[k.parity][k.rebuild ? 1 : 0][k.clocked ? 1 : 0], stream));
                dev_step++;
            } else {
                enqueue_kind(k);
            }
            advance(k);
            n--;
        }
    }

    int check_errors() {
        uint32_t e[4];
        HIP_TRY(hipMemcpy(e, err.p, sizeof(e), hipMemcpyDeviceToHost));
        if (e[0] == 0) return LIMA_OK;
        HIP_TRY(hipMemset(err.p, 0, 4 * sizeof(uint32_t)));
        char buf[256];
        if (e[0] & kErrCapacity) {
            snprintf(buf, sizeof(buf), "neighbour list capacity exceeded (needed: compound-compound %u, compound-solvent %u, solvent %u)",
                     e[1], e[2], e[3]);
            last_error = buf;
            return LIMA_E_CAPACITY;
        }
        if (e[0] & kErrCriticalForce) { last_error = "critical force |F| > 2e12 on a chain particle"; return LIMA_E_CRITICAL_FORCE; }
        if (e[0] & kErrNaN) { last_error = "NaN solvent position"; return LIMA_E_NAN; }
        last_error = "unknown device error";
        return LIMA_E_HIP;
    }
};

namespace {

// Smallest non-negative float x with sqrtf(x) >= t.  sqrtf is correctly rounded
// (here and on the device) and monotone, so  sqrtf(d2) < t  <=>  d2 < sqrt_bound(t):
// the neighbour tests compare squared distances with the exact same outcome.
float sqrt_bound(float t) {
    uint32_t lo = 0, hi = 0x7f800000u;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        float x;
        std::memcpy(&x, &mid, 4);
        if (std::sqrt(x) >= t) hi = mid;
        else lo = mid + 1;
    }
    float r;
    std::memcpy(&r, &lo, 4);
    return r;
}

// LJ parameters of one call site: sigma^2 and 24*eps rounded as the reference
// rounds sigma*sigma and 24.f*epsilon (Engine.cu:488-490), -2*(24 eps), and the
// mixed sigma itself (d sigma^2 / d sigma_type = sigma for the arithmetic sigma rule).
float4 pair_params(float sigma, float eps) {
    const float e24 = 24.f * eps;
    return make_float4(sigma * sigma, e24, -2.0f * e24, sigma);
}

// type tables of one system (mixing per call site, Engine.cu:643-646, 668-670,
// 689-691, 711-713, 724-727): [32] masses, [32] solvent <- type, [32] type <-
// solvent, [32*32] inter-compound, [32*32] intra-compound, [1] solvent-solvent
void fill_tables(const lima_particle_params* pp, float* mass, float4* tsc, float4* tcs, float4* tinter, float4* tintra,
                 float4* tss) {
    for (int t = 0; t < kMaxTypes; t++) {
        mass[t] = pp[t].mass;
        float sg = (pp[0].sigma + pp[t].sigma) * 0.5f;
        float ep = (pp[0].epsilon + pp[t].epsilon) * 0.5f;
        tsc[t] = pair_params(sg, ep);
        float epg = sqrtf(pp[t].epsilon * pp[0].epsilon);
        tcs[t] = pair_params(sg, epg);
        for (int u = 0; u < kMaxTypes; u++) {
            float su = (pp[t].sigma + pp[u].sigma) * 0.5f;
            float ea = (pp[t].epsilon + pp[u].epsilon) * 0.5f;
            float eg = sqrtf(pp[t].epsilon * pp[u].epsilon);
            tinter[t * 32 + u] = pair_params(su, ea);
            tintra[t * 32 + u] = pair_params(su, eg);
        }
    }
    *tss = pair_params(pp[0].sigma, pp[0].epsilon);
}

void validate_and_build(lima_engine* E, const lima_system* sys, int nsys) {
    const lima_run_params& P = E->params;
    std::vector<float> cs_bound;
    std::vector<int> cell_rep_h;
    if (nsys <= 0) throw LimaError(LIMA_E_ARG, "n_systems must be >= 1");
    if (!(P.cutoff > 0)) throw LimaError(LIMA_E_ARG, "cutoff must be > 0");
    if (P.thermostat < LIMA_THERMOSTAT_NONE || P.thermostat > LIMA_THERMOSTAT_REFERENCE)
        throw LimaError(LIMA_E_ARG, "unknown thermostat");
    if (P.max_compound_neighbors <= 0 || P.max_compound_solvents <= 0 || P.max_solvent_neighbors <= 0)
        throw LimaError(LIMA_E_ARG, "capacities must be > 0");
    std::vector<float4> cpos, cprev, spos, sprev;
    std::vector<uint32_t> imask;
    std::vector<int> xoff{0};
    std::vector<int2> xent;
    std::vector<int> comp_off{0}, comp_rep, comp_key;
    std::vector<float> comp_sphere;
    std::vector<int> bond_off{0}, angle_off{0}, dih_off{0}, inc_off{0};
    std::vector<BondD> bonds;
    std::vector<AngleD> angles;
    std::vector<DihD> dihs;
    std::vector<uint32_t> inc;
    std::vector<int> br_off{0}, br_glob, br_rep, brb_off{0}, bra_off{0}, brd_off{0};
    std::vector<BondD> brb;
    std::vector<AngleD> bra;
    std::vector<DihD> brd;
    std::vector<int> solv_rep;
    std::vector<float> mass(nsys * kMaxTypes);
    std::vector<float4> tsc(nsys * kMaxTypes), tcs(nsys * kMaxTypes), tinter(nsys * 1024), tintra(nsys * 1024), tss(nsys);
    E->reps_h.clear();
    E->ff_h.clear();
    E->sys_chain_begin.clear(); E->sys_solv_begin.clear(); E->sys_comp_begin.clear();
    std::vector<int> bridge_owner;

    for (int r = 0; r < nsys; r++) {
        const lima_system& s = sys[r];
        if (!(s.box_len > 0) || !(s.dt > 0)) throw LimaError(LIMA_E_ARG, "box_len and dt must be > 0");
        if (s.n_compounds < 0 || s.n_solvents < 0 || s.n_bridges < 0) throw LimaError(LIMA_E_ARG, "negative counts");
        ReplicaInfo ri{};
        ri.L = s.box_len;
        ri.half = s.box_len * 0.5f;
        ri.dt = s.dt;
        ri.tau = P.thermostat_tau > 0 ? P.thermostat_tau : 1.0;
        ri.T0 = s.thermostat_target;
        const int cb = (int)comp_rep.size();
        const int chb = (int)cpos.size();
        const int sb = (int)spos.size();
        E->sys_comp_begin.push_back(cb);
        E->sys_chain_begin.push_back(chb);
        E->sys_solv_begin.push_back(sb);
        ri.comp_begin = cb; ri.comp_end = cb + s.n_compounds;
        const int nchain = s.n_compounds > 0 ? s.compound_offset[s.n_compounds] : 0;
        ri.chain_begin = chb; ri.chain_end = chb + nchain;
        ri.solv_begin = sb; ri.solv_end = sb + s.n_solvents;
        fill_tables(s.forcefield.particle_parameters, &mass[r * kMaxTypes], &tsc[r * kMaxTypes], &tcs[r * kMaxTypes],
                    &tinter[r * 1024], &tintra[r * 1024], &tss[r]);
        E->ff_h.push_back(s.forcefield);
        // compounds
        for (int c = 0; c < s.n_compounds; c++) {
            const int o0 = s.compound_offset[c], o1 = s.compound_offset[c + 1];
            const int n = o1 - o0;
            if (n <= 0 || n > kCompoundBlock) throw LimaError(LIMA_E_CAPACITY, "compound size must be 1..128 particles");
            if (s.key_particle_index[c] < 0 || s.key_particle_index[c] >= n) throw LimaError(LIMA_E_ARG, "key_particle_index out of range");
            comp_off.push_back(comp_off.back() + n);
            comp_rep.push_back(r);
            comp_key.push_back(s.key_particle_index[c]);
            comp_sphere.push_back(s.confining_particle_sphere[c]);
            // threshold of the compound-solvent test, C + (r + C) (Engine.cu:182, 231)
            cs_bound.push_back(sqrt_bound(P.cutoff + (s.confining_particle_sphere[c] + P.cutoff)));
            const int gbase = chb + o0;
            const int nbd = s.bond_offset[c + 1] - s.bond_offset[c];
            const int nan = s.angle_offset[c + 1] - s.angle_offset[c];
            const int ndh = s.dihedral_offset[c + 1] - s.dihedral_offset[c];
            if (nbd > kMaxBondsPerCompound || nan > kMaxAnglesPerCompound || ndh > kMaxDihedralsPerCompound)
                throw LimaError(LIMA_E_CAPACITY, "too many bonded terms in a compound");
            E->max_cn = std::max(E->max_cn, n);
            E->max_cb = std::max(E->max_cb, nbd);
            E->max_ca = std::max(E->max_ca, nan);
            E->max_cd = std::max(E->max_cd, ndh);
            std::vector<std::vector<uint32_t>> pinc(n);
            for (int t = 0; t < nbd; t++) {
                const lima_pair_bond& b = s.bonds[s.bond_offset[c] + t];
                for (int k = 0; k < 2; k++) if ((int)b.atom_indexes[k] >= n) throw LimaError(LIMA_E_ARG, "bond index out of range");
                bonds.push_back(BondD{b.b0, b.kb, (int)b.atom_indexes[0], (int)b.atom_indexes[1]});
                for (int k = 0; k < 2; k++) pinc[b.atom_indexes[k]].push_back((0u << 30) | ((uint32_t)t << 2) | (uint32_t)k);
            }
            for (int t = 0; t < nan; t++) {
                const lima_angle_bond& b = s.angles[s.angle_offset[c] + t];
                for (int k = 0; k < 3; k++) if ((int)b.atom_indexes[k] >= n) throw LimaError(LIMA_E_ARG, "angle index out of range");
                angles.push_back(AngleD{b.theta_0, b.k_theta, (int)b.atom_indexes[0], (int)b.atom_indexes[1], (int)b.atom_indexes[2], 0});
                for (int k = 0; k < 3; k++) pinc[b.atom_indexes[k]].push_back((1u << 30) | ((uint32_t)t << 2) | (uint32_t)k);
            }
            for (int t = 0; t < ndh; t++) {
                const lima_dihedral_bond& b = s.dihedrals[s.dihedral_offset[c] + t];
                for (int k = 0; k < 4; k++) if ((int)b.atom_indexes[k] >= n) throw LimaError(LIMA_E_ARG, "dihedral index out of range");
                dihs.push_back(DihD{b.phi_0, b.k_phi, (int)b.n, (int)b.atom_indexes[0], (int)b.atom_indexes[1],
                                    (int)b.atom_indexes[2], (int)b.atom_indexes[3], 0});
                for (int k = 0; k < 4; k++) pinc[b.atom_indexes[k]].push_back((2u << 30) | ((uint32_t)t << 2) | (uint32_t)k);
            }
            bond_off.push_back((int)bonds.size());
            angle_off.push_back((int)angles.size());
            dih_off.push_back((int)dihs.size());
            for (int p = 0; p < n; p++) {
                for (uint32_t v : pinc[p]) inc.push_back(v);
                inc_off.push_back((int)inc.size());
            }
            for (int p = 0; p < n; p++) {
                const int li = o0 + p;
                const float* x = s.chain_positions + 3 * li;
                const float* xm = s.chain_prev_positions + 3 * li;
                const int t = s.chain_atom_types[li];
                if (t >= kMaxTypes) throw LimaError(LIMA_E_ARG, "atom type >= 32");
                float w;
                std::memcpy(&w, &t, 4);
                cpos.push_back(make_float4(x[0], x[1], x[2], w));
                cprev.push_back(make_float4(xm[0], xm[1], xm[2], 0.f));
                uint32_t m[4] = {0, 0, 0, 0};
                for (int e = s.ignore_offset[li]; e < s.ignore_offset[li + 1]; e++) {
                    const int oc = s.ignore_compound[e], ol = s.ignore_local[e];
                    if (oc < 0 || oc >= s.n_compounds) throw LimaError(LIMA_E_ARG, "ignore compound id out of range");
                    if (ol < 0 || ol >= kCompoundBlock) throw LimaError(LIMA_E_ARG, "ignore local id out of range");
                    if (oc == c) m[ol >> 5] |= 1u << (ol & 31);
                    else xent.push_back(make_int2(cb + oc, ol));
                }
                for (int k = 0; k < 4; k++) imask.push_back(m[k]);
                xoff.push_back((int)xent.size());
            }
            (void)gbase;
        }
        // bridges
        for (int b = 0; b < s.n_bridges; b++) {
            const int q0 = s.bridge_particle_offset[b], q1 = s.bridge_particle_offset[b + 1];
            if (q1 - q0 > 32 || q1 - q0 <= 0) throw LimaError(LIMA_E_CAPACITY, "bridge must have 1..32 particles");
            for (int k = q0; k < q1; k++) {
                const int cc = s.bridge_particle_compound[k], ll = s.bridge_particle_local[k];
                if (cc < 0 || cc >= s.n_compounds || ll < 0 || ll >= s.compound_offset[cc + 1] - s.compound_offset[cc])
                    throw LimaError(LIMA_E_ARG, "bridge particle reference out of range");
                br_glob.push_back(chb + s.compound_offset[cc] + ll);
            }
            br_off.push_back((int)br_glob.size());
            br_rep.push_back(r);
            const int np = q1 - q0;
            for (int t = s.bridge_bond_offset[b]; t < s.bridge_bond_offset[b + 1]; t++) {
                const lima_pair_bond& x = s.bridge_bonds[t];
                for (int k = 0; k < 2; k++) if ((int)x.atom_indexes[k] >= np) throw LimaError(LIMA_E_ARG, "bridge bond index out of range");
                brb.push_back(BondD{x.b0, x.kb, (int)x.atom_indexes[0], (int)x.atom_indexes[1]});
            }
            for (int t = s.bridge_angle_offset[b]; t < s.bridge_angle_offset[b + 1]; t++) {
                const lima_angle_bond& x = s.bridge_angles[t];
                for (int k = 0; k < 3; k++) if ((int)x.atom_indexes[k] >= np) throw LimaError(LIMA_E_ARG, "bridge angle index out of range");
                bra.push_back(AngleD{x.theta_0, x.k_theta, (int)x.atom_indexes[0], (int)x.atom_indexes[1], (int)x.atom_indexes[2], 0});
            }
            for (int t = s.bridge_dihedral_offset[b]; t < s.bridge_dihedral_offset[b + 1]; t++) {
                const lima_dihedral_bond& x = s.bridge_dihedrals[t];
                for (int k = 0; k < 4; k++) if ((int)x.atom_indexes[k] >= np) throw LimaError(LIMA_E_ARG, "bridge dihedral index out of range");
                brd.push_back(DihD{x.phi_0, x.k_phi, (int)x.n, (int)x.atom_indexes[0], (int)x.atom_indexes[1],
                                   (int)x.atom_indexes[2], (int)x.atom_indexes[3], 0});
            }
            brb_off.push_back((int)brb.size());
            bra_off.push_back((int)bra.size());
            brd_off.push_back((int)brd.size());
        }
        for (int k = 0; k < s.n_solvents; k++) {
            const float* x = s.solvent_positions + 3 * k;
            const float* xm = s.solvent_prev_positions + 3 * k;
            spos.push_back(make_float4(x[0], x[1], x[2], 0.f));
            sprev.push_back(make_float4(xm[0], xm[1], xm[2], 0.f));
            solv_rep.push_back(r);
        }
        E->reps_h.push_back(ri);
    }
    // a chain particle may belong to at most one bridge (the bridge force is a store)
    {
        std::vector<char> seen(cpos.size(), 0);
        for (int g : br_glob) {
            if (seen[g]) throw LimaError(LIMA_E_ARG, "a particle belongs to more than one bridge");
            seen[g] = 1;
        }
    }
    E->R = nsys;
    E->NC = (int)comp_rep.size();
    E->NCH = (int)cpos.size();
    E->NS = (int)spos.size();
    E->NB = (int)br_rep.size();

    // ---- solvent tiles, cell grids (the solvent<-chain work list: build_scwork) ----
    E->comp_off_h = comp_off;
    std::vector<int4> stl;
    int cell_begin = 0;
    for (int r = 0; r < nsys; r++) {
        ReplicaInfo& ri = E->reps_h[r];
        ri.stile_begin = (int)stl.size();
        for (int t0 = ri.solv_begin; t0 < ri.solv_end; t0 += 256)
            stl.push_back(make_int4(r, t0, std::min(t0 + 256, ri.solv_end), (int)stl.size()));
        ri.stile_end = (int)stl.size();
        const float thr = P.cutoff + P.cutoff;   // Engine.cu:198, 264
        ri.ss_bound = sqrt_bound(thr);
        int ncell = (int)std::floor(ri.L / (thr * (1.0f + 1e-5f)));
        ncell = std::max(1, std::min(ncell, 256));
        ri.ncell = ncell;
        ri.cell_inv = (float)ncell / ri.L;
        ri.cell_begin = cell_begin;
        cell_begin += ncell * ncell * ncell;
        cell_rep_h.insert(cell_rep_h.end(), (size_t)ncell * ncell * ncell, r);
    }
    E->build_scwork();
    E->n_cells = std::max(cell_begin, 1);
    E->n_stiles = (int)stl.size();

    // ---- upload ----
    E->reps.upload(E->reps_h);
    for (int k = 0; k < 2; k++) E->cpos[k].upload(cpos);
    E->cprev.upload(cprev);
    E->cforce.alloc(std::max(E->NCH, 1)); E->cforce.zero();
    E->fbridge.alloc(std::max(E->NCH, 1)); E->fbridge.zero();
    E->intra_mask.upload(imask);
    E->xoff.upload(xoff);
    E->xent.upload(xent);
    E->comp_off.upload(comp_off);
    E->comp_rep.upload(comp_rep);
    E->comp_key.upload(comp_key);
    E->comp_sphere.upload(comp_sphere);
    E->bond_off.upload(bond_off); E->bonds.upload(bonds);
    E->angle_off.upload(angle_off); E->angles.upload(angles);
    E->dih_off.upload(dih_off); E->dihs.upload(dihs);
    E->inc_off.upload(inc_off); E->inc.upload(inc);
    E->br_off.upload(br_off); E->br_glob.upload(br_glob); E->br_rep.upload(br_rep);
    E->brb_off.upload(brb_off); E->brb.upload(brb);
    E->bra_off.upload(bra_off); E->bra.upload(bra);
    E->brd_off.upload(brd_off); E->brd.upload(brd);
    E->bonds_h = bonds; E->angles_h = angles; E->brb_h = brb; E->bra_h = bra;
    E->bond_off_h = bond_off; E->angle_off_h = angle_off; E->brb_off_h = brb_off; E->bra_off_h = bra_off;
    E->br_rep_h = br_rep;
    for (int k = 0; k < 2; k++) E->spos[k].upload(spos);
    E->sprev.upload(sprev);
    E->sforce.alloc(std::max(E->NS, 1)); E->sforce.zero();
    E->solv_rep.upload(solv_rep);
    E->stiles.upload(stl);
    E->mass.upload(mass);
    E->tab_sc.upload(tsc); E->tab_cs.upload(tcs); E->tab_inter.upload(tinter); E->tab_intra.upload(tintra); E->tab_ss.upload(tss);
    E->tab_inter_ab.alloc(tinter.size());
    const size_t NCx = std::max(E->NC, 1), NSx = std::max(E->NS, 1), NCHx = std::max(E->NCH, 1);
    E->cc_cnt.alloc(NCx); E->cc_cnt.zero();
    E->cc_ids.alloc(NCx * P.max_compound_neighbors);
    E->cs_cnt.alloc(NCx); E->cs_cnt.zero();
    E->max_segs = 1;
    for (const ReplicaInfo& ri : E->reps_h)
        E->max_segs = std::max(E->max_segs, (ri.solv_end - ri.solv_begin + kSegSolvents - 1) / kSegSolvents);
    E->seg_cnt.alloc(NCx * E->max_segs); E->seg_cnt.zero();
    E->cs_ids.alloc(NCx * P.max_compound_solvents);
    E->ss_cnt.alloc(NSx); E->ss_cnt.zero();
    if ((NSx + 256) * (size_t)E->ss_ld() >= (size_t(1) << 32))
        throw LimaError(LIMA_E_ARG, "solvent lists exceed 2^32 entries (n_solvents x max_solvent_neighbors)");
    E->ss_ids.alloc((NSx + 256) * E->ss_ld());   // spare rows: solvent_listed reads whole 64-row groups
    E->cw_off.alloc(NCx + 1); E->cw_off.zero();
    E->cw_total.alloc(1); E->cw_total.zero();
    E->max_slots = NCx * (size_t)(P.max_compound_neighbors + (P.max_compound_solvents + kSolventChunk - 1) / kSolventChunk);
    E->item_comp.alloc(std::max<size_t>(E->max_slots, 1)); E->item_comp.zero();
    E->run_off.alloc(NCx + 1); E->run_off.zero();
    E->run_cs_off.alloc(NCx + 1); E->run_cs_off.zero();
    E->run_total.alloc(2); E->run_total.zero();
    E->run_comp.alloc(std::max<size_t>(E->max_slots, 1)); E->run_comp.zero();
    if (P.accumulation == LIMA_ACCUM_FIXED64) {
        E->partC_fx.alloc(E->max_slots * kCompoundBlock);
        E->partC.alloc(1);
        E->cloc_fx.alloc(NCHx);
        E->cloc.alloc(1);
    } else {
        E->partC.alloc(E->max_slots * kCompoundBlock);
        E->partC_fx.alloc(1);
        E->cloc.alloc(NCHx);
        E->cloc_fx.alloc(1);
    }
    std::vector<float> ones(nsys, 1.0f);
    E->lambda.upload(ones);
    E->lambda_applied.upload(ones);
    E->ke_comp.alloc(NCx); E->ke_comp.zero();
    E->ke_stile.alloc(std::max(E->n_stiles, 1)); E->ke_stile.zero();
    E->temperature.alloc(nsys); E->temperature.zero();
    E->step_count.alloc(nsys); E->step_count.zero();
    {
        // the thermostat runs in the integrators' last block of each system when every system has
        // integrator blocks of the kind that runs after the all-pairs kernel (its solvent tiles,
        // when solvents are on: the thermostat of system 0 also closes that kernel's clock span)
        const bool solv = E->params.enable_solvents && E->NS > 0;
        std::vector<int> total(nsys), total_comp(nsys);
        bool ok = true, ok_comp = E->NC > 0;
        for (int r = 0; r < nsys; r++) {
            const ReplicaInfo& ri = E->reps_h[r];
            const int tiles = solv ? ri.stile_end - ri.stile_begin : 0;
            total[r] = (E->NC > 0 ? ri.comp_end - ri.comp_begin : 0) + tiles;
            total_comp[r] = ri.comp_end - ri.comp_begin;
            ok = ok && total[r] > 0 && (!solv || tiles > 0);
            ok_comp = ok_comp && total_comp[r] > 0;
        }
        E->thermo_total.upload(total);
        E->thermo_total_comp.upload(total_comp);
        E->thermo_ok_comp = ok_comp;
        E->thermo_cnt.alloc(nsys); E->thermo_cnt.zero();
        E->thermo_ok = ok;
    }
    E->energy.alloc((size_t)nsys * kEnergyTerms); E->energy.zero();
    E->err.alloc(4); E->err.zero();
    E->cell_cnt.alloc(E->n_cells); E->cell_start.alloc(E->n_cells + 1); E->cell_cursor.alloc(E->n_cells);
    E->cell_ids.alloc(NSx); E->solv_cell.alloc(NSx);
    E->cell_rep.upload(cell_rep_h);
    E->comp_cs_bound.upload(cs_bound);
    E->ss_part.alloc(NSx); E->ss_part.zero();
    E->ss_img.alloc(NSx); E->ss_img.zero();
}

// ---- reverse mode ------------------------------------------------------------
// Frames k = 0..K-1 start at the current state.  A forward pass stores a state
// checkpoint at the start of every segment (segments begin on list rebuilds,
// every `checkpoint_interval` frames or more) together with the compound lists
// the segment's first rebuild reads (its hysteresis input), and evaluates the
// loss.  The backward pass visits the segments last to first: it restores the
// checkpoint, rebuilds the segment's first lists from the stored compound lists
// (the same kernels on the same input give the same lists), recomputes the
// segment's frames keeping positions and force residuals (and a full copy of
// the lists before each rebuild inside the segment), then runs the adjoint
// recurrence over the segment's frames in reverse, restoring each period's lists.

size_t list_words(const lima_engine* E) {
    const lima_run_params& P = E->params;
    return (size_t)E->NC * (2 + P.max_compound_neighbors + P.max_compound_solvents) +
           (size_t)E->NS * (1 + E->ss_ld());
}

void d2d(void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (bytes) HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
}

// full lists <-> one slot of words
void save_lists(lima_engine* E, int* dst) {
    const lima_run_params& P = E->params;
    const size_t nc = E->NC, ns = E->NS;
    hipStream_t st = E->stream;
    d2d(dst, E->cc_cnt.p, nc * 4, st); dst += nc;
    d2d(dst, E->cc_ids.p, nc * P.max_compound_neighbors * 4, st); dst += nc * P.max_compound_neighbors;
    d2d(dst, E->cs_cnt.p, nc * 4, st); dst += nc;
    d2d(dst, E->cs_ids.p, nc * P.max_compound_solvents * 4, st); dst += nc * P.max_compound_solvents;
    d2d(dst, E->ss_cnt.p, ns * 4, st); dst += ns;
    d2d(dst, E->ss_ids.p, ns * E->ss_ld() * 4, st);
}
void load_lists(lima_engine* E, const int* src) {
    const lima_run_params& P = E->params;
    const size_t nc = E->NC, ns = E->NS;
    hipStream_t st = E->stream;
    d2d(E->cc_cnt.p, src, nc * 4, st); src += nc;
    d2d(E->cc_ids.p, src, nc * P.max_compound_neighbors * 4, st); src += nc * P.max_compound_neighbors;
    d2d(E->cs_cnt.p, src, nc * 4, st); src += nc;
    d2d(E->cs_ids.p, src, nc * P.max_compound_solvents * 4, st); src += nc * P.max_compound_solvents;
    d2d(E->ss_cnt.p, src, ns * 4, st); src += ns;
    d2d(E->ss_ids.p, src, ns * E->ss_ld() * 4, st);
    E->launch_work_list(E->nlist_args(E->cur), st);   // listed work list of these lists
}
size_t cc_words(const lima_engine* E) { return (size_t)E->NC * (1 + E->params.max_compound_neighbors); }
void save_cc(lima_engine* E, int* dst) {
    d2d(dst, E->cc_cnt.p, (size_t)E->NC * 4, E->stream);
    d2d(dst + E->NC, E->cc_ids.p, (size_t)E->NC * E->params.max_compound_neighbors * 4, E->stream);
}
void load_cc(lima_engine* E, const int* src) {
    d2d(E->cc_cnt.p, src, (size_t)E->NC * 4, E->stream);
    d2d(E->cc_ids.p, src + E->NC, (size_t)E->NC * E->params.max_compound_neighbors * 4, E->stream);
}
// state (positions of the current parity and the previous positions) <-> a checkpoint
void save_state(lima_engine* E, float4* dst) {
    const size_t nch = E->NCH, ns = E->NS;
    d2d(dst, E->cpos[E->cur].p, nch * 16, E->stream);
    d2d(dst + nch, E->cprev.p, nch * 16, E->stream);
    d2d(dst + 2 * nch, E->spos[E->cur].p, ns * 16, E->stream);
    d2d(dst + 2 * nch + ns, E->sprev.p, ns * 16, E->stream);
}
void load_state(lima_engine* E, const float4* src) {
    const size_t nch = E->NCH, ns = E->NS;
    d2d(E->cpos[E->cur].p, src, nch * 16, E->stream);
    d2d(E->cprev.p, src + nch, nch * 16, E->stream);
    d2d(E->spos[E->cur].p, src + 2 * nch, ns * 16, E->stream);
    d2d(E->sprev.p, src + 2 * nch + ns, ns * 16, E->stream);
}

template <typename T> void ensure(DBuf<T>& b, size_t n) {
    if (b.n < n) b.alloc(n);
}

// per-particle constants of the adjoint (once per engine)
void adj_setup(lima_engine* E) {
    AdjBufs& B = E->adjb;
    if (B.ready) return;
    const int N = E->NCH + E->NS;
    std::vector<float> cfac(std::max(N, 1));
    std::vector<int> prep(std::max(N, 1));
    std::vector<float> mass_h((size_t)E->R * kMaxTypes);
    HIP_TRY(hipMemcpy(mass_h.data(), E->mass.p, mass_h.size() * 4, hipMemcpyDeviceToHost));
    std::vector<float4> cpos_h(std::max(E->NCH, 1));
    if (E->NCH) HIP_TRY(hipMemcpy(cpos_h.data(), E->cpos[E->cur].p, (size_t)E->NCH * 16, hipMemcpyDeviceToHost));
    int maxc = 1, maxs = 1;
    std::vector<int> mask_off(std::max(E->NC, 1), 0);
    int words = 0;
    for (int r = 0; r < E->R; r++) {
        const ReplicaInfo& ri = E->reps_h[r];
        const double dt = ri.dt;
        for (int i = ri.chain_begin; i < ri.chain_end; i++) {
            int t;
            std::memcpy(&t, &cpos_h[i].w, 4);
            cfac[i] = (float)(dt / (double)mass_h[r * kMaxTypes + t] * dt);
            prep[i] = r;
        }
        for (int s = ri.solv_begin; s < ri.solv_end; s++) {
            cfac[E->NCH + s] = (float)(dt / (double)mass_h[r * kMaxTypes] * dt);
            prep[E->NCH + s] = r;
        }
        maxc = std::max(maxc, ri.comp_end - ri.comp_begin);
        maxs = std::max(maxs, ri.solv_end - ri.solv_begin);
        const int wr = (ri.solv_end - ri.solv_begin + 31) / 32;
        for (int c = ri.comp_begin; c < ri.comp_end; c++) {
            mask_off[c] = words;
            words += wr;
        }
    }
    // split both all-pairs passes into at least ~2048 blocks: solvent tiles x compound chunks of at most
    // 8 compounds, and compounds x solvent slices of a multiple of 256
    const int tiles = std::max(E->n_stiles, 1), comps = std::max(E->NC, 1);
    int nch = std::max((maxc + kAdjCompChunk - 1) / kAdjCompChunk, (2048 + tiles - 1) / tiles);
    nch = std::min(nch, maxc);
    B.comp_chunk = (maxc + nch - 1) / nch;
    B.n_cchunks = (maxc + B.comp_chunk - 1) / B.comp_chunk;
    const int nsl = std::max(1, (2048 + comps - 1) / comps);
    int slice = (maxs + nsl - 1) / nsl;
    slice = std::max(kSolventChunk, (slice + kSolventChunk - 1) / kSolventChunk * kSolventChunk);
    B.solv_slice = slice;
    B.n_sslices = (maxs + slice - 1) / slice;
    B.gpart.alloc(std::max<size_t>(E->max_slots, 1) * kCompoundBlock);
    B.cfac.upload(cfac);
    B.prep.upload(prep);
    B.mask_off.upload(mask_off);
    B.mask.alloc(std::max(words, 1));
    B.adj.alloc((size_t)3 * std::max(N, 1));
    B.w.alloc(std::max(N, 1));
    B.gxc.alloc((size_t)(1 + B.n_sslices) * std::max(E->NCH, 1));
    B.gxs.alloc((size_t)(B.n_cchunks + 1) * std::max(E->NS, 1));
    B.beta.alloc(E->R);
    B.loss.alloc(E->R);
    B.lacc.alloc(std::max(N, 1));
    B.gpar.alloc((size_t)E->R * kAdjStride);
    B.ready = true;
}

void adj_build_mask(lima_engine* E) {
    AdjBufs& B = E->adjb;
    if (E->NC == 0 || E->NS == 0) return;
    adj_mask_zero_kernel<<<(int)((B.mask.n + 255) / 256), 256, 0, E->stream>>>(B.mask.p, (int)B.mask.n);
    adj_mask_set_kernel<<<E->NC, 256, 0, E->stream>>>(E->step_args(E->cur), B.mask.p, B.mask_off.p);
}

// the residuals 2 beta (F - F*) of the frame just stepped into Rs (and the loss when given);
// targets are the frame's rows of the uploaded chain / solvent target arrays, or none
void adj_residuals(lima_engine* E, float4* rs, const float* tc, const float* ts, bool with_loss, bool solv) {
    AdjBufs& B = E->adjb;
    const int nch = E->NCH, ns = E->NS;
    if (nch > 0)
        adj_residual_kernel<<<(nch + 255) / 256, 256, 0, E->stream>>>(E->cforce.p, tc, B.prep.p, B.beta.p, rs,
                                                                       with_loss ? B.lacc.p : nullptr, nch);
    if (ns > 0) {
        if (solv)
            adj_residual_kernel<<<(ns + 255) / 256, 256, 0, E->stream>>>(E->sforce.p, ts, B.prep.p + nch, B.beta.p,
                                                                          rs + nch, with_loss ? B.lacc.p + nch : nullptr, ns);
        else
            HIP_TRY(hipMemsetAsync(rs + nch, 0, (size_t)ns * 16, E->stream));
    }
}

// all K frames of the targets, once per call, as given (x, y, z per particle)
void adj_upload_targets(lima_engine* E, const lima_adjoint_spec& sp) {
    AdjBufs& B = E->adjb;
    const size_t K = sp.n_frames;
    if (sp.chain_force_target && E->NCH) {
        ensure(B.Tc, K * E->NCH * 3);
        HIP_TRY(hipMemcpy(B.Tc.p, sp.chain_force_target, K * E->NCH * 3 * sizeof(float), hipMemcpyHostToDevice));
    }
    if (sp.solvent_force_target && E->NS) {
        ensure(B.Ts, K * E->NS * 3);
        HIP_TRY(hipMemcpy(B.Ts.p, sp.solvent_force_target, K * E->NS * 3 * sizeof(float), hipMemcpyHostToDevice));
    }
}

// one step of the trajectory (the engine's step, optionally with FP64 energies)
void adj_step(lima_engine* E, bool energy) {
    if (energy) HIP_TRY(hipMemsetAsync(E->energy.p, 0, E->energy.n * sizeof(double), E->stream));
    E->launch_step(E->cur, energy);
    E->cur ^= 1;
    E->step++;
    E->since_rebuild++;
}

// one frame of the adjoint recurrence: positions X, residual Rs, energy weights gE
void adj_frame(lima_engine* E, int j, const float4* X, const float4* rs, const float* gE, bool solv) {
    AdjBufs& B = E->adjb;
    const int nch = E->NCH, ns = E->NS, N = nch + ns;
    double4* a_out = B.adj.p + (size_t)(j % 3) * N;
    const double4* a1 = B.adj.p + (size_t)((j + 1) % 3) * N;
    const double4* a2 = B.adj.p + (size_t)((j + 2) % 3) * N;
    adj_weight_kernel<<<(N + 255) / 256, 256, 0, E->stream>>>(rs, a1, B.cfac.p, B.w.p, N);
    AdjArgs A{};
    A.s = E->step_args(E->cur);
    A.s.cpos = X;
    A.s.spos = X + nch;
    A.wc = B.w.p;
    A.ws = B.w.p + nch;
    A.gE = gE;
    A.gxc = B.gxc.p;
    A.gxs = B.gxs.p;
    A.gpar = B.gpar.p;
    A.cs_mask = B.mask.p;
    A.mask_off = B.mask_off.p;
    A.n_chain = nch;
    A.gpart = B.gpart.p;
    A.comp_chunk = B.comp_chunk;
    A.n_cchunks = B.n_cchunks;
    A.solv_slice = B.solv_slice;
    const bool sv = solv && ns > 0;
    if (sv) adj_solvent_kernel<<<dim3(E->n_stiles, B.n_cchunks + 1), 256, 0, E->stream>>>(A);
    if (nch > 0) {
        if (sv) adj_chain_solvent_kernel<<<dim3(E->NC, B.n_sslices), kCompoundBlock, 0, E->stream>>>(A);
        adj_listed_kernel<<<2048, kCompoundBlock, 0, E->stream>>>(A);
        adj_compound_kernel<<<E->NC, kCompoundBlock, 0, E->stream>>>(A);
        if (E->NB > 0) adj_bridge_kernel<<<(E->NB + 63) / 64, 64, 0, E->stream>>>(A);
    }
    adj_update_kernel<<<(N + 255) / 256, 256, 0, E->stream>>>(B.gxc.p, nch > 0 ? 1 + (sv ? B.n_sslices : 0) : 0, B.gxs.p,
                                                              sv ? B.n_cchunks + 1 : 0, a1, a2, a_out, nch, ns);
}

struct SavedRun {
    int cur;
    long long step, since, nreb;
};

int adjoint_run(lima_engine* E, const lima_adjoint_spec& sp, double* loss_out, double* grad_out) {
    const lima_run_params& P = E->params;
    if (P.thermostat != LIMA_THERMOSTAT_NONE)
        throw LimaError(LIMA_E_UNSUPPORTED, "the adjoint differentiates the plain Verlet step: use LIMA_THERMOSTAT_NONE");
    const int K = sp.n_frames;
    if (K < 1) throw LimaError(LIMA_E_ARG, "n_frames must be >= 1");
    if (!loss_out) throw LimaError(LIMA_E_ARG, "null loss output");
    const int R = E->R, NCH = E->NCH, NS = E->NS, N = NCH + NS, NC = E->NC;
    const bool solv = P.enable_solvents && NS > 0;
    const int iv = P.nlist_interval;
    if (sp.checkpoint_interval < 0) throw LimaError(LIMA_E_ARG, "checkpoint_interval must be >= 0");
    HIP_TRY(hipSetDevice(E->device));
    HIP_TRY(hipStreamSynchronize(E->stream));
    adj_setup(E);
    AdjBufs& B = E->adjb;
    const long long since0 = E->since_rebuild;
    auto rebuild_after = [&](int j) { return iv > 0 && (since0 + j + 1) % iv == 0; };
    const size_t LW = std::max<size_t>(list_words(E), 1), CW = std::max<size_t>(cc_words(E), 1);
    // checkpoint_interval 0: one segment (nothing recomputed) when its frames, residuals and lists
    // take at most half of the free device memory, else segments of 20 frames
    int cpi = sp.checkpoint_interval;
    if (cpi == 0) {
        int rebuilds = 0;
        for (int j = 0; j + 1 < K; j++) rebuilds += rebuild_after(j) ? 1 : 0;
        const size_t need = (size_t)K * std::max(N, 1) * 2 * sizeof(float4) + (size_t)(1 + rebuilds) * LW * sizeof(int);
        const size_t held = B.X.n * sizeof(float4) + B.Rs.n * sizeof(float4) + B.lists.n * sizeof(int);
        size_t free_b = 0, total_b = 0;
        HIP_TRY(hipMemGetInfo(&free_b, &total_b));
        cpi = need <= held + free_b / 2 ? K : 20;
    }

    // the symmetric-pair terms need symmetric compound lists (the engine builds them so)
    if (NC > 0) {
        std::vector<int> cnt(NC), ids((size_t)NC * P.max_compound_neighbors);
        HIP_TRY(hipMemcpy(cnt.data(), E->cc_cnt.p, NC * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(ids.data(), E->cc_ids.p, ids.size() * 4, hipMemcpyDeviceToHost));
        std::vector<std::vector<int>> sets(NC);
        for (int c = 0; c < NC; c++) sets[c].assign(&ids[(size_t)c * P.max_compound_neighbors], &ids[(size_t)c * P.max_compound_neighbors] + cnt[c]);
        for (int c = 0; c < NC; c++)
            for (int nb : sets[c])
                if (nb < 0 || nb >= NC || std::find(sets[nb].begin(), sets[nb].end(), c) == sets[nb].end())
                    throw LimaError(LIMA_E_ARG, "the adjoint needs symmetric compound neighbour lists");
    }

    const SavedRun saved{E->cur, E->step, E->since_rebuild, E->n_rebuilds};
    // segment starts
    std::vector<int> starts{0};
    for (int j = 0; j + 1 < K; j++) {
        const bool boundary = iv > 0 ? rebuild_after(j) : true;
        if (boundary && j + 1 - starts.back() >= cpi) starts.push_back(j + 1);
    }
    const int nseg = (int)starts.size();
    int max_len = 0, max_inner = 0;
    for (int s = 0; s < nseg; s++) {
        const int k0 = starts[s], k1 = s + 1 < nseg ? starts[s + 1] : K;
        max_len = std::max(max_len, k1 - k0);
        int inner = 0;
        for (int j = k0; j + 1 < k1; j++) inner += rebuild_after(j) ? 1 : 0;
        max_inner = std::max(max_inner, inner);
    }
    ensure(B.ck, (size_t)nseg * 2 * std::max(N, 1));
    ensure(B.ck_cc, (size_t)nseg * CW);
    ensure(B.lists, (size_t)(1 + max_inner) * LW);
    ensure(B.X, (size_t)max_len * std::max(N, 1));
    ensure(B.Rs, (size_t)max_len * std::max(N, 1));
    ensure(B.efr, (size_t)K * R * kEnergyTerms);
    ensure(B.gE, (size_t)K * R);

    std::vector<double> beta(R, 0.0), alpha(R, 0.0);
    for (int r = 0; r < R; r++) {
        if (sp.force_weight) beta[r] = sp.force_weight[r];
        if (sp.energy_weight) alpha[r] = sp.energy_weight[r];
    }
    HIP_TRY(hipMemcpy(B.beta.p, beta.data(), R * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemsetAsync(B.lacc.p, 0, (size_t)N * 8, E->stream));
    adj_upload_targets(E, sp);
    auto tgt_c = [&](int j) -> const float* {
        return sp.chain_force_target && NCH ? B.Tc.p + (size_t)j * NCH * 3 : nullptr;
    };
    auto tgt_s = [&](int j) -> const float* {
        return sp.solvent_force_target && NS ? B.Ts.p + (size_t)j * NS * 3 : nullptr;
    };

    // ---- forward pass: checkpoints, energies, force loss ----
    // The last segment is the first one reversed: when a gradient is wanted, its frames,
    // residuals and interior lists are kept here instead of being recomputed.
    std::vector<int> kept_period_end;
    save_lists(E, B.lists.p);
    for (int s = 0; s < nseg; s++) {
        const int k0 = starts[s], k1 = s + 1 < nseg ? starts[s + 1] : K;
        const bool keep = grad_out && s == nseg - 1;
        save_state(E, B.ck.p + (size_t)s * 2 * N);
        for (int j = k0; j < k1; j++) {
            if (keep) {
                float4* X = B.X.p + (size_t)(j - k0) * N;
                d2d(X, E->cpos[E->cur].p, (size_t)NCH * 16, E->stream);
                d2d(X + NCH, E->spos[E->cur].p, (size_t)NS * 16, E->stream);
            }
            adj_step(E, true);
            d2d(B.efr.p + (size_t)j * R * kEnergyTerms, E->energy.p, (size_t)R * kEnergyTerms * 8, E->stream);
            adj_residuals(E, keep ? B.Rs.p + (size_t)(j - k0) * N : B.Rs.p, tgt_c(j), tgt_s(j), true, solv);
            if (j + 1 < K && rebuild_after(j)) {
                if (s + 1 < nseg && starts[s + 1] == j + 1) save_cc(E, B.ck_cc.p + (size_t)(s + 1) * CW);
                if (keep) {
                    save_lists(E, B.lists.p + (size_t)(1 + kept_period_end.size()) * LW);
                    kept_period_end.push_back(j);
                }
                E->launch_rebuild(E->cur);
                E->since_rebuild = 0;
            }
        }
    }
    adj_loss_reduce_kernel<<<R, 256, 0, E->stream>>>(E->reps.p, B.lacc.p, NCH, B.loss.p);
    HIP_TRY(hipStreamSynchronize(E->stream));
    int st = E->check_errors();
    std::vector<double> efr((size_t)K * R * kEnergyTerms), floss(R);
    HIP_TRY(hipMemcpy(efr.data(), B.efr.p, efr.size() * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(floss.data(), B.loss.p, R * 8, hipMemcpyDeviceToHost));
    std::vector<float> gE((size_t)K * R);
    for (int r = 0; r < R; r++) loss_out[r] = floss[r];
    for (int k = 0; k < K; k++)
        for (int r = 0; r < R; r++) {
            const double* v = &efr[((size_t)k * R + r) * kEnergyTerms];
            const double pot = v[kE_bond] + v[kE_angle] + v[kE_dih] +
                               0.5 * (v[kE_intra] + v[kE_inter] + v[kE_cs] + v[kE_sc] + v[kE_ss]);
            const double tgt = sp.energy_target ? sp.energy_target[(size_t)k * R + r] : 0.0;
            loss_out[r] += alpha[r] * (pot - tgt) * (pot - tgt);
            gE[(size_t)k * R + r] = (float)(2.0 * alpha[r] * (pot - tgt));
        }

    // ---- backward pass ----
    if (st == LIMA_OK && grad_out) {
        HIP_TRY(hipMemcpy(B.gE.p, gE.data(), gE.size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemsetAsync(B.adj.p, 0, (size_t)3 * N * sizeof(double4), E->stream));
        HIP_TRY(hipMemsetAsync(B.gpar.p, 0, (size_t)R * kAdjStride * 8, E->stream));
        for (int s = nseg - 1; s >= 0; s--) {
            const int k0 = starts[s], k1 = s + 1 < nseg ? starts[s + 1] : K;
            std::vector<int> period_end;
            if (s == nseg - 1) {
                // kept by the forward pass; the device holds the segment's last lists
                period_end = kept_period_end;
            } else {
                load_state(E, B.ck.p + (size_t)s * 2 * N);
                if (k0 == 0 || iv <= 0) {
                    load_lists(E, B.lists.p);
                    E->since_rebuild = iv > 0 ? (since0 + k0) % iv : 0;
                } else {
                    load_cc(E, B.ck_cc.p + (size_t)s * CW);
                    E->launch_rebuild(E->cur);
                    E->since_rebuild = 0;
                }
            }
            // recompute the segment
            for (int j = k0; j < k1 && s != nseg - 1; j++) {
                float4* X = B.X.p + (size_t)(j - k0) * N;
                d2d(X, E->cpos[E->cur].p, (size_t)NCH * 16, E->stream);
                d2d(X + NCH, E->spos[E->cur].p, (size_t)NS * 16, E->stream);
                adj_step(E, false);
                adj_residuals(E, B.Rs.p + (size_t)(j - k0) * N, tgt_c(j), tgt_s(j), false, solv);
                if (j + 1 < k1 && rebuild_after(j)) {
                    save_lists(E, B.lists.p + (size_t)(1 + period_end.size()) * LW);
                    period_end.push_back(j);
                    E->launch_rebuild(E->cur);
                    E->since_rebuild = 0;
                }
            }
            // reverse over the segment's frames
            int q = (int)period_end.size();
            bool mask_dirty = true;
            for (int j = k1 - 1; j >= k0; j--) {
                if (q > 0 && j == period_end[q - 1]) {
                    load_lists(E, B.lists.p + (size_t)q * LW);
                    q--;
                    mask_dirty = true;
                }
                if (mask_dirty) {
                    adj_build_mask(E);
                    mask_dirty = false;
                }
                adj_frame(E, j, B.X.p + (size_t)(j - k0) * N, B.Rs.p + (size_t)(j - k0) * N, B.gE.p + (size_t)j * R, solv);
            }
        }
        HIP_TRY(hipStreamSynchronize(E->stream));
        st = E->check_errors();
        std::vector<double> gp((size_t)R * kAdjStride);
        HIP_TRY(hipMemcpy(gp.data(), B.gpar.p, gp.size() * 8, hipMemcpyDeviceToHost));
        for (int r = 0; r < R; r++) {
            const double* g = &gp[(size_t)r * kAdjStride];
            double* o = grad_out + (size_t)r * LIMA_N_PARAMETERS;
            const lima_particle_params* pp = E->ff_h[r].particle_parameters;
            for (int t = 0; t < kMaxTypes; t++) {
                o[t] = g[t];
                const double eps = pp[t].epsilon;
                o[kMaxTypes + t] = 12.0 * g[kMaxTypes + t] + (eps > 0 ? g[2 * kMaxTypes + t] / (2.0 * eps) : 0.0);
            }
            for (int k = 0; k < 4; k++) o[2 * kMaxTypes + k] = g[kAdjTypeBins + k];
        }
    }

    // ---- leave the engine as it was ----
    E->cur = saved.cur;
    load_state(E, B.ck.p);
    load_lists(E, B.lists.p);
    HIP_TRY(hipStreamSynchronize(E->stream));
    E->step = saved.step;
    E->since_rebuild = saved.since;
    E->n_rebuilds = saved.nreb;
    return st;
}

// forces and potential energies of frames 0..K-1 from the current state
int record_frames(lima_engine* E, int K, float* cf, float* sf, double* pot) {
    if (K < 1) throw LimaError(LIMA_E_ARG, "n_frames must be >= 1");
    const int R = E->R, NCH = E->NCH, NS = E->NS, N = NCH + NS;
    const int iv = E->params.nlist_interval;
    HIP_TRY(hipSetDevice(E->device));
    HIP_TRY(hipStreamSynchronize(E->stream));
    AdjBufs& B = E->adjb;
    ensure(B.ck, (size_t)2 * std::max(N, 1));
    ensure(B.lists, std::max<size_t>(list_words(E), 1));
    const SavedRun saved{E->cur, E->step, E->since_rebuild, E->n_rebuilds};
    save_state(E, B.ck.p);
    save_lists(E, B.lists.p);
    // the thermostat scalars are part of the state handed back
    std::vector<float> lam(R), lam_applied(R);
    HIP_TRY(hipMemcpy(lam.data(), E->lambda.p, R * sizeof(float), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(lam_applied.data(), E->lambda_applied.p, R * sizeof(float), hipMemcpyDeviceToHost));
    E->sync_step_count();
    std::vector<float4> hc(std::max(NCH, 1)), hs(std::max(NS, 1));
    std::vector<double> e((size_t)R * kEnergyTerms);
    int st = LIMA_OK;
    for (int j = 0; j < K && st == LIMA_OK; j++) {
        adj_step(E, true);
        HIP_TRY(hipStreamSynchronize(E->stream));
        st = E->check_errors();
        if (cf && NCH) {
            HIP_TRY(hipMemcpy(hc.data(), E->cforce.p, (size_t)NCH * 16, hipMemcpyDeviceToHost));
            for (int i = 0; i < NCH; i++) {
                float* o = cf + ((size_t)j * NCH + i) * 3;
                o[0] = hc[i].x; o[1] = hc[i].y; o[2] = hc[i].z;
            }
        }
        if (sf && NS) {
            const bool solv = E->params.enable_solvents;
            if (solv) HIP_TRY(hipMemcpy(hs.data(), E->sforce.p, (size_t)NS * 16, hipMemcpyDeviceToHost));
            for (int i = 0; i < NS; i++) {
                float* o = sf + ((size_t)j * NS + i) * 3;
                o[0] = solv ? hs[i].x : 0.f; o[1] = solv ? hs[i].y : 0.f; o[2] = solv ? hs[i].z : 0.f;
            }
        }
        if (pot) {
            HIP_TRY(hipMemcpy(e.data(), E->energy.p, e.size() * 8, hipMemcpyDeviceToHost));
            for (int r = 0; r < R; r++) {
                const double* v = &e[(size_t)r * kEnergyTerms];
                pot[(size_t)j * R + r] = v[kE_bond] + v[kE_angle] + v[kE_dih] +
                                         0.5 * (v[kE_intra] + v[kE_inter] + v[kE_cs] + v[kE_sc] + v[kE_ss]);
            }
        }
        if (iv > 0 && E->since_rebuild >= iv) {
            E->launch_rebuild(E->cur);
            E->since_rebuild = 0;
        }
    }
    E->cur = saved.cur;
    load_state(E, B.ck.p);
    load_lists(E, B.lists.p);
    HIP_TRY(hipStreamSynchronize(E->stream));
    HIP_TRY(hipMemcpy(E->lambda.p, lam.data(), R * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(E->lambda_applied.p, lam_applied.data(), R * sizeof(float), hipMemcpyHostToDevice));
    E->step = saved.step;
    E->since_rebuild = saved.since;
    E->n_rebuilds = saved.nreb;
    return st;
}

// parameter vector of one system <-> force field and shared bonded constants
void get_parameters(lima_engine* E, int r, double* o) {
    const lima_particle_params* pp = E->ff_h[r].particle_parameters;
    for (int t = 0; t < kMaxTypes; t++) {
        o[t] = pp[t].sigma;
        o[kMaxTypes + t] = pp[t].epsilon;
    }
    const ReplicaInfo& ri = E->reps_h[r];
    const int b0 = E->bond_off_h[ri.comp_begin], b1 = E->bond_off_h[ri.comp_end];
    const int a0 = E->angle_off_h[ri.comp_begin], a1 = E->angle_off_h[ri.comp_end];
    o[64] = o[65] = o[66] = o[67] = 0.0;
    const BondD* bd = b1 > b0 ? &E->bonds_h[b0] : nullptr;
    const AngleD* ad = a1 > a0 ? &E->angles_h[a0] : nullptr;
    for (int b = 0; b < E->NB && (!bd || !ad); b++) {
        if (E->br_rep_h[b] != r) continue;
        if (!bd && E->brb_off_h[b + 1] > E->brb_off_h[b]) bd = &E->brb_h[E->brb_off_h[b]];
        if (!ad && E->bra_off_h[b + 1] > E->bra_off_h[b]) ad = &E->bra_h[E->bra_off_h[b]];
    }
    if (bd) { o[64] = bd->kb; o[65] = bd->b0; }
    if (ad) { o[66] = ad->k; o[67] = ad->t0; }
}

// Calls f(kb, b0) for every pair bond of system r (kind 0), or f(k_theta, theta_0) for every
// angle (kind 1), compound terms first, then the bridges' in bridge order.
template <typename F> void for_each_bonded(lima_engine* E, int r, int kind, F&& f) {
    const ReplicaInfo& ri = E->reps_h[r];
    if (kind == 0) {
        for (int b = E->bond_off_h[ri.comp_begin]; b < E->bond_off_h[ri.comp_end]; b++) f(E->bonds_h[b].kb, E->bonds_h[b].b0);
    } else {
        for (int a = E->angle_off_h[ri.comp_begin]; a < E->angle_off_h[ri.comp_end]; a++) f(E->angles_h[a].k, E->angles_h[a].t0);
    }
    for (int b = 0; b < E->NB; b++) {
        if (E->br_rep_h[b] != r) continue;
        if (kind == 0) {
            for (int t = E->brb_off_h[b]; t < E->brb_off_h[b + 1]; t++) f(E->brb_h[t].kb, E->brb_h[t].b0);
        } else {
            for (int t = E->bra_off_h[b]; t < E->bra_off_h[b + 1]; t++) f(E->bra_h[t].k, E->bra_h[t].t0);
        }
    }
}

void set_parameters(lima_engine* E, int r, const double* v) {
    for (int i = 0; i < LIMA_N_PARAMETERS; i++)
        if (!std::isfinite(v[i])) throw LimaError(LIMA_E_ARG, "parameters must be finite");
    // a type slot the force field leaves empty (sigma 0) may be passed back as it is; a new value
    // must be a valid one
    const lima_particle_params* pp_cur = E->ff_h[r].particle_parameters;
    for (int t = 0; t < kMaxTypes; t++) {
        const bool same = (float)v[t] == pp_cur[t].sigma && (float)v[kMaxTypes + t] == pp_cur[t].epsilon;
        if (!same && (v[t] <= 0 || v[kMaxTypes + t] < 0)) throw LimaError(LIMA_E_ARG, "sigma must be > 0 and epsilon >= 0");
    }
    // Entries 64-67 are one constant pair per kind shared by every term.  A system whose terms
    // carry their own constants (an ingested molecule) keeps them when the entries are passed
    // back unchanged, and refuses a change, which would replace the whole bonded force field.
    double cur[LIMA_N_PARAMETERS];
    get_parameters(E, r, cur);
    bool write_kind[2] = {false, false};
    for (int kind = 0; kind < 2; kind++) {
        const float k_new = (float)v[64 + 2 * kind], x_new = (float)v[65 + 2 * kind];
        if (k_new == (float)cur[64 + 2 * kind] && x_new == (float)cur[65 + 2 * kind]) continue;
        bool uniform = true;
        const float k0 = (float)cur[64 + 2 * kind], x0 = (float)cur[65 + 2 * kind];
        for_each_bonded(E, r, kind, [&](float& k, float& x) { uniform = uniform && k == k0 && x == x0; });
        if (!uniform)
            throw LimaError(LIMA_E_UNSUPPORTED, kind == 0 ? "the system's pair bonds differ per term: kb, b0 (entries 64, 65) cannot be set"
                                                          : "the system's angles differ per term: k_theta, theta_0 (entries 66, 67) cannot be set");
        write_kind[kind] = true;
    }
    HIP_TRY(hipSetDevice(E->device));
    HIP_TRY(hipStreamSynchronize(E->stream));
    lima_particle_params* pp = E->ff_h[r].particle_parameters;
    for (int t = 0; t < kMaxTypes; t++) {
        pp[t].sigma = (float)v[t];
        pp[t].epsilon = (float)v[kMaxTypes + t];
    }
    std::vector<float> mass(kMaxTypes);
    std::vector<float4> tsc(kMaxTypes), tcs(kMaxTypes), tinter(1024), tintra(1024), tss(1);
    fill_tables(pp, mass.data(), tsc.data(), tcs.data(), tinter.data(), tintra.data(), tss.data());
    HIP_TRY(hipMemcpy(E->tab_sc.p + r * kMaxTypes, tsc.data(), kMaxTypes * 16, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(E->tab_cs.p + r * kMaxTypes, tcs.data(), kMaxTypes * 16, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(E->tab_inter.p + (size_t)r * 1024, tinter.data(), 1024 * 16, hipMemcpyHostToDevice));
    E->fill_ab_tables();
    HIP_TRY(hipMemcpy(E->tab_intra.p + (size_t)r * 1024, tintra.data(), 1024 * 16, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(E->tab_ss.p + r, tss.data(), 16, hipMemcpyHostToDevice));
    for (int kind = 0; kind < 2; kind++) {
        if (!write_kind[kind]) continue;
        const float k_new = (float)v[64 + 2 * kind], x_new = (float)v[65 + 2 * kind];
        for_each_bonded(E, r, kind, [&](float& k, float& x) { k = k_new; x = x_new; });
    }
    auto put = [](auto& dbuf, const auto& host) {
        if (!host.empty()) HIP_TRY(hipMemcpy(dbuf.p, host.data(), host.size() * sizeof(host[0]), hipMemcpyHostToDevice));
    };
    put(E->bonds, E->bonds_h);
    put(E->angles, E->angles_h);
    put(E->brb, E->brb_h);
    put(E->bra, E->bra_h);
}

thread_local std::string g_last_error;

int fail(lima_engine* E, const LimaError& e) {
    if (E) E->last_error = e.msg;
    g_last_error = e.msg;
    return e.code;
}

template <typename F> int guarded(lima_engine* E, F&& f) {
    if (!E) return LIMA_E_ARG;
    try {
        return f();
    } catch (const LimaError& e) {
        return fail(E, e);
    }
}

// ---- autotune cache, keyed by the histogram of replica shapes ----
std::mutex g_tune_mu;
std::map<std::string, LaunchConfig> g_tune_cache;
bool g_tune_loaded = false;

std::string shape_key(const lima_engine* E) {
    std::map<std::pair<int, int>, int> hist;   // (chain particles / 128, solvents / 512) -> replicas
    for (const ReplicaInfo& ri : E->reps_h)
        hist[{(ri.chain_end - ri.chain_begin + 64) / 128, (ri.solv_end - ri.solv_begin + 256) / 512}]++;
    std::ostringstream os;
    os << "acc" << E->params.accumulation << "_sol" << E->params.enable_solvents;
    for (const auto& kv : hist) os << "_" << kv.first.first << "x" << kv.first.second << ":" << kv.second;
    return os.str();
}

std::string g_tune_path;

// reads the cache file once per path (a process may point LIMA_AUTOTUNE_CACHE elsewhere later)
void load_tune_file() {
    const char* path = std::getenv("LIMA_AUTOTUNE_CACHE");
    const std::string p = path ? path : "";
    if (g_tune_loaded && p == g_tune_path) return;
    g_tune_loaded = true;
    g_tune_path = p;
    if (!path) return;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        // key spt chunk_particles listed_grid [listed_lanes [listed_packed [fused_thermostat]]]
        std::istringstream is(line);
        std::string key;
        LaunchConfig c;
        if (!(is >> key >> c.spt >> c.chunk_particles >> c.listed_grid)) continue;
        if (!(is >> c.listed_lanes)) c.listed_lanes = 1;
        if (!(is >> c.listed_packed)) c.listed_packed = 1;
        if (!(is >> c.fused_thermostat)) c.fused_thermostat = 0;
        // only shapes the autotuner itself can pick: any other line is skipped
        const bool ok = (c.spt == 1 || c.spt == 2 || c.spt == 4) &&
                        (c.chunk_particles == 256 || c.chunk_particles == 512 || c.chunk_particles == 1024 ||
                         c.chunk_particles == 2048) && c.chunk_particles <= kChunkParticles &&
                        (c.listed_grid == 1024 || c.listed_grid == 2048 || c.listed_grid == 4096 || c.listed_grid == 8192) &&
                        (c.listed_lanes == 1 || c.listed_lanes == kSplitLanes) && (c.listed_packed == 0 || c.listed_packed == 1 || c.listed_packed == 2) &&
                        (c.fused_thermostat >= 0 && c.fused_thermostat <= 2);
        if (!ok) continue;
        g_tune_cache[key] = c;
    }
}

void save_tune_entry(const std::string& key, const LaunchConfig& c) {
    const char* path = std::getenv("LIMA_AUTOTUNE_CACHE");
    if (!path) return;
    std::ofstream f(path, std::ios::app);
    f << key << " " << c.spt << " " << c.chunk_particles << " " << c.listed_grid << " " << c.listed_lanes << " "
      << c.listed_packed << " " << c.fused_thermostat << "\n";
}

// one system's energy record from the device sums (half of every directed-pair LJ sum)
void fill_energies(const lima_engine* E, int r, const double* v, float lam, lima_energies& o) {
    const ReplicaInfo& ri = E->reps_h[r];
    std::memset(&o, 0, sizeof(o));
    o.bond = v[kE_bond]; o.angle = v[kE_angle]; o.dihedral = v[kE_dih];
    o.lj_intra = 0.5 * v[kE_intra]; o.lj_inter = 0.5 * v[kE_inter];
    o.lj_compound_solvent = 0.5 * v[kE_cs]; o.lj_solvent_compound = 0.5 * v[kE_sc];
    o.lj_solvent_solvent = 0.5 * v[kE_ss];
    o.potential = o.bond + o.angle + o.dihedral + o.lj_intra + o.lj_inter + o.lj_compound_solvent +
                  o.lj_solvent_compound + o.lj_solvent_solvent;
    o.kinetic = 0.5 * v[kE_kin] / (ri.dt * ri.dt);
    const double ntot = (double)((ri.chain_end - ri.chain_begin) + (ri.solv_end - ri.solv_begin));
    o.temperature = ntot > 0 ? v[kE_kin] / (ri.dt * ri.dt) / (3.0 * ntot * 8.3145) : 0.0;
    o.virial = 0.5 * v[kE_virial];
    o.thermostat_scalar = lam;
}

void get3(const std::vector<float4>& v, float* out) {
    for (size_t i = 0; i < v.size(); i++) { out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z; }
}

// The steps of lima_step_observed: record steps launched directly with their energy sums, each
// record's sums and lambda copied into series [record][system][kEnergyTerms] and lser
// [record][system] (every slot of a record is written before it is read), the plain steps
// between records as graph replays.  allow_capture false keeps to launches and replays of
// graphs already built (no capture or instantiation).
void observed_steps(lima_engine* E, int n, int every, double* series, float* lser, bool allow_capture) {
    const int R = E->R;
    int rec = 0;
    E->sync_step_count();
    for (int s = 0; s < n;) {
        // the recorded step: the step itself with its energy sums (Engine::step + the Analyzer's sums)
        HIP_TRY(hipMemsetAsync(E->energy.p, 0, E->energy.n * sizeof(double), E->stream));
        E->launch_step(E->cur, true);
        HIP_TRY(hipMemcpyAsync(series + (size_t)rec * R * kEnergyTerms, E->energy.p, (size_t)R * kEnergyTerms * 8,
                               hipMemcpyDeviceToDevice, E->stream));
        HIP_TRY(hipMemcpyAsync(lser + (size_t)rec * R, E->lambda_applied.p, (size_t)R * 4, hipMemcpyDeviceToDevice,
                               E->stream));
        E->cur ^= 1;
        E->step++;
        E->since_rebuild++;
        if (E->params.nlist_interval > 0 && E->since_rebuild >= E->params.nlist_interval) {
            E->launch_rebuild(E->cur);
            E->since_rebuild = 0;
            E->n_rebuilds++;
        }
        rec++;
        s++;
        // the plain steps up to the next record (graph replays where the cycle aligns)
        const int run = std::min(every - 1, n - s);
        if (run > 0) E->enqueue_steps(run, allow_capture);
        s += run;
    }
}

}  // namespace

extern "C" {

const char* lima_version(void) { return "lima-mi355x 0.1.0 (gfx950)"; }

int lima_device_count(int* count) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (count) *count = n;
    return LIMA_OK;
}

int lima_create(const lima_system* systems, int32_t n_systems, const lima_run_params* params, lima_engine** out) {
    if (!out || !systems || !params) return LIMA_E_ARG;
    *out = nullptr;
    lima_engine* E = new lima_engine();
    E->params = *params;
    try {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw LimaError(LIMA_E_HIP, "no HIP device available");
        if (params->device < 0 || params->device >= ndev) throw LimaError(LIMA_E_ARG, "device ordinal out of range");
        E->device = params->device;
        HIP_TRY(hipSetDevice(E->device));
        E->scratch_d = scratch_address(E->device);
        E->create_streams();
        validate_and_build(E, systems, n_systems);
        E->fill_ab_tables();
        // Engine::Engine -> handleNLISTS(force) (Engine.cu:35)
        E->launch_rebuild(E->cur);
        HIP_TRY(hipStreamSynchronize(E->stream));
        int st = E->check_errors();
        if (st != LIMA_OK) {
            g_last_error = E->last_error;
            delete E;
            return st;
        }
    } catch (const LimaError& e) {
        g_last_error = e.msg;
        delete E;
        return e.code;
    }
    *out = E;
    return LIMA_OK;
}

void lima_destroy(lima_engine* E) {
    if (!E) return;
    (void)hipSetDevice(E->device);
    (void)hipStreamSynchronize(E->stream);
    delete E;
}

const char* lima_last_error(const lima_engine* E) { return E ? E->last_error.c_str() : g_last_error.c_str(); }

int lima_step_async(lima_engine* E, int32_t n) {
    return guarded(E, [&] {
        if (n < 0) throw LimaError(LIMA_E_ARG, "n_steps < 0");
        HIP_TRY(hipSetDevice(E->device));
        auto t0 = std::chrono::steady_clock::now();
        if (!E->async_open && n > 0) {
            // the device time of queued steps is read at the next lima_sync
            HIP_TRY(hipEventRecord(E->ev_t0, E->stream));
            E->async_open = true;
        }
        E->enqueue_steps(n);
        E->t_host_us += std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
        return LIMA_OK;
    });
}

int lima_sync(lima_engine* E) {
    return guarded(E, [&] {
        HIP_TRY(hipSetDevice(E->device));
        if (E->async_open) {
            HIP_TRY(hipEventRecord(E->ev_t1, E->stream));
            HIP_TRY(hipEventSynchronize(E->ev_t1));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, E->ev_t0, E->ev_t1));
            E->t_device_us += ms * 1000.0;
            E->async_open = false;
        }
        HIP_TRY(hipStreamSynchronize(E->stream));
        return E->check_errors();
    });
}

int lima_step(lima_engine* E, int32_t n) {
    return guarded(E, [&] {
        if (n < 0) throw LimaError(LIMA_E_ARG, "n_steps < 0");
        HIP_TRY(hipSetDevice(E->device));
        if (E->async_open) {   // steps queued by lima_step_async are timed up to here
            int st = lima_sync(E);
            if (st != LIMA_OK) return st;
        }
        auto t0 = std::chrono::steady_clock::now();
        HIP_TRY(hipEventRecord(E->ev_t0, E->stream));
        E->enqueue_steps(n);
        HIP_TRY(hipEventRecord(E->ev_t1, E->stream));
        auto t1 = std::chrono::steady_clock::now();
        HIP_TRY(hipEventSynchronize(E->ev_t1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, E->ev_t0, E->ev_t1));
        E->t_device_us += ms * 1000.0;
        E->t_host_us += std::chrono::duration<double, std::micro>(t1 - t0).count();
        return E->check_errors();
    });
}

int lima_rebuild_neighbors(lima_engine* E) {
    return guarded(E, [&] {
        HIP_TRY(hipSetDevice(E->device));
        E->launch_rebuild(E->cur);
        E->since_rebuild = 0;
        E->n_rebuilds++;
        HIP_TRY(hipStreamSynchronize(E->stream));
        return E->check_errors();
    });
}

int lima_get_state(lima_engine* E, float* cx, float* cxm, float* sx, float* sxm) {
    return guarded(E, [&] {
        HIP_TRY(hipStreamSynchronize(E->stream));
        std::vector<float4> h;
        auto pull = [&](DBuf<float4>& b, int n, float* out) {
            if (!out || n == 0) return;
            h.resize(n);
            HIP_TRY(hipMemcpy(h.data(), b.p, n * sizeof(float4), hipMemcpyDeviceToHost));
            get3(h, out);
        };
        pull(E->cpos[E->cur], E->NCH, cx);
        pull(E->cprev, E->NCH, cxm);
        pull(E->spos[E->cur], E->NS, sx);
        pull(E->sprev, E->NS, sxm);
        return LIMA_OK;
    });
}

int lima_set_state(lima_engine* E, const float* cx, const float* cxm, const float* sx, const float* sxm) {
    return guarded(E, [&] {
        HIP_TRY(hipStreamSynchronize(E->stream));
        std::vector<float4> h;
        auto push = [&](DBuf<float4>& b, int n, const float* in, bool keep_w) {
            if (!in || n == 0) return;
            h.resize(n);
            HIP_TRY(hipMemcpy(h.data(), b.p, n * sizeof(float4), hipMemcpyDeviceToHost));
            for (int i = 0; i < n; i++) h[i] = make_float4(in[3 * i], in[3 * i + 1], in[3 * i + 2], keep_w ? h[i].w : 0.f);
            HIP_TRY(hipMemcpy(b.p, h.data(), n * sizeof(float4), hipMemcpyHostToDevice));
        };
        push(E->cpos[E->cur], E->NCH, cx, true);
        push(E->cprev, E->NCH, cxm, false);
        push(E->spos[E->cur], E->NS, sx, false);
        push(E->sprev, E->NS, sxm, false);
        return LIMA_OK;
    });
}

int lima_get_forces(lima_engine* E, float* cf, float* sf) {
    return guarded(E, [&] {
        HIP_TRY(hipStreamSynchronize(E->stream));
        std::vector<float4> h;
        if (cf && E->NCH) { h.resize(E->NCH); HIP_TRY(hipMemcpy(h.data(), E->cforce.p, E->NCH * sizeof(float4), hipMemcpyDeviceToHost)); get3(h, cf); }
        if (sf && E->NS) { h.resize(E->NS); HIP_TRY(hipMemcpy(h.data(), E->sforce.p, E->NS * sizeof(float4), hipMemcpyDeviceToHost)); get3(h, sf); }
        return LIMA_OK;
    });
}

int lima_compute_forces(lima_engine* E, float* cf, float* sf) {
    return guarded(E, [&] {
        HIP_TRY(hipSetDevice(E->device));
        if (E->params.accumulation == LIMA_ACCUM_FIXED64) E->launch_forces<false, true, false>(E->cur);
        else E->launch_forces<false, false, false>(E->cur);
        HIP_TRY(hipStreamSynchronize(E->stream));
        int st = E->check_errors();
        if (st != LIMA_OK) return st;
        return lima_get_forces(E, cf, sf);
    });
}

int lima_get_energies(lima_engine* E, lima_energies* out) {
    return guarded(E, [&] {
        if (!out) throw LimaError(LIMA_E_ARG, "null output");
        HIP_TRY(hipSetDevice(E->device));
        HIP_TRY(hipMemsetAsync(E->energy.p, 0, E->energy.n * sizeof(double), E->stream));
        if (E->params.accumulation == LIMA_ACCUM_FIXED64) E->launch_forces<false, true, true>(E->cur);
        else E->launch_forces<false, false, true>(E->cur);
        HIP_TRY(hipStreamSynchronize(E->stream));
        std::vector<double> e(E->energy.n);
        HIP_TRY(hipMemcpy(e.data(), E->energy.p, e.size() * sizeof(double), hipMemcpyDeviceToHost));
        std::vector<float> lam(E->R);
        HIP_TRY(hipMemcpy(lam.data(), E->lambda_applied.p, E->R * sizeof(float), hipMemcpyDeviceToHost));
        for (int r = 0; r < E->R; r++) fill_energies(E, r, &e[(size_t)r * kEnergyTerms], lam[r], out[r]);
        return E->check_errors();
    });
}

// Steps with the reference's per-step logs of one system: Engine::offloadLoggingData /
// offloadTrainData (Engine.cu:288-310) move device rings to the host every 100 steps
// (STEPS_PER_TRAINDATATRANSFER, Constants.cuh:99); here the rings are copied into the
// caller's arrays at the same cadence.
constexpr int kLogRing = 100;

int lima_step_logged(lima_engine* E, int32_t n, int32_t system, int32_t slots, float* logdata, float* traindata) {
    return guarded(E, [&] {
        if (n < 0 || system < 0 || system >= E->R) throw LimaError(LIMA_E_ARG, "n_steps < 0 or system out of range");
        if (n > 0 && !logdata && !traindata) throw LimaError(LIMA_E_ARG, "no output");
        const ReplicaInfo& ri = E->reps_h[system];
        const int nc = ri.comp_end - ri.comp_begin;
        int maxp = 0;
        for (int c = ri.comp_begin; c < ri.comp_end; c++) maxp = std::max(maxp, E->comp_off_h[c + 1] - E->comp_off_h[c]);
        if (traindata && (slots < std::max(maxp, 1) || slots > kCompoundBlock))
            throw LimaError(LIMA_E_ARG, "slots_per_compound must cover the largest compound and be at most 128");
        if (nc == 0) throw LimaError(LIMA_E_ARG, "the system has no compounds to log");
        HIP_TRY(hipSetDevice(E->device));
        HIP_TRY(hipStreamSynchronize(E->stream));
        const size_t per_step = (size_t)nc * std::max(slots, 1) * 9;
        if (E->log_data.n < (size_t)kLogRing * 10) E->log_data.alloc((size_t)kLogRing * 10);
        if (traindata && E->train_data.n < kLogRing * per_step) E->train_data.alloc(kLogRing * per_step);
        // the rings enter the captured step through its arguments: recapture with them, and again without
        auto reset_graph = [&] {
            E->drop_graphs();
        };
        struct Off {   // logging is off again however the call ends
            lima_engine* E;
            ~Off() {
                E->log_rep = -1;
                E->log_train = false;
                E->drop_graphs();
            }
        } off{E};
        E->log_rep = system;
        E->log_train = traindata != nullptr;
        E->log_slots = traindata ? slots : 0;
        E->log_ring = kLogRing;
        reset_graph();
        std::vector<float> hl((size_t)kLogRing * 10), ht(traindata ? kLogRing * per_step : 0);
        int st = LIMA_OK;
        for (int done = 0; done < n && st == LIMA_OK;) {
            const int chunk = std::min(kLogRing, n - done);
            const long long step0 = E->step;
            E->enqueue_steps(chunk);
            HIP_TRY(hipStreamSynchronize(E->stream));
            st = E->check_errors();
            HIP_TRY(hipMemcpy(hl.data(), E->log_data.p, hl.size() * 4, hipMemcpyDeviceToHost));
            if (traindata) HIP_TRY(hipMemcpy(ht.data(), E->train_data.p, ht.size() * 4, hipMemcpyDeviceToHost));
            for (int i = 0; i < chunk; i++) {
                const size_t ring = (size_t)((step0 + i) % kLogRing);
                if (logdata) std::memcpy(logdata + (size_t)(done + i) * 10, hl.data() + ring * 10, 40);
                if (traindata) std::memcpy(traindata + (size_t)(done + i) * per_step, ht.data() + ring * per_step, per_step * 4);
            }
            done += chunk;
        }
        return st;
    });
}

int lima_step_observed(lima_engine* E, int32_t n, int32_t every, lima_energies* out) {
    return guarded(E, [&] {
        if (n < 0 || every <= 0) throw LimaError(LIMA_E_ARG, "n_steps < 0 or every <= 0");
        if (!out && n > 0) throw LimaError(LIMA_E_ARG, "null output");
        HIP_TRY(hipSetDevice(E->device));
        const int R = E->R;
        const int nrec = (n + every - 1) / every;
        std::vector<double> e((size_t)std::max(nrec, 1) * R * kEnergyTerms);
        std::vector<float> lam((size_t)std::max(nrec, 1) * R);
        DBuf<double> series;
        DBuf<float> lser;
        series.alloc(e.size());
        lser.alloc(lam.size());
        observed_steps(E, n, every, series.p, lser.p, true);
        HIP_TRY(hipStreamSynchronize(E->stream));
        if (nrec > 0) {
            HIP_TRY(hipMemcpy(e.data(), series.p, (size_t)nrec * R * kEnergyTerms * 8, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(lam.data(), lser.p, (size_t)nrec * R * 4, hipMemcpyDeviceToHost));
        }
        for (int k = 0; k < nrec; k++)
            for (int r = 0; r < R; r++)
                fill_energies(E, r, &e[((size_t)k * R + r) * kEnergyTerms], lam[(size_t)k * R + r], out[(size_t)k * R + r]);
        return E->check_errors();
    });
}

int lima_step_observed_noalloc(lima_engine* E, int32_t n, int32_t every, lima_energies* out) {
    return guarded(E, [&] {
        if (n < 0 || every <= 0) throw LimaError(LIMA_E_ARG, "n_steps < 0 or every <= 0");
        if (!out && n > 0) throw LimaError(LIMA_E_ARG, "null output");
        HIP_TRY(hipSetDevice(E->device));
        const int R = E->R;
        const size_t nrec = (size_t)((n + every - 1) / every);
        const size_t ser_bytes = nrec * R * kEnergyTerms * sizeof(double);
        const size_t lam_off = (ser_bytes + 255) & ~(size_t)255;
        if (lam_off + nrec * R * sizeof(float) > kObservedScratchBytes) {
            char buf[200];
            snprintf(buf, sizeof(buf), "%zu records x %d systems exceed the no-alloc staging of %zu bytes (%zu bytes per record and system)",
                     nrec, R, kObservedScratchBytes, kEnergyTerms * sizeof(double) + sizeof(float));
            throw LimaError(LIMA_E_ARG, buf);
        }
        unsigned char* base = E->scratch_d;
        if (!base) throw LimaError(LIMA_E_UNSUPPORTED, "no-alloc staging is not set up on this device");
        double* series = reinterpret_cast<double*>(base);
        float* lser = reinterpret_cast<float*>(base + lam_off);
        observed_steps(E, n, every, series, lser, false);
        HIP_TRY(hipStreamSynchronize(E->stream));
        // out of the staging in fixed host chunks (no host allocation either)
        constexpr size_t kChunk = 64;
        double e[kChunk * kEnergyTerms];
        float lam[kChunk];
        const size_t total = nrec * R;
        for (size_t i = 0; i < total; i += kChunk) {
            const size_t m = std::min(kChunk, total - i);
            HIP_TRY(hipMemcpy(e, series + i * kEnergyTerms, m * kEnergyTerms * sizeof(double), hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(lam, lser + i, m * sizeof(float), hipMemcpyDeviceToHost));
            for (size_t j = 0; j < m; j++) fill_energies(E, (int)((i + j) % R), &e[j * kEnergyTerms], lam[j], out[i + j]);
        }
        return E->check_errors();
    });
}

int lima_get_neighbor_counts(lima_engine* E, int32_t* ccc, int32_t* csc, int32_t* ssc) {
    return guarded(E, [&] {
        HIP_TRY(hipStreamSynchronize(E->stream));
        if (ccc && E->NC) HIP_TRY(hipMemcpy(ccc, E->cc_cnt.p, E->NC * sizeof(int), hipMemcpyDeviceToHost));
        if (csc && E->NC) HIP_TRY(hipMemcpy(csc, E->cs_cnt.p, E->NC * sizeof(int), hipMemcpyDeviceToHost));
        if (ssc && E->NS) HIP_TRY(hipMemcpy(ssc, E->ss_cnt.p, E->NS * sizeof(int), hipMemcpyDeviceToHost));
        return LIMA_OK;
    });
}

int lima_get_neighbors(lima_engine* E, int32_t* cci, int32_t* csi, int32_t* ssi) {
    return guarded(E, [&] {
        HIP_TRY(hipStreamSynchronize(E->stream));
        const lima_run_params& P = E->params;
        if (cci && E->NC) HIP_TRY(hipMemcpy(cci, E->cc_ids.p, (size_t)E->NC * P.max_compound_neighbors * sizeof(int), hipMemcpyDeviceToHost));
        if (csi && E->NC) HIP_TRY(hipMemcpy(csi, E->cs_ids.p, (size_t)E->NC * P.max_compound_solvents * sizeof(int), hipMemcpyDeviceToHost));
        if (ssi && E->NS) {
            const size_t ld = E->ss_ld();
            std::vector<int> t((size_t)E->NS * ld);
            HIP_TRY(hipMemcpy(t.data(), E->ss_ids.p, t.size() * sizeof(int), hipMemcpyDeviceToHost));
            for (int s = 0; s < E->NS; s++)
                for (int k = 0; k < P.max_solvent_neighbors; k++) ssi[(size_t)s * P.max_solvent_neighbors + k] = t[(size_t)s * ld + k];
        }
        return LIMA_OK;
    });
}

int lima_set_neighbors(lima_engine* E, const int32_t* ccc, const int32_t* cci, const int32_t* csc, const int32_t* csi,
                       const int32_t* ssc, const int32_t* ssi) {
    return guarded(E, [&] {
        HIP_TRY(hipStreamSynchronize(E->stream));
        const lima_run_params& P = E->params;
        // every id must name an entity of the same replica (the kernels index per-replica ranges)
        std::vector<int> comp_rep_h(E->NC), solv_rep_h(E->NS);
        for (int r = 0; r < E->R; r++) {
            for (int c = E->reps_h[r].comp_begin; c < E->reps_h[r].comp_end; c++) comp_rep_h[c] = r;
            for (int s = E->reps_h[r].solv_begin; s < E->reps_h[r].solv_end; s++) solv_rep_h[s] = r;
        }
        if (ccc && cci && E->NC) {
            for (int c = 0; c < E->NC; c++) {
                if (ccc[c] < 0 || ccc[c] > P.max_compound_neighbors) throw LimaError(LIMA_E_CAPACITY, "compound list count exceeds capacity");
                const ReplicaInfo& ri = E->reps_h[comp_rep_h[c]];
                for (int k = 0; k < ccc[c]; k++) {
                    const int j = cci[(size_t)c * P.max_compound_neighbors + k];
                    if (j < ri.comp_begin || j >= ri.comp_end || j == c) throw LimaError(LIMA_E_ARG, "compound neighbour id outside the compound's system");
                }
            }
            HIP_TRY(hipMemcpy(E->cc_cnt.p, ccc, E->NC * sizeof(int), hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(E->cc_ids.p, cci, (size_t)E->NC * P.max_compound_neighbors * sizeof(int), hipMemcpyHostToDevice));
        }
        if (csc && csi && E->NC) {
            for (int c = 0; c < E->NC; c++) {
                if (csc[c] < 0 || csc[c] > P.max_compound_solvents) throw LimaError(LIMA_E_CAPACITY, "compound solvent list count exceeds capacity");
                const ReplicaInfo& ri = E->reps_h[comp_rep_h[c]];
                for (int k = 0; k < csc[c]; k++) {
                    const int s = csi[(size_t)c * P.max_compound_solvents + k];
                    if (s < ri.solv_begin || s >= ri.solv_end) throw LimaError(LIMA_E_ARG, "solvent id outside the compound's system");
                }
            }
            HIP_TRY(hipMemcpy(E->cs_cnt.p, csc, E->NC * sizeof(int), hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(E->cs_ids.p, csi, (size_t)E->NC * P.max_compound_solvents * sizeof(int), hipMemcpyHostToDevice));
        }
        if (ssc && ssi && E->NS) {
            for (int s = 0; s < E->NS; s++) {
                if (ssc[s] < 0 || ssc[s] > P.max_solvent_neighbors) throw LimaError(LIMA_E_CAPACITY, "solvent list count exceeds capacity");
                const ReplicaInfo& ri = E->reps_h[solv_rep_h[s]];
                for (int k = 0; k < ssc[s]; k++) {
                    const int j = ssi[(size_t)s * P.max_solvent_neighbors + k];
                    if (j < ri.solv_begin || j >= ri.solv_end || j == s) throw LimaError(LIMA_E_ARG, "solvent neighbour id outside the solvent's system");
                }
            }
            const size_t ld = E->ss_ld();
            std::vector<int> t((size_t)E->NS * ld, 0);
            for (int s = 0; s < E->NS; s++)
                for (int k = 0; k < ssc[s]; k++) t[(size_t)s * ld + k] = ssi[(size_t)s * P.max_solvent_neighbors + k];
            HIP_TRY(hipMemcpy(E->ss_cnt.p, ssc, E->NS * sizeof(int), hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(E->ss_ids.p, t.data(), t.size() * sizeof(int), hipMemcpyHostToDevice));
        }
        if (E->NC) {
            E->launch_work_list(E->nlist_args(E->cur), E->stream);
            HIP_TRY(hipStreamSynchronize(E->stream));
        }
        return LIMA_OK;
    });
}

int lima_get_timings(lima_engine* E, double* out5) {
    return guarded(E, [&] {
        if (!out5) throw LimaError(LIMA_E_ARG, "null output");
        out5[0] = E->t_device_us;
        out5[1] = E->t_rebuild_us;
        out5[2] = E->t_host_us;
        out5[3] = (double)E->step;
        out5[4] = (double)E->n_rebuilds;
        return LIMA_OK;
    });
}

int lima_time_kernels(lima_engine* E, int32_t reps, double* out8) {
    return guarded(E, [&] {
        if (!out8 || reps <= 0) throw LimaError(LIMA_E_ARG, "bad arguments");
        double* out6 = out8;
        HIP_TRY(hipSetDevice(E->device));
        HIP_TRY(hipStreamSynchronize(E->stream));
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        StepArgs a = E->step_args(E->cur);
        a.kclock = nullptr;   // these launches are not steps: they stay out of the in-step clock
        const bool fx = E->params.accumulation == LIMA_ACCUM_FIXED64;
        auto timeit = [&](auto&& launch) -> double {
            launch();   // warm
            HIP_TRY(hipEventRecord(e0, E->stream));
            for (int i = 0; i < reps; i++) launch();
            HIP_TRY(hipEventRecord(e1, E->stream));
            HIP_TRY(hipEventSynchronize(e1));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            return ms * 1000.0 / reps;
        };
        for (int k = 0; k < 8; k++) out8[k] = 0.0;
        const int nw = (int)E->scwork_h.size();
        if (E->params.enable_solvents && nw > 0) {
            out6[0] = timeit([&] {
                if (fx) E->launch_solvent_chain<true, false>(a, E->stream);
                else E->launch_solvent_chain<false, false>(a, E->stream);
            });
        }
        if (E->params.enable_solvents && E->NS > 0)
            out8[6] = timeit([&] {
                if (!fx && E->cfg.listed_lanes == kSplitLanes)
                    solvent_listed_split_kernel<false><<<xcd_grid(E->n_stiles * kSplitItems), 64, 0, E->stream>>>(a, E->n_stiles * kSplitItems);
                else
                    solvent_listed_kernel<false><<<xcd_grid(E->n_stiles * kListedItems), 64, 0, E->stream>>>(a, E->n_stiles * kListedItems);
            });
        if (E->NC > 0) {
            const int grid = E->cfg.listed_grid > 0 ? E->cfg.listed_grid : 4096;
            out6[1] = timeit([&] {
                if (fx) E->launch_compound_listed<true, false>(a, grid, E->stream);
                else E->launch_compound_listed<false, false>(a, grid, E->stream);
            });
            out6[2] = timeit([&] {
                if (fx) compound_integrate_kernel<false, true, false><<<E->NC, kCompoundBlock, 0, E->stream>>>(a);
                else compound_integrate_kernel<false, false, false><<<E->NC, kCompoundBlock, 0, E->stream>>>(a);
            });
            out8[7] = timeit([&] {
                if (fx) E->launch_compound_local<true, false>(a, E->stream);
                else E->launch_compound_local<false, false>(a, E->stream);
            });
        }
        if (E->params.enable_solvents && E->NS > 0) {
            out6[3] = timeit([&] {
                if (fx) solvent_integrate_kernel<false, true, false><<<E->n_stiles, 256, 0, E->stream>>>(a);
                else solvent_integrate_kernel<false, false, false><<<E->n_stiles, 256, 0, E->stream>>>(a);
            });
        }
        if (E->NB > 0) out6[4] = timeit([&] { bridge_kernel<false><<<E->NB, 64, 0, E->stream>>>(a); });
        out6[5] = timeit([&] { E->launch_rebuild(E->cur); });
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return E->check_errors();
    });
}

int lima_get_config(lima_engine* E, int32_t* out, int32_t n_out) {
    return guarded(E, [&] {
        if (!out || n_out < 0) throw LimaError(LIMA_E_ARG, "null output or negative count");
        const int32_t v[12] = {E->cfg.spt, E->cfg.chunk_particles, (int32_t)E->scwork_h.size(), E->kmax, E->cfg.use_graph,
                               E->n_cells, E->n_stiles, E->R, E->cfg.listed_grid > 0 ? E->cfg.listed_grid : 4096,
                               E->params.accumulation == LIMA_ACCUM_FIXED64 ? 1 : E->cfg.listed_lanes,
                               !E->pack_ppl() ? 0 : E->cfg.listed_packed == 2 && E->params.accumulation != LIMA_ACCUM_FIXED64 ? 2
                                                  : E->cfg.listed_packed == 1 ? 1 : 0,
                               E->thermo_form()};
        for (int i = 0; i < std::min<int32_t>(n_out, 12); i++) out[i] = v[i];
        return LIMA_OK;
    });
}

// Autotuner: picks the all-pairs kernel's solvents per thread and chunk size, and
// the listed kernel's grid, by timing each candidate alone on the current state
// (state and lists are not modified).  Results are cached per shape histogram in
// the process and, when LIMA_AUTOTUNE_CACHE names a file, across processes.
int lima_set_kernel_clock(lima_engine* E, int32_t on) {
    return guarded(E, [&] {
        HIP_TRY(hipSetDevice(E->device));
        HIP_TRY(hipStreamSynchronize(E->stream));
        const bool want = on != 0;
        if (want) {
            ensure(E->kclock, (size_t)4);
            HIP_TRY(hipMemset(E->kclock.p, 0xFF, sizeof(unsigned long long)));
            HIP_TRY(hipMemset(E->kclock.p + 1, 0, 3 * sizeof(unsigned long long)));
        }
        if (want != E->kclock_on) {
            E->kclock_on = want;
            E->drop_graphs();
        }
        return LIMA_OK;
    });
}

int lima_get_kernel_clock(lima_engine* E, double* total_us, int64_t* launches) {
    return guarded(E, [&] {
        if (!total_us || !launches) throw LimaError(LIMA_E_ARG, "null output");
        if (!E->kclock_on) throw LimaError(LIMA_E_ARG, "the kernel clock is off (lima_set_kernel_clock)");
        HIP_TRY(hipSetDevice(E->device));
        HIP_TRY(hipStreamSynchronize(E->stream));
        unsigned long long v[4];
        HIP_TRY(hipMemcpy(v, E->kclock.p, sizeof v, hipMemcpyDeviceToHost));
        int khz = 0;
        HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, E->device));
        if (khz <= 0) throw LimaError(LIMA_E_UNSUPPORTED, "the device reports no wall-clock rate");
        *total_us = (double)v[2] * 1000.0 / khz;
        *launches = (int64_t)v[3];
        return LIMA_OK;
    });
}

int lima_autotune(lima_engine* E, int32_t reps) {
    return guarded(E, [&] {
        if (reps <= 0) reps = 10;
        HIP_TRY(hipSetDevice(E->device));
        HIP_TRY(hipStreamSynchronize(E->stream));
        const std::string key = shape_key(E);
        LaunchConfig best = E->cfg;
        bool have = false;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            load_tune_file();
            auto it = g_tune_cache.find(key);
            if (it != g_tune_cache.end()) { best = it->second; have = true; }
        }
        if (!have) {
            // Each shape is timed on the whole step: captured cycles of all streams, rebuilds
            // included, from the same saved state, which is restored afterwards.  Isolated kernel
            // timings ranked the all-pairs chunk sizes differently from the step they run in.
            hipEvent_t e0, e1;
            HIP_TRY(hipEventCreate(&e0));
            HIP_TRY(hipEventCreate(&e1));
            const int NCH = E->NCH, NS = E->NS, R = E->R;
            DBuf<float4> st_save;
            DBuf<int> lists_save;
            st_save.alloc((size_t)2 * std::max(NCH + NS, 1));
            lists_save.alloc(std::max<size_t>(list_words(E), 1));
            save_state(E, st_save.p);
            save_lists(E, lists_save.p);
            std::vector<float> lam(R), lam_applied(R);
            HIP_TRY(hipStreamSynchronize(E->stream));
            HIP_TRY(hipMemcpy(lam.data(), E->lambda.p, R * sizeof(float), hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(lam_applied.data(), E->lambda_applied.p, R * sizeof(float), hipMemcpyDeviceToHost));
            const int cur0 = E->cur;
            const long long step0 = E->step, since0 = E->since_rebuild, nreb0 = E->n_rebuilds;
            auto restore = [&] {
                HIP_TRY(hipStreamSynchronize(E->stream));
                E->cur = cur0;
                load_state(E, st_save.p);
                load_lists(E, lists_save.p);
                HIP_TRY(hipStreamSynchronize(E->stream));
                HIP_TRY(hipMemcpy(E->lambda.p, lam.data(), R * sizeof(float), hipMemcpyHostToDevice));
                HIP_TRY(hipMemcpy(E->lambda_applied.p, lam_applied.data(), R * sizeof(float), hipMemcpyHostToDevice));
                E->step = step0;
                E->since_rebuild = since0;
                E->n_rebuilds = nreb0;
            };
            const int cyc = E->cycle_steps();
            int timed = std::max(reps, cyc);
            // ms per step of the current shape: one cycle to capture and warm, then the timed steps
            auto trial = [&]() -> double {
                E->drop_graphs();
                E->build_scwork();
                E->reps.upload(E->reps_h);
                E->enqueue_steps(cyc);
                HIP_TRY(hipEventRecord(e0, E->stream));
                E->enqueue_steps(timed);
                HIP_TRY(hipEventRecord(e1, E->stream));
                HIP_TRY(hipEventSynchronize(e1));
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
                const int st = E->check_errors();
                restore();
                return st == LIMA_OK ? (double)ms / timed : 1e30;
            };
            // trials of at least 8 ms each, in whole list cycles, sized from one probe of the
            // current shape: 10-step trials (2 ms on the production system) picked a shape 27%
            // slower than the best on a freshly started device
            {
                const double ms = trial();
                if (ms > 0.0 && ms < 1e29) {
                    const int want = (int)std::ceil(8.0 / ms);
                    timed = std::max(timed, (want + cyc - 1) / cyc * cyc);
                }
            }
            // three sweeps, the fastest time of each shape (clocks ramp after idle periods)
            constexpr int kSweeps = 3;
            // the all-pairs shape, swept again after the listed kernels' choices below (a shape
            // picked while a slow listed form dominates the step can be the wrong one once it is gone)
            auto sweep_all_pairs = [&] {
                if (!(E->params.enable_solvents && E->NS > 0 && E->NCH > 0)) return;
                const int spts[3] = {1, 2, 4}, chunks[4] = {256, 512, 1024, 2048};
                double tmin[3][4];
                for (auto& row : tmin)
                    for (double& v : row) v = 1e30;
                for (int sweep = 0; sweep < kSweeps; sweep++) {
                    for (int i = 0; i < 3; i++) {
                        for (int j = 0; j < 4; j++) {
                            E->cfg.spt = spts[i];
                            E->cfg.chunk_particles = chunks[j];
                            tmin[i][j] = std::min(tmin[i][j], trial());
                        }
                    }
                }
                double tbest = 1e30;
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 4; j++)
                        if (tmin[i][j] < tbest) { tbest = tmin[i][j]; best.spt = spts[i]; best.chunk_particles = chunks[j]; }
                // shapes within 2% of the best step are ties at the trials' resolution (repeated
                // whole-step trials of one shape spread by about that much).  Among them, those
                // whose all-pairs kernel alone is within 5% of the fastest (a shape with too few
                // work items to fill the device alone leaves the step to the concurrent streams
                // at this size and may not at another), and of those the one writing the fewest
                // per-chunk partials (less HBM traffic, fewer blocks beside the other streams;
                // shapes that differ only in chunking run within a few per cent alone), then the
                // fastest alone
                double kms[3][4];
                long long parts[3][4];
                double kbest = 1e30;
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 4; j++) {
                        kms[i][j] = 1e30;
                        parts[i][j] = 0;
                        if (tmin[i][j] > 1.02 * tbest) continue;
                        E->cfg.spt = spts[i];
                        E->cfg.chunk_particles = chunks[j];
                        E->build_scwork();
                        E->reps.upload(E->reps_h);
                        for (int r = 0; r < E->R; r++)
                            parts[i][j] += (long long)E->reps_h[r].chunk_count *
                                           (E->reps_h[r].solv_end - E->reps_h[r].solv_begin);
                        StepArgs ka = E->step_args(E->cur);
                        ka.kclock = nullptr;
                        auto launch = [&] {
                            if (E->params.accumulation == LIMA_ACCUM_FIXED64) E->launch_solvent_chain<true, false>(ka, E->stream);
                            else E->launch_solvent_chain<false, false>(ka, E->stream);
                        };
                        launch();
                        HIP_TRY(hipEventRecord(e0, E->stream));
                        for (int k = 0; k < 5; k++) launch();
                        HIP_TRY(hipEventRecord(e1, E->stream));
                        HIP_TRY(hipEventSynchronize(e1));
                        float ms = 0;
                        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
                        kms[i][j] = ms;
                        kbest = std::min(kbest, (double)ms);
                    }
                long long pbest = -1;
                double kpick = 1e30;
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 4; j++) {
                        if (kms[i][j] > 1.05 * kbest) continue;   // not a step tie, or slow alone
                        if (pbest < 0 || parts[i][j] < pbest || (parts[i][j] == pbest && kms[i][j] < kpick)) {
                            pbest = parts[i][j];
                            kpick = kms[i][j];
                            best.spt = spts[i];
                            best.chunk_particles = chunks[j];
                        }
                    }
                restore();
                E->cfg.spt = best.spt;
                E->cfg.chunk_particles = best.chunk_particles;
            };
            sweep_all_pairs();
            if (E->NC > 0) {
                // grid sizes, with the packed and the plain listed kernel where compounds allow both
                // (they give the same sums: a small system fills the device better with the plain one)
                // forms: packed (1), plain (0), plain with split small items (2, FP32 only)
                const int grids[4] = {1024, 2048, 4096, 8192};
                const int forms[3] = {1, 0, 2};
                const int nform = !E->pack_ppl() ? 1 : (E->params.accumulation == LIMA_ACCUM_FIXED64 ? 2 : 3);
                double gmin[3][4];
                for (auto& row : gmin)
                    for (double& v : row) v = 1e30;
                for (int sweep = 0; sweep < kSweeps; sweep++) {
                    for (int k = 0; k < nform; k++) {
                        E->cfg.listed_packed = forms[k];
                        for (int j = 0; j < 4; j++) {
                            E->cfg.listed_grid = grids[j];
                            gmin[k][j] = std::min(gmin[k][j], trial());
                        }
                    }
                }
                double gbest = 1e30;
                for (int k = 0; k < nform; k++)
                    for (int j = 0; j < 4; j++)
                        if (gmin[k][j] < gbest) { gbest = gmin[k][j]; best.listed_grid = grids[j]; best.listed_packed = forms[k]; }
                E->cfg.listed_grid = best.listed_grid;
                E->cfg.listed_packed = best.listed_packed;
            }
            if (E->params.enable_solvents && E->NS > 0 && E->params.accumulation != LIMA_ACCUM_FIXED64) {
                // one or four lanes per solvent in the listed solvent kernel (FP32 only: the four
                // partial sums group the floats differently)
                const int lanes[2] = {1, kSplitLanes};
                double lmin[2] = {1e30, 1e30};
                for (int sweep = 0; sweep < kSweeps; sweep++) {
                    for (int j = 0; j < 2; j++) {
                        E->cfg.listed_lanes = lanes[j];
                        lmin[j] = std::min(lmin[j], trial());
                    }
                }
                best.listed_lanes = lmin[1] < lmin[0] ? kSplitLanes : 1;
                E->cfg.listed_lanes = best.listed_lanes;
            }
            if (E->thermo_ok || E->thermo_ok_comp) {
                // the thermostat as its own kernel after the integrators, or in their last blocks: that
                // saves a launch, but each counting block then releases its partial at device scope,
                // which writes back its XCD's L2 while other kernels may still run (form 1, both
                // integrators count); form 2 counts in the compound integrator only, run after the
                // solvent one
                const bool offer[3] = {true, E->thermo_ok, E->thermo_ok_comp};
                double fmin[3] = {1e30, 1e30, 1e30};
                for (int sweep = 0; sweep < kSweeps; sweep++) {
                    for (int j = 0; j < 3; j++) {
                        if (!offer[j]) continue;
                        E->cfg.fused_thermostat = j;
                        fmin[j] = std::min(fmin[j], trial());
                    }
                }
                int jb = 0;
                for (int j = 1; j < 3; j++)
                    if (fmin[j] < fmin[jb]) jb = j;
                best.fused_thermostat = jb;
                E->cfg.fused_thermostat = best.fused_thermostat;
            }
            sweep_all_pairs();
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune_cache[key] = best;
            save_tune_entry(key, best);
        }
        E->cfg = best;
        E->build_scwork();
        E->reps.upload(E->reps_h);
        E->drop_graphs();
        return E->check_errors();
    });
}

int lima_get_parameters(lima_engine* E, int32_t system, double* params) {
    return guarded(E, [&] {
        if (!params || system < 0 || system >= E->R) throw LimaError(LIMA_E_ARG, "system index out of range or null output");
        get_parameters(E, system, params);
        return LIMA_OK;
    });
}

int lima_set_parameters(lima_engine* E, int32_t system, const double* params) {
    return guarded(E, [&] {
        if (!params || system < 0 || system >= E->R) throw LimaError(LIMA_E_ARG, "system index out of range or null input");
        set_parameters(E, system, params);
        return LIMA_OK;
    });
}

int lima_adjoint_run(lima_engine* E, const lima_adjoint_spec* spec, double* loss, double* grad) {
    return guarded(E, [&] {
        if (!spec) throw LimaError(LIMA_E_ARG, "null spec");
        return adjoint_run(E, *spec, loss, grad);
    });
}

int lima_record_frames(lima_engine* E, int32_t n_frames, float* chain_forces, float* solvent_forces, double* potential) {
    return guarded(E, [&] { return record_frames(E, n_frames, chain_forces, solvent_forces, potential); });
}

}  // extern "C"
