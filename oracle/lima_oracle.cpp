// lima_oracle.cpp — CPU restatement of LIMA_MD's per-step path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP implementation is
// compared against; it is loaded by tests/, by __graft_entry__.smoke() and by
// bench.py's cpu_baseline leg, and by nothing in the product path.
//
// It restates, in C++ templated on the arithmetic type T (float = reference
// FP32 semantics, double = the FP64 oracle), the reference's:
//   * Float3 arithmetic incl. the FP64-promoting operator*(double)      QuantomTypes.cuh:33-168
//   * applyHyperpos / applyPBC                                          Engine.cuh:24-30, Engine.cu:461-466
//   * calcLJForce, bonded forces, integratePosition                     Engine.cu:475-637, 849-869
//   * the call sites and their mixing rules                             Engine.cu:640-730
//   * compoundBridgeKernel / compoundKernel / solventForceKernel        Engine.cu:1168-1223, 880-1057, 1070-1161
//   * the host neighbour-list update (cull, then add)                   Engine.cu:146-287
// Kernel loops are executed sequentially per entity, in the order the
// reference accumulates (thread-order serialized scatter = term index order,
// Engine.cu:744-847).  Where the reference has a data race or reads stale
// shared memory, the intended semantics are restated and noted inline.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp).  Exact FP32
// operation order matters: no FMA contraction, correctly rounded sqrtf.
//
// Parity pins (tests/test_oracle_*.py): closed forms (LJ zero at 2^(1/6)σ and
// the FP64 LJ scalar of VirtualPathMaker.cpp:160-170; bond force = -dU/dx;
// angle/dihedral zero-force and constructed geometries; integrator constant
// velocity), invariants (translation, box shift, Newton's third law on the
// symmetric call sites) and brute-force neighbour sets with hand-built
// hysteresis cases.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <limits>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/lima.h"

namespace {

template <typename T> struct V3 {
    T x = 0, y = 0, z = 0;
    V3() {}
    explicit V3(T a) : x(a), y(a), z(a) {}
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    // Float3::operator*(const double) multiplies in FP64 and narrows (QuantomTypes.cuh:39)
    V3 muld(double a) const { return V3(T(x * a), T(y * a), T(z * a)); }
    V3 operator+(const V3& a) const { return V3(x + a.x, y + a.y, z + a.z); }
    V3 operator-(const V3& a) const { return V3(x - a.x, y - a.y, z - a.z); }
    void operator+=(const V3& a) { x += a.x; y += a.y; z += a.z; }
    T lenSquared() const { return x * x + y * y + z * z; }                 // QuantomTypes.cuh:61
    T len() const { return std::sqrt(x * x + y * y + z * z); }             // QuantomTypes.cuh:60
    V3 norm() const {                                                      // QuantomTypes.cuh:50-55
        T l = len();
        if (l) return muld((double)(T(1) / l));
        return V3(0, 0, 0);
    }
    V3 cross(const V3& a) const {                                          // QuantomTypes.cuh:89
        return V3(y * a.z - z * a.y, z * a.x - x * a.z, x * a.y - y * a.x);
    }
    T dot(const V3& a) const { return x * a.x + y * a.y + z * a.z; }       // QuantomTypes.cuh:90
    T& at(int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

template <typename T> inline T t_acos(T v) { return std::acos(v); }
template <typename T> inline T t_sin(T v) { return std::sin(v); }
template <typename T> inline T t_cos(T v) { return std::cos(v); }

// Float3::getAngle (QuantomTypes.cuh:74-83): clamp then acos, in T precision.
template <typename T> T get_angle(const V3<T>& v1, const V3<T>& v2) {
    T val = v1.dot(v2) / (v1.len() * v2.len());
    val = val > T(1) ? T(1) : val;
    val = val < T(-1) ? T(-1) : val;
    return t_acos(val);
}

// LIMAENG::applyHyperpos (Engine.cuh:24-30): per dimension, shift `m` by +L, then
// by -L using the already shifted value; branch-free as in the reference.
template <typename T> inline void hyperpos(const V3<T>& s, V3<T>& m, T L) {
    const T half = L * T(0.5);
    m.x += L * (T)((s.x - m.x) > half);
    m.x -= L * (T)((s.x - m.x) < -half);
    m.y += L * (T)((s.y - m.y) > half);
    m.y -= L * (T)((s.y - m.y) < -half);
    m.z += L * (T)((s.z - m.z) > half);
    m.z -= L * (T)((s.z - m.z) < -half);
}

// applyPBC (Engine.cu:461-466)
template <typename T> inline void apply_pbc(V3<T>& p, T L) {
    p.x += L * (T)(p.x < 0);
    p.x -= L * (T)(p.x > L);
    p.y += L * (T)(p.y < 0);
    p.y -= L * (T)(p.y > L);
    p.z += L * (T)(p.z < 0);
    p.z -= L * (T)(p.z > L);
}

// calcLJForce (Engine.cu:475-522): force on p0; no cutoff, no energy.
// When `energy`/`virial` are given, also accumulates 4ε(s²-s) and d·F in FP64
// (the build's energy report; the reference never adds LJ to potE).
template <typename T>
inline V3<T> lj_force(const V3<T>& p0, const V3<T>& p1, T sigma, T epsilon,
                      double* energy = nullptr, double* virial = nullptr) {
    V3<T> d = p1 - p0;
    T dist_sq = d.lenSquared();
    T s = sigma * sigma / dist_sq;
    s = s * s * s;
    T force_scalar = T(24) * epsilon * s / dist_sq * (T(1) - T(2) * s);
    if (energy) {
        double sd = (double)sigma * (double)sigma / (double)dist_sq;
        double s6 = sd * sd * sd;
        *energy += 4.0 * (double)epsilon * (s6 * s6 - s6);
        *virial += (double)dist_sq * (double)force_scalar;
    }
    return d.muld((double)force_scalar);
}

// calcPairbondForces (Engine.cu:525-548)
template <typename T>
inline void pairbond(const V3<T>& a, const V3<T>& b, T b0, T kb, V3<T>* res, double* potE) {
    V3<T> difference = a - b;
    double error = (double)(T)(difference.len() - b0);
    *potE += 0.5 * (double)kb * (error * error);
    difference = difference.norm();
    double force_scalar = -(double)kb * error;
    res[0] = difference.muld(force_scalar);
    res[1] = difference.muld(force_scalar).muld(-1.0);
}

// calcAnglebondForces (Engine.cu:552-578); the reference's energy carries 0.25.
template <typename T>
inline void anglebond(const V3<T>& l, const V3<T>& m, const V3<T>& r, T theta0, T k,
                      V3<T>* res, double* potE) {
    V3<T> v1 = l - m;
    V3<T> v2 = r - m;
    V3<T> normal1 = v1.cross(v2);
    V3<T> normal2 = v2.cross(v1);
    V3<T> dir1 = (v1.cross(normal2)).norm();
    V3<T> dir2 = (v2.cross(normal1)).norm();
    double angle = (double)get_angle(v1, v2);
    double error = angle - (double)theta0;
    *potE += 0.5 * (double)k * error * error * 0.5;
    double force_scalar = (double)k * error;
    res[0] = dir1.muld(force_scalar);
    res[2] = dir2.muld(force_scalar);
    res[1] = (res[0] + res[2]).muld(-1.0);
}

// calcDihedralbondForces (Engine.cu:580-637).  The reference assigns potE
// (Engine.cu:612); the energy report accumulates it instead.
template <typename T>
inline void dihedral(const V3<T>& l, const V3<T>& lm, const V3<T>& rm, const V3<T>& r,
                     T phi0, T k, int n, V3<T>* res, double* potE) {
    V3<T> normal1 = ((l - lm).cross(rm - lm)).norm();
    V3<T> normal2 = ((lm - rm).cross(r - rm)).norm();
    T torsion = get_angle(normal1, normal2);
    if (((l + normal1) - r).len() < ((l - normal1) - r).len())
        normal2 = V3<T>(-normal2.x, -normal2.y, -normal2.z);
    else
        normal1 = V3<T>(-normal1.x, -normal1.y, -normal1.z);
    T arg = T(n) * torsion - phi0;
    T force_scalar = k * t_sin(arg);
    *potE += (double)(k * (T(1) - t_cos(arg)));
    res[0] = normal1.muld((double)force_scalar);
    res[3] = normal2.muld((double)force_scalar);
    res[1] = (res[0] + res[3]).muld(-1.0).muld(0.5);
    res[2] = (res[0] + res[3]).muld(-1.0).muld(0.5);
}

// integratePosition (Engine.cu:849-869): position (Störmer) Verlet with the
// displacement rescaled by the thermostat scalar.  Returns the displacement
// x_new - x_old used for the kinetic energy.
template <typename T>
inline V3<T> integrate(V3<T>& pos, V3<T>& pos_tsub1, const V3<T>& force, double mass,
                       double dt, T lambda, T L) {
    V3<T> temp = pos;
    hyperpos(pos, pos_tsub1, L);
    double masstime_scaling = dt / mass * dt;
    pos = (pos.muld(2.0) - pos_tsub1) + force.muld(masstime_scaling);
    pos_tsub1 = temp;
    V3<T> delta_pos = pos - pos_tsub1;
    pos = pos_tsub1 + delta_pos.muld((double)lambda);
    return pos - pos_tsub1;
}

// LIMAENG::calcKineticEnergy (Engine.cuh:31-44): positions two steps apart, the
// older re-imaged next to the newer; v = |a - b| * 0.5 / dt and KE = 0.5 m v^2,
// each narrowed to T.
template <typename T>
inline T ref_kinetic(const V3<T>& a, V3<T> b, double mass, double dt, T L) {
    hyperpos(a, b, L);
    T vel = (T)((double)((a - b).len() * T(0.5)) / dt);
    return (T)(0.5 * mass * (double)vel * (double)vel);
}

// Engine::neighborWithinCutoff (Engine.cu:146-151)
template <typename T>
inline bool within(const V3<T>& a, const V3<T>& b, T offset, T cutoff, T L) {
    V3<T> bt = b;
    hyperpos(a, bt, L);
    T dist = (a - bt).len();
    return dist < (cutoff + offset);
}

struct Term2 { float p0, p1; int a[2]; };
struct Term3 { float p0, p1; int a[3]; };
struct Term4 { float p0, p1; int n; int a[4]; };

constexpr double GAS_R = 8.3145;  // Engine.cu:356

template <typename T> struct Oracle {
    // ---- system ----
    T L;
    double dt;
    double T0;
    lima_forcefield ff;
    int nc = 0, nchain = 0, ns = 0, nbr = 0;
    std::vector<int> coff, key;
    std::vector<T> sphere;
    std::vector<V3<T>> cx, cxm, sx, sxm;
    std::vector<uint8_t> ctype;
    std::vector<int> bond_off, angle_off, dih_off;
    std::vector<Term2> bonds; std::vector<Term3> angles; std::vector<Term4> dihs;
    std::vector<int> ign_off, ign_comp, ign_local;
    std::vector<int> br_poff, br_pglobal, br_bond_off, br_angle_off, br_dih_off;
    std::vector<Term2> br_bonds; std::vector<Term3> br_angles; std::vector<Term4> br_dihs;
    // ---- params ----
    T cutoff;
    int nlist_interval;
    int thermostat;
    double tau;
    bool enable_solvents, enable_dihedrals;
    int cap_cc, cap_cs, cap_ss;
    // ---- state ----
    std::vector<std::vector<int>> cc, cs, sc, ss;   // NeighborList contents, in list order
    std::vector<V3<T>> fbridge, cforce, sforce;
    std::vector<V3<T>> csforce;   // per chain particle: computeSolventToCompoundLJForces' result, -1s when not run
    T lambda = T(1);
    long step = 0;
    int err = LIMA_OK;
    double last_ke = 0;
    double ke_ref = 0;   // sum of ref_kinetic over the last advanced step (the reference cadence thermostat)
    double ref_temperature = 0;   // the last temperature that thermostat evaluated

    bool ignore(int p_global, int other_compound, int other_local) const {
        // LJ_Ignores::ignore (Bodies.cuh:191-197), full-width ids
        for (int e = ign_off[p_global]; e < ign_off[p_global + 1]; e++)
            if (ign_local[e] == other_local && ign_comp[e] == other_compound) return true;
        return false;
    }
    T sig(int t) const { return (T)ff.particle_parameters[t].sigma; }
    T eps(int t) const { return (T)ff.particle_parameters[t].epsilon; }
    double mass(int t) const { return (double)ff.particle_parameters[t].mass; }

    // Bonded forces of one entity (compound or bridge), accumulated per particle
    // in term-index order as the serialized scatter does (Engine.cu:744-847).
    void bonded(const V3<T>* pos, int np, const Term2* b, int nb, const Term3* a, int na,
                const Term4* d, int nd, V3<T>* out, double* eb, double* ea, double* ed) const {
        std::vector<V3<T>> ps(np), as(np), ds(np);
        for (int i = 0; i < nb; i++) {
            V3<T> r[2];
            pairbond(pos[b[i].a[0]], pos[b[i].a[1]], (T)b[i].p0, (T)b[i].p1, r, eb);
            ps[b[i].a[0]] += r[0];
            ps[b[i].a[1]] += r[1];
        }
        for (int i = 0; i < na; i++) {
            V3<T> r[3];
            anglebond(pos[a[i].a[0]], pos[a[i].a[1]], pos[a[i].a[2]], (T)a[i].p0, (T)a[i].p1, r, ea);
            for (int k = 0; k < 3; k++) as[a[i].a[k]] += r[k];
        }
        if (enable_dihedrals) {
            for (int i = 0; i < nd; i++) {
                V3<T> r[4];
                dihedral(pos[d[i].a[0]], pos[d[i].a[1]], pos[d[i].a[2]], pos[d[i].a[3]],
                         (T)d[i].p0, (T)d[i].p1, d[i].n, r, ed);
                for (int k = 0; k < 4; k++) ds[d[i].a[k]] += r[k];
            }
        }
        for (int p = 0; p < np; p++) {
            out[p] += ps[p];
            out[p] += as[p];
            if (enable_dihedrals) out[p] += ds[p];
        }
    }

    // compoundBridgeKernel (Engine.cu:1168-1223): forces are stored, not added.
    void bridge_pass(lima_energies* E) {
        for (int b = 0; b < nbr; b++) {
            int p0 = br_poff[b], np = br_poff[b + 1] - p0;
            std::vector<V3<T>> pos(np), f(np);
            for (int k = 0; k < np; k++) pos[k] = cx[br_pglobal[p0 + k]];
            for (int k = 0; k < np; k++) hyperpos(pos[0], pos[k], L);
            double eb = 0, ea = 0, ed = 0;
            bonded(pos.data(), np, br_bonds.data() + br_bond_off[b], br_bond_off[b + 1] - br_bond_off[b],
                   br_angles.data() + br_angle_off[b], br_angle_off[b + 1] - br_angle_off[b],
                   br_dihs.data() + br_dih_off[b], br_dih_off[b + 1] - br_dih_off[b], f.data(), &eb, &ea, &ed);
            for (int k = 0; k < np; k++) fbridge[br_pglobal[p0 + k]] = f[k];
            if (E) { E->bond += eb; E->angle += ea; E->dihedral += ed; }
        }
    }

    // compoundKernel (Engine.cu:880-1057) for every compound.  When `advance`
    // is false only forces are produced (state untouched).
    void compound_pass(bool advance, lima_energies* E, double* ke, std::vector<V3<T>>& cx_next) {
        const int ts = 0;  // ATOMTYPE_SOL
        double e_b = 0, e_a = 0, e_d = 0, e_intra = 0, e_inter = 0, e_cs = 0, vir = 0, kesum = 0, kref = 0;
        int local_err = LIMA_OK;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : e_b, e_a, e_d, e_intra, e_inter, e_cs, vir, kesum, kref)
        for (int c = 0; c < nc; c++) {
            const int base = coff[c], n = coff[c + 1] - coff[c];
            std::vector<V3<T>> pos(cx.begin() + base, cx.begin() + base + n);
            std::vector<V3<T>> force(n), stage;
            for (int p = 0; p < n; p++) hyperpos(pos[0], pos[p], L);           // Engine.cu:923
            for (int p = 0; p < n; p++) force[p] = fbridge[base + p];           // Engine.cu:918
            double eb = 0, ea = 0, ed = 0;
            bonded(pos.data(), n, bonds.data() + bond_off[c], bond_off[c + 1] - bond_off[c],
                   angles.data() + angle_off[c], angle_off[c + 1] - angle_off[c],
                   dihs.data() + dih_off[c], dih_off[c + 1] - dih_off[c], force.data(), &eb, &ea, &ed);
            e_b += eb; e_a += ea; e_d += ed;
            double ei = 0, vi = 0;
            double* pei = E ? &ei : nullptr;
            double* pvi = E ? &vi : nullptr;
            // computeIntramolecularLJForces (Engine.cu:678-700): geometric ε
            for (int p = 0; p < n; p++) {
                int tp = ctype[base + p];
                V3<T> f;
                for (int i = 0; i < n; i++) {
                    if (i == p || ignore(base + p, c, i)) continue;
                    int ti = ctype[base + i];
                    f += lj_force(pos[p], pos[i], (sig(tp) + sig(ti)) * T(0.5), (T)std::sqrt(eps(tp) * eps(ti)), pei, pvi);
                }
                force[p] += f;
            }
            e_intra += ei;
            // computerIntermolecularLJForces (Engine.cu:653-676) over the compound list: arithmetic ε
            double en = 0;
            double* pen = E ? &en : nullptr;
            for (int nbid : cc[c]) {
                const int nb_base = coff[nbid], nn = coff[nbid + 1] - coff[nbid];
                stage.assign(cx.begin() + nb_base, cx.begin() + nb_base + nn);
                for (int j = 0; j < nn; j++) hyperpos(pos[0], stage[j], L);    // Engine.cu:949
                for (int p = 0; p < n; p++) {
                    int tp = ctype[base + p];
                    V3<T> f;
                    for (int j = 0; j < nn; j++) {
                        if (ignore(base + p, nbid, j)) continue;
                        int tj = ctype[nb_base + j];
                        f += lj_force(pos[p], stage[j], (sig(tp) + sig(tj)) * T(0.5), (T)((eps(tp) + eps(tj)) * T(0.5)), pen, pvi);
                    }
                    force[p] += f;
                }
            }
            e_inter += en;
            // computeSolventToCompoundLJForces over the compound's solvent list (Engine.cu:964-974, 721-730).
            // The reference stages only every other id (Bodies.cuh:280-283); all ids are staged here.
            double es = 0;
            double* pes = E ? &es : nullptr;
            if (enable_solvents) {
                stage.resize(cs[c].size());
                for (size_t k = 0; k < cs[c].size(); k++) {
                    stage[k] = sx[cs[c][k]];
                    hyperpos(pos[0], stage[k], L);
                }
                for (int p = 0; p < n; p++) {
                    int tp = ctype[base + p];
                    T sgm = (sig(tp) + sig(ts)) * T(0.5);
                    T ep = (T)std::sqrt(eps(tp) * eps(ts));
                    V3<T> f;
                    for (size_t k = 0; k < stage.size(); k++) f += lj_force(pos[p], stage[k], sgm, ep, pes, pvi);
                    force[p] += f;
                    csforce[base + p] = f;
                }
            }
            e_cs += es;
            vir += vi;
            for (int p = 0; p < n; p++) {
                cforce[base + p] = force[p];
                if (force[p].len() > T(2e12)) {
#pragma omp critical
                    local_err = LIMA_E_CRITICAL_FORCE;
                }
            }
            if (!advance) continue;
            std::vector<V3<T>> older(n);
            for (int p = 0; p < n; p++) {
                V3<T> prev = cxm[base + p];
                older[p] = prev;
                V3<T> disp = integrate(pos[p], prev, force[p], mass(ctype[base + p]), dt, lambda, L);
                cxm[base + p] = prev;
                kesum += mass(ctype[base + p]) * ((double)disp.x * disp.x + (double)disp.y * disp.y + (double)disp.z * disp.z);
            }
            // PBC on particle 0, then every particle re-imaged against the wrapped
            // particle 0 (Engine.cu:993-996; the barrier the reference lacks is implied).
            apply_pbc(pos[0], L);
            for (int p = 0; p < n; p++) hyperpos(pos[0], pos[p], L);
            for (int p = 0; p < n; p++) cx_next[base + p] = pos[p];
            // the logged positions (Engine.cu:1013) against those two steps back (Engine.cu:333-337)
            for (int p = 0; p < n; p++) kref += (double)ref_kinetic(pos[p], older[p], mass(ctype[base + p]), dt, L);
        }
        if (local_err) err = local_err;
        if (E) {
            E->bond += e_b; E->angle += e_a; E->dihedral += e_d;
            E->lj_intra += 0.5 * e_intra; E->lj_inter += 0.5 * e_inter; E->lj_compound_solvent += 0.5 * e_cs;
            E->virial += 0.5 * vir;
        }
        if (ke) *ke += kesum;
        if (advance) ke_ref += kref;
    }

    // solventForceKernel (Engine.cu:1070-1161) for every solvent.
    void solvent_pass(bool advance, lima_energies* E, double* ke, std::vector<V3<T>>& sx_next) {
        const int ts = 0;
        double e_sc = 0, e_ss = 0, vir = 0, kesum = 0, kref = 0;
        int local_err = LIMA_OK;
#pragma omp parallel for schedule(static) reduction(+ : e_sc, e_ss, vir, kesum, kref)
        for (int s = 0; s < ns; s++) {
            V3<T> sp = sx[s], spm = sxm[s], force;
            double esc = 0, ess = 0, v = 0;
            double* pe1 = E ? &esc : nullptr;
            double* pe2 = E ? &ess : nullptr;
            double* pv = E ? &v : nullptr;
            for (int i = 0; i < nc; i++) {
                const int b = coff[i], n = coff[i + 1] - coff[i];
                hyperpos(cx[b], sp, L);                                         // Engine.cu:1111
                V3<T> f;                                                        // computeLJForces (Engine.cu:640-651)
                for (int j = 0; j < n; j++) {
                    int tj = ctype[b + j];
                    f += lj_force(sp, cx[b + j], (sig(ts) + sig(tj)) * T(0.5), (T)((eps(ts) + eps(tj)) * T(0.5)), pe1, pv);
                }
                force += f;
            }
            V3<T> f;                                                            // Engine.cu:706-720
            for (int id : ss[s]) {
                V3<T> nbp = sx[id];
                hyperpos(nbp, sp, L);
                f += lj_force(sp, nbp, sig(ts), eps(ts), pe2, pv);
            }
            force += f;
            e_sc += esc; e_ss += ess; vir += v;
            sforce[s] = force;
            if (!advance) continue;
            const V3<T> older = spm;
            V3<T> disp = integrate(sp, spm, force, mass(ts), dt, lambda, L);
            apply_pbc(sp, L);
            kref += (double)ref_kinetic(sp, older, mass(ts), dt, L);           // Engine.cu:347-351
            kesum += mass(ts) * ((double)disp.x * disp.x + (double)disp.y * disp.y + (double)disp.z * disp.z);
            if (sp.x != sp.x) {
#pragma omp critical
                local_err = LIMA_E_NAN;
            }
            sx_next[s] = sp;
            sxm[s] = spm;
        }
        if (local_err && !err) err = local_err;
        if (E) { E->lj_solvent_compound += 0.5 * e_sc; E->lj_solvent_solvent += 0.5 * e_ss; E->virial += 0.5 * vir; }
        if (ke) *ke += kesum;
        if (advance) ke_ref += kref;
    }

    // Engine::hostMaster's cadence (Engine.cu:49-75) with handleBoxtemp (Engine.cu:362-393)
    // and getBoxTemperature (Engine.cu:314-359): with s steps done, s % 40 == 0 evaluates
    // T = avg KE * 2 / (3 R) from the two-step kinetic energies; when s > 20 the scalar
    // T0 / T, clamped to [0.9, 1.1], applies to the next step only (reset at s % 40 == 1).
    void update_lambda_reference(long s) {
        lambda = T(1);
        if (s % 40 != 0) return;
        double avg = ke_ref / (double)(nchain + ns);
        float temp = (float)(avg * 2.f / (3.f * GAS_R));
        temp = temp == 0.f ? 1.f : temp;
        ref_temperature = temp;
        float sc = (float)T0 / temp;
        sc = std::min(sc, 1.1f);
        sc = std::max(sc, 0.9f);
        if (s > 20) lambda = (T)sc;
    }

    void update_lambda(double kesum) {
        // Berendsen coupling on the displacement (the thermostat_scalar hook,
        // Simulation.cuh:70), with the reference's clamp and T=0 guard (Engine.cu:368-374).
        if (thermostat != LIMA_THERMOSTAT_BERENDSEN) { lambda = T(1); return; }
        double ntot = (double)(nchain + ns);
        double temp = kesum / (dt * dt) / (3.0 * ntot * GAS_R);
        if (temp == 0.0) temp = 1.0;
        double lam = std::sqrt(1.0 + dt / tau * (T0 / temp - 1.0));
        lam = std::min(lam, 1.1);
        lam = std::max(lam, 0.9);
        lambda = (T)(float)lam;
    }

    int step_once() {
        std::vector<V3<T>> cx_next(cx.size()), sx_next(sx.size());
        double ke = 0;
        ke_ref = 0;
        bridge_pass(nullptr);
        compound_pass(true, nullptr, &ke, cx_next);
        if (enable_solvents) solvent_pass(true, nullptr, &ke, sx_next);
        else sx_next = sx;
        cx.swap(cx_next);                                                      // Engine.cu:425-432
        sx.swap(sx_next);
        last_ke = ke;
        step++;
        if (thermostat == LIMA_THERMOSTAT_REFERENCE) update_lambda_reference(step);
        else update_lambda(ke);
        if (nlist_interval > 0 && step % nlist_interval == 0) update_neighbors();
        return err;
    }

    void compute_forces() {
        std::vector<V3<T>> dummy_c(cx.size()), dummy_s(sx.size());
        bridge_pass(nullptr);
        compound_pass(false, nullptr, nullptr, dummy_c);
        if (enable_solvents) solvent_pass(false, nullptr, nullptr, dummy_s);
    }

    void energies(lima_energies* E) {
        std::memset(E, 0, sizeof(*E));
        std::vector<V3<T>> dummy_c(cx.size()), dummy_s(sx.size());
        bridge_pass(E);
        compound_pass(false, E, nullptr, dummy_c);
        if (enable_solvents) solvent_pass(false, E, nullptr, dummy_s);
        double ke = 0;
        for (int i = 0; i < nchain; i++) {
            V3<T> pm = cxm[i];
            hyperpos(cx[i], pm, L);
            V3<T> d = cx[i] - pm;
            ke += mass(ctype[i]) * ((double)d.x * d.x + (double)d.y * d.y + (double)d.z * d.z);
        }
        for (int s = 0; s < ns; s++) {
            V3<T> pm = sxm[s];
            hyperpos(sx[s], pm, L);
            V3<T> d = sx[s] - pm;
            ke += mass(0) * ((double)d.x * d.x + (double)d.y * d.y + (double)d.z * d.z);
        }
        E->kinetic = 0.5 * ke / (dt * dt);
        E->temperature = ke / (dt * dt) / (3.0 * (double)(nchain + ns) * GAS_R);
        E->potential = E->bond + E->angle + E->dihedral + E->lj_intra + E->lj_inter +
                       E->lj_compound_solvent + E->lj_solvent_compound + E->lj_solvent_solvent;
        E->thermostat_scalar = (double)lambda;
    }

    // NeighborList::addId / removeId (Bodies.cuh:215-271), runtime capacity
    bool add_id(std::vector<int>& list, int id, int cap) {
        if ((int)list.size() >= cap) { err = LIMA_E_CAPACITY; return false; }
        list.push_back(id);
        return true;
    }
    static void remove_id(std::vector<int>& list, int id) {
        for (size_t i = 0; i < list.size(); i++)
            if (list[i] == id) { list[i] = list.back(); list.pop_back(); return; }
    }

    // Engine::updateNeighborLists (Engine.cu:208-287), run synchronously.  The
    // HashTable dedupe (QuantomTypes.cuh:316-354) is set membership.
    void update_neighbors() {
        std::vector<V3<T>> kp(nc);
        for (int c = 0; c < nc; c++) kp[c] = cx[coff[c] + key[c]];             // preparePositionData (Engine.cuh:82-89)
        const T C = cutoff;
        // cullDistantNeighbors (Engine.cu:155-206)
        for (int self = 0; self < nc; self++) {
            T radd_self = sphere[self];
            for (int j = 0; j < (int)cc[self].size(); j++) {
                int nb = cc[self][j];
                if (self < nb && !within(kp[self], kp[nb], radd_self + sphere[nb] + C, C, L)) {
                    remove_id(cc[self], nb);
                    remove_id(cc[nb], self);
                    j--;
                }
            }
            for (int j = 0; j < (int)cs[self].size(); j++) {
                int nb = cs[self][j];
                if (!within(kp[self], sx[nb], radd_self + C, C, L)) {
                    remove_id(cs[self], nb);
                    remove_id(sc[nb], self);
                    j--;
                }
            }
        }
        for (int self = 0; self < ns; self++) {
            for (int j = 0; j < (int)ss[self].size(); j++) {
                int nb = ss[self][j];
                if (!within(sx[self], sx[nb], C, C, L)) {
                    remove_id(ss[self], nb);
                    remove_id(ss[nb], self);
                    j--;
                }
            }
        }
        // add compound->solvent and compound->compound (Engine.cu:219-253)
        std::vector<char> mark_s(ns, 0), mark_c(nc, 0);
        for (int self = 0; self < nc; self++) {
            T radd_self = sphere[self];
            for (int id : cc[self]) mark_c[id] = 1;
            for (int id : cs[self]) mark_s[id] = 1;
            if (enable_solvents) {
                std::vector<char> hit(ns, 0);
#pragma omp parallel for schedule(static)
                for (int cand = 0; cand < ns; cand++) hit[cand] = within(kp[self], sx[cand], radd_self + C, C, L);
                for (int cand = 0; cand < ns; cand++) {
                    if (hit[cand] && !mark_s[cand]) {
                        mark_s[cand] = 1;
                        add_id(cs[self], cand, cap_cs);
                        add_id(sc[cand], self, cap_cc);
                    }
                }
            }
            for (int cand = self + 1; cand < nc; cand++) {
                T radd_cand = sphere[self];  // the reference reads the self radius here (Engine.cu:243)
                if (within(kp[self], kp[cand], radd_self + radd_cand + C, C, L) && !mark_c[cand]) {
                    mark_c[cand] = 1;
                    add_id(cc[self], cand, cap_cc);
                    add_id(cc[cand], self, cap_cc);
                }
            }
            for (int id : cc[self]) mark_c[id] = 0;
            for (int id : cs[self]) mark_s[id] = 0;
        }
        // add solvent->solvent (Engine.cu:256-275); distance tests in parallel,
        // list mutations in the reference's sequential order.
        if (enable_solvents && ns > 0) {
            std::vector<std::vector<int>> hits(ns);
#pragma omp parallel for schedule(dynamic, 64)
            for (int self = 0; self < ns; self++)
                for (int cand = self + 1; cand < ns; cand++)
                    if (within(sx[self], sx[cand], C, C, L)) hits[self].push_back(cand);
            for (int self = 0; self < ns; self++) {
                for (int id : ss[self]) mark_s[id] = 1;
                for (int cand : hits[self]) {
                    if (!mark_s[cand]) {
                        mark_s[cand] = 1;
                        add_id(ss[self], cand, cap_ss);
                        add_id(ss[cand], self, cap_ss);
                    }
                }
                for (int id : ss[self]) mark_s[id] = 0;
            }
        }
    }
};

template <typename T> int build(Oracle<T>& o, const lima_system* s, const lima_run_params* p) {
    o.L = (T)s->box_len;
    o.dt = s->dt;
    o.T0 = s->thermostat_target;
    o.ff = s->forcefield;
    o.nc = s->n_compounds;
    o.ns = s->n_solvents;
    o.nbr = s->n_bridges;
    o.coff.assign(s->compound_offset, s->compound_offset + o.nc + 1);
    if (o.nc == 0) o.coff.assign(1, 0);
    o.nchain = o.coff[o.nc];
    o.key.assign(s->key_particle_index, s->key_particle_index + o.nc);
    o.sphere.resize(o.nc);
    for (int c = 0; c < o.nc; c++) o.sphere[c] = (T)s->confining_particle_sphere[c];
    auto v3 = [](const float* a, int n) {
        std::vector<V3<T>> v(n);
        for (int i = 0; i < n; i++) v[i] = V3<T>((T)a[3 * i], (T)a[3 * i + 1], (T)a[3 * i + 2]);
        return v;
    };
    o.cx = v3(s->chain_positions, o.nchain);
    o.cxm = v3(s->chain_prev_positions, o.nchain);
    o.ctype.assign(s->chain_atom_types, s->chain_atom_types + o.nchain);
    o.sx = v3(s->solvent_positions, o.ns);
    o.sxm = v3(s->solvent_prev_positions, o.ns);
    auto offs = [](const int32_t* a, int n) { return n > 0 ? std::vector<int>(a, a + n + 1) : std::vector<int>(1, 0); };
    o.bond_off = offs(s->bond_offset, o.nc);
    o.angle_off = offs(s->angle_offset, o.nc);
    o.dih_off = offs(s->dihedral_offset, o.nc);
    for (int i = 0; i < o.bond_off.back(); i++)
        o.bonds.push_back({s->bonds[i].b0, s->bonds[i].kb, {(int)s->bonds[i].atom_indexes[0], (int)s->bonds[i].atom_indexes[1]}});
    for (int i = 0; i < o.angle_off.back(); i++) {
        const auto& a = s->angles[i];
        o.angles.push_back({a.theta_0, a.k_theta, {(int)a.atom_indexes[0], (int)a.atom_indexes[1], (int)a.atom_indexes[2]}});
    }
    for (int i = 0; i < o.dih_off.back(); i++) {
        const auto& d = s->dihedrals[i];
        o.dihs.push_back({d.phi_0, d.k_phi, (int)d.n, {(int)d.atom_indexes[0], (int)d.atom_indexes[1], (int)d.atom_indexes[2], (int)d.atom_indexes[3]}});
    }
    o.ign_off = offs(s->ignore_offset, o.nchain);
    o.ign_comp.assign(s->ignore_compound, s->ignore_compound + o.ign_off.back());
    o.ign_local.assign(s->ignore_local, s->ignore_local + o.ign_off.back());
    o.br_poff = offs(s->bridge_particle_offset, o.nbr);
    for (int k = 0; k < o.br_poff.back(); k++)
        o.br_pglobal.push_back(o.coff[s->bridge_particle_compound[k]] + s->bridge_particle_local[k]);
    o.br_bond_off = offs(s->bridge_bond_offset, o.nbr);
    o.br_angle_off = offs(s->bridge_angle_offset, o.nbr);
    o.br_dih_off = offs(s->bridge_dihedral_offset, o.nbr);
    for (int i = 0; i < o.br_bond_off.back(); i++)
        o.br_bonds.push_back({s->bridge_bonds[i].b0, s->bridge_bonds[i].kb, {(int)s->bridge_bonds[i].atom_indexes[0], (int)s->bridge_bonds[i].atom_indexes[1]}});
    for (int i = 0; i < o.br_angle_off.back(); i++) {
        const auto& a = s->bridge_angles[i];
        o.br_angles.push_back({a.theta_0, a.k_theta, {(int)a.atom_indexes[0], (int)a.atom_indexes[1], (int)a.atom_indexes[2]}});
    }
    for (int i = 0; i < o.br_dih_off.back(); i++) {
        const auto& d = s->bridge_dihedrals[i];
        o.br_dihs.push_back({d.phi_0, d.k_phi, (int)d.n, {(int)d.atom_indexes[0], (int)d.atom_indexes[1], (int)d.atom_indexes[2], (int)d.atom_indexes[3]}});
    }
    o.cutoff = (T)p->cutoff;
    o.nlist_interval = p->nlist_interval;
    o.thermostat = p->thermostat;
    o.tau = p->thermostat_tau;
    o.enable_solvents = p->enable_solvents != 0;
    o.enable_dihedrals = p->enable_dihedrals != 0;
    o.cap_cc = p->max_compound_neighbors;
    o.cap_cs = p->max_compound_solvents;
    o.cap_ss = p->max_solvent_neighbors;
    o.cc.assign(o.nc, {}); o.cs.assign(o.nc, {}); o.sc.assign(o.ns, {}); o.ss.assign(o.ns, {});
    o.fbridge.assign(o.nchain, V3<T>());
    o.cforce.assign(o.nchain, V3<T>());
    o.csforce.assign(o.nchain, V3<T>(T(-1), T(-1), T(-1)));
    o.sforce.assign(o.ns, V3<T>());
    o.update_neighbors();                                                       // Engine.cu:35
    return o.err;
}

struct Handle {
    int precision;
    Oracle<float>* f = nullptr;
    Oracle<double>* d = nullptr;
};

template <typename T> void put3(const std::vector<V3<T>>& v, double* out) {
    if (!out) return;
    for (size_t i = 0; i < v.size(); i++) { out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z; }
}
template <typename T> void get3(std::vector<V3<T>>& v, const double* in) {
    if (!in) return;
    for (size_t i = 0; i < v.size(); i++) v[i] = V3<T>((T)in[3 * i], (T)in[3 * i + 1], (T)in[3 * i + 2]);
}

}  // namespace

#define DISPATCH(h, expr)                      \
    do {                                       \
        Handle* H = (Handle*)(h);              \
        if (H->f) { auto& o = *H->f; expr; }   \
        else { auto& o = *H->d; expr; }        \
    } while (0)

extern "C" {

void* oracle_create(const lima_system* sys, const lima_run_params* p, int precision, int* status) {
    Handle* h = new Handle();
    h->precision = precision;
    int st;
    if (precision == 64) { h->d = new Oracle<double>(); st = build(*h->d, sys, p); }
    else { h->f = new Oracle<float>(); st = build(*h->f, sys, p); }
    if (status) *status = st;
    return h;
}

void oracle_destroy(void* h) {
    Handle* H = (Handle*)h;
    delete H->f;
    delete H->d;
    delete H;
}

int oracle_step(void* h, int n) {
    int st = 0;
    DISPATCH(h, for (int i = 0; i < n && !st; i++) st = o.step_once());
    return st;
}

int oracle_compute_forces(void* h, double* cf, double* sf) {
    int st = 0;
    DISPATCH(h, { o.compute_forces(); put3(o.cforce, cf); put3(o.sforce, sf); st = o.err; });
    return st;
}

int oracle_get_forces(void* h, double* cf, double* sf) {
    DISPATCH(h, { put3(o.cforce, cf); put3(o.sforce, sf); });
    return 0;
}

// computeSolventToCompoundLJForces of the last force pass, per chain particle
int oracle_get_cs_forces(void* h, double* out) {
    int st = LIMA_OK;
    DISPATCH(h, {
        for (int i = 0; i < o.nchain; i++) {
            out[3 * i] = (double)o.csforce[i].x; out[3 * i + 1] = (double)o.csforce[i].y; out[3 * i + 2] = (double)o.csforce[i].z;
        }
        st = o.err;
    });
    return st;
}

int oracle_rebuild_neighbors(void* h) {
    int st = 0;
    DISPATCH(h, { o.update_neighbors(); st = o.err; });
    return st;
}

int oracle_get_state(void* h, double* cx, double* cxm, double* sx, double* sxm) {
    DISPATCH(h, { put3(o.cx, cx); put3(o.cxm, cxm); put3(o.sx, sx); put3(o.sxm, sxm); });
    return 0;
}

int oracle_set_state(void* h, const double* cx, const double* cxm, const double* sx, const double* sxm) {
    DISPATCH(h, { get3(o.cx, cx); get3(o.cxm, cxm); get3(o.sx, sx); get3(o.sxm, sxm); });
    return 0;
}

int oracle_get_energies(void* h, lima_energies* out) {
    DISPATCH(h, o.energies(out));
    return 0;
}

double oracle_get_lambda(void* h) {
    double v = 0;
    DISPATCH(h, v = (double)o.lambda);
    return v;
}

double oracle_get_ref_temperature(void* h) {
    double v = 0;
    DISPATCH(h, v = o.ref_temperature);
    return v;
}

void oracle_set_lambda(void* h, double lam) {
    DISPATCH(h, o.lambda = (decltype(o.lambda))(float)lam);
}

long oracle_get_step(void* h) {
    long v = 0;
    DISPATCH(h, v = o.step);
    return v;
}

// kind: 0 = compound->compound, 1 = compound->solvent, 2 = solvent->compound, 3 = solvent->solvent
int oracle_get_neighbor_counts(void* h, int kind, int32_t* counts) {
    DISPATCH(h, {
        auto& L = kind == 0 ? o.cc : kind == 1 ? o.cs : kind == 2 ? o.sc : o.ss;
        for (size_t i = 0; i < L.size(); i++) counts[i] = (int32_t)L[i].size();
    });
    return 0;
}

int oracle_get_neighbors(void* h, int kind, int32_t* ids, int stride) {
    DISPATCH(h, {
        auto& L = kind == 0 ? o.cc : kind == 1 ? o.cs : kind == 2 ? o.sc : o.ss;
        for (size_t i = 0; i < L.size(); i++)
            for (size_t k = 0; k < L[i].size() && (int)k < stride; k++) ids[i * stride + k] = L[i][k];
    });
    return 0;
}

int oracle_set_neighbors(void* h, int kind, const int32_t* counts, const int32_t* ids, int stride) {
    DISPATCH(h, {
        auto& L = kind == 0 ? o.cc : kind == 1 ? o.cs : kind == 2 ? o.sc : o.ss;
        for (size_t i = 0; i < L.size(); i++) L[i].assign(ids + i * stride, ids + i * stride + counts[i]);
    });
    return 0;
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// ---- single routines, for closed-form tests ----
#define V3F(a) V3<float>(a[0], a[1], a[2])
#define V3D(a) V3<double>(a[0], a[1], a[2])
#define OUT3(v, o) do { (o)[0] = (v).x; (o)[1] = (v).y; (o)[2] = (v).z; } while (0)

void oracle_lj_force_f32(const float* p0, const float* p1, float sigma, float eps, float* out) {
    OUT3(lj_force(V3F(p0), V3F(p1), sigma, eps), out);
}
void oracle_lj_force_f64(const double* p0, const double* p1, double sigma, double eps, double* out) {
    OUT3(lj_force(V3D(p0), V3D(p1), sigma, eps), out);
}
void oracle_pairbond_f32(const float* a, const float* b, float b0, float kb, float* out, double* potE) {
    V3<float> r[2];
    pairbond(V3F(a), V3F(b), b0, kb, r, potE);
    OUT3(r[0], out); OUT3(r[1], out + 3);
}
void oracle_pairbond_f64(const double* a, const double* b, double b0, double kb, double* out, double* potE) {
    V3<double> r[2];
    pairbond(V3D(a), V3D(b), b0, kb, r, potE);
    OUT3(r[0], out); OUT3(r[1], out + 3);
}
void oracle_anglebond_f32(const float* l, const float* m, const float* r, float t0, float k, float* out, double* potE) {
    V3<float> res[3];
    anglebond(V3F(l), V3F(m), V3F(r), t0, k, res, potE);
    for (int i = 0; i < 3; i++) OUT3(res[i], out + 3 * i);
}
void oracle_anglebond_f64(const double* l, const double* m, const double* r, double t0, double k, double* out, double* potE) {
    V3<double> res[3];
    anglebond(V3D(l), V3D(m), V3D(r), t0, k, res, potE);
    for (int i = 0; i < 3; i++) OUT3(res[i], out + 3 * i);
}
void oracle_dihedral_f32(const float* l, const float* lm, const float* rm, const float* r, float phi0, float k, int n,
                         float* out, double* potE) {
    V3<float> res[4];
    dihedral(V3F(l), V3F(lm), V3F(rm), V3F(r), phi0, k, n, res, potE);
    for (int i = 0; i < 4; i++) OUT3(res[i], out + 3 * i);
}
void oracle_dihedral_f64(const double* l, const double* lm, const double* rm, const double* r, double phi0, double k,
                         int n, double* out, double* potE) {
    V3<double> res[4];
    dihedral(V3D(l), V3D(lm), V3D(rm), V3D(r), phi0, k, n, res, potE);
    for (int i = 0; i < 4; i++) OUT3(res[i], out + 3 * i);
}
void oracle_integrate_f32(float* pos, float* prev, const float* force, double mass, double dt, float lambda, float L) {
    V3<float> p = V3F(pos), pm = V3F(prev);
    integrate(p, pm, V3F(force), mass, dt, lambda, L);
    OUT3(p, pos); OUT3(pm, prev);
}
void oracle_hyperpos_f32(const float* s, float* m, float L) {
    V3<float> mm = V3F(m);
    hyperpos(V3F(s), mm, L);
    OUT3(mm, m);
}
void oracle_pbc_f32(float* p, float L) {
    V3<float> pp = V3F(p);
    apply_pbc(pp, L);
    OUT3(pp, p);
}
float oracle_get_angle_f32(const float* a, const float* b) { return get_angle(V3F(a), V3F(b)); }
int oracle_within_f32(const float* a, const float* b, float offset, float cutoff, float L) {
    return within(V3F(a), V3F(b), offset, cutoff, L);
}

}  // extern "C"
