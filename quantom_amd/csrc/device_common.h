// device_common.h — shared device-side types and arithmetic for the gfx950 kernels.
//
// The FP32 helpers restate reference arithmetic where results must be
// bit-identical to the CPU restatement (neighbour-list distances, PBC,
// re-imaging, integration): they use explicitly rounded intrinsics so that no
// FMA contraction or reassociation can change the result.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LIMA_DEV __device__ __forceinline__

namespace lima {

constexpr int kMaxTypes = 32;
constexpr int kCompoundBlock = 128;      // max particles per compound (2 wavefronts)
constexpr int kSolventChunk = 256;       // solvent-list entries per compound work item
constexpr int kFixedShift = 21;          // int64 fixed point: F_fixed = rint(F * 2^21)
constexpr int kMaxBondsPerCompound = 256;
constexpr int kMaxAnglesPerCompound = 384;
constexpr int kMaxDihedralsPerCompound = 512;

// device error flags (replace exit()/critical_error_encountered)
enum : uint32_t {
    kErrCapacity = 1u,
    kErrCriticalForce = 2u,
    kErrNaN = 4u,
};

struct ReplicaInfo {
    float L, half;
    double dt;
    double tau;
    float T0;
    int comp_begin, comp_end;     // global compound ids
    int chain_begin, chain_end;   // global chain particle ids
    int solv_begin, solv_end;     // global solvent ids
    int stile_begin, stile_end;   // solvent integrate tiles (KE partials)
    int chunk_count;              // solvent<-chain compound chunks of this replica
    int cell_begin;               // first cell of this replica's solvent grid
    int ncell;                    // cells per dimension
    float cell_inv;               // ncell / L
    float ss_bound;               // squared-distance bound of the solvent-solvent test (C + C)
};

// work item of the all-pairs solvent<-chain kernel
struct SCWork {
    int replica;
    int solv_begin, solv_end;
    int comp_begin, comp_end;     // compound range of this chunk
    int chunk;                    // partial slot index within the replica
    int pad0, pad1;
};

// ---- exact FP32 arithmetic: every operation rounded on its own --------------
// HIP's __fadd_rn family are plain operators, which the compiler may fuse into
// FMAs with a neighbouring product; with contraction off here they never are.
LIMA_DEV float fadd(float a, float b) {
#pragma clang fp contract(off)
    return a + b;
}
LIMA_DEV float fsub(float a, float b) {
#pragma clang fp contract(off)
    return a - b;
}
LIMA_DEV float fmul(float a, float b) {
#pragma clang fp contract(off)
    return a * b;
}

// Float3::len (QuantomTypes.cuh:60): sqrtf((x*x + y*y) + z*z), correctly rounded
LIMA_DEV float exact_len(float x, float y, float z) {
    return __fsqrt_rn(fadd(fadd(fmul(x, x), fmul(y, y)), fmul(z, z)));
}

// LIMAENG::applyHyperpos (Engine.cuh:24-30) on one coordinate
// m + L*flag with flag 0 or 1: the product is exact, so one fused operation
// rounds exactly as the separate multiply and add do
LIMA_DEV float hyper1(float s, float m, float L, float half) {
    m = __builtin_fmaf(L, (float)(fsub(s, m) > half), m);
    m = __builtin_fmaf(-L, (float)(fsub(s, m) < -half), m);
    return m;
}
LIMA_DEV void hyperpos(float sx, float sy, float sz, float& mx, float& my, float& mz, float L, float half) {
    mx = hyper1(sx, mx, L, half);
    my = hyper1(sy, my, L, half);
    mz = hyper1(sz, mz, L, half);
}

// applyPBC (Engine.cu:461-466) on one coordinate
LIMA_DEV float pbc1(float p, float L) {
    p = __builtin_fmaf(L, (float)(p < 0.0f), p);   // exact product, as in hyper1
    p = __builtin_fmaf(-L, (float)(p > L), p);
    return p;
}

// Engine::neighborWithinCutoff (Engine.cu:146-151)
LIMA_DEV bool within(float ax, float ay, float az, float bx, float by, float bz, float offset, float cutoff,
                     float L, float half) {
    hyperpos(ax, ay, az, bx, by, bz, L, half);
    float d = exact_len(fsub(ax, bx), fsub(ay, by), fsub(az, bz));
    return d < fadd(cutoff, offset);
}

// the squared length inside Engine::neighborWithinCutoff: b re-imaged against a,
// then (x*x + y*y) + z*z of a - b, exactly rounded.  Because sqrtf is correctly
// rounded and monotone, sqrtf(d2) < t  <=>  d2 < B(t) with B(t) the smallest
// float whose sqrtf is >= t (computed on the host), so the comparison is exact.
LIMA_DEV float dist2_reimaged(float ax, float ay, float az, float bx, float by, float bz, float L, float half) {
    hyperpos(ax, ay, az, bx, by, bz, L, half);
    const float dx = fsub(ax, bx), dy = fsub(ay, by), dz = fsub(az, bz);
    return fadd(fadd(fmul(dx, dx), fmul(dy, dy)), fmul(dz, dz));
}

// LIMAENG::calcKineticEnergy (Engine.cuh:31-44): positions two steps apart,
// the older one re-imaged next to the newer; v = |a - b| * 0.5f / dt narrowed to
// FP32, KE = 0.5 m v^2 in FP64 narrowed to FP32.
LIMA_DEV float ref_kinetic(float ax, float ay, float az, float bx, float by, float bz, double mass, double dt, float L,
                           float half) {
    hyperpos(ax, ay, az, bx, by, bz, L, half);
    const float d = exact_len(fsub(ax, bx), fsub(ay, by), fsub(az, bz));
    const float vel = (float)((double)fmul(d, 0.5f) / dt);
    return (float)(0.5 * mass * (double)vel * (double)vel);
}

// integratePosition (Engine.cu:849-869); returns the displacement after rescale.
LIMA_DEV void integrate(float& px, float& py, float& pz, float& qx, float& qy, float& qz, float fx, float fy,
                        float fz, double mts, float lambda, float L, float half, float& dx, float& dy, float& dz) {
    float tx = px, ty = py, tz = pz;
    hyperpos(px, py, pz, qx, qy, qz, L, half);
    float nx = fadd(fsub(fmul(px, 2.0f), qx), (float)((double)fx * mts));
    float ny = fadd(fsub(fmul(py, 2.0f), qy), (float)((double)fy * mts));
    float nz = fadd(fsub(fmul(pz, 2.0f), qz), (float)((double)fz * mts));
    qx = tx; qy = ty; qz = tz;
    float ex = fsub(nx, qx), ey = fsub(ny, qy), ez = fsub(nz, qz);
    px = fadd(qx, fmul(ex, lambda));
    py = fadd(qy, fmul(ey, lambda));
    pz = fadd(qz, fmul(ez, lambda));
    dx = fsub(px, qx); dy = fsub(py, qy); dz = fsub(pz, qz);
}

// ---- LJ pair (calcLJForce, Engine.cu:475-522), fast form -------------------
// The reference's scalar 24 eps s^3 (1 - 2 s^3) / r^2, s = sigma^2 / r^2
// (Engine.cu:488-490), is evaluated as inv^4 (A + B inv^3) with inv = 1/r^2 and
// the per-type-pair constants A = 24 eps sigma^6, B = -48 eps sigma^12 (lj_ab,
// from sigma^2 and 24 eps rounded as the reference rounds them).  1/r^2 is the
// hardware reciprocal (1 ulp) instead of two IEEE divisions.  Five multiplies or
// FMAs after the reciprocal, one fewer than through s.
LIMA_DEV float2 lj_ab(float sig2, float e24) {
    const float s6 = sig2 * sig2 * sig2;
    return make_float2(e24 * s6, -2.0f * e24 * s6 * s6);
}

LIMA_DEV float lj_scalar(float r2, float A, float B, float& inv_out) {
    const float inv = __builtin_amdgcn_rcpf(r2);
    const float i2 = inv * inv;
    const float i3 = i2 * inv;
    const float i4 = i2 * i2;
    inv_out = inv;
    return i4 * __builtin_fmaf(B, i3, A);
}

// pair energy 4 eps (s^6 - s^3) = -(inv^3 / 6) (A + B inv^3 / 2) of one directed pair, from the
// force's reciprocal; evaluated in FP32 and returned for the FP64 sums
LIMA_DEV double lj_energy(float inv, float A, float B) {
    const float i3 = inv * inv * inv;
    return (double)(-0.16666667f * i3 * __builtin_fmaf(0.5f * B, i3, A));
}

// |d|^2, f += d fs and the virial sum with a fixed fused-multiply-add association: left to
// the compiler, which products fuse depends on the surrounding code, and so would the rounding
LIMA_DEV float pair_r2(float dx, float dy, float dz) { return __builtin_fmaf(dz, dz, __builtin_fmaf(dx, dx, dy * dy)); }
LIMA_DEV void pair_add(float dx, float dy, float dz, float fs, float& fx, float& fy, float& fz) {
    fx = __builtin_fmaf(dx, fs, fx);
    fy = __builtin_fmaf(dy, fs, fy);
    fz = __builtin_fmaf(dz, fs, fz);
}
LIMA_DEV void pair_virial(float r2, float fs, double& w) { w = __builtin_fma((double)r2, (double)fs, w); }

LIMA_DEV void lj_acc(float px, float py, float pz, float qx, float qy, float qz, float A, float B, float& fx,
                     float& fy, float& fz) {
    const float dx = qx - px, dy = qy - py, dz = qz - pz;
    const float r2 = pair_r2(dx, dy, dz);
    float inv;
    const float fs = lj_scalar(r2, A, B, inv);
    pair_add(dx, dy, dz, fs, fx, fy, fz);
}

// same pair with FP64 energy and virial d.F accumulation
LIMA_DEV void lj_acc_energy(float px, float py, float pz, float qx, float qy, float qz, float A, float B, float& fx,
                            float& fy, float& fz, double& e, double& w) {
    const float dx = qx - px, dy = qy - py, dz = qz - pz;
    const float r2 = pair_r2(dx, dy, dz);
    float inv;
    const float fs = lj_scalar(r2, A, B, inv);
    pair_add(dx, dy, dz, fs, fx, fy, fz);
    e += lj_energy(inv, A, B);
    pair_virial(r2, fs, w);
}

// A partial force enters the int64 sums only while |F| < 2^40 J/(mol nm) (2^61 in fixed
// point), so that sums of a few such partials stay inside int64; the critical force the
// reference stops at is 2e12 (Engine.cu:1050), just under 2^41.  A NaN or infinite partial
// raises kErrNaN and contributes 0; a finite partial at or beyond the bound raises
// kErrCriticalForce and saturates at +-2^60.  The float-to-int64 conversion itself is
// undefined for such values, so the test comes first.
constexpr float kFixedPartialBound = 1099511627776.0f;   // 2^40
LIMA_DEV long long to_fixed(float v, uint32_t& bad) {
    const float m = fabsf(v);
    if (!(m < kFixedPartialBound)) {
        if (m != m || m == INFINITY) {
            bad |= kErrNaN;
            return 0;
        }
        bad |= kErrCriticalForce;
        return v > 0.f ? (1ll << 60) : -(1ll << 60);
    }
    return __float2ll_rn(v * (float)(1 << kFixedShift));
}
LIMA_DEV float from_fixed(long long v) {
    return (float)((double)v * (1.0 / (double)(1 << kFixedShift)));
}

// wavefront reduction of a double (wave64)
LIMA_DEV double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

}  // namespace lima
