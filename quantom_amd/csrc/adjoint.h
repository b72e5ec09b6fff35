// adjoint.h — reverse-mode (adjoint) kernels of the LIMA_MD step for parameter
// gradients of trajectory losses.
//
// Forward model (thermostat off, lambda = 1, neighbour lists piecewise constant):
//     x_{k+1} = 2 x_k - x_{k-1} + c * F(x_k; theta),   c = dt^2 / m   (Engine.cu:849-869)
// Loss over frames k = 0..K-1:
//     L = sum_k [ alpha_r (E_k - E*_k)^2 + beta_r sum_i |F_k,i - F*_k,i|^2 ]
// Backward recurrence (a_k = dL/dx_k):
//     w_k = a_{k+1} * c + 2 beta (F_k - F*_k),  g_k = 2 alpha (E_k - E*_k)
//     a_k = G_x(x_k; w_k, g_k) + 2 a_{k+1} - a_{k+2},   dL/dtheta += G_theta(x_k; w_k, g_k)
// where G(w, g) is the gradient of  w . F(x) + g E(x)  with respect to x and theta.
//
// LJ pairs (calcLJForce, Engine.cu:475-522): f = d phi(q), d = x_j - x_i, q = |d|^2,
//     phi = 24e s3 (1 - 2 s3)/q,  phi' = dphi/dq = 24e s3 (14 s3 - 4)/q^2,  s3 = (sigma^2/q)^3,
//     d(w.f)/dd = phi w + 2 phi' (w.d) d.
// The energy of one directed pair is half the pair energy, U/2 = (24e/12)(s3^2 - s3).
// Position gradients are gathered only: symmetric sites (solvent-solvent,
// inter- and intra-compound) combine both directed pairs per thread; the
// asymmetric solvent<->compound sites are evaluated from both sides, in one pass
// per side.  Parameter derivatives of every directed pair are credited by each
// endpoint's thread to its own type only, which gives exactly the mixing-rule
// chain rule (both types receive d/dsigma^2 * sigma_m; arithmetic eps 12 x, geometric
// eps e24/(2 eps_t) x).  Bonded terms are differentiated in forward mode with dual
// numbers over their 6/9/12 coordinates and their two constants.
#pragma once
#include "kernels.h"

namespace lima {

constexpr int kAdjTypeBins = 3 * kMaxTypes;   // per type: sigma, eps (arithmetic sites), eps*e24 (geometric sites)
constexpr int kAdjParams = 2 * kMaxTypes + 4; // sigma[32], eps[32], kb, b0, k_theta, theta0

// ---- dual numbers -----------------------------------------------------------
template <int N> struct Dual {
    float v;
    float d[N];
};

template <int N> LIMA_DEV Dual<N> dconst(float v) {
    Dual<N> r;
    r.v = v;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = 0.f;
    return r;
}
template <int N> LIMA_DEV Dual<N> dvar(float v, int k) {
    Dual<N> r = dconst<N>(v);
    r.d[k] = 1.f;
    return r;
}
template <int N> LIMA_DEV Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int N> LIMA_DEV Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int N> LIMA_DEV Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r;
    r.v = -a.v;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = -a.d[i];
    return r;
}
template <int N> LIMA_DEV Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int N> LIMA_DEV Dual<N> operator*(const Dual<N>& a, float s) {
    Dual<N> r;
    r.v = a.v * s;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] * s;
    return r;
}
template <int N> LIMA_DEV Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    const float inv = 1.0f / b.v;
    r.v = a.v * inv;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
    return r;
}
template <int N> LIMA_DEV Dual<N> dsqrt(const Dual<N>& a) {
    Dual<N> r;
    r.v = sqrtf(a.v);
    const float g = r.v > 0.f ? 0.5f / r.v : 0.f;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] * g;
    return r;
}
template <int N> LIMA_DEV Dual<N> dfun(const Dual<N>& a, float v, float g) {
    Dual<N> r;
    r.v = v;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] * g;
    return r;
}
LIMA_DEV float fval(float a) { return a; }
template <int N> LIMA_DEV float fval(const Dual<N>& a) { return a.v; }

template <class S> struct V3d {
    S x, y, z;
};
template <class S> LIMA_DEV V3d<S> vadd(const V3d<S>& a, const V3d<S>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class S> LIMA_DEV V3d<S> vsub(const V3d<S>& a, const V3d<S>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class S> LIMA_DEV V3d<S> vneg(const V3d<S>& a) { return {-a.x, -a.y, -a.z}; }
template <class S> LIMA_DEV V3d<S> vscale(const V3d<S>& a, const S& s) { return {a.x * s, a.y * s, a.z * s}; }
template <class S> LIMA_DEV V3d<S> vscalef(const V3d<S>& a, float s) { return {a.x * s, a.y * s, a.z * s}; }
template <class S> LIMA_DEV S vdot(const V3d<S>& a, const V3d<S>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class S> LIMA_DEV V3d<S> vcross(const V3d<S>& a, const V3d<S>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <int N> LIMA_DEV Dual<N> vlen(const V3d<Dual<N>>& a) { return dsqrt(vdot(a, a)); }
template <int N> LIMA_DEV V3d<Dual<N>> vnorm(const V3d<Dual<N>>& a) {
    Dual<N> l = vlen(a);
    if (l.v != 0.f) return vscale(a, dconst<N>(1.f) / l);
    return {dconst<N>(0.f), dconst<N>(0.f), dconst<N>(0.f)};
}
// Float3::getAngle with its clamp (QuantomTypes.cuh:74-83)
template <int N> LIMA_DEV Dual<N> dangle(const V3d<Dual<N>>& a, const V3d<Dual<N>>& b) {
    Dual<N> val = vdot(a, b) / (vlen(a) * vlen(b));
    if (val.v >= 1.f) return dconst<N>(0.f);
    if (val.v <= -1.f) return dconst<N>(3.14159265358979f);
    return dfun(val, acosf(val.v), -1.0f / sqrtf(1.0f - val.v * val.v));
}

template <int N> LIMA_DEV V3d<Dual<N>> dpos(float4 p, int k0) {
    return {dvar<N>(p.x, k0), dvar<N>(p.y, k0 + 1), dvar<N>(p.z, k0 + 2)};
}
template <int N> LIMA_DEV void dot_w(const V3d<Dual<N>>& f, float4 w, float* out) {
#pragma unroll
    for (int i = 0; i < N; i++) out[i] += w.x * f.x.d[i] + w.y * f.y.d[i] + w.z * f.z.d[i];
}

// Gradient of  sum_atoms w.F + g E  for one bond (calcPairbondForces, Engine.cu:525-548):
// out[0..5] positions, out[6] b0, out[7] kb
LIMA_DEV void bond_vjp(float4 a, float4 b, float b0, float kb, float4 wa, float4 wb, float g, float* out) {
    constexpr int N = 8;
    auto A = dpos<N>(a, 0), B = dpos<N>(b, 3);
    Dual<N> B0 = dvar<N>(b0, 6), KB = dvar<N>(kb, 7);
    auto diff = vsub(A, B);
    Dual<N> err = vlen(diff) - B0;
    Dual<N> E = KB * err * err * 0.5f;
    auto fa = vscale(vnorm(diff), -(KB * err));
    for (int i = 0; i < N; i++) out[i] = 0.f;
    dot_w(fa, wa, out);
    dot_w(vneg(fa), wb, out);
#pragma unroll
    for (int i = 0; i < N; i++) out[i] += g * E.d[i];
}

// one angle (calcAnglebondForces, Engine.cu:552-578): out[0..8] positions, out[9] theta0, out[10] k
LIMA_DEV void angle_vjp(float4 l, float4 m, float4 r, float t0, float k, float4 wl, float4 wm, float4 wr, float g,
                        float* out) {
    constexpr int N = 11;
    auto L = dpos<N>(l, 0), M = dpos<N>(m, 3), R = dpos<N>(r, 6);
    Dual<N> T0 = dvar<N>(t0, 9), K = dvar<N>(k, 10);
    auto v1 = vsub(L, M), v2 = vsub(R, M);
    auto dir1 = vnorm(vcross(v1, vcross(v2, v1)));
    auto dir2 = vnorm(vcross(v2, vcross(v1, v2)));
    Dual<N> err = dangle(v1, v2) - T0;
    Dual<N> E = K * err * err * 0.25f;
    Dual<N> fsc = K * err;
    auto fl = vscale(dir1, fsc), fr = vscale(dir2, fsc);
    auto fm = vneg(vadd(fl, fr));
    for (int i = 0; i < N; i++) out[i] = 0.f;
    dot_w(fl, wl, out);
    dot_w(fm, wm, out);
    dot_w(fr, wr, out);
#pragma unroll
    for (int i = 0; i < N; i++) out[i] += g * E.d[i];
}

// one dihedral (calcDihedralbondForces, Engine.cu:580-637): out[0..11] positions
LIMA_DEV void dihedral_vjp(float4 l, float4 lm, float4 rm, float4 r, float phi0, float k, int n, float4 wl, float4 wlm,
                           float4 wrm, float4 wr, float g, float* out) {
    constexpr int N = 12;
    auto L = dpos<N>(l, 0), LM = dpos<N>(lm, 3), RM = dpos<N>(rm, 6), R = dpos<N>(r, 9);
    auto n1 = vnorm(vcross(vsub(L, LM), vsub(RM, LM)));
    auto n2 = vnorm(vcross(vsub(LM, RM), vsub(R, RM)));
    Dual<N> tors = dangle(n1, n2);
    const float lp = fval(vlen(vsub(vadd(L, n1), R))), lmn = fval(vlen(vsub(vsub(L, n1), R)));
    if (lp < lmn) n2 = vneg(n2);
    else n1 = vneg(n1);
    Dual<N> arg = tors * (float)n - dconst<N>(phi0);
    Dual<N> fs = dfun(arg, k * sinf(arg.v), k * cosf(arg.v));
    Dual<N> E = dfun(arg, k * (1.0f - cosf(arg.v)), k * sinf(arg.v));
    auto fl = vscale(n1, fs), fr = vscale(n2, fs);
    auto fmid = vscalef(vadd(fl, fr), -0.5f);
    for (int i = 0; i < N; i++) out[i] = 0.f;
    dot_w(fl, wl, out);
    dot_w(fmid, wlm, out);
    dot_w(fmid, wrm, out);
    dot_w(fr, wr, out);
#pragma unroll
    for (int i = 0; i < N; i++) out[i] += g * E.d[i];
}

// ---- LJ pair derivatives -----------------------------------------------------
struct PairDeriv {
    float phi, dphi;          // phi(q), dphi/dq
    float s3, inv;
};
LIMA_DEV PairDeriv pair_deriv(float q, float sig2, float e24) {
    PairDeriv p;
    p.inv = __builtin_amdgcn_rcpf(q);   // 1 ulp, as the force's reciprocal
    const float s = sig2 * p.inv;
    p.s3 = s * s * s;
    p.phi = e24 * p.s3 * p.inv * (1.0f - 2.0f * p.s3);
    p.dphi = e24 * p.s3 * p.inv * p.inv * (14.0f * p.s3 - 4.0f);
    return p;
}
// d(w.f)/dd for f = d phi(q)
LIMA_DEV float3 pair_h(const PairDeriv& p, float3 d, float3 w) {
    const float wd = w.x * d.x + w.y * d.y + w.z * d.z;
    const float c = 2.0f * p.dphi * wd;
    return make_float3(p.phi * w.x + c * d.x, p.phi * w.y + c * d.y, p.phi * w.z + c * d.z);
}
// derivatives of  wd * phi + gU * U  with respect to e24 and sigma^2, where U is the
// full pair energy (e24/6)(s3^2 - s3) scaled by gU
LIMA_DEV void pair_param(const PairDeriv& p, float wd, float gU, float sig2, float e24, float& de24, float& dsig2) {
    de24 = wd * p.s3 * (1.0f - 2.0f * p.s3) * p.inv + gU * (p.s3 * p.s3 - p.s3) * (1.0f / 6.0f);
    dsig2 = (wd * e24 * p.s3 * (3.0f - 12.0f * p.s3) * p.inv + gU * e24 * p.s3 * (2.0f * p.s3 - 1.0f) * 0.5f) *
            __builtin_amdgcn_rcpf(sig2);
}

constexpr int kAdjStride = kAdjTypeBins + 4;   // per replica: type bins, then kb, b0, k_theta, theta0
constexpr int kAdjCompChunk = 8;               // most compounds per block of the solvent-side pass

struct AdjArgs {
    StepArgs s;              // the frame's positions (cpos, spos), its lists and the type tables
    const float4* wc;        // [NCH] weight of the chain forces
    const float4* ws;        // [NS] weight of the solvent forces
    const float* gE;         // [R] weight of the frame's potential energy
    float4* gxc;             // [1 + solvent slices][NCH]: slot 0 listed/intra/bonded (+ bridges)
    float4* gxs;             // [compound chunks][NS]
    double* gpar;            // [R][kAdjStride]
    const uint32_t* cs_mask; // per compound, bitmask over its replica's solvents: listed by the compound
    const int* mask_off;     // [n_comp] first word of each compound's mask
    int n_chain;             // NCH: stride of the chain parts
    float4* gpart;           // [work item][128]: listed-term parts of the compound side, per step work item
    int comp_chunk;          // compounds per block of the solvent-side pass (<= kAdjCompChunk)
    int n_cchunks;           // compound chunks; grid row n_cchunks takes the solvent-solvent pairs
    int solv_slice;          // solvents per block of the chain-side all-pairs pass (multiple of kSolventChunk)
};

// One directed LJ site seen from particle P: d = x_other - x_P, u = the force
// weight as seen from P (w_P when P receives the force, -w_receiver when P is
// the source), gU = weight of the site's energy.  Adds the position gradient of
// u.f + gU U to acc and the derivatives with respect to (sigma^2, 24 eps).
// Sums over a particle's thousands of pairs cancel heavily, so the gathers
// accumulate in FP64 (the per-pair terms stay FP32).
struct D3 {
    double x, y, z;
};
LIMA_DEV void site(float3 d, float3 u, float gU, float sig2, float e24, D3& acc, double& dsig2, double& de24) {
    const PairDeriv p = pair_deriv(d.x * d.x + d.y * d.y + d.z * d.z, sig2, e24);
    const float3 h = pair_h(p, d, u);
    const float ge = gU * p.phi;
    acc.x -= (double)(h.x + ge * d.x);
    acc.y -= (double)(h.y + ge * d.y);
    acc.z -= (double)(h.z + ge * d.z);
    float de, dsg;
    pair_param(p, u.x * d.x + u.y * d.y + u.z * d.z, gU, sig2, e24, de, dsg);
    dsig2 += (double)dsg;
    de24 += (double)de;
}

LIMA_DEV void bins_zero(double* bins) {
    for (int i = threadIdx.x; i < kAdjTypeBins; i += blockDim.x) bins[i] = 0.0;
    __syncthreads();
}
// type t's credits: sigma, arithmetic 24eps, geometric 24eps (already times 24eps)
LIMA_DEV void bins_credit(double* bins, int t, double s, double ea, double eg) {
    if (s != 0.0) atomicAdd(&bins[t], s);
    if (ea != 0.0) atomicAdd(&bins[kMaxTypes + t], ea);
    if (eg != 0.0) atomicAdd(&bins[2 * kMaxTypes + t], eg);
}
LIMA_DEV void bins_flush(const double* bins, double* g) {
    __syncthreads();
    for (int i = threadIdx.x; i < kAdjTypeBins; i += blockDim.x)
        if (bins[i] != 0.0) atomicAdd(&g[i], bins[i]);
}
// flush the block's bins and clear them for the next replica (each thread clears the entries it flushed)
LIMA_DEV void bins_flush_reset(double* bins, double* g) {
    __syncthreads();
    for (int i = threadIdx.x; i < kAdjTypeBins; i += blockDim.x) {
        if (bins[i] != 0.0) atomicAdd(&g[i], bins[i]);
        bins[i] = 0.0;
    }
    __syncthreads();
}
LIMA_DEV float3 xyz(float4 v) { return make_float3(v.x, v.y, v.z); }
LIMA_DEV float3 neg3(float3 v) { return make_float3(-v.x, -v.y, -v.z); }

// ---------------------------------------------------------------------------
// Solvent side.  Block = (solvent tile of 256, chunk of comp_chunk <= 8 compounds).
// For every chain particle j of the chunk: the all-pairs site "solvent <- j"
// (arithmetic eps, Engine.cu:1099-1117) and, when the solvent is in j's
// compound list, the transpose of "j <- solvent" (geometric eps,
// Engine.cu:964-974).  The grid's last row takes the solvent-solvent pairs, both
// directions of each symmetric pair (Engine.cu:1119-1128).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void adj_solvent_kernel(AdjArgs A) {
    const StepArgs& a = A.s;
    __shared__ float4 cp[kAdjCompChunk * kCompoundBlock];
    __shared__ float4 cw[kAdjCompChunk * kCompoundBlock];
    __shared__ double bins[kAdjTypeBins];
    bins_zero(bins);
    const int4 tile = a.stiles[blockIdx.x];
    const int r = tile.x;
    const ReplicaInfo rp = a.reps[r];
    const float L = rp.L, half = rp.half;
    const int s = tile.y + threadIdx.x;
    const bool active = s < tile.z;
    const int si = active ? s : tile.y;
    const float4 me = a.spos[si];
    const float3 w = xyz(A.ws[si]);
    const float g = A.gE[r];
    const int sl = si - rp.solv_begin;
    D3 acc{0.0, 0.0, 0.0};
    double bs = 0.0, bea = 0.0, beg = 0.0;
    const int c0 = rp.comp_begin + blockIdx.y * A.comp_chunk;
    const int c1 = min(c0 + A.comp_chunk, rp.comp_end);
    if (c0 < c1) {
        const int pb = a.comp_off[c0], np = a.comp_off[c1] - pb;
        for (int k = threadIdx.x; k < np; k += blockDim.x) {
            cp[k] = a.cpos[pb + k];
            cw[k] = A.wc[pb + k];
        }
        __syncthreads();
        for (int c = c0; c < c1; c++) {
            const int lb = a.comp_off[c] - pb, le = a.comp_off[c + 1] - pb;
            const float4 p0 = cp[lb];
            float ix = me.x, iy = me.y, iz = me.z;
            hyperpos(p0.x, p0.y, p0.z, ix, iy, iz, L, half);
            const bool listed = a.enable_solvents && ((A.cs_mask[A.mask_off[c] + (sl >> 5)] >> (sl & 31)) & 1u);
            for (int j = lb; j < le; j++) {
                float4 q = cp[j];
                const int tj = ftype(q);
                const float4 pa = a.tab_sc[r * kMaxTypes + tj];
                double ds = 0.0, de = 0.0;
                site(make_float3(q.x - ix, q.y - iy, q.z - iz), w, 0.5f * g, pa.x, pa.y, acc, ds, de);
                bs += ds * pa.w;
                bea += de;
                if (listed) {
                    hyperpos(p0.x, p0.y, p0.z, q.x, q.y, q.z, L, half);
                    const float4 pg = a.tab_cs[r * kMaxTypes + tj];
                    ds = 0.0; de = 0.0;
                    site(make_float3(q.x - ix, q.y - iy, q.z - iz), neg3(xyz(cw[j])), 0.5f * g, pg.x, pg.y, acc, ds, de);
                    bs += ds * pg.w;
                    beg += de * pg.y;
                }
            }
        }
    }
    if (blockIdx.y == A.n_cchunks) {
        const float4 pss = a.tab_ss[r];
        const int cnt = active ? a.ss_cnt[si] : 0;
        for (int k = 0; k < cnt; k++) {
            const int j = a.ss_ids[(size_t)si * a.ss_ld + k];
            float4 q = a.spos[j];
            hyperpos(me.x, me.y, me.z, q.x, q.y, q.z, L, half);
            const float3 wj = xyz(A.ws[j]);
            double ds = 0.0, de = 0.0;
            site(make_float3(q.x - me.x, q.y - me.y, q.z - me.z), make_float3(w.x - wj.x, w.y - wj.y, w.z - wj.z), g,
                 pss.x, pss.y, acc, ds, de);
            bs += ds * pss.w;
            bea += de;
        }
    }
    if (active) {
        A.gxs[(size_t)blockIdx.y * a.n_solv + s] = make_float4((float)acc.x, (float)acc.y, (float)acc.z, 0.f);
        bins_credit(bins, 0, bs, bea, beg);
    }
    bins_flush(bins, A.gpar + (size_t)r * kAdjStride);
}

// ---------------------------------------------------------------------------
// Chain side of the all-pairs site.  Block = (compound, slice of solv_slice
// solvents); one thread per chain particle: the transpose of "solvent <- j" with
// the solvent imaged against the compound's particle 0 and j's stored position,
// as the forward pass pairs them.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(128) void adj_chain_solvent_kernel(AdjArgs A) {
    const StepArgs& a = A.s;
    __shared__ float4 sp[kSolventChunk];
    __shared__ float4 sw[kSolventChunk];
    __shared__ double bins[kAdjTypeBins];
    bins_zero(bins);
    const int c = blockIdx.x, p = threadIdx.x;
    const int r = a.comp_rep[c];
    const ReplicaInfo rp = a.reps[r];
    const float L = rp.L, half = rp.half;
    const int base = a.comp_off[c], n = a.comp_off[c + 1] - base;
    const bool active = p < n;
    const int gi = base + (active ? p : 0);
    const float4 me = a.cpos[gi];
    const float4 p0 = a.cpos[base];
    const int tp = ftype(me);
    const float4 pa = a.tab_sc[r * kMaxTypes + tp];
    const float g = A.gE[r];
    D3 acc{0.0, 0.0, 0.0};
    double bs = 0.0, bea = 0.0;
    const int s0 = rp.solv_begin + blockIdx.y * A.solv_slice;
    const int s1 = min(s0 + A.solv_slice, rp.solv_end);
    for (int t0 = s0; t0 < s1; t0 += kSolventChunk) {
        const int ns = min(kSolventChunk, s1 - t0);
        __syncthreads();
        for (int k = p; k < ns; k += blockDim.x) {
            float4 q = a.spos[t0 + k];
            hyperpos(p0.x, p0.y, p0.z, q.x, q.y, q.z, L, half);
            sp[k] = q;
            sw[k] = A.ws[t0 + k];
        }
        __syncthreads();
        for (int k = 0; k < ns; k++) {
            const float4 q = sp[k];
            site(make_float3(q.x - me.x, q.y - me.y, q.z - me.z), neg3(xyz(sw[k])), 0.5f * g, pa.x, pa.y, acc, bs, bea);
        }
    }
    if (active) {
        A.gxc[(size_t)(1 + blockIdx.y) * A.n_chain + gi] = make_float4((float)acc.x, (float)acc.y, (float)acc.z, 0.f);
        bins_credit(bins, tp, bs * pa.w, bea, 0.0);
    }
    bins_flush(bins, A.gpar + (size_t)r * kAdjStride);
}

// ---------------------------------------------------------------------------
// Compound side of the listed terms, on the step's work list (compound, slot) as
// compound_listed_kernel walks it: a slot below the compound's neighbour count
// is one neighbour compound, both directions of each pair (arithmetic eps,
// Engine.cu:653-676; the pair enters p's force unless p excludes j and j's
// force unless j excludes p); a later slot is 256 entries of the compound's
// solvent list, "particle <- solvent" (geometric eps, Engine.cu:964-974).  One
// part per work item; parameter credits per replica.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(128) void adj_listed_kernel(AdjArgs A) {
    const StepArgs& a = A.s;
    __shared__ float4 sp[kSolventChunk];
    __shared__ float4 sw[kSolventChunk];
    __shared__ uint32_t nbmask[kCompoundBlock * 4];   // the neighbour's exclusions of this compound's particles
    __shared__ double bins[kAdjTypeBins];
    bins_zero(bins);
    const int p = threadIdx.x;
    const int total = *a.cw_total;
    int cur_r = -1;
    for (int wi = blockIdx.x; wi < total; wi += gridDim.x) {
        const int c = find_compound(a.cw_off, a.n_comp, wi);
        const int k = wi - a.cw_off[c];
        const int r = a.comp_rep[c];
        if (r != cur_r) {
            if (cur_r >= 0) bins_flush_reset(bins, A.gpar + (size_t)cur_r * kAdjStride);
            cur_r = r;
        }
        const ReplicaInfo rp = a.reps[r];
        const float L = rp.L, half = rp.half;
        const int base = a.comp_off[c], n = a.comp_off[c + 1] - base;
        const bool active = p < n;
        const int gi = base + (active ? p : 0);
        float4 me = a.cpos[gi];
        const float4 p0 = a.cpos[base];
        hyperpos(p0.x, p0.y, p0.z, me.x, me.y, me.z, L, half);
        const float3 w = xyz(A.wc[gi]);
        const float g = A.gE[r];
        const int tp = ftype(me);
        D3 acc{0.0, 0.0, 0.0};
        double bs = 0.0, bea = 0.0, beg = 0.0;
        const int ncc = a.cc_cnt[c];
        __syncthreads();
        if (k < ncc) {
            const int nb = a.cc_ids[(size_t)c * a.cap_cc + k];
            const int nbase = a.comp_off[nb], nn = a.comp_off[nb + 1] - nbase;
            if (p < nn) {
                float4 q = a.cpos[nbase + p];
                hyperpos(p0.x, p0.y, p0.z, q.x, q.y, q.z, L, half);
                sp[p] = q;
                sw[p] = A.wc[nbase + p];
                uint32_t m[4] = {0u, 0u, 0u, 0u};
                for (int e = a.xoff[nbase + p]; e < a.xoff[nbase + p + 1]; e++) {
                    const int2 x = a.xent[e];
                    if (x.x == c) m[x.y >> 5] |= 1u << (x.y & 31);
                }
                for (int q4 = 0; q4 < 4; q4++) nbmask[p * 4 + q4] = m[q4];
            }
            __syncthreads();
            if (active) {
                uint32_t m[4] = {0u, 0u, 0u, 0u};
                for (int e = a.xoff[gi]; e < a.xoff[gi + 1]; e++) {
                    const int2 x = a.xent[e];
                    if (x.x == nb) m[x.y >> 5] |= 1u << (x.y & 31);
                }
                for (int j = 0; j < nn; j++) {
                    const bool in_p = !((m[j >> 5] >> (j & 31)) & 1u);
                    const bool in_j = !((nbmask[j * 4 + (p >> 5)] >> (p & 31)) & 1u);
                    if (!in_p && !in_j) continue;
                    const float4 q = sp[j];
                    const float4 pr = a.tab_inter[(size_t)r * kMaxTypes * kMaxTypes + ftype(q) * kMaxTypes + tp];
                    const float3 wj = xyz(sw[j]);
                    const float3 u = make_float3((in_p ? w.x : 0.f) - (in_j ? wj.x : 0.f),
                                                 (in_p ? w.y : 0.f) - (in_j ? wj.y : 0.f),
                                                 (in_p ? w.z : 0.f) - (in_j ? wj.z : 0.f));
                    double ds = 0.0, de = 0.0;
                    site(make_float3(q.x - me.x, q.y - me.y, q.z - me.z), u, 0.5f * g * (float)((int)in_p + (int)in_j),
                         pr.x, pr.y, acc, ds, de);
                    bs += ds * pr.w;
                    bea += de;
                }
            }
        } else {
            const int j0 = (k - ncc) * kSolventChunk;
            const int ns = min(kSolventChunk, a.cs_cnt[c] - j0);
            for (int e = p; e < ns; e += blockDim.x) {
                float4 q = a.spos[a.cs_ids[(size_t)c * a.cap_cs + j0 + e]];
                hyperpos(p0.x, p0.y, p0.z, q.x, q.y, q.z, L, half);
                sp[e] = q;
            }
            __syncthreads();
            if (active) {
                const float4 pg = a.tab_cs[r * kMaxTypes + tp];
                double ds = 0.0, de = 0.0;
                for (int e = 0; e < ns; e++) {
                    const float4 q = sp[e];
                    site(make_float3(q.x - me.x, q.y - me.y, q.z - me.z), w, 0.5f * g, pg.x, pg.y, acc, ds, de);
                }
                bs += ds * pg.w;
                beg += de * pg.y;
            }
        }
        if (active) {
            A.gpart[(size_t)wi * kCompoundBlock + p] = make_float4((float)acc.x, (float)acc.y, (float)acc.z, 0.f);
            bins_credit(bins, tp, bs, bea, beg);
        }
    }
    if (cur_r >= 0) bins_flush_reset(bins, A.gpar + (size_t)cur_r * kAdjStride);
}

// ---------------------------------------------------------------------------
// Compound side, one block per compound: intra-compound pairs (geometric eps,
// both directions, exclusions of both particles), the bonded terms (dual
// numbers; per-term coordinate gradients gathered per particle in term order)
// and the listed-term parts of adj_listed_kernel.  Bonded-constant derivatives
// are summed over terms.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(128) void adj_compound_kernel(AdjArgs A) {
    const StepArgs& a = A.s;
    __shared__ float4 pos[kCompoundBlock];
    __shared__ float4 wv[kCompoundBlock];
    __shared__ double bins[kAdjTypeBins];
    __shared__ float3 gb[kMaxBondsPerCompound * 2];
    __shared__ float3 ga[kMaxAnglesPerCompound * 3];
    __shared__ float3 gd[kMaxDihedralsPerCompound * 4];
    __shared__ float rpar[4][kCompoundBlock];
    bins_zero(bins);
    const int c = blockIdx.x, p = threadIdx.x;
    const int r = a.comp_rep[c];
    const ReplicaInfo rp = a.reps[r];
    const float L = rp.L, half = rp.half;
    const int base = a.comp_off[c], n = a.comp_off[c + 1] - base;
    const bool active = p < n;
    const int gi = base + (active ? p : 0);
    float4 me = a.cpos[gi];
    const float4 p0 = a.cpos[base];
    hyperpos(p0.x, p0.y, p0.z, me.x, me.y, me.z, L, half);
    const float3 w = xyz(A.wc[gi]);
    const float g = A.gE[r];
    const int tp = ftype(me);
    pos[p] = me;
    wv[p] = make_float4(w.x, w.y, w.z, 0.f);
    D3 acc{0.0, 0.0, 0.0};
    double bs = 0.0, bea = 0.0, beg = 0.0;
    __syncthreads();
    // (3) intra-compound pairs, geometric eps (Engine.cu:678-700), exclusions of both particles
    if (active) {
        const uint32_t* mk = a.intra_mask + (size_t)gi * 4;
        const uint32_t m0 = mk[0], m1 = mk[1], m2 = mk[2], m3 = mk[3];
        for (int j = 0; j < n; j++) {
            if (j == p) continue;
            const uint32_t mw = (j >> 5) == 0 ? m0 : (j >> 5) == 1 ? m1 : (j >> 5) == 2 ? m2 : m3;
            const bool in_p = !((mw >> (j & 31)) & 1u);
            const bool in_j = !((a.intra_mask[(size_t)(base + j) * 4 + (p >> 5)] >> (p & 31)) & 1u);
            if (!in_p && !in_j) continue;
            const float4 q = pos[j];
            const float4 pr = a.tab_intra[(size_t)r * kMaxTypes * kMaxTypes + ftype(q) * kMaxTypes + tp];
            const float3 wj = xyz(wv[j]);
            const float3 u = make_float3((in_p ? w.x : 0.f) - (in_j ? wj.x : 0.f), (in_p ? w.y : 0.f) - (in_j ? wj.y : 0.f),
                                         (in_p ? w.z : 0.f) - (in_j ? wj.z : 0.f));
            double ds = 0.0, de = 0.0;
            site(make_float3(q.x - me.x, q.y - me.y, q.z - me.z), u, 0.5f * g * (float)((int)in_p + (int)in_j), pr.x,
                 pr.y, acc, ds, de);
            bs += ds * pr.w;
            beg += de * pr.y;
        }
    }
    // (4) bonded terms (Engine.cu:525-637), one thread per term
    float kb_g = 0.f, b0_g = 0.f, kt_g = 0.f, t0_g = 0.f;
    const int b0i = a.bond_off[c], nbd = a.bond_off[c + 1] - b0i;
    const int a0i = a.angle_off[c], nan_ = a.angle_off[c + 1] - a0i;
    const int d0i = a.dih_off[c], ndh = a.enable_dihedrals ? a.dih_off[c + 1] - d0i : 0;
    for (int t = p; t < nbd; t += blockDim.x) {
        const BondD bd = a.bonds[b0i + t];
        float o[8];
        bond_vjp(pos[bd.a0], pos[bd.a1], bd.b0, bd.kb, wv[bd.a0], wv[bd.a1], g, o);
        gb[t * 2] = make_float3(o[0], o[1], o[2]);
        gb[t * 2 + 1] = make_float3(o[3], o[4], o[5]);
        b0_g += o[6];
        kb_g += o[7];
    }
    for (int t = p; t < nan_; t += blockDim.x) {
        const AngleD ad = a.angles[a0i + t];
        float o[11];
        angle_vjp(pos[ad.a0], pos[ad.a1], pos[ad.a2], ad.t0, ad.k, wv[ad.a0], wv[ad.a1], wv[ad.a2], g, o);
        for (int q = 0; q < 3; q++) ga[t * 3 + q] = make_float3(o[3 * q], o[3 * q + 1], o[3 * q + 2]);
        t0_g += o[9];
        kt_g += o[10];
    }
    for (int t = p; t < ndh; t += blockDim.x) {
        const DihD dd = a.dihs[d0i + t];
        float o[12];
        dihedral_vjp(pos[dd.a0], pos[dd.a1], pos[dd.a2], pos[dd.a3], dd.p0, dd.k, dd.n, wv[dd.a0], wv[dd.a1], wv[dd.a2],
                     wv[dd.a3], g, o);
        for (int q = 0; q < 4; q++) gd[t * 4 + q] = make_float3(o[3 * q], o[3 * q + 1], o[3 * q + 2]);
    }
    rpar[0][p] = kb_g;
    rpar[1][p] = b0_g;
    rpar[2][p] = kt_g;
    rpar[3][p] = t0_g;
    __syncthreads();
    if (active) {
        for (int e = a.inc_off[gi]; e < a.inc_off[gi + 1]; e++) {
            const uint32_t v = a.inc[e];
            const uint32_t kind = v >> 30, slot = v & 3u, term = (v >> 2) & 0xFFFFFFFu;
            if (kind == 2 && !a.enable_dihedrals) continue;
            const float3 x = kind == 0 ? gb[term * 2 + slot] : kind == 1 ? ga[term * 3 + slot] : gd[term * 4 + slot];
            acc.x += x.x;
            acc.y += x.y;
            acc.z += x.z;
        }
        // listed terms (neighbour compounds and listed solvents), one part per work item in slot order
        for (int wi = a.cw_off[c]; wi < a.cw_off[c + 1]; wi++) {
            const float4 v = A.gpart[(size_t)wi * kCompoundBlock + p];
            acc.x += v.x;
            acc.y += v.y;
            acc.z += v.z;
        }
        A.gxc[gi] = make_float4((float)acc.x, (float)acc.y, (float)acc.z, 0.f);
        bins_credit(bins, tp, bs, bea, beg);
    }
    double* gr = A.gpar + (size_t)r * kAdjStride;
    if (p < 4) {
        double sum = 0.0;
        for (int i = 0; i < kCompoundBlock; i++) sum += (double)rpar[p][i];
        if (sum != 0.0) atomicAdd(&gr[kAdjTypeBins + p], sum);
    }
    bins_flush(bins, gr);
}

// Bridges (Engine.cu:897-916): one thread per bridge; the coordinate gradients go
// to slot 0 of the chain parts after the compound pass (a particle belongs to at
// most one bridge, so the updates do not collide).
__global__ __launch_bounds__(64) void adj_bridge_kernel(AdjArgs A) {
    const StepArgs& a = A.s;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.n_bridges) return;
    const int r = a.br_rep[b];
    const ReplicaInfo rp = a.reps[r];
    const float g = A.gE[r];
    const int q0 = a.br_off[b], np = min(a.br_off[b + 1] - q0, 32);
    float4 pos[32], wv[32];
    float3 gx[32];
    for (int k = 0; k < np; k++) {
        pos[k] = a.cpos[a.br_glob[q0 + k]];
        wv[k] = A.wc[a.br_glob[q0 + k]];
        gx[k] = make_float3(0.f, 0.f, 0.f);
    }
    for (int k = 1; k < np; k++) hyperpos(pos[0].x, pos[0].y, pos[0].z, pos[k].x, pos[k].y, pos[k].z, rp.L, rp.half);
    float kb = 0.f, b0 = 0.f, kt = 0.f, t0 = 0.f;
    auto add = [&](int k, const float* o) {
        gx[k].x += o[0];
        gx[k].y += o[1];
        gx[k].z += o[2];
    };
    for (int t = a.brb_off[b]; t < a.brb_off[b + 1]; t++) {
        const BondD bd = a.brb[t];
        float o[8];
        bond_vjp(pos[bd.a0], pos[bd.a1], bd.b0, bd.kb, wv[bd.a0], wv[bd.a1], g, o);
        add(bd.a0, o);
        add(bd.a1, o + 3);
        b0 += o[6];
        kb += o[7];
    }
    for (int t = a.bra_off[b]; t < a.bra_off[b + 1]; t++) {
        const AngleD ad = a.bra[t];
        float o[11];
        angle_vjp(pos[ad.a0], pos[ad.a1], pos[ad.a2], ad.t0, ad.k, wv[ad.a0], wv[ad.a1], wv[ad.a2], g, o);
        add(ad.a0, o);
        add(ad.a1, o + 3);
        add(ad.a2, o + 6);
        t0 += o[9];
        kt += o[10];
    }
    if (a.enable_dihedrals) {
        for (int t = a.brd_off[b]; t < a.brd_off[b + 1]; t++) {
            const DihD dd = a.brd[t];
            float o[12];
            dihedral_vjp(pos[dd.a0], pos[dd.a1], pos[dd.a2], pos[dd.a3], dd.p0, dd.k, dd.n, wv[dd.a0], wv[dd.a1],
                         wv[dd.a2], wv[dd.a3], g, o);
            add(dd.a0, o);
            add(dd.a1, o + 3);
            add(dd.a2, o + 6);
            add(dd.a3, o + 9);
        }
    }
    for (int k = 0; k < np; k++) {
        float4* dst = A.gxc + a.br_glob[q0 + k];
        float4 v = *dst;
        *dst = make_float4(v.x + gx[k].x, v.y + gx[k].y, v.z + gx[k].z, 0.f);
    }
    double* gr = A.gpar + (size_t)r * kAdjStride;
    if (kb != 0.f) atomicAdd(&gr[kAdjTypeBins + 0], (double)kb);
    if (b0 != 0.f) atomicAdd(&gr[kAdjTypeBins + 1], (double)b0);
    if (kt != 0.f) atomicAdd(&gr[kAdjTypeBins + 2], (double)kt);
    if (t0 != 0.f) atomicAdd(&gr[kAdjTypeBins + 3], (double)t0);
}

// bitmask of each compound's solvent list over its replica's solvents
__global__ void adj_mask_zero_kernel(uint32_t* m, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) m[i] = 0u;
}
__global__ __launch_bounds__(256) void adj_mask_set_kernel(StepArgs a, uint32_t* mask, const int* mask_off) {
    const int c = blockIdx.x;
    const ReplicaInfo rp = a.reps[a.comp_rep[c]];
    const int cnt = a.cs_cnt[c];
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
        const int sl = a.cs_ids[(size_t)c * a.cap_cs + k] - rp.solv_begin;
        atomicOr(&mask[mask_off[c] + (sl >> 5)], 1u << (sl & 31));
    }
}

// Per particle: the force residual r = 2 beta (F - F*) of the frame (targets as
// x, y, z triples, or none) and, when `lacc` is given, beta |F - F*|^2 added to the
// particle's own FP64 loss accumulator (one owner per entry: no atomics).
__global__ __launch_bounds__(256) void adj_residual_kernel(const float4* f, const float* ftarget, const int* prep,
                                                           const double* beta, float4* rout, double* lacc, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double b = beta[prep[i]];
    const float4 F = f[i];
    const float3 T = ftarget ? make_float3(ftarget[3 * i], ftarget[3 * i + 1], ftarget[3 * i + 2])
                             : make_float3(0.f, 0.f, 0.f);
    const float dx = F.x - T.x, dy = F.y - T.y, dz = F.z - T.z;
    const float tb = (float)(2.0 * b);
    rout[i] = make_float4(tb * dx, tb * dy, tb * dz, 0.f);
    if (lacc) lacc[i] += b * ((double)dx * dx + (double)dy * dy + (double)dz * dz);
}

// per system: the sum of its particles' loss accumulators (chain range, then solvent range)
__global__ __launch_bounds__(256) void adj_loss_reduce_kernel(const ReplicaInfo* reps, const double* lacc, int nch,
                                                              double* loss) {
    __shared__ double red[4];
    const ReplicaInfo rp = reps[blockIdx.x];
    double s = 0.0;
    for (int i = rp.chain_begin + threadIdx.x; i < rp.chain_end; i += blockDim.x) s += lacc[i];
    for (int i = rp.solv_begin + threadIdx.x; i < rp.solv_end; i += blockDim.x) s += lacc[nch + i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) loss[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// w_k = (dt^2/m) a_{k+1} + r_k
__global__ __launch_bounds__(256) void adj_weight_kernel(const float4* resid, const double4* a_next, const float* cfac,
                                                         float4* w, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 r = resid[i];
    const double4 an = a_next[i];
    const double c = cfac[i];
    w[i] = make_float4((float)(an.x * c + r.x), (float)(an.y * c + r.y), (float)(an.z * c + r.z), 0.f);
}

// a_k = G_x + 2 a_{k+1} - a_{k+2} in FP64 (the recurrence runs over the whole
// trajectory), G_x summed over its parts in slot order
__global__ __launch_bounds__(256) void adj_update_kernel(const float4* gxc, int nc_slots, const float4* gxs, int ns_slots,
                                                         const double4* a1, const double4* a2, double4* out, int nch,
                                                         int ns) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nch + ns) return;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (i < nch) {
        for (int k = 0; k < nc_slots; k++) {
            const float4 v = gxc[(size_t)k * nch + i];
            gx += v.x; gy += v.y; gz += v.z;
        }
    } else {
        for (int k = 0; k < ns_slots; k++) {
            const float4 v = gxs[(size_t)k * ns + (i - nch)];
            gx += v.x; gy += v.y; gz += v.z;
        }
    }
    const double4 x = a1[i], y = a2[i];
    out[i] = make_double4(gx + 2.0 * x.x - y.x, gy + 2.0 * x.y - y.y, gz + 2.0 * x.z - y.z, 0.0);
}

}  // namespace lima
