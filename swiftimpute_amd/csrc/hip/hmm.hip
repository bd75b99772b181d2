// Li-Stephens forward-backward, posterior dosage and Philox sampling for gfx950.
//
// Reference semantics (src/kernels/forward_backward.cu:41-317,
// src/kernels/sampling.cu:52-131): with the -999 floor of every log-sum-exp
// fold written out, the quantities are
//   forward   a_m[j] ~ (e^-999 + sum_i a_{m-1}[i] t_ij) e^{E_m[j]}   (rescaled per marker)
//   backward  b_m[i] =  e^-999 + sum_j t_ij e^{E_{m+1}[j]} b_{m+1}[j] (never rescaled)
//   posterior p_m[k] = a_m[k] b_m[k] / sum_l a_m[l] b_m[l]
// with t_ij = e^To + (e^Td - e^To) [i == j]: rank one plus diagonal, so a
// step costs O(L), not O(L^2) log-sum-exps, and no matrix unit is involved.
//
// Two numeric forms share the same structure (template parameter LOG):
//   * linear FP32 (default): forward sum-normalized, a'_j = (t_o + u a_j) eps_j
//     (the floor term is below FP32 resolution whenever To >= -80, checked on
//     the host); backward b_m = e^{Bl_m} v_m with the log scale Bl_m in FP64,
//     so the floor enters exactly as f = exp(-999 - Bl_m) once b reaches it.
//     exp/log run once per (sample, marker), not per state.
//   * log-space FP32: forward log vector normalized to maximum 0 (as the
//     reference normalizes, so its -999 floor weighs the same), its log mass
//     ls = LSE(alpha) carried beside it; backward log vector relative to an
//     FP64 offset; every step is a three-term log-sum-exp
//     LSE(-999, To + ls, U + alpha_j) per state, U = log(e^Td - e^To).
//     Valid for any To, including the -999 of coincident markers.
// Both forms take emissions relative to the best class present at the marker
// (so a marker where every state mismatches does not push the FP32 values
// towards -1000, where their resolution is 6e-5) and carry that offset in
// FP64, in the log-likelihood and the backward scale.
//
// Two passes per job (all samples of the job in both):
//   chain pass: one lane group per (sample, direction) walks all M markers,
//     forward groups store the forward vector at each block's first marker,
//     backward groups the backward vector (and its scale) at each block's last
//     marker (block = kBlock markers);
//   block pass: one lane group per (sample, block) replays the forward vector
//     through the block into registers, then walks the backward vector down
//     the block, forming posteriors, dosages and the sampled haplotypes.
// A lane group is LPS = Lp/4 lanes of one wavefront; each lane holds 4 states.
#include <algorithm>

#include "si_device.hpp"
#include "si_internal.hpp"

namespace si {

constexpr int kFbThreads = 128;
constexpr float kFloor = -999.0f;

// Emission table staged in LDS: row[code * 8 + class mask] holds the four
// per-class values (class 3 = padding state), off[] the row's log offset,
// off = max of gl over the classes present (0 if none):
//   linear: row = (exp(gl_c - off), ..., 0), absent classes 0
//   log:    row = (gl_c - off, ..., -inf),   absent classes -inf
struct EmisLds {
    float4 row[kMaxPalette * 8];
    float off[kMaxPalette * 8];
};

template <bool LOG>
__device__ __forceinline__ void load_tables(const FbTables& tb, EmisLds& t) {
    const uint32_t n = tb.P * 8;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        if constexpr (LOG) {
            const float* g = tb.palette + 3 * (i / 8);
            const uint32_t c = i % 8;
            const float o = tb.emax[i];
            t.row[i] = make_float4((c & 1u) ? g[0] - o : -INFINITY, (c & 2u) ? g[1] - o : -INFINITY,
                                   (c & 4u) ? g[2] - o : -INFINITY, -INFINITY);
            t.off[i] = o;
        } else {
            t.row[i] = reinterpret_cast<const float4*>(tb.eps)[i];
            t.off[i] = tb.emax[i];
        }
    }
}

// Emission values of this lane's four states (class bits 4-5 of each byte).
__device__ __forceinline__ void lane_emission(float4 e3, uint32_t sa4, float e[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t cls = (sa4 >> (8 * i + 4)) & 3u;
        const float x = (cls & 1u) ? e3.y : e3.x;
        const float y = (cls & 1u) ? e3.w : e3.z;
        e[i] = (cls & 2u) ? y : x;
    }
}

// The class-present mask of a marker rides in bit 7 of the first three state
// bytes of every lane's word (see the state table), so the emission row needs
// no load of its own.
__device__ __forceinline__ uint32_t class_mask(uint32_t w) {
    return ((w >> 7) & 1u) | ((w >> 14) & 2u) | ((w >> 21) & 4u);
}

// Emission values of this lane's four states at a marker with palette code
// `code` and state word w; returns the row's log offset.  Codes at or beyond
// the palette size are read as the last palette entry, so no input can index
// past the table.
__device__ __forceinline__ float emission(const EmisLds& tab, uint32_t P, uint32_t code, uint32_t w,
                                          float e[4]) {
    const uint32_t row = min(code, P - 1u) * 8u + class_mask(w);
    lane_emission(tab.row[row], w, e);
    return tab.off[row];
}

// Component-wise select (keeps a select of two register vectors from being
// turned into a select of their addresses).
__device__ __forceinline__ float4 sel4(bool c, float4 a, float4 b) {
    return make_float4(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.w : b.w);
}

// One 16-marker block of a lane's inputs: the sample's 16 genotype codes, per
// marker the lane's four state bytes and (per-marker transitions) the step
// coefficients: (t_o, u) for the linear form, (lto, lu) for the log form.
// Loaded whole (indices clamped in bounds past the last marker), so every
// load of the block is in flight at once and the steps that follow read
// registers only.
template <bool LOG, bool PM>
struct BlockIn {
    uint4 cw;
    uint32_t sa[kBlock];
    float2 k[PM ? kBlock : 1];
    __device__ __forceinline__ void load(const JobView& J, const uint8_t* codes, uint32_t lp4,
                                         uint32_t r, uint32_t c) {
        const uint32_t m0 = c * kBlock, ml = J.M - 1u;
        cw = *reinterpret_cast<const uint4*>(codes + m0);
        const uint32_t* sa32 = reinterpret_cast<const uint32_t*>(J.sa);
#pragma unroll
        for (uint32_t t = 0; t < kBlock; ++t) sa[t] = sa32[size_t(min(m0 + t, ml)) * lp4 + r];
        if constexpr (PM) {
#pragma unroll
            for (uint32_t t = 0; t < kBlock; ++t) {
                const float4 q = J.steps[min(m0 + t, ml)];
                k[t] = LOG ? make_float2(q.z, q.w) : make_float2(q.x, q.y);
            }
        } else {
            k[0] = LOG ? make_float2(J.step.z, J.step.w) : make_float2(J.step.x, J.step.y);
        }
    }
    __device__ __forceinline__ uint32_t code(uint32_t t) const {
        return byte_of((&cw.x)[t >> 2], t & 3);
    }
    __device__ __forceinline__ float2 coef(uint32_t t) const { return k[PM ? t : 0]; }
};

// What a block's 16 steps consume: this lane's emissions (read from the LDS
// table together, before any step needs them) and the step coefficients.
// Filling it frees the BlockIn for the next block's loads, which are then in
// flight while the steps run.
template <bool LOG, bool PM>
struct BlockEm {
    float e[kBlock][4];
    float off[kBlock];
    float2 k[PM ? kBlock : 1];
    __device__ __forceinline__ void fill(const EmisLds& tab, uint32_t P, const BlockIn<LOG, PM>& in) {
#pragma unroll
        for (uint32_t t = 0; t < kBlock; ++t) off[t] = emission(tab, P, in.code(t), in.sa[t], e[t]);
#pragma unroll
        for (uint32_t t = 0; t < (PM ? kBlock : 1); ++t) k[t] = in.k[t];
    }
    __device__ __forceinline__ float2 coef(uint32_t t) const { return k[PM ? t : 0]; }
};

// Orders this wavefront's LDS accesses: lane groups exchange data through LDS
// only within their own wavefront.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int LPS>
__device__ __forceinline__ float group_lse(float x0, float x1, float x2, float x3) {
    const float mx = group_max<LPS>(fmaxf(fmaxf(x0, x1), fmaxf(x2, x3)));
    const float s = group_sum<LPS>((fast_exp(x0 - mx) + fast_exp(x1 - mx)) +
                                   (fast_exp(x2 - mx) + fast_exp(x3 - mx)));
    return mx + fast_ln(s);
}

__device__ __forceinline__ float lse3(float a, float b, float c) {
    const float mx = fmaxf(fmaxf(a, b), c);
    return mx + fast_ln((fast_exp(a - mx) + fast_exp(b - mx)) + fast_exp(c - mx));
}

// Log mass LSE_i(a_i) of a max-normalized log vector (a value in [0, ln Lp]).
template <int LPS>
__device__ __forceinline__ float log_mass(float4 a) {
    return fast_ln(group_sum<LPS>((fast_exp(a.x) + fast_exp(a.y)) + (fast_exp(a.z) + fast_exp(a.w))));
}

// ---------------------------------------------------------------- forward --
// Marker 0.  Returns the log normalizer (log-likelihood increment); ls is the
// log mass of the normalized log-space vector (unused by the linear form,
// whose vector sums to 1).
template <int LPS, bool LOG>
__device__ __forceinline__ double fwd_init(float4& a, float& ls, const float e[4], float off) {
    if constexpr (LOG) {
        const float c = group_max<LPS>(fmaxf(fmaxf(e[0], e[1]), fmaxf(e[2], e[3])));
        a = make_float4(e[0] - c, e[1] - c, e[2] - c, e[3] - c);
        ls = log_mass<LPS>(a);
        return double(off) + double(c);
    } else {
        const float S = group_sum<LPS>((e[0] + e[1]) + (e[2] + e[3]));
        const float r = __builtin_amdgcn_rcpf(S);
        a = make_float4(e[0] * r, e[1] * r, e[2] * r, e[3] * r);
        return double(off) + double(fast_ln(S));
    }
}

// Step into a marker with emission e / offset off and coefficients k.
template <int LPS, bool LOG>
__device__ __forceinline__ double fwd_step(float4& a, float& ls, const float e[4], float off,
                                           float2 k) {
    if constexpr (LOG) {
        // k = (lto, lu); LSE_i(a_i + To) = To + ls
        const float lo = k.x + ls;
        const float x0 = lse3(kFloor, lo, k.y + a.x) + e[0];
        const float x1 = lse3(kFloor, lo, k.y + a.y) + e[1];
        const float x2 = lse3(kFloor, lo, k.y + a.z) + e[2];
        const float x3 = lse3(kFloor, lo, k.y + a.w) + e[3];
        const float c = group_max<LPS>(fmaxf(fmaxf(x0, x1), fmaxf(x2, x3)));
        a = make_float4(x0 - c, x1 - c, x2 - c, x3 - c);
        ls = log_mass<LPS>(a);
        return double(off) + double(c);
    } else {
        // k = (t_o, u); a sums to 1, so sum_i a_i t_o = t_o.
        const float x0 = fmaf(k.y, a.x, k.x) * e[0];
        const float x1 = fmaf(k.y, a.y, k.x) * e[1];
        const float x2 = fmaf(k.y, a.z, k.x) * e[2];
        const float x3 = fmaf(k.y, a.w, k.x) * e[3];
        const float S = group_sum<LPS>((x0 + x1) + (x2 + x3));
        const float r = __builtin_amdgcn_rcpf(S);
        a = make_float4(x0 * r, x1 * r, x2 * r, x3 * r);
        return double(off) + double(fast_ln(S));
    }
}

// --------------------------------------------------------------- backward --
// From the vector at marker n (emission e, offset off) to n - 1.  The true
// backward vector is e^{Bl} v (linear) or Bl + v (log).  With
// nat = Bl + off + log W the new scale is max(nat, -999): when the floor
// dominates (nat < -999) the scale pins at -999 and the non-floor part is
// shifted down by sh = -999 - nat, so no value ever leaves a safe range.
template <int LPS, bool LOG>
__device__ __forceinline__ void bwd_step(float4& v, double& Bl, const float e[4], float off,
                                         float2 k) {
    if constexpr (LOG) {
        const float y0 = e[0] + v.x, y1 = e[1] + v.y, y2 = e[2] + v.z, y3 = e[3] + v.w;
        const float lw = group_lse<LPS>(y0, y1, y2, y3);
        const double nat = Bl + double(off) + double(lw);
        const double arg = -999.0 - nat;
        const bool floor_dominates = arg > 0.0;
        const float fl = floor_dominates ? 0.0f : float(arg);
        const float sh = floor_dominates ? float(arg) : 0.0f;
        Bl = floor_dominates ? -999.0 : nat;
        const float bo = k.x - sh, bu = k.y - lw - sh;
        v = make_float4(lse3(fl, bo, bu + y0), lse3(fl, bo, bu + y1), lse3(fl, bo, bu + y2),
                        lse3(fl, bo, bu + y3));
    } else {
        const float w0 = e[0] * v.x, w1 = e[1] * v.y, w2 = e[2] * v.z, w3 = e[3] * v.w;
        const float W = group_sum<LPS>((w0 + w1) + (w2 + w3));
        const double nat = Bl + double(off) + double(fast_ln(W));
        const double arg = -999.0 - nat;
        const float x = fast_exp(float(-fabs(arg)));
        const bool floor_dominates = arg > 0.0;
        const float f = floor_dominates ? 1.0f : x;
        const float g = floor_dominates ? x : 1.0f;
        Bl = floor_dominates ? -999.0 : nat;
        const float gr = g * k.y * __builtin_amdgcn_rcpf(W);
        const float base = fmaf(g, k.x, f);
        v = make_float4(fmaf(gr, w0, base), fmaf(gr, w1, base), fmaf(gr, w2, base),
                        fmaf(gr, w3, base));
    }
}

// Initial backward vector at the last marker (b = 1).
template <bool LOG>
__device__ __forceinline__ float4 bwd_init() {
    return LOG ? make_float4(0.f, 0.f, 0.f, 0.f) : make_float4(1.f, 1.f, 1.f, 1.f);
}

// -------------------------------------------------------------- posterior --
template <int LPS, bool LOG>
__device__ __forceinline__ float4 posterior(float4 a, float4 v) {
    if constexpr (LOG) {
        const float q0 = a.x + v.x, q1 = a.y + v.y, q2 = a.z + v.z, q3 = a.w + v.w;
        const float mx = group_max<LPS>(fmaxf(fmaxf(q0, q1), fmaxf(q2, q3)));
        const float w0 = fast_exp(q0 - mx), w1 = fast_exp(q1 - mx), w2 = fast_exp(q2 - mx),
                    w3 = fast_exp(q3 - mx);
        const float rq = 1.0f / group_sum<LPS>((w0 + w1) + (w2 + w3));
        return make_float4(w0 * rq, w1 * rq, w2 * rq, w3 * rq);
    } else {
        const float4 q = make_float4(a.x * v.x, a.y * v.y, a.z * v.z, a.w * v.w);
        const float rq = 1.0f / group_sum<LPS>((q.x + q.y) + (q.z + q.w));
        return make_float4(q.x * rq, q.y * rq, q.z * rq, q.w * rq);
    }
}

// ---------------------------------------------------------------------------
// One chain (forward or backward) of sample s of job J, by one lane group.
// The next block's inputs are fetched while the current block is stepped;
// every step is computed and committed by a select where it lies inside the
// chain, so the block's 16 steps are straight-line code.
// ---------------------------------------------------------------------------
template <int LPS, bool LOG, bool PM>
__device__ __forceinline__ void run_chain(const JobView& J, const EmisLds& tab, uint32_t P,
                                          uint32_t Lp, uint32_t s, bool fwd, uint32_t r) {
    const uint8_t* codes = J.codes + size_t(s) * J.pitch;
    const uint32_t lp4 = Lp / 4;
    BlockIn<LOG, PM> cur;
    BlockEm<LOG, PM> em;
    if (fwd) {
        cur.load(J, codes, lp4, r, 0);
        float4 a;
        float ls = 0.0f;
        double ll;
        {
            float e[4];
            const float off = emission(tab, P, cur.code(0), cur.sa[0], e);
            ll = fwd_init<LPS, LOG>(a, ls, e, off);
        }
        for (uint32_t c = 0; c < J.nblk; ++c) {
            em.fill(tab, P, cur);
            if (c + 1 < J.nblk) cur.load(J, codes, lp4, r, c + 1);
            const uint32_t m0 = c * kBlock;
#pragma unroll
            for (uint32_t t = 0; t < kBlock; ++t) {
                const uint32_t m = m0 + t;
                float4 an = a;
                float lsn = ls;
                const double inc = fwd_step<LPS, LOG>(an, lsn, em.e[t], em.off[t], em.coef(t));
                const bool live = m > 0 && m < J.M;
                a = sel4(live, an, a);
                ls = live ? lsn : ls;
                ll += live ? inc : 0.0;
                if (t == 0)
                    reinterpret_cast<float4*>(J.ck_alpha)[(size_t(s) * J.nblk + c) * lp4 + r] = a;
            }
        }
        if (LOG) ll += double(ls);  // log mass of the last normalized vector
        if (r == 0 && J.loglik) J.loglik[s] = ll;
    } else {
        float4 v = bwd_init<LOG>();
        double Bl = 0.0;
        const size_t last = size_t(s) * J.nblk + (J.nblk - 1);
        reinterpret_cast<float4*>(J.ck_beta)[last * lp4 + r] = v;
        if (r == 0) J.ck_bl[last] = Bl;
        cur.load(J, codes, lp4, r, J.nblk - 1);
        for (int c = int(J.nblk) - 1; c >= 0; --c) {
            em.fill(tab, P, cur);
            if (c > 0) cur.load(J, codes, lp4, r, uint32_t(c - 1));
            const uint32_t m0 = uint32_t(c) * kBlock;
#pragma unroll
            for (int t = int(kBlock) - 1; t >= 0; --t) {
                const uint32_t n = m0 + uint32_t(t);
                float4 vn = v;
                double Bln = Bl;
                bwd_step<LPS, LOG>(vn, Bln, em.e[t], em.off[t], em.coef(uint32_t(t)));
                const bool live = n > 0 && n < J.M;
                v = sel4(live, vn, v);
                Bl = live ? Bln : Bl;
                if (t == 0 && c > 0) {  // v is the vector at marker m0 - 1, last of block c - 1
                    const size_t ck = size_t(s) * J.nblk + uint32_t(c - 1);
                    reinterpret_cast<float4*>(J.ck_beta)[ck * lp4 + r] = v;
                    if (r == 0) J.ck_bl[ck] = Bl;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Block c of sample s of job J, by one lane group: replay the forward vector
// through the block into registers, then walk the backward vector down the
// block forming posteriors (into the group's LDS slice) and the allele
// probabilities (group sums).  Then sample: lane r takes markers r, r + LPS,
// ... of the block and walks the states in order with a sequential FP32
// cumulative sum, which the host replays exactly.  `active` groups do the
// work; every lane of the wavefront passes the wave_sync.
// ---------------------------------------------------------------------------
template <int LPS, bool LOG, bool PM>
__device__ __forceinline__ void run_block(const JobView& J, const EmisLds& tab, uint32_t P,
                                          uint32_t Lp, bool active, uint32_t s, uint32_t c,
                                          uint32_t r, float (*pbuf)[4 * LPS],
                                          uint32_t (*sbuf)[LPS], uint4* rng) {
    constexpr uint32_t kPer = kBlock / LPS;  // markers per lane in the sampling phase
    const uint32_t m0 = c * kBlock;
    const uint32_t nm = active ? min(kBlock, J.M - m0) : 0u;
    const uint32_t lp4 = Lp / 4;
    float ap0r[kPer], ap1r[kPer];
    if (active) {
        BlockIn<LOG, PM> in;
        in.load(J, J.codes + size_t(s) * J.pitch, lp4, r, c);
        BlockEm<LOG, PM> em;
        em.fill(tab, P, in);
#pragma unroll
        for (uint32_t t = 0; t < kBlock; ++t) sbuf[t][r] = in.sa[t];
        // forward vectors of the block, into this lane's part of the LDS slice
        float4 a = reinterpret_cast<const float4*>(J.ck_alpha)[(size_t(s) * J.nblk + c) * lp4 + r];
        *reinterpret_cast<float4*>(&pbuf[0][4 * r]) = a;
        // the chain pass derived ls from the same vector by the same function
        float ls = LOG ? log_mass<LPS>(a) : 0.0f;
#pragma unroll
        for (uint32_t t = 1; t < kBlock; ++t) {
            float4 an = a;
            float lsn = ls;
            fwd_step<LPS, LOG>(an, lsn, em.e[t], em.off[t], em.coef(t));
            const bool live = t < nm;
            a = sel4(live, an, a);
            ls = live ? lsn : ls;
            *reinterpret_cast<float4*>(&pbuf[t][4 * r]) = a;
        }
        const size_t ck = size_t(s) * J.nblk + c;
        float4 v = reinterpret_cast<const float4*>(J.ck_beta)[ck * lp4 + r];
        double Bl = J.ck_bl[ck];
#pragma unroll
        for (int t = int(kBlock) - 1; t >= 0; --t) {
            const float4 p =
                posterior<LPS, LOG>(*reinterpret_cast<const float4*>(&pbuf[t][4 * r]), v);
            const uint32_t w = in.sa[t];
            const float x0 = ((w & 1u) ? p.x : 0.0f) + ((w & 0x100u) ? p.y : 0.0f) +
                             ((w & 0x10000u) ? p.z : 0.0f) + ((w & 0x1000000u) ? p.w : 0.0f);
            const float x1 = ((w & 2u) ? p.x : 0.0f) + ((w & 0x200u) ? p.y : 0.0f) +
                             ((w & 0x20000u) ? p.z : 0.0f) + ((w & 0x2000000u) ? p.w : 0.0f);
            const float ap0 = group_sum<LPS>(x0), ap1 = group_sum<LPS>(x1);
            if (uint32_t(t) < nm) *reinterpret_cast<float4*>(&pbuf[t][4 * r]) = p;
            if (uint32_t(t) % LPS == r) {
                ap0r[uint32_t(t) / LPS] = ap0;
                ap1r[uint32_t(t) / LPS] = ap1;
            }
            if (t > 0) {
                float4 vn = v;
                double Bln = Bl;
                bwd_step<LPS, LOG>(vn, Bln, em.e[t], em.off[t], em.coef(uint32_t(t)));
                const bool live = uint32_t(t) < nm;
                v = sel4(live, vn, v);
                Bl = live ? Bln : Bl;
            }
        }
        const uint2 key = make_uint2(uint32_t(J.seed), uint32_t(J.seed >> 32));
        for (uint32_t j = r; j < kBlock / 4; j += LPS)
            rng[j] = philox4x32_10(make_uint4((m0 >> 2) + j, J.sample_base + s, J.rng_stream, 0u),
                                   key);
    }
    wave_sync();
    if (!active) return;

    const bool argmax = (J.flags & SI_FB_ARGMAX) != 0;
    const uint32_t L = J.L;
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
        const uint32_t t = r + q * LPS;
        if (t >= nm) continue;
        const uint32_t m = m0 + t;
        const float* pt = pbuf[t];
        uint32_t ks = 0;
        if (argmax) {  // first maximum
            float best = -1.0f;
            for (uint32_t k = 0; k < L; ++k)
                if (pt[k] > best) {
                    best = pt[k];
                    ks = k;
                }
        } else {  // first k with u < cumsum, else 0
            const float u = u24((&rng[t >> 2].x)[t & 3]);
            float cs = 0.0f;
            uint32_t kd = 0xffffffffu;
            for (uint32_t k = 0; k < L; ++k) {
                cs += pt[k];
                if (kd == 0xffffffffu && u < cs) kd = k;
            }
            ks = kd == 0xffffffffu ? 0u : kd;
        }
        const size_t o = size_t(s) * J.M + m;
        if (J.ds) J.ds[o] = ap0r[q] + ap1r[q];
        if (J.ap) reinterpret_cast<float2*>(J.ap)[o] = make_float2(ap0r[q], ap1r[q]);
        if (J.hap) {
            const uint32_t b = reinterpret_cast<const uint8_t*>(sbuf[t])[ks];
            J.hap[size_t(s) * 2 * J.M + m] = uint8_t((b & 4u) ? 255u : (b & 1u));
            J.hap[size_t(s) * 2 * J.M + J.M + m] = uint8_t((b & 8u) ? 255u : ((b >> 1) & 1u));
        }
        if (J.prob)
            for (uint32_t k = 0; k < L; ++k) J.prob[o * L + k] = pt[k];
    }
    wave_sync();  // the slices are reused by the group's next item
}

// ---------------------------------------------------------------------------
// Plain form: one job; grid = groups.
// ---------------------------------------------------------------------------
template <int LPS, bool LOG>
__global__ __launch_bounds__(kFbThreads) void fb_chain_kernel(JobView J, FbTables tb,
                                                              uint32_t Bpad) {
    __shared__ EmisLds tab;
    load_tables<LOG>(tb, tab);
    __syncthreads();
    const uint32_t gtid = blockIdx.x * kFbThreads + threadIdx.x;
    const uint32_t grp = gtid / LPS, r = gtid % LPS;
    const bool fwd = grp < Bpad;
    const uint32_t s = fwd ? grp : grp - Bpad;
    if (s >= J.B) return;  // whole lane groups leave together
    if (J.flags & SI_FB_PER_MARKER) run_chain<LPS, LOG, true>(J, tab, tb.P, 4 * LPS, s, fwd, r);
    else run_chain<LPS, LOG, false>(J, tab, tb.P, 4 * LPS, s, fwd, r);
}

template <int LPS, bool LOG>
__global__ __launch_bounds__(kFbThreads) void fb_block_kernel(JobView J, FbTables tb) {
    constexpr uint32_t G = kFbThreads / LPS;  // lane groups per workgroup
    __shared__ EmisLds tab;
    __shared__ float pbuf[G][kBlock][4 * LPS];
    __shared__ uint32_t sbuf[G][kBlock][LPS];
    __shared__ uint4 rng[G][kBlock / 4];
    load_tables<LOG>(tb, tab);
    __syncthreads();
    const uint32_t gtid = blockIdx.x * kFbThreads + threadIdx.x;
    const uint32_t grp = gtid / LPS, r = gtid % LPS, gl = threadIdx.x / LPS;
    const bool active = grp < J.B * J.nblk;
    const uint32_t s = active ? grp / J.nblk : 0u;
    const uint32_t c = active ? grp % J.nblk : 0u;
    if (J.flags & SI_FB_PER_MARKER)
        run_block<LPS, LOG, true>(J, tab, tb.P, 4 * LPS, active, s, c, r, pbuf[gl], sbuf[gl], rng[gl]);
    else
        run_block<LPS, LOG, false>(J, tab, tb.P, 4 * LPS, active, s, c, r, pbuf[gl], sbuf[gl], rng[gl]);
}

// ---------------------------------------------------------------------------
// Batched form: a persistent grid; each wavefront repeatedly takes the next
// 64/LPS items of the class's queue (one atomic per wavefront) until the
// queue is exhausted.  Items are ordered job by job (jobs longest first for
// the chain pass), so a wavefront's groups usually share a job.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t find_job(const uint64_t* off, uint32_t n, uint64_t item) {
    uint32_t lo = 0, hi = n;  // off[lo] <= item < off[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= item) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint64_t wave_take(uint32_t* counter, uint32_t n) {
    uint32_t base = 0;
    if (__lane_id() == 0) base = atomicAdd(counter, n);
    return uint64_t(__builtin_amdgcn_readfirstlane(base));
}

template <int LPS, bool LOG>
__global__ __launch_bounds__(kFbThreads) void fb_chain_batch_kernel(BatchClass bc, FbTables tb) {
    __shared__ EmisLds tab;
    load_tables<LOG>(tb, tab);
    __syncthreads();
    constexpr uint32_t GPW = kWave / LPS;
    const uint32_t lane = __lane_id(), gw = lane / LPS, r = lane % LPS;
    for (;;) {
        const uint64_t base = wave_take(&bc.counters[0], GPW);
        if (base >= bc.chain_items) break;
        const uint64_t item = base + gw;
        if (item < bc.chain_items) {
            const uint32_t j = find_job(bc.chain_off, bc.n_jobs, item);
            const JobView J = bc.jobs[j];
            const uint64_t k = item - bc.chain_off[j];
            if (J.flags & SI_FB_PER_MARKER)
                run_chain<LPS, LOG, true>(J, tab, tb.P, bc.Lp, uint32_t(k >> 1), (k & 1) == 0, r);
            else
                run_chain<LPS, LOG, false>(J, tab, tb.P, bc.Lp, uint32_t(k >> 1), (k & 1) == 0, r);
        }
    }
}

template <int LPS, bool LOG>
__global__ __launch_bounds__(kFbThreads) void fb_block_batch_kernel(BatchClass bc, FbTables tb) {
    constexpr uint32_t G = kFbThreads / LPS;
    constexpr uint32_t GPW = kWave / LPS;
    __shared__ EmisLds tab;
    __shared__ float pbuf[G][kBlock][4 * LPS];
    __shared__ uint32_t sbuf[G][kBlock][LPS];
    __shared__ uint4 rng[G][kBlock / 4];
    load_tables<LOG>(tb, tab);
    __syncthreads();
    const uint32_t lane = __lane_id(), gw = lane / LPS, r = lane % LPS;
    const uint32_t gl = threadIdx.x / LPS;
    for (;;) {
        const uint64_t base = wave_take(&bc.counters[1], GPW);
        if (base >= bc.block_items) break;
        const uint64_t item = base + gw;
        const bool active = item < bc.block_items;
        uint32_t j = 0, s = 0, c = 0;
        if (active) {
            j = find_job(bc.block_off, bc.n_jobs, item);
            const uint64_t k = item - bc.block_off[j];
            const uint32_t nblk = bc.jobs[j].nblk;
            s = uint32_t(k / nblk);
            c = uint32_t(k % nblk);
        }
        const JobView J = bc.jobs[j];
        if (J.flags & SI_FB_PER_MARKER)
            run_block<LPS, LOG, true>(J, tab, tb.P, bc.Lp, active, s, c, r, pbuf[gl], sbuf[gl], rng[gl]);
        else
            run_block<LPS, LOG, false>(J, tab, tb.P, bc.Lp, active, s, c, r, pbuf[gl], sbuf[gl], rng[gl]);
    }
}

__global__ void zero_words_kernel(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
size_t fb_workspace_bytes_one(uint32_t B, uint32_t M, uint32_t Lp) {
    const size_t nblk = (size_t(M) + kBlock - 1) / kBlock;
    const size_t vec = size_t(B) * nblk * Lp * sizeof(float);
    const size_t vec_al = (vec + 255) & ~size_t(255);
    const size_t bl = size_t(B) * nblk * sizeof(double);
    return 2 * vec_al + ((bl + 255) & ~size_t(255));
}

void fb_bind_workspace(JobView& v, uint32_t Lp, char* ws) {
    const size_t vec = size_t(v.B) * v.nblk * Lp * sizeof(float);
    const size_t vec_al = (vec + 255) & ~size_t(255);
    v.ck_alpha = reinterpret_cast<float*>(ws);
    v.ck_beta = reinterpret_cast<float*>(ws + vec_al);
    v.ck_bl = reinterpret_cast<double*>(ws + 2 * vec_al);
}

JobView fb_job_view(const si_fb_job& j) {
    JobView v{};
    v.B = j.n_samples;
    v.M = j.n_markers;
    v.L = j.n_states;
    v.flags = j.flags;
    v.pitch = j.code_pitch;
    v.rng_stream = j.rng_stream;
    v.sample_base = j.sample_base;
    v.nblk = (j.n_markers + kBlock - 1) / kBlock;
    v.seed = j.seed;
    v.step = make_float4(j.step.t_o, j.step.u, j.step.lto, j.step.lu);
    v.steps = reinterpret_cast<const float4*>(j.steps);
    v.sa = j.sa;
    v.mmask = j.mmask;
    v.codes = j.codes;
    v.ds = j.ds;
    v.ap = j.ap;
    v.hap = j.hap;
    v.prob = j.prob;
    v.loglik = j.loglik;
    return v;
}

template <int LPS, bool LOG>
static int launch_plain(const JobView& v, const FbTables& t, uint32_t passes, hipStream_t stream) {
    const uint32_t gpw = kWave / LPS;  // groups per wavefront
    const uint32_t Bpad = (v.B + gpw - 1) / gpw * gpw;
    const size_t chain_threads = size_t(Bpad + v.B) * LPS;
    const size_t block_threads = size_t(v.B) * v.nblk * LPS;
    const size_t g1 = (chain_threads + kFbThreads - 1) / kFbThreads;
    const size_t g2 = (block_threads + kFbThreads - 1) / kFbThreads;
    if (g1 > 0x7fffffffu || g2 > 0x7fffffffu) return SI_ERR_UNSUPPORTED;
    if (passes & SI_FB_PASS_CHAIN) {
        hipLaunchKernelGGL((fb_chain_kernel<LPS, LOG>), dim3(uint32_t(g1)), dim3(kFbThreads), 0,
                           stream, v, t, Bpad);
        if (hipGetLastError() != hipSuccess) return SI_ERR_DEVICE;
    }
    if (passes & SI_FB_PASS_BLOCK) {
        hipLaunchKernelGGL((fb_block_kernel<LPS, LOG>), dim3(uint32_t(g2)), dim3(kFbThreads), 0,
                           stream, v, t);
        if (hipGetLastError() != hipSuccess) return SI_ERR_DEVICE;
    }
    return 0;
}

template <int LPS, bool LOG>
static int launch_batch(const BatchClass& c, const FbTables& t, uint32_t passes,
                        hipStream_t stream) {
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, stream, c.counters, 2u);
    if (hipGetLastError() != hipSuccess) return SI_ERR_DEVICE;
    if ((passes & SI_FB_PASS_CHAIN) && c.chain_items) {
        hipLaunchKernelGGL((fb_chain_batch_kernel<LPS, LOG>), dim3(c.grid_chain), dim3(kFbThreads),
                           0, stream, c, t);
        if (hipGetLastError() != hipSuccess) return SI_ERR_DEVICE;
    }
    if ((passes & SI_FB_PASS_BLOCK) && c.block_items) {
        hipLaunchKernelGGL((fb_block_batch_kernel<LPS, LOG>), dim3(c.grid_block), dim3(kFbThreads),
                           0, stream, c, t);
        if (hipGetLastError() != hipSuccess) return SI_ERR_DEVICE;
    }
    return 0;
}

template <bool LOG>
static int dispatch_plain(const JobView& v, const FbTables& t, uint32_t Lp, uint32_t passes,
                          hipStream_t stream) {
    switch (Lp) {
        case 4: return launch_plain<1, LOG>(v, t, passes, stream);
        case 8: return launch_plain<2, LOG>(v, t, passes, stream);
        case 16: return launch_plain<4, LOG>(v, t, passes, stream);
        case 32: return launch_plain<8, LOG>(v, t, passes, stream);
        case 64: return launch_plain<16, LOG>(v, t, passes, stream);
        default: return SI_ERR_UNSUPPORTED;
    }
}

template <bool LOG>
static int dispatch_batch(const BatchClass& c, const FbTables& t, uint32_t passes,
                          hipStream_t stream) {
    switch (c.Lp) {
        case 4: return launch_batch<1, LOG>(c, t, passes, stream);
        case 8: return launch_batch<2, LOG>(c, t, passes, stream);
        case 16: return launch_batch<4, LOG>(c, t, passes, stream);
        case 32: return launch_batch<8, LOG>(c, t, passes, stream);
        case 64: return launch_batch<16, LOG>(c, t, passes, stream);
        default: return SI_ERR_UNSUPPORTED;
    }
}

int launch_fb(const JobView& v, const FbTables& t, uint32_t Lp, uint32_t passes,
              hipStream_t stream) {
    if (v.B == 0 || v.M == 0) return 0;
    return (v.flags & SI_FB_LOGSPACE) ? dispatch_plain<true>(v, t, Lp, passes, stream)
                                      : dispatch_plain<false>(v, t, Lp, passes, stream);
}

// Persistent grid: as many workgroups as can be resident at once, capped by
// the work available.
template <int LPS, bool LOG>
static void batch_grid_lps(BatchClass& c) {
    int dev = 0, cus = 0, occ_c = 0, occ_b = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_c, fb_chain_batch_kernel<LPS, LOG>,
                                                       kFbThreads, 0);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_b, fb_block_batch_kernel<LPS, LOG>,
                                                       kFbThreads, 0);
    const uint64_t gpb = kFbThreads / LPS;  // lane groups per workgroup
    const uint64_t need_c = (c.chain_items + gpb - 1) / gpb, need_b = (c.block_items + gpb - 1) / gpb;
    const uint64_t ncu = uint64_t(cus > 0 ? cus : 1);
    const uint64_t cap_c = ncu * uint64_t(occ_c > 0 ? occ_c : 1);
    const uint64_t cap_b = ncu * uint64_t(occ_b > 0 ? occ_b : 1);
    c.grid_chain = uint32_t(std::max<uint64_t>(1, std::min(need_c, cap_c)));
    c.grid_block = uint32_t(std::max<uint64_t>(1, std::min(need_b, cap_b)));
}

template <bool LOG>
static void batch_grid_space(BatchClass& c) {
    switch (c.Lp) {
        case 4: batch_grid_lps<1, LOG>(c); break;
        case 8: batch_grid_lps<2, LOG>(c); break;
        case 16: batch_grid_lps<4, LOG>(c); break;
        case 32: batch_grid_lps<8, LOG>(c); break;
        case 64: batch_grid_lps<16, LOG>(c); break;
        default: c.grid_chain = c.grid_block = 1;
    }
}

void fb_batch_grid(BatchClass& c) {
    if (c.flags & SI_FB_LOGSPACE) batch_grid_space<true>(c);
    else batch_grid_space<false>(c);
}

int launch_fb_batch(const BatchClass& c, const FbTables& t, uint32_t passes, hipStream_t stream) {
    return (c.flags & SI_FB_LOGSPACE) ? dispatch_batch<true>(c, t, passes, stream)
                                      : dispatch_batch<false>(c, t, passes, stream);
}

}  // namespace si
