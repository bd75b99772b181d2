// Li-Stephens forward-backward, posterior dosage and Philox sampling for gfx950:
// the device code, templated on the lane shape.  Each (padded states Lp,
// lanes per sample G) pair is instantiated in a translation unit of its own
// (fb_inst.hip); hmm.hip holds the host side.
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
// Two numeric forms share the structure (template parameter LOG):
//   * linear FP32: forward sum-normalized, a'_j = (t_o + u a_j) eps_j (the
//     floor term is below FP32 resolution whenever To >= -80, checked on the
//     host); backward b_m = e^{Bl_m} v_m with the log scale Bl_m in FP64, so
//     the floor enters exactly as f = exp(-999 - Bl_m) once b reaches it.
//     exp/log run once per (sample, marker), not per state.
//   * log-space FP32: forward log vector normalized to maximum 0 (as the
//     reference normalizes, so its -999 floor weighs the same); backward log
//     vector relative to an FP64 offset; every step is a three-term
//     log-sum-exp LSE(-999, To + ls, U + alpha_j) per state (ls = LSE(alpha),
//     U = log(e^Td - e^To)), whose first two terms are shared by all states:
//     with the exponentials of the step's own log-mass sum it costs one log
//     per state (an exp and a log where the shared terms underflow).  Valid
//     for any To, including the -999 of coincident markers.
// Both take emissions relative to the best class present at the marker (so a
// marker where every state mismatches does not push FP32 values towards
// -1000, where their resolution is 6e-5) and carry that offset in FP64.
//
// Work split.  A lane group of G adjacent lanes of one wavefront owns one
// sample's vector, R = Lp / G states per lane (lane r holds states
// [r R, r R + R)).  Every sum over states is the same binary tree over the Lp
// states in index order (pairwise within the lane, then lane pairs by DPP),
// and every maximum is exact, so results are bit-identical for every (G, R):
// the lane shape is a pure performance choice, made by the autotuner.
//
// Two passes per job (all samples of the job in both):
//   chain pass: one lane group per (sample, direction) walks all M markers;
//     forward groups store the forward vector at the first marker of every
//     checkpoint block (CB markers), backward groups the backward vector and
//     its FP64 scale at every block's last marker;
//   block pass: one lane group per (sample, checkpoint block).  The block is
//     handled as CB/16 sub-blocks of 16 markers, last first: the forward
//     vector is replayed from the block checkpoint to the sub-block (kept in
//     LDS for its 16 markers), then the backward vector walks down the
//     sub-block forming posteriors, dosages and allele probabilities; lanes
//     then sample the sub-block's markers from the LDS posteriors.  CB = 32
//     halves checkpoint memory for 16 extra forward steps per block.
#pragma once

#include <stdint.h>

#include <type_traits>

#include "si_device.hpp"
#include "si_internal.hpp"

namespace si {

constexpr int kFbThreads = 128;
constexpr float kFloor = -999.0f;
constexpr uint32_t kSub = 16;  // markers per load block and per block-pass sub-block

// Lane shape: G lanes per sample, R states per lane.
template <int G_, int R_>
struct Shape {
    static constexpr int G = G_, R = R_, Lp = G_ * R_;
    static constexpr int NW = R_ >= 4 ? R_ / 4 : 1;  // state-table words a lane loads per marker
    static constexpr int GPW = kWave / G_;           // lane groups per wavefront
    static constexpr int GPB = kFbThreads / G_;      // lane groups per workgroup
    static_assert(G_ >= 1 && G_ <= 64 && R_ >= 2 && R_ <= 8, "lane shape");
};

// Block-pass LDS of one lane group: posteriors [16][T] floats (row t at
// t*T, group stride PS = 16 T + E floats), 16 Philox outputs (in registers
// for one-lane groups) and the sampler's state bits (below).  Its size sets
// how many workgroups a CU holds, which bounds these latency-bound kernels,
// so up to 16 states the rows carry no padding (T = Lp).  The group padding
// E makes the walks' vector writes and reads free of bank conflicts on 32-
// and 64-bank models for 4 and 8 states; 16 states go without it (2-way
// conflicts) so that 4 workgroups of 16 states on 4 lanes fit a CU.  The
// sampling phase's 16-byte row reads keep 2- to 4-way conflicts (measured,
// cheaper than padding or rotating the rows).  With 32 or 64 states
// T = Lp + 4, conflict-free for the sampling reads too.
__host__ __device__ constexpr int pbuf_T(int Lp) { return Lp <= 16 ? Lp : Lp + 4; }
__host__ __device__ constexpr int pbuf_E(int Lp) {
    return Lp == 4 ? 4 : Lp == 8 ? 8 : Lp == 32 ? 32 : 0;
}
template <class S>
struct BlockLds {
    static constexpr int T = pbuf_T(S::Lp);
    static constexpr int PS = int(kSub) * T + pbuf_E(S::Lp);  // floats per group
    // State bytes [16][Lp] of the group in LDS for the sampler (SW words per
    // group).  The plain kernel drops them where that lets one more
    // workgroup be resident (16 states on 4 lanes: 3 -> 4 per CU); there the
    // sampler reads the sampled state's byte from the state table, which the
    // item has just read.  Elsewhere, and in the batched kernel (whose
    // classes share the caches), the load latency costs more than it gains.
    static constexpr bool kSbPlain = !(S::Lp == 16 && S::G == 4);
    // 4 bits per state (alleles, missing flags): 16 Lp nibbles = 2 Lp words
    static constexpr int SW = int(kSub) * S::Lp / 8;
    // Philox outputs of a one- or two-lane group of 4 states per lane stay in
    // its registers (two lanes swap theirs by DPP); with 2 states per lane
    // the register budget is the tighter limit
    static constexpr bool kRngLds = S::G > 2 || S::R < 4;
    static_assert(T % 4 == 0 && PS % 4 == 0, "16-byte aligned rows");
};

// ------------------------------------------------------------ emissions --
// Emission table staged in LDS (dynamic): row = code * 8 + class mask, five
// floats per row: the four per-class values (class 3 = padding state) and the
// row's log offset off = max of gl over the classes present (0 if none):
//   linear: (exp(gl_c - off), ..., 0, off), absent classes 0
//   log:    (gl_c - off, ..., -inf, off),   absent classes -inf
// A state's value is one LDS read at row * 5 + class.
// The table's last word holds the untyped code (below), checked against the
// table itself: kNoUntyped unless that code's row for all three classes is
// emission 1 (log: 0) with offset 0, i.e. its likelihoods are (0, 0, 0).
struct Tab {
    const float* f;
    const uint32_t* uc;  // LDS word, read after the table's barrier
};
constexpr uint32_t kTabRow = 5;

template <bool LOG>
__device__ __forceinline__ Tab load_tables(const FbTables& tb) {
    extern __shared__ float4 fb_dyn_lds[];
    float* f = reinterpret_cast<float*>(fb_dyn_lds);
    uint32_t* uc = reinterpret_cast<uint32_t*>(f + tb.P * 8 * kTabRow);
    const uint32_t n = tb.P * 8;
    const uint32_t urow = tb.ucode < tb.P ? tb.ucode * 8 + 7 : 0xffffffffu;
    if (urow == 0xffffffffu && threadIdx.x == 0) *uc = kNoUntyped;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        float* row = f + i * kTabRow;
        const float o = tb.emax[i];
        if constexpr (LOG) {
            const float* g = tb.palette + 3 * (i / 8);
            const uint32_t c = i % 8;
            row[0] = (c & 1u) ? g[0] - o : -INFINITY;
            row[1] = (c & 2u) ? g[1] - o : -INFINITY;
            row[2] = (c & 4u) ? g[2] - o : -INFINITY;
            row[3] = -INFINITY;
            if (i == urow) *uc = (g[0] == 0.0f && g[1] == 0.0f && g[2] == 0.0f && o == 0.0f) ? tb.ucode : kNoUntyped;
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) row[k] = tb.eps[4 * i + k];
            if (i == urow)
                *uc = (row[0] == 1.0f && row[1] == 1.0f && row[2] == 1.0f && o == 0.0f) ? tb.ucode : kNoUntyped;
        }
        row[4] = o;
    }
    return Tab{f, uc};
}

__host__ __device__ constexpr size_t fb_table_lds_bytes(uint32_t P) {
    return size_t(P) * 8 * kTabRow * sizeof(float) + 16;
}

// The class-present mask of a marker rides in bit 7 of the first three state
// bytes of every aligned state-table word, so the emission row needs no load
// of its own.  The 24-bit product moves bits 7, 15, 23 to 21, 22, 23 (the
// other partial products land on distinct bits, so nothing carries).
__device__ __forceinline__ uint32_t class_mask(uint32_t w) {
    return (__umul24(w & 0x808080u, 0x4081u) >> 21) & 7u;
}

// The lane's state bytes of one marker: NW aligned words; for R < 4 the
// lane's bytes start at byte `sub` of its word.
template <class S>
struct LaneWords {
    uint32_t w[S::NW];
    __device__ __forceinline__ uint32_t byte(int i, uint32_t sub) const {
        if constexpr (S::R >= 4) return (w[i >> 2] >> (8 * (i & 3))) & 0xffu;
        else return (w[0] >> (8 * (sub + uint32_t(i)))) & 0xffu;
    }
    // The emission class of state i (bits 2-3 of its byte).
    __device__ __forceinline__ uint32_t cls(int i, uint32_t sub) const {
        if constexpr (S::R >= 4) return (w[i >> 2] >> (8 * (i & 3) + 2)) & 3u;
        else return (w[0] >> (8u * (sub + uint32_t(i)) + 2u)) & 3u;
    }
    // Bit b of state byte i as an all-ones / all-zeros mask.
    __device__ __forceinline__ uint32_t bitmask(int i, uint32_t sub, int b) const {
        const uint32_t x = S::R >= 4 ? w[i >> 2] : w[0];
        const uint32_t pos = S::R >= 4 ? uint32_t(8 * (i & 3) + b) : 8u * (sub + uint32_t(i)) + b;
        return uint32_t(__builtin_amdgcn_sbfe(int32_t(x), pos, 1u));  // one signed bit-field extract
    }
};

template <class S>
__device__ __forceinline__ uint32_t lane_sub(uint32_t r) {
    return S::R >= 4 ? 0u : (r * S::R) & 3u;
}

// Emission values of the lane's R states at a marker with palette code `code`;
// codes at or beyond the palette size read as the last entry.  Returns the
// row's log offset.
template <class S>
__device__ __forceinline__ float lane_emission(const Tab& tab, uint32_t P, uint32_t code,
                                               const LaneWords<S>& sw, uint32_t sub,
                                               float (&e)[S::R]) {
    const float* rp = tab.f + (min(code, P - 1u) * 8u + class_mask(sw.w[0])) * kTabRow;
#pragma unroll
    for (int i = 0; i < S::R; ++i) e[i] = rp[sw.cls(i, sub)];
    return rp[4];
}

// ------------------------------------------------------------ reductions --
template <int N>
__device__ __forceinline__ float tree_sum(const float* x) {
    if constexpr (N == 1) {
        return x[0];
    } else {
        float h[N / 2];
#pragma unroll
        for (int i = 0; i < N / 2; ++i) h[i] = x[2 * i] + x[2 * i + 1];
        return tree_sum<N / 2>(h);
    }
}

template <int N>
__device__ __forceinline__ float tree_max(const float* x) {
    float m = x[0];
#pragma unroll
    for (int i = 1; i < N; ++i) m = fmaxf(m, x[i]);
    return m;
}

// Sum / max over the lane group (G lanes), every lane receiving the result.
template <int G>
__device__ __forceinline__ float gsum(float v) {
    if constexpr (G <= 16) {
        return group_sum<G>(v);
    } else {
        v = group_sum<16>(v);
        v += __shfl_xor(v, 16, kWave);
        if constexpr (G >= 64) v += __shfl_xor(v, 32, kWave);
        return v;
    }
}
template <int G>
__device__ __forceinline__ float gmax(float v) {
    if constexpr (G <= 16) {
        return group_max<G>(v);
    } else {
        v = group_max<16>(v);
        v = fmaxf(v, __shfl_xor(v, 16, kWave));
        if constexpr (G >= 64) v = fmaxf(v, __shfl_xor(v, 32, kWave));
        return v;
    }
}

template <class S>
__device__ __forceinline__ float vsum(const float (&x)[S::R]) {
    return gsum<S::G>(tree_sum<S::R>(x));
}
// Two independent group sums, level by level, so each cross-lane add has the
// other to hide its latency behind (the same tree as gsum).
template <int G>
__device__ __forceinline__ void gsum2(float& a, float& b) {
    if constexpr (G >= 2) { a += dpp<0xB1>(a); b += dpp<0xB1>(b); }
    if constexpr (G >= 4) { a += dpp<0x4E>(a); b += dpp<0x4E>(b); }
    if constexpr (G >= 8) { a += dpp<0x141>(a); b += dpp<0x141>(b); }
    if constexpr (G >= 16) { a += dpp<0x140>(a); b += dpp<0x140>(b); }
    if constexpr (G >= 32) {
        a += __shfl_xor(a, 16, kWave);
        b += __shfl_xor(b, 16, kWave);
    }
    if constexpr (G >= 64) {
        a += __shfl_xor(a, 32, kWave);
        b += __shfl_xor(b, 32, kWave);
    }
}
template <class S>
__device__ __forceinline__ float vmax(const float (&x)[S::R]) {
    return gmax<S::G>(tree_max<S::R>(x));
}

// LSE(a, b, c_i) for a pair (a, b) shared by every state and a per-state c_i:
// the pair is folded once per step into (M0, S0) with M0 = max(a, b),
// S0 = e^{a-M0} + e^{b-M0} in [1, 2]; each state then costs one exp and one
// log: with d = c - M0,
//   d <= 0:  M0 + ln(S0 + e^d)
//   d >  0:  c  + ln(1 + S0 e^-d)
struct LsePair {
    float M0, S0;
};
__device__ __forceinline__ LsePair lse_pair(float a, float b) {
    const float m = fmaxf(a, b);
    return LsePair{m, fast_exp(a - m) + fast_exp(b - m)};
}
__device__ __forceinline__ float lse_with(const LsePair& q, float c) {
    const float d = c - q.M0;
    const float e = fast_exp(-fabsf(d));
    return d <= 0.0f ? q.M0 + fast_ln(q.S0 + e) : c + fast_ln(fmaf(q.S0, e, 1.0f));
}

// Log mass LSE_i(a_i) of a max-normalized log vector (a value in [0, ln Lp]).
template <class S>
__device__ __forceinline__ float log_mass(const float (&a)[S::R]) {
    float w[S::R];
#pragma unroll
    for (int i = 0; i < S::R; ++i) w[i] = fast_exp(a[i]);
    return fast_ln(vsum<S>(w));
}

template <int R>
__device__ __forceinline__ void vsel(bool c, float (&dst)[R], const float (&src)[R]) {
#pragma unroll
    for (int i = 0; i < R; ++i) dst[i] = c ? src[i] : dst[i];
}

// ---------------------------------------------------------------- forward --
// Marker 0.  Returns the log normalizer (log-likelihood increment).
template <class S, bool LOG>
__device__ __forceinline__ double fwd_init(float (&a)[S::R], const float (&e)[S::R], float off) {
    if constexpr (LOG) {
        const float c = vmax<S>(e);
#pragma unroll
        for (int i = 0; i < S::R; ++i) a[i] = e[i] - c;
        return double(off) + double(c);
    } else {
        const float Sm = vsum<S>(e);
        const float r = __builtin_amdgcn_rcpf(Sm);
#pragma unroll
        for (int i = 0; i < S::R; ++i) a[i] = e[i] * r;
        return double(off) + double(fast_ln(Sm));
    }
}

// Step coefficients: (t_o, u) for the linear form, (t_o, u, lto, lu) for the
// log form (t_o = e^lto, u = e^lu).
template <bool LOG>
struct CoefT {
    using type = float2;
};
template <>
struct CoefT<true> {
    using type = float4;
};

// The log form's three-term LSE(floor, switch term, u e^{a_i}) over the
// states is ln(F + K w_i), with w_i = e^{a_i - max} the exponentials the
// step forms anyway for its log mass, whenever the state-independent part F
// is at least 2^-30: the two terms are then added in linear space (no
// cancellation; a w_i that underflows, a_i < -87, drops a term smaller than
// 2^-96 F) and each state costs one log instead of an exp and a log.  Below
// that (coincident markers, where t_o underflows) the exact three-term form
// is used.
constexpr float kLseLinMin = 9.3132257e-10f;  // 2^-30

// Step into a marker with emission e / offset off and coefficients k.  In
// the log form the vector a is max-normalized (its largest entry is 0).
template <class S, bool LOG>
__device__ __forceinline__ double fwd_step(float (&a)[S::R], const float (&e)[S::R], float off,
                                           typename CoefT<LOG>::type k) {
    float x[S::R];
    if constexpr (LOG) {
        float w[S::R];
#pragma unroll
        for (int i = 0; i < S::R; ++i) w[i] = fast_exp(a[i]);
        const float Wm = vsum<S>(w);  // e^{log mass}
        const float F = k.x * Wm;     // LSE_i(a_i + To), in linear space
        if (F >= kLseLinMin) {
#pragma unroll
            for (int i = 0; i < S::R; ++i) x[i] = fast_ln(fmaf(k.y, w[i], F)) + e[i];
        } else {
            const LsePair q = lse_pair(kFloor, k.z + fast_ln(Wm));
#pragma unroll
            for (int i = 0; i < S::R; ++i) x[i] = lse_with(q, k.w + a[i]) + e[i];
        }
        const float c = vmax<S>(x);
#pragma unroll
        for (int i = 0; i < S::R; ++i) a[i] = x[i] - c;
        return double(off) + double(c);
    } else {
        // a sums to 1, so sum_i a_i t_o = t_o
#pragma unroll
        for (int i = 0; i < S::R; ++i) x[i] = fmaf(k.y, a[i], k.x) * e[i];
        const float Sm = vsum<S>(x);
        const float r = __builtin_amdgcn_rcpf(Sm);
#pragma unroll
        for (int i = 0; i < S::R; ++i) a[i] = x[i] * r;
        return double(off) + double(fast_ln(Sm));
    }
}

// --------------------------------------------------------------- backward --
// From the vector at marker n (emission e, offset off) to n - 1.  The true
// backward vector is e^{Bl} v (linear) or Bl + v (log).  With
// nat = Bl + off + log W the new scale is max(nat, -999): when the floor
// dominates (nat < -999) the scale pins at -999 and the non-floor part is
// shifted down by sh = -999 - nat, so no value ever leaves a safe range.
template <class S, bool LOG>
__device__ __forceinline__ void bwd_step(float (&v)[S::R], double& Bl, const float (&e)[S::R],
                                         float off, typename CoefT<LOG>::type k) {
    float y[S::R];
    if constexpr (LOG) {
#pragma unroll
        for (int i = 0; i < S::R; ++i) y[i] = e[i] + v[i];
        const float mx = vmax<S>(y);
        float w[S::R];
#pragma unroll
        for (int i = 0; i < S::R; ++i) w[i] = fast_exp(y[i] - mx);
        const float Sw = vsum<S>(w);
        const float lw = mx + fast_ln(Sw);
        const double nat = Bl + double(off) + double(lw);
        const double arg = -999.0 - nat;
        const bool floor_dominates = arg > 0.0;
        Bl = floor_dominates ? -999.0 : nat;
        // v_i = LSE(fl, To - sh, U - sh + y_i - lw) with (fl, sh) = (arg, 0),
        // or (0, arg) when the floor dominates; with z = e^{-|arg|} that is
        // ln(F + K w_i), F = z + t_o or 1 + t_o z, K = u / Sw or u z / Sw.
        const float z = fast_exp(float(-fabs(arg)));
        const float F = floor_dominates ? fmaf(k.x, z, 1.0f) : z + k.x;
        if (F >= kLseLinMin) {
            const float K = (floor_dominates ? k.y * z : k.y) * __builtin_amdgcn_rcpf(Sw);
#pragma unroll
            for (int i = 0; i < S::R; ++i) v[i] = fast_ln(fmaf(K, w[i], F));
        } else {
            const float fl = floor_dominates ? 0.0f : float(arg);
            const float sh = floor_dominates ? float(arg) : 0.0f;
            const LsePair q = lse_pair(fl, k.z - sh);
            const float bu = k.w - lw - sh;
#pragma unroll
            for (int i = 0; i < S::R; ++i) v[i] = lse_with(q, bu + y[i]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < S::R; ++i) y[i] = e[i] * v[i];
        const float W = vsum<S>(y);
        const double nat = Bl + double(off) + double(fast_ln(W));
        const double arg = -999.0 - nat;
        const float x = fast_exp(float(-fabs(arg)));
        const bool floor_dominates = arg > 0.0;
        const float f = floor_dominates ? 1.0f : x;
        const float g = floor_dominates ? x : 1.0f;
        Bl = floor_dominates ? -999.0 : nat;
        const float gr = g * k.y * __builtin_amdgcn_rcpf(W);
        const float base = fmaf(g, k.x, f);
#pragma unroll
        for (int i = 0; i < S::R; ++i) v[i] = fmaf(gr, y[i], base);
    }
}

template <class S, bool LOG>
__device__ __forceinline__ void bwd_init(float (&v)[S::R]) {
#pragma unroll
    for (int i = 0; i < S::R; ++i) v[i] = LOG ? 0.0f : 1.0f;
}

// -------------------------------------------------------------- posterior --
template <class S, bool LOG>
__device__ __forceinline__ void posterior(const float (&a)[S::R], const float (&v)[S::R],
                                          float (&p)[S::R]) {
    if constexpr (LOG) {
        float q[S::R];
#pragma unroll
        for (int i = 0; i < S::R; ++i) q[i] = a[i] + v[i];
        const float mx = vmax<S>(q);
#pragma unroll
        for (int i = 0; i < S::R; ++i) p[i] = fast_exp(q[i] - mx);
    } else {
#pragma unroll
        for (int i = 0; i < S::R; ++i) p[i] = a[i] * v[i];
    }
    const float rq = __builtin_amdgcn_rcpf(vsum<S>(p));
#pragma unroll
    for (int i = 0; i < S::R; ++i) p[i] *= rq;
}

// ---------------------------------------------------------- untyped steps --
// At an untyped marker (genotype likelihoods (0, 0, 0): every real state's
// emission is 1, offset 0) a step is the same affine map for every sample:
//   forward, sum-normalized:  a -> (t_o 1_L + u a) / S,  S = L t_o + u,
//                             log-likelihood + ln S
//   backward:                 b -> e^-999 1 + t_o (sum_L b) 1 + u b
// (1_L: ones on the L real states).  Under constant transitions k of them in
// a row have a closed form (c = u / S, geometric sums):
//   forward   a -> c^k a + (1 - c^k)/L 1_L,    log-likelihood + k ln S
//   backward  b = e^B v -> e^B' (x 1 + y v), with D = B + k ln S + 999 and
//             f = e^(-999 - B), W = sum_L v:
//     D > 0 (mass above the floor):  B' = B + k ln S,  y = c^k,
//                                    x = f (S^-k - 1)/(1 - S) + (W/L)(1 - c^k)
//     D <= 0 (the floor dominates):  B' = -999,  y = c^k e^D,
//                                    x = (1 - S^k)/(1 - S) + (W/L)(1 - c^k) e^D
//   (differences of 1 and a power as expm1, exact near 1).
// The chain pass therefore does nothing at an untyped marker of a fast load
// block: the vector stays at an anchor marker and is brought forward in one
// closed-form run just before every other step and at each 16-aligned marker
// (the checkpoints).  Where runs end depends on the markers only, so the
// chain and block passes carry the same values and every lane shape and
// checkpoint block gives the same bits.  The block pass also needs the
// vectors at untyped markers for their posteriors: it forms them with a
// per-marker recurrence from the anchor that is never carried.  Fast load
// blocks: constant transitions (URun::ok), no windows, the block's whole
// 32-marker super-block inside the job and past marker 31, so a checkpoint
// block is fast or not as a whole.  The log form takes untyped steps only
// where t_o >= 2^-30: its own step then adds the terms in linear space too
// (the forward's -999 floor is below resolution).

// Bit t set where code byte t of a load block equals uc (< 256).
__device__ __forceinline__ uint32_t untyped_bits(const uint4& cw, uint32_t uc) {
    const uint32_t rep = uc * 0x01010101u;
    const uint32_t wd[4] = {cw.x, cw.y, cw.z, cw.w};
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t x = wd[k] ^ rep;
        // bit 7 of each byte set iff the byte is 0 (no carries between bytes)
        const uint32_t z = ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
        // bits 7, 15, 23 -> 21, 22, 23 (distinct partial products, no carries); bit 31 -> 3
        const uint32_t b = ((__umul24((z >> 7) & 0x10101u, 0x204081u) >> 21) & 7u) | ((z >> 28) & 8u);
        m |= b << (4 * k);
    }
    return m;
}

// Whether the untyped code's row is a missing genotype (likelihoods 0, 0, 0:
// emission 1 in every class, offset 0), the predicate load_tables records in
// the table's last word.  If not, the code is ignored: every marker counts
// as typed.
__device__ __forceinline__ bool untyped_row_ok(const FbTables& tb, bool log) {
    if (tb.ucode >= tb.P) return false;
    const uint32_t i = tb.ucode * 8u + 7u;  // the row with every class present
    const float o = tb.emax[i];
    if (log) {
        const float* g = tb.palette + 3u * tb.ucode;
        return g[0] == 0.0f && g[1] == 0.0f && g[2] == 0.0f && o == 0.0f;
    }
    return tb.eps[4u * i] == 1.0f && tb.eps[4u * i + 1u] == 1.0f && tb.eps[4u * i + 2u] == 1.0f && o == 0.0f;
}

// The job's typed-marker mask: bit t of tmask[c] set where marker 16c + t
// (< M) has a code other than uc in some sample (every marker when ok is
// false), so a marker is untyped for the job when no sample has a genotype
// there; the decision is the same for every sample of the job (wavefront-
// uniform control in both passes) and does not depend on which samples share
// a wavefront.  A workgroup covers
// kMaskBlocks load blocks with kMaskLanes lanes each striding over the
// samples (consecutive lanes read consecutive 16-byte code words of a row);
// 32 lanes give a 100,000-marker job 196 workgroups of 16 wavefronts (8
// lanes: 4), enough loads in flight to stream the codes near the HBM rate.
constexpr uint32_t kMaskBlocks = kTmaskBlocksPerGroup, kMaskLanes = 32,
                   kMaskThreads = kMaskBlocks * kMaskLanes;
__device__ __forceinline__ void tmask_tile(const uint8_t* codes, uint32_t B, uint32_t M, uint32_t pitch,
                                           uint32_t uc, bool ok, uint32_t c0, uint32_t* tmask) {
    __shared__ uint32_t red[kMaskThreads];
    const uint32_t ci = threadIdx.x % kMaskBlocks, sl = threadIdx.x / kMaskBlocks;
    const uint32_t c = c0 + ci;
    const uint32_t nlb = (M + kSub - 1) / kSub;
    uint32_t m = 0;
    if (c < nlb) {
        if (!ok) m = 0xffffu;
        else
#pragma unroll 4
            for (uint32_t s = sl; s < B; s += kMaskLanes)
                m |= untyped_bits(*reinterpret_cast<const uint4*>(codes + size_t(s) * pitch + kSub * c), uc) ^
                     0xffffu;
        if (kSub * c + kSub > M) m &= (1u << (M - kSub * c)) - 1u;  // markers past the job
    }
    red[threadIdx.x] = m;
    __syncthreads();
    if (sl == 0 && c < nlb) {
#pragma unroll
        for (uint32_t k = 1; k < kMaskLanes; ++k) m |= red[k * kMaskBlocks + ci];
        tmask[c] = m;
    }
}

// Real-state mask of the lane's R states (1 on states < L).
template <class S>
__device__ __forceinline__ void real_mask(uint32_t r, uint32_t L, float (&rm)[S::R]) {
#pragma unroll
    for (int i = 0; i < S::R; ++i) rm[i] = (r * uint32_t(S::R) + uint32_t(i) < L) ? 1.0f : 0.0f;
}

// e^x - 1, accurate near 0 (series below 1/8: truncation < 1e-10 relative).
__device__ __forceinline__ float em1(float x) {
    const float p =
        x * (1.0f + x * (0.5f + x * (0.16666667f + x * (0.041666668f + x * (0.0083333338f + x * 0.0013888889f)))));
    return fabsf(x) < 0.125f ? p : fast_exp(x) - 1.0f;
}

// The forward brought k untyped steps past its anchor (no-op for k = 0);
// returns the log-likelihood increment.  Log form: the max-normalized log
// vector, taken to linear space for the run.
template <class S, bool LOG>
__device__ __forceinline__ double fwd_run(float (&a)[S::R], const float (&rm)[S::R], uint32_t k,
                                          const URun& q) {
    if (k == 0u) return 0.0;
    const float kf = float(k);
    const float al = __builtin_amdgcn_exp2f(kf * q.lc1);  // c^k
    const float be = -em1(kf * q.lnc1) * q.rL;           // (1 - c^k) / L
    double inc = double(k) * q.lnSd;
    if constexpr (LOG) {
        float w[S::R];
#pragma unroll
        for (int i = 0; i < S::R; ++i) w[i] = fast_exp(a[i]);
        const float Wm = vsum<S>(w);
        const float aw = al * __builtin_amdgcn_rcpf(Wm);
#pragma unroll
        for (int i = 0; i < S::R; ++i) a[i] = fast_ln(fmaf(aw, w[i], be * rm[i]));
        const float c = vmax<S>(a);
#pragma unroll
        for (int i = 0; i < S::R; ++i) a[i] -= c;
        inc += double(fast_ln(Wm)) + double(c);
    } else {
#pragma unroll
        for (int i = 0; i < S::R; ++i) a[i] = fmaf(al, a[i], be * rm[i]);
    }
    return inc;
}

// The backward's anchor as a linear base: v itself (linear) or e^(v - max)
// over the real states with the max folded into the scale (log; a scale
// below -999 is rescaled to the floor, dropping less than 2^-24 of it).
// Returns the base's sum over the real states; B = its scale.
template <class S, bool LOG>
__device__ __forceinline__ float bwd_base(const float (&v)[S::R], double Bl, const float (&rm)[S::R],
                                          float (&base)[S::R], double& B) {
    B = Bl;
    if constexpr (LOG) {
        float vr[S::R];
#pragma unroll
        for (int i = 0; i < S::R; ++i) vr[i] = rm[i] != 0.0f ? v[i] : -INFINITY;
        const float mv = vmax<S>(vr);
        B = Bl + double(mv);
        float sh = 0.0f;
        if (B < -999.0) {
            sh = float(B + 999.0);
            B = -999.0;
        }
#pragma unroll
        for (int i = 0; i < S::R; ++i) base[i] = fast_exp(v[i] - mv + sh);
    } else {
#pragma unroll
        for (int i = 0; i < S::R; ++i) base[i] = v[i];
    }
    float t[S::R];
#pragma unroll
    for (int i = 0; i < S::R; ++i) t[i] = base[i] * rm[i];
    return vsum<S>(t);
}

// The backward brought k untyped steps below its anchor (no-op for k = 0).
template <class S, bool LOG>
__device__ __forceinline__ void bwd_run(float (&v)[S::R], double& Bl, const float (&rm)[S::R], uint32_t k,
                                        const URun& q) {
    if (k == 0u) return;
    float base[S::R];
    double B;
    const float W0 = bwd_base<S, LOG>(v, Bl, rm, base, B);
    const float kf = float(k);
    const double Bn = fma(double(k), q.lnSd, B);
    const float D = float(Bn + 999.0);
    const float ck = __builtin_amdgcn_exp2f(kf * q.lc1);
    const float wl = W0 * q.rL * -em1(kf * q.lnc1);  // (W/L)(1 - c^k)
    float x, y;
    if (D > 0.0f) {
        const float f0 = fast_exp(float(-999.0 - B));
        const float g = -kf * q.lnS;  // f (S^-k - 1) = e^-D - f
        const float fl = g < 0.125f ? f0 * em1(g) : fast_exp(-D) - f0;
        x = fmaf(fl, q.gS, wl);
        y = ck;
        Bl = Bn;
    } else {
        const float E = fast_exp(D);
        x = fmaf(-em1(kf * q.lnS), q.gS, wl * E);
        y = ck * E;
        Bl = -999.0;
    }
#pragma unroll
    for (int i = 0; i < S::R; ++i) v[i] = LOG ? fast_ln(fmaf(y, base[i], x)) : fmaf(y, base[i], x);
}

// ------------------------------------------------------------ block inputs --
// One 16-marker load block of a lane's inputs: the sample's 16 genotype
// codes, per marker the lane's state-table words and (per-marker
// transitions) the step coefficients.  Loaded whole, with indices clamped
// inside the job (markers past the end repeat the last one, a load block
// past the end reads the last whole one), so every load of the block is in
// flight at once and the steps that follow read registers only.
template <class S, bool LOG, bool PM>
struct BlockIn {
    uint4 cw;
    LaneWords<S> sa[kSub];
    typename CoefT<LOG>::type k[PM ? kSub : 1];
    __device__ __forceinline__ void load(const JobView& J, const uint8_t* codes, uint32_t r,
                                         uint32_t m0) {
        const uint32_t ml = J.M - 1u;
        m0 = min(m0, ml & ~(kSub - 1u));
        cw = *reinterpret_cast<const uint4*>(codes + m0);
        const uint32_t wpr = uint32_t(S::Lp) / 4u;       // words per state-table row
        const uint32_t w0 = (r * uint32_t(S::R)) / 4u;   // the lane's first word
        const uint32_t* sa32 = reinterpret_cast<const uint32_t*>(J.sa);
#pragma unroll
        for (uint32_t t = 0; t < kSub; ++t) {
            const uint32_t* p = sa32 + size_t(min(m0 + t, ml)) * wpr + w0;
#pragma unroll
            for (int w = 0; w < S::NW; ++w) sa[t].w[w] = p[w];  // the table is 4-byte aligned
        }
        if constexpr (PM) {
#pragma unroll
            for (uint32_t t = 0; t < kSub; ++t) {
                const float4 q = J.steps[min(m0 + t, ml)];
                if constexpr (LOG) k[t] = q;
                else k[t] = make_float2(q.x, q.y);
            }
        } else {
            if constexpr (LOG) k[0] = J.step;
            else k[0] = make_float2(J.step.x, J.step.y);
        }
    }
    __device__ __forceinline__ uint32_t code(uint32_t t) const {
        return byte_of((&cw.x)[t >> 2], t & 3);
    }
};

// What a load block's 16 steps consume: the lane's emissions (read from the
// LDS table together, before any step needs them) and the coefficients.
// Filling it frees the BlockIn for the next block's loads.
template <class S, bool LOG, bool PM>
struct BlockEm {
    float e[kSub][S::R];
    float off[kSub];
    typename CoefT<LOG>::type k[PM ? kSub : 1];
    __device__ __forceinline__ void fill(const Tab& tab, uint32_t P, const BlockIn<S, LOG, PM>& in,
                                         uint32_t sub) {
#pragma unroll
        for (uint32_t t = 0; t < kSub; ++t)
            off[t] = lane_emission<S>(tab, P, in.code(t), in.sa[t], sub, e[t]);
#pragma unroll
        for (uint32_t t = 0; t < (PM ? kSub : 1); ++t) k[t] = in.k[t];
    }
    __device__ __forceinline__ typename CoefT<LOG>::type coef(uint32_t t) const {
        return k[PM ? t : 0];
    }
};

// R consecutive floats of a checkpoint vector.
template <int R>
__device__ __forceinline__ void store_vec(float* p, const float (&x)[R]) {
    if constexpr (R == 2) {
        *reinterpret_cast<float2*>(p) = make_float2(x[0], x[1]);
    } else {
#pragma unroll
        for (int i = 0; i < R; i += 4)
            *reinterpret_cast<float4*>(p + i) = make_float4(x[i], x[i + 1], x[i + 2], x[i + 3]);
    }
}
template <int R>
__device__ __forceinline__ void load_vec(const float* p, float (&x)[R]) {
    if constexpr (R == 2) {
        const float2 q = *reinterpret_cast<const float2*>(p);
        x[0] = q.x;
        x[1] = q.y;
    } else {
#pragma unroll
        for (int i = 0; i < R; i += 4) {
            const float4 q = *reinterpret_cast<const float4*>(p + i);
            x[i] = q.x;
            x[i + 1] = q.y;
            x[i + 2] = q.z;
            x[i + 3] = q.w;
        }
    }
}

// Orders this wavefront's LDS accesses: lane groups exchange data through LDS
// only within their own wavefront.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// One chain (forward or backward) of sample s of job J, by one lane group.
// The next load block's inputs are fetched while the current one is stepped;
// every step is computed and committed by a select where it lies inside the
// chain, so a load block's 16 steps are straight-line code.  J.nblk counts
// checkpoint blocks of CB markers.
//
// Windows (J.W > 0, each window [lo, hi]): the forward chain of window w > 0
// first runs from the carried vector (the previous window's last one; the
// step into lo is not taken) to hi, adding to the log-likelihood, which gives
// the window's last vector F_w; it then runs the window again from F_w,
// storing the checkpoints, and carries F_w into window w + 1.  Window 0 runs
// once.  The backward chain does not take the step out of a window start.
// Without windows the whole job is window 0.
// ---------------------------------------------------------------------------
template <class S, bool LOG, bool PM, int CB>
__device__ __forceinline__ void run_chain(const JobView& J, const Tab& tab, uint32_t P, uint32_t s,
                                          bool fwd, uint32_t r) {
    constexpr uint32_t NS = uint32_t(CB) / kSub;  // load blocks per checkpoint block
    const uint8_t* codes = J.codes + size_t(s) * J.pitch;
    const uint32_t nlb = (J.M + kSub - 1) / kSub;
    const uint32_t sub = lane_sub<S>(r);
    const size_t vo = size_t(r) * S::R;  // the lane's offset in a checkpoint vector
    BlockIn<S, LOG, PM> cur;
    BlockEm<S, LOG, PM> em;
    if (fwd) {
        const uint32_t W = J.W ? J.W : J.M;
        const uint32_t nw = J.W ? J.nwin : 1u;
        cur.load(J, codes, r, 0);
        float a[S::R];
        double ll;
        {
            float e[S::R];
            const float off = lane_emission<S>(tab, P, cur.code(0), cur.sa[0], sub, e);
            ll = fwd_init<S, LOG>(a, e, off);
        }
        // One flat walk over the load blocks of every (window, pass), so the
        // next block's inputs are always in flight: pass 0 (windows after the
        // first) is the carried chain, pass 1 the checkpointed run from the
        // window's last vector.
        uint32_t w = 0, pass = 1, lo = 0, hi = min(W, J.M) - 1u, c = 0;
        for (;;) {
            em.fill(tab, P, cur, sub);
            const uint32_t c1 = hi / kSub;
            // the load block after this one
            uint32_t nc = c + 1u;
            const bool pass_end = c == c1;
            const bool last = pass_end && pass == 1u && w + 1u == nw;
            if (pass_end) nc = pass == 0u ? lo / kSub : (hi + 1u) / kSub;
            if (!last) cur.load(J, codes, r, nc * kSub);
            const uint32_t m0 = c * kSub;
            const bool keep = pass == 1u;
            const bool sum_ll = pass == 0u || w == 0u;
            const bool ck = keep && (c % NS) == 0 && m0 >= lo;
            float* ckp = J.ck_alpha + (size_t(s) * J.nblk + m0 / CB) * S::Lp + vo;
            if (lo < m0 && m0 + kSub - 1u <= hi) {  // every step inside: no selects
#pragma unroll
                for (uint32_t t = 0; t < kSub; ++t) {
                    const double inc = fwd_step<S, LOG>(a, em.e[t], em.off[t], em.coef(t));
                    if (sum_ll) ll += inc;
                    if (t == 0 && ck) store_vec<S::R>(ckp, a);
                }
            } else {
#pragma unroll
                for (uint32_t t = 0; t < kSub; ++t) {
                    const uint32_t m = m0 + t;
                    float an[S::R];
#pragma unroll
                    for (int i = 0; i < S::R; ++i) an[i] = a[i];
                    const double inc = fwd_step<S, LOG>(an, em.e[t], em.off[t], em.coef(t));
                    const bool live = m > lo && m <= hi;
                    vsel(live, a, an);
                    ll += (live && sum_ll) ? inc : 0.0;
                    if (t == 0 && ck) store_vec<S::R>(ckp, a);
                }
            }
            if (last) break;
            if (pass_end) {
                float* fwp = J.ck_win + (size_t(s) * J.nwin + w) * S::Lp + vo;
                if (pass == 0u) {  // F_w: the restart of pass 1 and of the block pass
                    store_vec<S::R>(fwp, a);
                    pass = 1u;
                } else {  // carry F_w into the next window
                    if (w > 0u) load_vec<S::R>(fwp, a);
                    ++w;
                    pass = 0u;
                    lo = w * W;
                    hi = min(lo + W, J.M) - 1u;
                }
            }
            c = nc;
        }
        if (LOG) ll += double(log_mass<S>(a));  // log mass of the last normalized vector
        if (r == 0 && J.loglik) J.loglik[s] = ll;
    } else {
        float v[S::R];
        bwd_init<S, LOG>(v);
        double Bl = 0.0;
        const size_t last = size_t(s) * J.nblk + (J.nblk - 1);
        store_vec<S::R>(J.ck_beta + last * S::Lp + vo, v);
        if (r == 0) J.ck_bl[last] = Bl;
        // the highest window start at or below the current load block's end
        // (0 without windows: the step out of marker 0 is never taken)
        uint32_t ws = J.W ? (J.M - 1u) / J.W * J.W : 0u;
        cur.load(J, codes, r, (nlb - 1) * kSub);
        for (int c = int(nlb) - 1; c >= 0; --c) {
            em.fill(tab, P, cur, sub);
            if (c > 0) cur.load(J, codes, r, uint32_t(c - 1) * kSub);
            const uint32_t m0 = uint32_t(c) * kSub;
            const bool ck = c > 0 && (uint32_t(c) % NS) == 0;
            const bool has_ws = ws >= m0;
            // v at marker m0 - 1 is the last of checkpoint block m0 / CB - 1
            const size_t b = size_t(s) * J.nblk + (m0 / CB) - 1u;
            if (m0 > 0 && m0 + kSub <= J.M && !has_ws) {  // every step inside: no selects
#pragma unroll
                for (int t = int(kSub) - 1; t >= 0; --t) {
                    bwd_step<S, LOG>(v, Bl, em.e[t], em.off[t], em.coef(uint32_t(t)));
                    if (t == 0 && ck) {
                        store_vec<S::R>(J.ck_beta + b * S::Lp + vo, v);
                        if (r == 0) J.ck_bl[b] = Bl;
                    }
                }
            } else {
#pragma unroll
                for (int t = int(kSub) - 1; t >= 0; --t) {
                    const uint32_t n = m0 + uint32_t(t);
                    float vn[S::R];
#pragma unroll
                    for (int i = 0; i < S::R; ++i) vn[i] = v[i];
                    double Bln = Bl;
                    bwd_step<S, LOG>(vn, Bln, em.e[t], em.off[t], em.coef(uint32_t(t)));
                    const bool live = n > 0 && n < J.M && n != ws;
                    vsel(live, v, vn);
                    Bl = live ? Bln : Bl;
                    if (t == 0 && ck) {
                        store_vec<S::R>(J.ck_beta + b * S::Lp + vo, v);
                        if (r == 0) J.ck_bl[b] = Bl;
                    }
                }
            }
            if (has_ws && ws >= J.W) ws -= J.W;  // windows are longer than a load block
        }
    }
}

// The sampler's 4 bits of a state byte (bits 0-1 alleles, 4-5 missing
// flags) as a nibble: alleles in bits 0-1, missing flags in bits 2-3.  nib4
// packs the 4 state bytes of a word into 16 bits, nib2 the 2 of a half-word
// into 8 (state k's nibble at bit 4k).
__device__ __forceinline__ uint32_t nib4(uint32_t w) {
    uint32_t x = (w & 0x03030303u) | ((w >> 2) & 0x0C0C0C0Cu);
    x |= x >> 4;
    return (x & 0xFFu) | ((x >> 8) & 0xFF00u);
}
__device__ __forceinline__ uint32_t nib2(uint32_t h) {
    uint32_t x = (h & 0x0303u) | ((h >> 2) & 0x0C0Cu);
    x |= x >> 4;
    return x & 0xFFu;
}

// Sampling phase of a sub-block (both block-pass forms): Philox outputs for
// its 16 markers, then lane r samples markers r, r + G, ... from the group's
// LDS posterior rows and writes dosage, allele probabilities, haplotypes and
// (tests) probabilities.  Every lane of the wavefront passes the wave_syncs.
template <class S, bool SB, uint32_t kPer>
__device__ __forceinline__ void block_sample(const JobView& J, bool active, uint32_t s, uint32_t m0,
                                             uint32_t base, uint32_t nm, uint32_t r, const float* pb,
                                             const uint32_t* sb, uint4* rng, const float (&ap0r)[kPer],
                                             const float (&ap1r)[kPer]) {
    using Lds = BlockLds<S>;
    const bool argmax = (J.flags & SI_FB_ARGMAX) != 0;
    const uint2 key = make_uint2(uint32_t(J.seed), uint32_t(J.seed >> 32));
    uint4 rl[Lds::kRngLds ? 1 : kSub / 4];  // Philox outputs of a one- or two-lane group
    if (active) {
        if constexpr (Lds::kRngLds) {
            for (uint32_t q = r; q < kSub / 4; q += S::G)
                rng[q] = philox4x32_10(
                    make_uint4(((m0 + base) >> 2) + q, J.sample_base + s, J.rng_stream, 0u), key);
        } else if constexpr (S::G == 1) {
#pragma unroll
            for (uint32_t q = 0; q < kSub / 4; ++q)
                rl[q] = philox4x32_10(
                    make_uint4(((m0 + base) >> 2) + q, J.sample_base + s, J.rng_stream, 0u), key);
        } else {
            // lane r forms blocks r and r + 2, then takes its partner's
            // (blocks 1 - r and 3 - r): rl[q] holds block q on both lanes
#pragma unroll
            for (uint32_t h = 0; h < 2; ++h) {
                const uint4 own = philox4x32_10(
                    make_uint4(((m0 + base) >> 2) + r + 2u * h, J.sample_base + s, J.rng_stream, 0u),
                    key);
                const uint4 oth = make_uint4(dpp_u32<0xB1>(own.x), dpp_u32<0xB1>(own.y),
                                             dpp_u32<0xB1>(own.z), dpp_u32<0xB1>(own.w));
                rl[2 * h] = r == 0 ? own : oth;
                rl[2 * h + 1] = r == 0 ? oth : own;
            }
        }
    }
    wave_sync();
    if (active) {
        const uint32_t L = J.L;
#pragma unroll
        for (uint32_t q = 0; q < kPer; ++q) {
            const uint32_t t = r + q * S::G;
            const uint32_t tt = base + t;
            if (t >= kSub || tt >= nm) continue;
            const uint32_t m = m0 + tt;
            const float* pt = pb + t * Lds::T;
            uint32_t ks = 0;
            if (argmax) {  // first maximum
                float best = -1.0f;
#pragma unroll
                for (int k4 = 0; k4 < S::Lp; k4 += 4) {
                    const float4 x = *reinterpret_cast<const float4*>(pt + k4);
                    const float xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const bool gt = xs[i] > best;
                        best = gt ? xs[i] : best;
                        ks = gt ? uint32_t(k4 + i) : ks;
                    }
                }
            } else {  // first k with u < cumsum, else 0 (padding states add 0)
                uint32_t rv;
                if constexpr (Lds::kRngLds) {
                    rv = (&rng[t >> 2].x)[t & 3];
                } else {  // block t >> 2, with t = q (one lane) or r + 2q (two lanes)
                    const uint4 rb = rl[S::G == 1 ? (q >> 2) : (q >> 1)];
                    const uint32_t c = t & 3u;
                    rv = c == 0u ? rb.x : c == 1u ? rb.y : c == 2u ? rb.z : rb.w;
                }
                // the cumulative sums never decrease (p >= 0), so the first k
                // with u < cs_k is the number of k with cs_k <= u
                const float uu = u24(rv);
                float cs = 0.0f;
                uint32_t below = 0;
#pragma unroll
                for (int k4 = 0; k4 < S::Lp; k4 += 4) {
                    const float4 x = *reinterpret_cast<const float4*>(pt + k4);
                    const float xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        cs += xs[i];
                        below += (uu < cs) ? 0u : 1u;
                    }
                }
                ks = below < uint32_t(S::Lp) ? below : 0u;
            }
            const size_t o = size_t(s) * J.M + m;
            if (J.ds) J.ds[o] = ap0r[q] + ap1r[q];
            if (J.ap) reinterpret_cast<float2*>(J.ap)[o] = make_float2(ap0r[q], ap1r[q]);
            if (J.hap) {
                // bits: alleles at 0-1, missing flags at 4-5 (the state byte's)
                uint32_t b;
                if (SB) {
                    const uint32_t nb =
                        (reinterpret_cast<const uint8_t*>(sb)[t * (S::Lp / 2) + (ks >> 1)] >> (4 * (ks & 1))) & 15u;
                    b = (nb & 3u) | ((nb & 12u) << 2);
                } else {
                    b = J.sa[size_t(m) * S::Lp + ks];
                }
                J.hap[size_t(s) * 2 * J.M + m] = uint8_t((b & 16u) ? 255u : (b & 1u));
                J.hap[size_t(s) * 2 * J.M + J.M + m] = uint8_t((b & 32u) ? 255u : ((b >> 1) & 1u));
            }
            if (J.prob)
                #pragma unroll 1
                for (uint32_t k = 0; k < L; ++k) J.prob[o * L + k] = pt[k];
        }
    }
    wave_sync();  // the group's LDS rows are reused by the next sub-block
}

// ---------------------------------------------------------------------------
// Checkpoint block c of sample s of job J, by one lane group.  For each
// 16-marker sub-block j (last first): replay the forward vector from the
// block checkpoint through sub-blocks 0..j-1 (registers only) and through j
// (into the group's LDS rows), walk the backward vector down j forming the
// posteriors (in place in LDS) and the allele probabilities (group sums),
// then sample: lane r takes markers r, r + G, ... of the sub-block and walks
// the states in order with a sequential FP32 cumulative sum, which the host
// replays exactly.  Every lane of the wavefront passes the wave_syncs.
// With windows (at least CB markers long), a window start inside the block
// restarts the forward vector from the window's last vector F_w (stored by
// the chain pass) and the backward step out of it is not taken.
// ---------------------------------------------------------------------------
template <class S, bool LOG, bool PM, int CB, bool SB>
__device__ __forceinline__ void run_block(const JobView& J, const Tab& tab, uint32_t P, bool active,
                                          uint32_t s, uint32_t c, uint32_t r, float* pb,
                                          uint32_t* sb, uint4* rng) {
    using Lds = BlockLds<S>;
    constexpr uint32_t NS = uint32_t(CB) / kSub;
    float* const prow = pb + r * S::R;  // the lane's R states of row 0
    // markers per lane in a sampling phase (with 32 lanes, half the lanes sample)
    constexpr uint32_t kPer = (kSub + uint32_t(S::G) - 1) / uint32_t(S::G);
    const uint32_t m0 = c * CB;
    const uint32_t nm = active ? min(uint32_t(CB), J.M - m0) : 0u;
    const uint32_t sub = lane_sub<S>(r);
    const uint8_t* codes = J.codes + size_t(s) * J.pitch;
    const size_t ck = size_t(s) * J.nblk + c;
    // offset in the block of a window start after its first marker (or none),
    // and the window's last forward vector, the restart there (in registers,
    // so the step loops issue no loads of their own)
    uint32_t tws = 0xffffffffu;
    float fwv[S::R];
#pragma unroll
    for (int i = 0; i < S::R; ++i) fwv[i] = 0.0f;
    if (J.W) {
        const uint32_t wi = m0 / J.W + 1u;
        if (wi * J.W < m0 + nm) {
            tws = wi * J.W - m0;
            load_vec<S::R>(J.ck_win + (size_t(s) * J.nwin + wi) * S::Lp + size_t(r) * S::R, fwv);
        }
    }
    float v[S::R];
    double Bl = 0.0;
    if (active) {
        load_vec<S::R>(J.ck_beta + ck * S::Lp + size_t(r) * S::R, v);
        Bl = J.ck_bl[ck];
    }
    BlockIn<S, LOG, PM> in;
    BlockEm<S, LOG, PM> em;
    for (int j = int(NS) - 1; j >= 0; --j) {
        const uint32_t base = uint32_t(j) * kSub;  // sub-block start within the block
        float ap0r[kPer], ap1r[kPer];
        if (active) {
            float a[S::R];
            load_vec<S::R>(J.ck_alpha + ck * S::Lp + size_t(r) * S::R, a);
            in.load(J, codes, r, m0);
            // replay through the sub-blocks before j
            for (uint32_t u = 0; u < uint32_t(j); ++u) {
                em.fill(tab, P, in, sub);
                in.load(J, codes, r, m0 + (u + 1) * kSub);
#pragma unroll
                for (uint32_t t = 0; t < kSub; ++t) {
                    const uint32_t tt = u * kSub + t;
                    float an[S::R];
#pragma unroll
                    for (int i = 0; i < S::R; ++i) an[i] = a[i];
                    fwd_step<S, LOG>(an, em.e[t], em.off[t], em.coef(t));
                    const bool live = tt > 0 && tt < nm;
                    vsel(live, a, an);
                    vsel(tt == tws, a, fwv);
                }
            }
            em.fill(tab, P, in, sub);
            // Sub-block j: forward vectors into LDS (state bytes beside them),
            // then the backward walk.  A block inside the job (nm == CB) needs
            // no per-step selects: only its very first step is not taken.
            auto walk = [&](auto full_c) {
                constexpr bool FULL = decltype(full_c)::value;
#pragma unroll
                for (uint32_t t = 0; t < kSub; ++t) {
                    if constexpr (!SB) {
                    } else if constexpr (S::R == 4) {
                        reinterpret_cast<uint16_t*>(sb)[t * (S::Lp / 4) + r] = uint16_t(nib4(in.sa[t].w[0]));
                    } else {
                        static_assert(S::R == 2, "block pass lane shapes");
                        reinterpret_cast<uint8_t*>(sb)[t * (S::Lp / 2) + r] =
                            uint8_t(nib2(in.sa[t].w[0] >> (8 * sub)));
                    }
                    const uint32_t tt = base + t;
                    if (FULL && t > 0) {
                        fwd_step<S, LOG>(a, em.e[t], em.off[t], em.coef(t));
                    } else if (!(FULL && NS == 1)) {  // with one sub-block, step 0 is never taken
                        float an[S::R];
#pragma unroll
                        for (int i = 0; i < S::R; ++i) an[i] = a[i];
                        fwd_step<S, LOG>(an, em.e[t], em.off[t], em.coef(t));
                        const bool live = tt > 0 && tt < nm;
                        vsel(live, a, an);
                        if constexpr (!FULL) {
                            vsel(tt == tws, a, fwv);
                        }
                    }
                    store_vec<S::R>(prow + t * Lds::T, a);
                }
#pragma unroll
                for (int t = int(kSub) - 1; t >= 0; --t) {
                    const uint32_t tt = base + uint32_t(t);
                    float af[S::R], p[S::R];
                    load_vec<S::R>(prow + t * Lds::T, af);
                    posterior<S, LOG>(af, v, p);
                    float x0[S::R], x1[S::R];
#pragma unroll
                    for (int i = 0; i < S::R; ++i) {
                        const uint32_t pi = __float_as_uint(p[i]);
                        x0[i] = __uint_as_float(pi & in.sa[t].bitmask(i, sub, 0));
                        x1[i] = __uint_as_float(pi & in.sa[t].bitmask(i, sub, 1));
                    }
                    float ap0 = tree_sum<S::R>(x0), ap1 = tree_sum<S::R>(x1);
                    gsum2<S::G>(ap0, ap1);
                    if (FULL || tt < nm) store_vec<S::R>(prow + t * Lds::T, p);
                    if (uint32_t(t) % S::G == r) {
                        ap0r[uint32_t(t) / S::G] = ap0;
                        ap1r[uint32_t(t) / S::G] = ap1;
                    }
                    if (FULL) {
                        if (t > 0 || base > 0) bwd_step<S, LOG>(v, Bl, em.e[t], em.off[t], em.coef(uint32_t(t)));
                    } else if (tt > 0) {
                        float vn[S::R];
#pragma unroll
                        for (int i = 0; i < S::R; ++i) vn[i] = v[i];
                        double Bln = Bl;
                        bwd_step<S, LOG>(vn, Bln, em.e[t], em.off[t], em.coef(uint32_t(t)));
                        const bool live = tt < nm && tt != tws;
                        vsel(live, v, vn);
                        Bl = live ? Bln : Bl;
                    }
                }
            };
            if (nm == uint32_t(CB) && tws == 0xffffffffu) walk(std::true_type{});
            else walk(std::false_type{});
        }
        block_sample<S, SB, kPer>(J, active, s, m0, base, nm, r, pb, sb, rng, ap0r, ap1r);
    }
}

// ---------------------------------------------------------------------------
// Plain form: one job; grid = lane groups.  Chain: forward groups first,
// padded to whole wavefronts (Bpad), then backward groups.
// ---------------------------------------------------------------------------
template <int G, int R, bool LOG, int CB>
__global__ __launch_bounds__(kFbThreads) void fb_chain_kernel(JobView J, FbTables tb, uint32_t Bpad) {
    using S = Shape<G, R>;
    const Tab tab = load_tables<LOG>(tb);
    __syncthreads();
    const uint32_t gtid = blockIdx.x * kFbThreads + threadIdx.x;
    const uint32_t grp = gtid / G, r = gtid % G;
    const bool fwd = grp < Bpad;
    const uint32_t s = fwd ? grp : grp - Bpad;
    if (s >= J.B) return;  // whole lane groups leave together
    if (J.flags & SI_FB_PER_MARKER) run_chain<S, LOG, true, CB>(J, tab, tb.P, s, fwd, r);
    else run_chain<S, LOG, false, CB>(J, tab, tb.P, s, fwd, r);
}

// Wavefronts per SIMD the block kernels are compiled for: with two states per
// lane the live state fits three wavefronts' registers.
template <int R>
constexpr int kBlockWaves = R <= 2 ? 3 : 2;

template <int G, int R, bool LOG, int CB>
__global__ __launch_bounds__(kFbThreads, kBlockWaves<R>) void fb_block_kernel(JobView J, FbTables tb) {
    using S = Shape<G, R>;
    using Lds = BlockLds<S>;
    __shared__ __align__(16) float pbuf[S::GPB * Lds::PS];
    constexpr bool SB = Lds::kSbPlain;
    __shared__ uint32_t sbuf[S::GPB * (SB ? Lds::SW : 1)];
    __shared__ uint4 rng[Lds::kRngLds ? S::GPB * (kSub / 4) : 1];
    const Tab tab = load_tables<LOG>(tb);
    __syncthreads();
    const uint32_t gtid = blockIdx.x * kFbThreads + threadIdx.x;
    const uint32_t grp = gtid / G, r = gtid % G, gl = threadIdx.x / G;
    const bool active = grp < J.B * J.nblk;
    const uint32_t s = active ? grp / J.nblk : 0u;
    const uint32_t c = active ? grp % J.nblk : 0u;
    float* pb = pbuf + gl * Lds::PS;
    uint32_t* sb = sbuf + (SB ? gl * Lds::SW : 0u);
    uint4* rg = rng + (Lds::kRngLds ? gl * (kSub / 4) : 0u);
    if (J.flags & SI_FB_PER_MARKER)
        run_block<S, LOG, true, CB, SB>(J, tab, tb.P, active, s, c, r, pb, sb, rg);
    else
        run_block<S, LOG, false, CB, SB>(J, tab, tb.P, active, s, c, r, pb, sb, rg);
}

// ---------------------------------------------------------------------------
// Batched form: a persistent grid; each wavefront repeatedly takes the next
// GPW items of the class's queue (one atomic per wavefront) until the queue
// is exhausted.  Items are ordered job by job (jobs longest first), so a
// wavefront's groups usually share a job.
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

// Next chunk of a queue for this wavefront (guided self-scheduling): about a
// quarter of the remaining items spread over the grid's wavefronts, at least
// one round (GPW items) and at most max_rounds, in whole rounds, so that the
// queue counter sees few atomics while the tail still balances.  Chains are
// long items, taken one round at a time; block items are short.
struct Chunk {
    uint64_t base;
    uint32_t n;
};
__device__ __forceinline__ Chunk wave_take(unsigned long long* counter, uint64_t total, uint32_t gpw,
                                           uint32_t max_rounds) {
    uint32_t lo = 0, hi = 0, n = 0;
    if (__lane_id() == 0) {
        const uint64_t seen = __hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t waves = uint64_t(gridDim.x) * (kFbThreads / kWave);
        const uint64_t rem = seen < total ? total - seen : 0;
        uint64_t want = rem / (4 * waves) / gpw * gpw;
        const uint64_t cap = uint64_t(max_rounds) * gpw;
        want = want < gpw ? gpw : (want > cap ? cap : want);
        n = uint32_t(want);
        const uint64_t b = atomicAdd(counter, static_cast<unsigned long long>(want));
        lo = uint32_t(b);
        hi = uint32_t(b >> 32);
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    hi = __builtin_amdgcn_readfirstlane(hi);
    n = __builtin_amdgcn_readfirstlane(n);
    return Chunk{(uint64_t(hi) << 32) | lo, n};
}

// The job a whole wavefront works on, in scalar registers.
__device__ __forceinline__ JobView uniform_job(const JobView* p) {
    static_assert(sizeof(JobView) % 4 == 0, "JobView words");
    JobView J;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(p);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&J);
#pragma unroll
    for (uint32_t i = 0; i < sizeof(JobView) / 4; ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
    return J;
}

__device__ __forceinline__ uint32_t find_job_uniform(const uint64_t* off, uint32_t n, uint64_t item) {
    uint32_t lo = 0, hi = n;  // off[lo] <= item < off[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint64_t o = off[mid];
        const uint64_t ou = (uint64_t(__builtin_amdgcn_readfirstlane(uint32_t(o >> 32))) << 32) |
                            __builtin_amdgcn_readfirstlane(uint32_t(o));
        if (ou <= item) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <int G, int R, bool LOG, int CB>
__global__ __launch_bounds__(kFbThreads) void fb_chain_batch_kernel(BatchClass bc, FbTables tb) {
    using S = Shape<G, R>;
    const Tab tab = load_tables<LOG>(tb);
    __syncthreads();
    const uint32_t lane = __lane_id(), gw = lane / G, r = lane % G;
    const uint32_t m_max = bc.n_jobs ? bc.jobs[0].M : 0u;  // jobs are sorted longest first
    // chain_off counts two items (forward, backward) per sample.  The queue
    // hands out rounds of GPW samples (the class's samples in job order) in
    // one direction: round 2i runs the forward chains of samples
    // [i*GPW, (i+1)*GPW), round 2i + 1 their backward chains, so every group
    // of a wavefront runs the same direction's code.
    const uint64_t n_samples = bc.chain_items >> 1;
    const uint64_t total = 2 * ((n_samples + S::GPW - 1) / S::GPW) * S::GPW;
    for (;;) {
        const Chunk ch = wave_take(&bc.counters[0], total, S::GPW, 1);
        if (ch.base >= total) break;
        // The pass lasts as long as its longest chains, which the queue hands
        // out first: their wavefronts issue ahead of the short chains sharing
        // their SIMD, which fill the remaining issue slots.
        const uint64_t first = (ch.base / S::GPW >> 1) * S::GPW;  // < n_samples
        if (2u * bc.jobs[find_job(bc.chain_off, bc.n_jobs, 2 * first)].M >= m_max)
            __builtin_amdgcn_s_setprio(3);
        else
            __builtin_amdgcn_s_setprio(0);
        for (uint32_t k = 0; k < ch.n; k += S::GPW) {
            const uint64_t round = (ch.base + k) / S::GPW;
            const uint64_t gs = (round >> 1) * S::GPW + gw;  // sample within the class
            const bool fwd = (round & 1) == 0;
            if (gs < n_samples) {
                const uint32_t j = find_job(bc.chain_off, bc.n_jobs, 2 * gs);
                const JobView J = bc.jobs[j];
                const uint32_t q = uint32_t(gs - (bc.chain_off[j] >> 1));
                if (q >= J.B) continue;  // a job's samples are padded to a multiple of 64
                if (J.flags & SI_FB_PER_MARKER)
                    run_chain<S, LOG, true, CB>(J, tab, tb.P, q, fwd, r);
                else
                    run_chain<S, LOG, false, CB>(J, tab, tb.P, q, fwd, r);
            }
        }
    }
}

// Block items of each job are padded to whole wavefront rounds (the block
// offsets are multiples of 64), so the groups of one round share a job and
// its descriptor is wavefront-uniform; padding items are inactive.
template <int G, int R, bool LOG, int CB>
__global__ __launch_bounds__(kFbThreads, kBlockWaves<R>) void fb_block_batch_kernel(BatchClass bc,
                                                                                    FbTables tb) {
    using S = Shape<G, R>;
    using Lds = BlockLds<S>;
    __shared__ __align__(16) float pbuf[S::GPB * Lds::PS];
    __shared__ uint32_t sbuf[S::GPB * Lds::SW];
    __shared__ uint4 rng[Lds::kRngLds ? S::GPB * (kSub / 4) : 1];
    const Tab tab = load_tables<LOG>(tb);
    __syncthreads();
    const uint32_t lane = __lane_id(), gw = lane / G, r = lane % G;
    const uint32_t gl = threadIdx.x / G;
    float* pb = pbuf + gl * Lds::PS;
    uint32_t* sb = sbuf + gl * Lds::SW;
    uint4* rg = rng + (Lds::kRngLds ? gl * (kSub / 4) : 0u);
    uint64_t lo = 0, hi = 0;  // item range of the job in J (wavefront-uniform)
    JobView J{};
    for (;;) {
        const Chunk ch = wave_take(&bc.counters[1], bc.block_items, S::GPW, 32);
        if (ch.base >= bc.block_items) break;
        for (uint32_t k = 0; k < ch.n; k += S::GPW) {
            const uint64_t round = ch.base + k;  // wavefront-uniform
            if (round >= bc.block_items) break;
            if (round < lo || round >= hi) {  // another job: look it up once
                const uint32_t j = find_job_uniform(bc.block_off, bc.n_jobs, round);
                J = uniform_job(bc.jobs + j);
                lo = bc.block_off[j];
                hi = bc.block_off[j + 1];
                lo = (uint64_t(__builtin_amdgcn_readfirstlane(uint32_t(lo >> 32))) << 32) |
                     __builtin_amdgcn_readfirstlane(uint32_t(lo));
                hi = (uint64_t(__builtin_amdgcn_readfirstlane(uint32_t(hi >> 32))) << 32) |
                     __builtin_amdgcn_readfirstlane(uint32_t(hi));
            }
            const uint64_t q = round - lo + gw;
            const bool active = q < uint64_t(J.B) * J.nblk;
            const uint32_t s = active ? uint32_t(q / J.nblk) : 0u;
            const uint32_t c = active ? uint32_t(q % J.nblk) : 0u;
            if (J.flags & SI_FB_PER_MARKER)
                run_block<S, LOG, true, CB, true>(J, tab, tb.P, active, s, c, r, pb, sb, rg);
            else
                run_block<S, LOG, false, CB, true>(J, tab, tb.P, active, s, c, r, pb, sb, rg);
        }
    }
}

// ===========================================================================
// Fast jobs: constant transitions, no windows, an untyped code.  The chain
// pass walks only the job's typed markers and its 16-aligned anchors; the
// block pass takes an untyped marker in O(1) per lane apart from the
// posterior itself.  Checkpoints are block-major ([block][sample][Lp]) and
// block-pass items run samples fastest, so the groups of a wavefront share a
// checkpoint block and every branch on the job's typed markers is
// wavefront-uniform.
// ===========================================================================

// The job's typed markers 1..M-1 in order, from its mask: one workgroup of
// kListThreads per job, a workgroup-wide scan of the blocks' popcounts.
constexpr uint32_t kListThreads = 256;
__device__ __forceinline__ void tlist_build(const uint32_t* tmask, uint32_t M, uint32_t* tlist,
                                            uint32_t* tcount) {
    __shared__ uint32_t sc[kListThreads];
    const uint32_t nlb = (M + kSub - 1) / kSub;
    uint32_t base = 0;
    for (uint32_t c0 = 0; c0 < nlb; c0 += kListThreads) {
        const uint32_t c = c0 + threadIdx.x;
        uint32_t bits = c < nlb ? tmask[c] : 0u;
        if (c == 0) bits &= ~1u;  // marker 0 starts the forward, it is not a step
        const uint32_t n = __popc(bits);
        sc[threadIdx.x] = n;
        __syncthreads();
        for (uint32_t d = 1; d < kListThreads; d <<= 1) {
            const uint32_t x = threadIdx.x >= d ? sc[threadIdx.x - d] : 0u;
            __syncthreads();
            sc[threadIdx.x] += x;
            __syncthreads();
        }
        uint32_t o = base + sc[threadIdx.x] - n;
        while (bits) {
            tlist[o++] = kSub * c + uint32_t(__ffs(bits)) - 1u;
            bits &= bits - 1u;
        }
        base += sc[kListThreads - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) *tcount = base;
}

// A typed marker's inputs for one lane: its index, the sample's code and the
// lane's state-table words.
template <class S>
struct MarkIn {
    uint32_t m, code;
    LaneWords<S> sw;
};
template <class S>
__device__ __forceinline__ void mark_load(const JobView& J, const uint8_t* codes, uint32_t r, uint32_t m,
                                          MarkIn<S>& x) {
    x.m = m;
    x.code = codes[m];
    const uint32_t* p =
        reinterpret_cast<const uint32_t*>(J.sa) + size_t(m) * (uint32_t(S::Lp) / 4u) + (r * uint32_t(S::R)) / 4u;
#pragma unroll
    for (int w = 0; w < S::NW; ++w) x.sw.w[w] = p[w];
}

template <bool LOG>
__device__ __forceinline__ typename CoefT<LOG>::type const_coef(const JobView& J) {
    if constexpr (LOG) return J.step;
    else return make_float2(J.step.x, J.step.y);
}

// One chain of a fast job.  The vector stays at an anchor marker; a
// closed-form run brings it up to each typed marker (whose step is exact)
// and to each 16-aligned marker (forward: the marker itself; backward: the
// marker before it), where checkpoints are stored.  Inputs of the next K
// typed markers are in flight.
template <class S, bool LOG, int CB>
__device__ __forceinline__ void run_chain_fast(const JobView& J, const Tab& tab, uint32_t P, uint32_t s,
                                               bool fwd, uint32_t r) {
    constexpr uint32_t K = 8;
    const uint8_t* codes = J.codes + size_t(s) * J.pitch;
    const uint32_t sub = lane_sub<S>(r);
    const size_t vo = size_t(r) * S::R;
    const URun& q = J.ur;
    float rm[S::R];
    real_mask<S>(r, J.L, rm);
    const typename CoefT<LOG>::type k = const_coef<LOG>(J);
    const uint32_t T = __builtin_amdgcn_readfirstlane(*J.tcount);
    const uint32_t* tl = J.tlist;
    const uint32_t last = J.M - 1u;
    // Typed marker i (in walking order) sits in slot i % K: ring holds the
    // inputs of markers i..i+K-1, pos the indices of markers i+K..i+2K-1,
    // both loaded by vector loads (wavefront-uniform addresses) so that no
    // marker's index or inputs are waited for when its step comes.
    MarkIn<S> ring[K];
    uint32_t pos[K];
    auto ti = [&](uint32_t i) { return fwd ? i : T - 1u - i; };  // walking order -> list index
#pragma unroll
    for (uint32_t j = 0; j < K; ++j) {
        if (j < T) mark_load<S>(J, codes, r, tl[ti(j)], ring[j]);
        pos[j] = (K + j < T) ? tl[ti(K + j)] : 0u;
    }
    if (fwd) {
        float a[S::R];
        double ll;
        {
            MarkIn<S> x0;
            mark_load<S>(J, codes, r, 0, x0);
            float e[S::R];
            const float off = lane_emission<S>(tab, P, x0.code, x0.sw, sub, e);
            ll = fwd_init<S, LOG>(a, e, off);
        }
        store_vec<S::R>(J.ck_alpha + size_t(s) * J.nblk * S::Lp + vo, a);  // block 0
        uint32_t pa = 0, q16 = kSub;
        auto anchors_below = [&](uint32_t lim) {
            while (q16 < lim) {
                ll += fwd_run<S, LOG>(a, rm, q16 - pa, q);
                pa = q16;
                if (q16 % uint32_t(CB) == 0u)
                    store_vec<S::R>(J.ck_alpha + (size_t(s) * J.nblk + q16 / CB) * S::Lp + vo, a);
                q16 += kSub;
            }
        };
        for (uint32_t i0 = 0; i0 < T; i0 += K) {
#pragma unroll
            for (uint32_t j = 0; j < K; ++j) {
                const uint32_t i = i0 + j;
                if (i < T) {
                    const uint32_t mt = __builtin_amdgcn_readfirstlane(ring[j].m);
                    anchors_below(mt);
                    ll += fwd_run<S, LOG>(a, rm, mt - 1u - pa, q);
                    float e[S::R];
                    const float off = lane_emission<S>(tab, P, ring[j].code, ring[j].sw, sub, e);
                    ll += fwd_step<S, LOG>(a, e, off, k);
                    pa = mt;
                    if (i + K < T) mark_load<S>(J, codes, r, pos[j], ring[j]);
                    if (i + 2u * K < T) pos[j] = tl[i + 2u * K];
                }
            }
        }
        anchors_below(J.M);
        ll += fwd_run<S, LOG>(a, rm, last - pa, q);
        if (LOG) ll += double(log_mass<S>(a));  // log mass of the last normalized vector
        if (r == 0 && J.loglik) J.loglik[s] = ll;
    } else {
        float v[S::R];
        bwd_init<S, LOG>(v);
        double Bl = 0.0;
        {
            const size_t b = size_t(s) * J.nblk + (J.nblk - 1u);
            store_vec<S::R>(J.ck_beta + b * S::Lp + vo, v);
            if (r == 0) J.ck_bl[b] = Bl;
        }
        uint32_t pb = last, q16 = last / kSub * kSub;
        // anchors at q16 - 1 (after the step out of q16) for 16-aligned q16 > lim
        auto anchors_above = [&](uint32_t lim) {
            while (q16 > lim && q16 >= kSub) {
                bwd_run<S, LOG>(v, Bl, rm, pb - (q16 - 1u), q);
                pb = q16 - 1u;
                if (q16 % uint32_t(CB) == 0u) {
                    const size_t b = size_t(s) * J.nblk + (q16 / CB - 1u);
                    store_vec<S::R>(J.ck_beta + b * S::Lp + vo, v);
                    if (r == 0) J.ck_bl[b] = Bl;
                }
                q16 -= kSub;
            }
        };
        for (uint32_t i0 = 0; i0 < T; i0 += K) {
#pragma unroll
            for (uint32_t j = 0; j < K; ++j) {
                const uint32_t i = i0 + j;
                if (i < T) {
                    const uint32_t n = __builtin_amdgcn_readfirstlane(ring[j].m);
                    anchors_above(n);
                    bwd_run<S, LOG>(v, Bl, rm, pb - n, q);
                    float e[S::R];
                    const float off = lane_emission<S>(tab, P, ring[j].code, ring[j].sw, sub, e);
                    bwd_step<S, LOG>(v, Bl, e, off, k);
                    pb = n - 1u;
                    if (i + K < T) mark_load<S>(J, codes, r, pos[j], ring[j]);
                    if (i + 2u * K < T) pos[j] = tl[T - 1u - (i + 2u * K)];
                }
            }
        }
        anchors_above(0u);
    }
}

template <int G, int R, bool LOG, int CB>
__global__ __launch_bounds__(kFbThreads) void fb_chain_fast_kernel(JobView J, FbTables tb, uint32_t Bpad) {
    using S = Shape<G, R>;
    const Tab tab = load_tables<LOG>(tb);
    __syncthreads();
    const uint32_t gtid = blockIdx.x * kFbThreads + threadIdx.x;
    const uint32_t grp = gtid / G, r = gtid % G;
    const bool fwd = grp < Bpad;
    const uint32_t s = fwd ? grp : grp - Bpad;
    if (s >= J.B) return;  // whole lane groups leave together
    run_chain_fast<S, LOG, CB>(J, tab, tb.P, s, fwd, r);
}

template <int G, int R, bool LOG, int CB>
__global__ __launch_bounds__(kFbThreads) void fb_chain_fast_batch_kernel(BatchClass bc, FbTables tb) {
    using S = Shape<G, R>;
    const Tab tab = load_tables<LOG>(tb);
    __syncthreads();
    const uint32_t lane = __lane_id(), gw = lane / G, r = lane % G;
    const uint32_t m_max = bc.n_jobs ? bc.jobs[0].M : 0u;  // jobs are sorted longest first
    // as fb_chain_batch_kernel; a job's samples are padded to a multiple of
    // 64, so a round of GPW groups never spans two jobs
    const uint64_t n_samples = bc.chain_items >> 1;
    const uint64_t total = 2 * n_samples;
    for (;;) {
        const Chunk ch = wave_take(&bc.counters[0], total, S::GPW, 1);
        if (ch.base >= total) break;
        const uint64_t first = (ch.base / S::GPW >> 1) * S::GPW;
        const uint32_t j0 = find_job_uniform(bc.chain_off, bc.n_jobs, 2 * first);
        if (2u * bc.jobs[j0].M >= m_max)
            __builtin_amdgcn_s_setprio(3);
        else
            __builtin_amdgcn_s_setprio(0);
        for (uint32_t kk = 0; kk < ch.n; kk += S::GPW) {
            const uint64_t round = (ch.base + kk) / S::GPW;
            const uint64_t gs0 = (round >> 1) * S::GPW;
            const bool fwd = (round & 1) == 0;
            if (gs0 >= n_samples) continue;
            const uint32_t j = find_job_uniform(bc.chain_off, bc.n_jobs, 2 * gs0);
            const JobView J = uniform_job(bc.jobs + j);
            const uint32_t q = uint32_t(gs0 - (bc.chain_off[j] >> 1)) + gw;
            if (q < J.B) run_chain_fast<S, LOG, CB>(J, tab, tb.P, q, fwd, r);
        }
    }
}

// ---------------------------------------------------------------------------
// Banded log space (jobs with SI_FB_BANDED: log-space steps whose switch
// log-probability falls to the -999 floor, as at coincident markers, where
// the posteriors are set by values e^-999 apart).  A value is carried as the
// pair (r, n) = r - 999 n: an FP32 residual r and a small integer n per state
// (a byte), so values resting on the floor, or on a multiple of it, keep FP32
// precision relative to it instead of its 6.1e-5 spacing at |x| ~ 999.  A
// log-sum-exp takes its terms by value: the largest sets the band of the
// result and the others enter as e^d, d = (r_t - r_m) - 999 (n_t - n_m),
// which FP32 gives exactly for terms of one band (the ones that matter near
// a floor).  Vectors are normalized to a pivot within 0.5 of their largest
// value; the -999 floor is the term (0, 1).  Log-likelihood increments and
// the backward scale stay in FP64.  Padding states carry band kPadBand and
// residual -inf.  Reference semantics: the same quantities as the log form
// above (forward_backward.cu:41-317).  No windows (SI_FB_BANDED is ignored
// for windowed jobs), 16-marker checkpoint blocks.
// ---------------------------------------------------------------------------
constexpr uint32_t kPadBand = 255u;
constexpr uint32_t kMaxBand = 254u;
constexpr float kBandStep = 999.0f;

template <class S>
struct Bands {
    static constexpr int NW = (S::R + 3) / 4;
    uint32_t w[NW];
    __device__ __forceinline__ uint32_t get(int i) const { return (w[i >> 2] >> (8 * (i & 3))) & 0xffu; }
    __device__ __forceinline__ void set(int i, uint32_t n) {
        const uint32_t sh = 8u * uint32_t(i & 3);
        w[i >> 2] = (w[i >> 2] & ~(0xffu << sh)) | (n << sh);
    }
    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int k = 0; k < NW; ++k) w[k] = 0u;
    }
};

struct BT {
    float r;
    uint32_t n;
};
// Approximate value, for comparisons only.
__device__ __forceinline__ float bval(const BT& t) { return t.r - kBandStep * float(t.n); }
// Band form of a log value given in FP32 (x + 999 n is exact for |x| < 2048).
__device__ __forceinline__ BT band_f(float x) {
    const uint32_t n = x <= -499.5f ? (x <= -1498.5f ? 2u : 1u) : 0u;
    return BT{x + kBandStep * float(n), n};
}
// Band form of a log value given in FP64 (one rounding, of the residual).
__host__ __device__ inline BT band_d(double x) {
    double q = floor((499.5 - x) / 999.0);
    q = q < 0.0 ? 0.0 : (q > 254.0 ? 254.0 : q);
    return BT{float(x + 999.0 * q), uint32_t(q)};
}
// LSE of three band terms (absent terms: residual -inf).
__device__ __forceinline__ BT lse3b(const BT& a, const BT& b, const BT& c) {
    BT m = a;
    float vm = bval(a);
    const float vb = bval(b), vc = bval(c);
    if (vb > vm) {
        m = b;
        vm = vb;
    }
    if (vc > vm) m = c;
    if (!(m.r > -INFINITY)) return BT{-INFINITY, 0u};
    const float mn = float(m.n);
    const float s = fast_exp((a.r - m.r) - kBandStep * (float(a.n) - mn)) +
                    fast_exp((b.r - m.r) - kBandStep * (float(b.n) - mn)) +
                    fast_exp((c.r - m.r) - kBandStep * (float(c.n) - mn));
    return BT{m.r + fast_ln(s), m.n};
}

// Pivot of a band vector over the lane group: among the states within 0.5 of
// the largest value, the smallest band and, in it, the largest residual.
template <class S>
__device__ __forceinline__ BT band_pivot(const float (&x)[S::R], const Bands<S>& xb) {
    float v[S::R];
#pragma unroll
    for (int i = 0; i < S::R; ++i) v[i] = x[i] - kBandStep * float(xb.get(i));
    const float vm = vmax<S>(v);
    float c[S::R];
#pragma unroll
    for (int i = 0; i < S::R; ++i) c[i] = v[i] >= vm - 0.5f ? -float(xb.get(i)) : -1.0e9f;
    const float nb = -vmax<S>(c);
    float rr[S::R];
#pragma unroll
    for (int i = 0; i < S::R; ++i) rr[i] = (v[i] >= vm - 0.5f && float(xb.get(i)) == nb) ? x[i] : -INFINITY;
    return BT{vmax<S>(rr), uint32_t(nb)};
}

// x relative to the pivot p: band n - p.n; a state below the pivot's band
// (a residual far under its band's floor) folds into band 0.
template <class S>
__device__ __forceinline__ void band_shift(float (&x)[S::R], Bands<S>& xb, const BT& p) {
#pragma unroll
    for (int i = 0; i < S::R; ++i) {
        const uint32_t n = xb.get(i);
        if (n == kPadBand) continue;
        const float d = x[i] - p.r;
        if (n >= p.n) {
            x[i] = d;
            xb.set(i, n - p.n);
        } else {
            x[i] = d + kBandStep * float(p.n - n);
            xb.set(i, 0u);
        }
    }
}

template <class S>
__device__ __forceinline__ double band_normalize(float (&x)[S::R], Bands<S>& xb) {
    const BT p = band_pivot<S>(x, xb);
    band_shift<S>(x, xb, p);
    return double(p.r) - 999.0 * double(p.n);
}

// Emission table for the banded kernels (dynamic LDS): per (code, class
// mask) row, the band residuals of the four classes (gl_c - off, exact in
// FP64 before one rounding), the row's offset and the classes' bands in the
// bytes of a word (class 3, the padding state, and absent classes:
// kPadBand).
constexpr uint32_t kTabRowBand = 8;
__host__ __device__ constexpr size_t fb_table_lds_bytes_band(uint32_t P) {
    return size_t(P) * 8 * kTabRowBand * sizeof(float);
}
__device__ __forceinline__ const float* load_tables_band(const FbTables& tb) {
    extern __shared__ float4 fb_dyn_lds[];
    float* f = reinterpret_cast<float*>(fb_dyn_lds);
    const uint32_t n = tb.P * 8;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        float* row = f + i * kTabRowBand;
        const float o = tb.emax[i];
        const float* g = tb.palette + 3 * (i / 8);
        const uint32_t cm = i % 8;
        uint32_t bb = kPadBand << 24;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (cm & (1u << c)) {
                const BT t = band_d(double(g[c]) - double(o));
                row[c] = t.r;
                bb |= t.n << (8 * c);
            } else {
                row[c] = -INFINITY;
                bb |= kPadBand << (8 * c);
            }
        }
        row[3] = -INFINITY;
        row[4] = o;
        row[5] = __uint_as_float(bb);
        row[6] = 0.0f;
        row[7] = 0.0f;
    }
    return f;
}

template <class S>
__device__ __forceinline__ float lane_emission_band(const float* tab, uint32_t P, uint32_t code,
                                                    const LaneWords<S>& sw, uint32_t sub, float (&e)[S::R],
                                                    Bands<S>& eb) {
    const float* rp = tab + (min(code, P - 1u) * 8u + class_mask(sw.w[0])) * kTabRowBand;
    const uint32_t bb = __float_as_uint(rp[5]);
#pragma unroll
    for (int i = 0; i < S::R; ++i) {
        const uint32_t c = sw.cls(i, sub);
        e[i] = rp[c];
        eb.set(i, (bb >> (8u * c)) & 0xffu);
    }
    return rp[4];
}

// Marker 0: a = e, normalized.  Returns the log-likelihood increment.
template <class S>
__device__ __forceinline__ double fwd_init_band(float (&a)[S::R], Bands<S>& ab, const float (&e)[S::R],
                                                const Bands<S>& eb, float off) {
#pragma unroll
    for (int i = 0; i < S::R; ++i) a[i] = e[i];
#pragma unroll
    for (int k = 0; k < Bands<S>::NW; ++k) ab.w[k] = eb.w[k];
    return double(off) + band_normalize<S>(a, ab);
}

// ln sum_i e^{a_i} of a normalized band vector (states below band 0 add
// e^-999 or less, zero in FP32 as in FP64).
template <class S>
__device__ __forceinline__ float log_mass_band(const float (&a)[S::R], const Bands<S>& ab) {
    float w[S::R];
#pragma unroll
    for (int i = 0; i < S::R; ++i) w[i] = fast_exp(a[i] - kBandStep * float(ab.get(i)));
    return fast_ln(vsum<S>(w));
}

// Step into a marker: a_j <- LSE(-999, To + LSE(a), U + a_j) + e_j, with
// k = (t_o, u, To, U).
template <class S>
__device__ __forceinline__ double fwd_step_band(float (&a)[S::R], Bands<S>& ab, const float (&e)[S::R],
                                                const Bands<S>& eb, float off, float4 k) {
    const float lm = log_mass_band<S>(a, ab);
    const BT tfl{0.0f, 1u};
    BT tto = band_f(k.z);
    tto.r += lm;
    const BT tu = band_f(k.w);
    float x[S::R];
    Bands<S> xb;
    xb.clear();
#pragma unroll
    for (int i = 0; i < S::R; ++i) {
        const uint32_t n = ab.get(i), ne = eb.get(i);
        const bool pad = n == kPadBand || ne == kPadBand;
        const BT u{tu.r + a[i], min(tu.n + n, kMaxBand)};
        const BT t = lse3b(tfl, tto, pad ? BT{-INFINITY, 0u} : u);
        x[i] = pad ? -INFINITY : t.r + e[i];
        xb.set(i, pad ? kPadBand : min(t.n + ne, kMaxBand));
    }
#pragma unroll
    for (int i = 0; i < S::R; ++i) a[i] = x[i];
#pragma unroll
    for (int w = 0; w < Bands<S>::NW; ++w) ab.w[w] = xb.w[w];
    return double(off) + band_normalize<S>(a, ab);
}

// From the vector at marker n (emission e, offset off) to n - 1: the true
// vector is Bl + (v, vb); b_i <- LSE(-999, To + nat, U + Y_i) with
// Y_i = Bl + off + e_i + v_i and nat = LSE_i Y_i, and the new scale
// Bl' = max(nat, -999) (the floor pins it, as in the log form).
template <class S>
__device__ __forceinline__ void bwd_step_band(float (&v)[S::R], Bands<S>& vb, double& Bl, const float (&e)[S::R],
                                              const Bands<S>& eb, float off, float4 k) {
    float y[S::R];
    Bands<S> yb;
    yb.clear();
#pragma unroll
    for (int i = 0; i < S::R; ++i) {
        const uint32_t n = vb.get(i), ne = eb.get(i);
        const bool pad = n == kPadBand || ne == kPadBand;
        y[i] = pad ? -INFINITY : v[i] + e[i];
        yb.set(i, pad ? kPadBand : min(n + ne, kMaxBand));
    }
    const BT p = band_pivot<S>(y, yb);
    float w[S::R];
#pragma unroll
    for (int i = 0; i < S::R; ++i) {
        const uint32_t n = yb.get(i);
        w[i] = n == kPadBand ? 0.0f : fast_exp((y[i] - p.r) - kBandStep * (float(n) - float(p.n)));
    }
    const float lw = p.r + fast_ln(vsum<S>(w));  // LSE_i y_i = lw - 999 p.n
    const double nat = Bl + double(off) + double(lw) - 999.0 * double(p.n);
    const double Bn = nat > -999.0 ? nat : -999.0;
    const double sh = nat - Bn;  // 0, or below 0 where the floor dominates
    const BT tfl = band_d(-999.0 - Bn);
    const BT tto = band_d(double(k.z) + sh);
    const BT tus = band_d(double(k.w) + sh);
#pragma unroll
    for (int i = 0; i < S::R; ++i) {
        const uint32_t n = yb.get(i);
        const bool pad = n == kPadBand;
        // Y_i - Bl' = sh + (y_i - LSE y), as a band term
        float d = y[i] - lw;
        uint32_t nd = 0u;
        if (n >= p.n) nd = n - p.n;
        else d += kBandStep * float(p.n - n);
        const BT u{tus.r + d, min(tus.n + nd, kMaxBand)};
        const BT t = lse3b(tfl, tto, pad ? BT{-INFINITY, 0u} : u);
        v[i] = pad ? -INFINITY : t.r;
        vb.set(i, pad ? kPadBand : t.n);
    }
    Bl = Bn;
}

template <class S>
__device__ __forceinline__ void posterior_band(const float (&a)[S::R], const Bands<S>& ab, const float (&v)[S::R],
                                               const Bands<S>& vb, float (&p)[S::R]) {
    float q[S::R];
    Bands<S> qb;
    qb.clear();
#pragma unroll
    for (int i = 0; i < S::R; ++i) {
        const uint32_t na = ab.get(i), nv = vb.get(i);
        const bool pad = na == kPadBand || nv == kPadBand;
        q[i] = pad ? -INFINITY : a[i] + v[i];
        qb.set(i, pad ? kPadBand : min(na + nv, kMaxBand));
    }
    const BT pv = band_pivot<S>(q, qb);
#pragma unroll
    for (int i = 0; i < S::R; ++i) {
        const uint32_t n = qb.get(i);
        p[i] = n == kPadBand ? 0.0f : fast_exp((q[i] - pv.r) - kBandStep * (float(n) - float(pv.n)));
    }
    const float rq = __builtin_amdgcn_rcpf(vsum<S>(p));
#pragma unroll
    for (int i = 0; i < S::R; ++i) p[i] *= rq;
}

// A lane's band bytes of a vector in memory (R consecutive bytes).
template <class S>
__device__ __forceinline__ void store_bands(uint8_t* p, const Bands<S>& b) {
    if constexpr (S::R == 2) {
        *reinterpret_cast<uint16_t*>(p) = uint16_t(b.w[0]);
    } else {
#pragma unroll
        for (int k = 0; k < Bands<S>::NW; ++k) reinterpret_cast<uint32_t*>(p)[k] = b.w[k];
    }
}
template <class S>
__device__ __forceinline__ void load_bands(const uint8_t* p, Bands<S>& b) {
    if constexpr (S::R == 2) {
        b.w[0] = *reinterpret_cast<const uint16_t*>(p);
    } else {
#pragma unroll
        for (int k = 0; k < Bands<S>::NW; ++k) b.w[k] = reinterpret_cast<const uint32_t*>(p)[k];
    }
}

// Chain pass of a banded job (no windows; checkpoints at every 16-marker
// block), one lane group per (sample, direction), as run_chain.
template <class S, bool PM>
__device__ __forceinline__ void run_chain_band(const JobView& J, const float* tab, uint32_t P, uint32_t s,
                                               bool fwd, uint32_t r) {
    const uint8_t* codes = J.codes + size_t(s) * J.pitch;
    const uint32_t nlb = (J.M + kSub - 1) / kSub;
    const uint32_t sub = lane_sub<S>(r);
    const size_t vo = size_t(r) * S::R;
    BlockIn<S, true, PM> cur, nx;
    cur.load(J, codes, r, 0);
    if (fwd) {
        float a[S::R];
        Bands<S> ab;
        double ll;
        {
            float e[S::R];
            Bands<S> eb;
            const float off = lane_emission_band<S>(tab, P, cur.code(0), cur.sa[0], sub, e, eb);
            ll = fwd_init_band<S>(a, ab, e, eb, off);
        }
        for (uint32_t c = 0; c < nlb; ++c) {
            if (c + 1 < nlb) nx.load(J, codes, r, (c + 1) * kSub);
            const uint32_t m0 = c * kSub;
#pragma unroll
            for (uint32_t t = 0; t < kSub; ++t) {
                const uint32_t m = m0 + t;
                if (m >= 1 && m < J.M) {
                    float e[S::R];
                    Bands<S> eb;
                    const float off = lane_emission_band<S>(tab, P, cur.code(t), cur.sa[t], sub, e, eb);
                    ll += fwd_step_band<S>(a, ab, e, eb, off, cur.k[PM ? t : 0]);
                }
                if (t == 0) {
                    const size_t ck = size_t(s) * J.nblk + c;
                    store_vec<S::R>(J.ck_alpha + ck * S::Lp + vo, a);
                    store_bands<S>(J.ck_aband + ck * S::Lp + vo, ab);
                }
            }
            cur = nx;
        }
        ll += double(log_mass_band<S>(a, ab));
        if (r == 0 && J.loglik) J.loglik[s] = ll;
    } else {
        float v[S::R];
        Bands<S> vb;
#pragma unroll
        for (int i = 0; i < S::R; ++i) v[i] = 0.0f;
        vb.clear();
        double Bl = 0.0;
        const size_t last = size_t(s) * J.nblk + (J.nblk - 1);
        store_vec<S::R>(J.ck_beta + last * S::Lp + vo, v);
        store_bands<S>(J.ck_bband + last * S::Lp + vo, vb);
        if (r == 0) J.ck_bl[last] = Bl;
        cur.load(J, codes, r, (nlb - 1) * kSub);
        for (int c = int(nlb) - 1; c >= 0; --c) {
            if (c > 0) nx.load(J, codes, r, uint32_t(c - 1) * kSub);
            const uint32_t m0 = uint32_t(c) * kSub;
#pragma unroll
            for (int t = int(kSub) - 1; t >= 0; --t) {
                const uint32_t n = m0 + uint32_t(t);
                if (n >= 1 && n < J.M) {
                    float e[S::R];
                    Bands<S> eb;
                    const float off = lane_emission_band<S>(tab, P, cur.code(uint32_t(t)), cur.sa[t], sub, e, eb);
                    bwd_step_band<S>(v, vb, Bl, e, eb, off, cur.k[PM ? t : 0]);
                }
                if (t == 0 && c > 0) {  // v at m0 - 1 closes block c - 1
                    const size_t b = size_t(s) * J.nblk + uint32_t(c) - 1u;
                    store_vec<S::R>(J.ck_beta + b * S::Lp + vo, v);
                    store_bands<S>(J.ck_bband + b * S::Lp + vo, vb);
                    if (r == 0) J.ck_bl[b] = Bl;
                }
            }
            cur = nx;
        }
    }
}

// Block pass of a banded job: one lane group per (sample, 16-marker block),
// as run_block with one sub-block; the forward rows' bands sit beside them in
// LDS (bb: 16 rows of Lp bytes per group).
template <class S, bool PM, bool SB>
__device__ __forceinline__ void run_block_band(const JobView& J, const float* tab, uint32_t P, bool active,
                                               uint32_t s, uint32_t c, uint32_t r, float* pb, uint32_t* sb,
                                               uint8_t* bb, uint4* rng) {
    using Lds = BlockLds<S>;
    static_assert(S::R <= 4, "block pass lane shapes");
    float* const prow = pb + r * S::R;
    uint8_t* const brow = bb + r * S::R;
    constexpr uint32_t kPer = (kSub + uint32_t(S::G) - 1) / uint32_t(S::G);
    const uint32_t m0 = c * kSub;
    const uint32_t nm = active ? min(kSub, J.M - m0) : 0u;
    const uint32_t sub = lane_sub<S>(r);
    const uint8_t* codes = J.codes + size_t(s) * J.pitch;
    const size_t ck = size_t(s) * J.nblk + c;
    float ap0r[kPer], ap1r[kPer];
    if (active) {
        float a[S::R], v[S::R];
        Bands<S> ab, vb;
        load_vec<S::R>(J.ck_alpha + ck * S::Lp + size_t(r) * S::R, a);
        load_bands<S>(J.ck_aband + ck * S::Lp + size_t(r) * S::R, ab);
        load_vec<S::R>(J.ck_beta + ck * S::Lp + size_t(r) * S::R, v);
        load_bands<S>(J.ck_bband + ck * S::Lp + size_t(r) * S::R, vb);
        double Bl = J.ck_bl[ck];
        BlockIn<S, true, PM> in;
        in.load(J, codes, r, m0);
#pragma unroll
        for (uint32_t t = 0; t < kSub; ++t) {
            if constexpr (!SB) {
            } else if constexpr (S::R == 4) {
                reinterpret_cast<uint16_t*>(sb)[t * (S::Lp / 4) + r] = uint16_t(nib4(in.sa[t].w[0]));
            } else {
                reinterpret_cast<uint8_t*>(sb)[t * (S::Lp / 2) + r] = uint8_t(nib2(in.sa[t].w[0] >> (8 * sub)));
            }
            if (t > 0 && t < nm) {
                float e[S::R];
                Bands<S> eb;
                const float off = lane_emission_band<S>(tab, P, in.code(t), in.sa[t], sub, e, eb);
                fwd_step_band<S>(a, ab, e, eb, off, in.k[PM ? t : 0]);
            }
            store_vec<S::R>(prow + t * Lds::T, a);
            store_bands<S>(brow + t * S::Lp, ab);
        }
#pragma unroll
        for (int t = int(kSub) - 1; t >= 0; --t) {
            float af[S::R], p[S::R];
            Bands<S> afb;
            load_vec<S::R>(prow + t * Lds::T, af);
            load_bands<S>(brow + t * S::Lp, afb);
            posterior_band<S>(af, afb, v, vb, p);
            float x0[S::R], x1[S::R];
#pragma unroll
            for (int i = 0; i < S::R; ++i) {
                const uint32_t pi = __float_as_uint(p[i]);
                x0[i] = __uint_as_float(pi & in.sa[t].bitmask(i, sub, 0));
                x1[i] = __uint_as_float(pi & in.sa[t].bitmask(i, sub, 1));
            }
            float ap0 = tree_sum<S::R>(x0), ap1 = tree_sum<S::R>(x1);
            gsum2<S::G>(ap0, ap1);
            if (uint32_t(t) < nm) store_vec<S::R>(prow + t * Lds::T, p);
            if (uint32_t(t) % S::G == r) {
                ap0r[uint32_t(t) / S::G] = ap0;
                ap1r[uint32_t(t) / S::G] = ap1;
            }
            if (t > 0 && uint32_t(t) < nm) {
                float e[S::R];
                Bands<S> eb;
                const float off = lane_emission_band<S>(tab, P, in.code(uint32_t(t)), in.sa[t], sub, e, eb);
                bwd_step_band<S>(v, vb, Bl, e, eb, off, in.k[PM ? t : 0]);
            }
        }
    }
    block_sample<S, SB, kPer>(J, active, s, m0, 0u, nm, r, pb, sb, rng, ap0r, ap1r);
}

template <int G, int R>
__global__ __launch_bounds__(kFbThreads) void fb_chain_band_kernel(JobView J, FbTables tb, uint32_t Bpad) {
    using S = Shape<G, R>;
    const float* tab = load_tables_band(tb);
    __syncthreads();
    const uint32_t gtid = blockIdx.x * kFbThreads + threadIdx.x;
    const uint32_t grp = gtid / G, r = gtid % G;
    const bool fwd = grp < Bpad;
    const uint32_t s = fwd ? grp : grp - Bpad;
    if (s >= J.B) return;  // whole lane groups leave together
    if (J.flags & SI_FB_PER_MARKER) run_chain_band<S, true>(J, tab, tb.P, s, fwd, r);
    else run_chain_band<S, false>(J, tab, tb.P, s, fwd, r);
}

template <int G, int R>
__global__ __launch_bounds__(kFbThreads) void fb_block_band_kernel(JobView J, FbTables tb) {
    using S = Shape<G, R>;
    using Lds = BlockLds<S>;
    __shared__ __align__(16) float pbuf[S::GPB * Lds::PS];
    constexpr bool SB = Lds::kSbPlain;
    __shared__ uint32_t sbuf[S::GPB * (SB ? Lds::SW : 1)];
    __shared__ uint4 rng[Lds::kRngLds ? S::GPB * (kSub / 4) : 1];
    __shared__ __align__(16) uint8_t bbuf[S::GPB * kSub * S::Lp];
    const float* tab = load_tables_band(tb);
    __syncthreads();
    const uint32_t gtid = blockIdx.x * kFbThreads + threadIdx.x;
    const uint32_t grp = gtid / G, r = gtid % G, gl = threadIdx.x / G;
    const bool active = grp < J.B * J.nblk;
    const uint32_t s = active ? grp / J.nblk : 0u;
    const uint32_t c = active ? grp % J.nblk : 0u;
    float* pb = pbuf + gl * Lds::PS;
    uint32_t* sb = sbuf + (SB ? gl * Lds::SW : 0u);
    uint8_t* bb = bbuf + gl * kSub * S::Lp;
    uint4* rg = rng + (Lds::kRngLds ? gl * (kSub / 4) : 0u);
    if (J.flags & SI_FB_PER_MARKER) run_block_band<S, true, SB>(J, tab, tb.P, active, s, c, r, pb, sb, bb, rg);
    else run_block_band<S, false, SB>(J, tab, tb.P, active, s, c, r, pb, sb, bb, rg);
}

template <int G, int R>
__global__ __launch_bounds__(kFbThreads) void fb_chain_band_batch_kernel(BatchClass bc, FbTables tb) {
    using S = Shape<G, R>;
    const float* tab = load_tables_band(tb);
    __syncthreads();
    const uint32_t lane = __lane_id(), gw = lane / G, r = lane % G;
    // rounds of GPW samples in one direction, as fb_chain_batch_kernel
    const uint64_t n_samples = bc.chain_items >> 1;
    const uint64_t total = 2 * ((n_samples + S::GPW - 1) / S::GPW) * S::GPW;
    for (;;) {
        const Chunk ch = wave_take(&bc.counters[0], total, S::GPW, 1);
        if (ch.base >= total) break;
        for (uint32_t k = 0; k < ch.n; k += S::GPW) {
            const uint64_t round = (ch.base + k) / S::GPW;
            const uint64_t gs = (round >> 1) * S::GPW + gw;
            const bool fwd = (round & 1) == 0;
            if (gs < n_samples) {
                const uint32_t j = find_job(bc.chain_off, bc.n_jobs, 2 * gs);
                const JobView J = bc.jobs[j];
                const uint32_t q = uint32_t(gs - (bc.chain_off[j] >> 1));
                if (q >= J.B) continue;
                if (J.flags & SI_FB_PER_MARKER) run_chain_band<S, true>(J, tab, tb.P, q, fwd, r);
                else run_chain_band<S, false>(J, tab, tb.P, q, fwd, r);
            }
        }
    }
}

template <int G, int R>
__global__ __launch_bounds__(kFbThreads) void fb_block_band_batch_kernel(BatchClass bc, FbTables tb) {
    using S = Shape<G, R>;
    using Lds = BlockLds<S>;
    __shared__ __align__(16) float pbuf[S::GPB * Lds::PS];
    __shared__ uint32_t sbuf[S::GPB * Lds::SW];
    __shared__ uint4 rng[Lds::kRngLds ? S::GPB * (kSub / 4) : 1];
    __shared__ __align__(16) uint8_t bbuf[S::GPB * kSub * S::Lp];
    const float* tab = load_tables_band(tb);
    __syncthreads();
    const uint32_t lane = __lane_id(), gw = lane / G, r = lane % G;
    const uint32_t gl = threadIdx.x / G;
    float* pb = pbuf + gl * Lds::PS;
    uint32_t* sb = sbuf + gl * Lds::SW;
    uint8_t* bb = bbuf + gl * kSub * S::Lp;
    uint4* rg = rng + (Lds::kRngLds ? gl * (kSub / 4) : 0u);
    uint64_t lo = 0, hi = 0;  // item range of the job in J (wavefront-uniform)
    JobView J{};
    for (;;) {
        const Chunk ch = wave_take(&bc.counters[1], bc.block_items, S::GPW, 32);
        if (ch.base >= bc.block_items) break;
        for (uint32_t kk = 0; kk < ch.n; kk += S::GPW) {
            const uint64_t round = ch.base + kk;
            if (round >= bc.block_items) break;
            if (round < lo || round >= hi) {
                const uint32_t j = find_job_uniform(bc.block_off, bc.n_jobs, round);
                J = uniform_job(bc.jobs + j);
                lo = bc.block_off[j];
                hi = bc.block_off[j + 1];
                lo = (uint64_t(__builtin_amdgcn_readfirstlane(uint32_t(lo >> 32))) << 32) |
                     __builtin_amdgcn_readfirstlane(uint32_t(lo));
                hi = (uint64_t(__builtin_amdgcn_readfirstlane(uint32_t(hi >> 32))) << 32) |
                     __builtin_amdgcn_readfirstlane(uint32_t(hi));
            }
            const uint64_t q = round - lo + gw;
            const bool active = q < uint64_t(J.B) * J.nblk;
            const uint32_t s = active ? uint32_t(q / J.nblk) : 0u;
            const uint32_t c = active ? uint32_t(q % J.nblk) : 0u;
            if (J.flags & SI_FB_PER_MARKER)
                run_block_band<S, true, true>(J, tab, tb.P, active, s, c, r, pb, sb, bb, rg);
            else
                run_block_band<S, false, true>(J, tab, tb.P, active, s, c, r, pb, sb, bb, rg);
        }
    }
}

}  // namespace si
