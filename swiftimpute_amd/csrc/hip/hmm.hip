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
// This file keeps both vectors in linear FP32 space:
//   * forward: sum-normalized, a'_j = (t_o + u a_j) eps_j; the floor term is
//     below FP32 resolution whenever To >= -80 (checked on the host);
//   * backward: b_m = e^{Bl_m} v_m with the log scale Bl_m in FP64, so the
//     floor enters exactly as f = exp(-999 - Bl_m) once b reaches it.
// exp/log run once per (sample, marker), not per state.
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
#include "si_device.hpp"
#include "si_internal.hpp"

namespace si {

constexpr int kFbThreads = 128;

__device__ __forceinline__ void load_tables(const FbParams& p, float4* eps_l, float* emax_l) {
    const uint32_t n = p.P * 8;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        eps_l[i] = reinterpret_cast<const float4*>(p.eps)[i];
        emax_l[i] = p.emax[i];
    }
}

// Emission factors of this lane's four states: class 0/1/2 select
// eps.x/y/z, class 3 (padding state) gives 0.
__device__ __forceinline__ void lane_emission(float4 e3, uint32_t sa4, float e[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t cls = (sa4 >> (8 * i + 4)) & 3u;
        const float x = (cls & 1u) ? e3.y : e3.x;
        const float y = (cls & 1u) ? 0.0f : e3.z;
        e[i] = (cls & 2u) ? y : x;
    }
}

// Emission-table row for (palette code, class mask of marker m).  Codes at or
// beyond the palette size are read as the last palette entry, so no input can
// index past the table.
__device__ __forceinline__ uint32_t emis_row(const FbParams& p, uint32_t code, uint32_t m) {
    return min(code, p.P - 1u) * 8u + p.mmask[m];
}

__device__ __forceinline__ float2 step_tu(const FbParams& p, uint32_t m) {
    if (p.flags & SI_FB_PER_MARKER) {
        const float4 c = reinterpret_cast<const float4*>(p.steps)[m];
        return make_float2(c.x, c.y);
    }
    return make_float2(p.step.t_o, p.step.u);
}

// Forward: a <- normalize((t_o + u a) * e).  Returns the normalizer.
template <int LPS>
__device__ __forceinline__ float fwd_step(float4& a, const float e[4], float t_o, float u) {
    const float x0 = fmaf(u, a.x, t_o) * e[0];
    const float x1 = fmaf(u, a.y, t_o) * e[1];
    const float x2 = fmaf(u, a.z, t_o) * e[2];
    const float x3 = fmaf(u, a.w, t_o) * e[3];
    const float S = group_sum<LPS>((x0 + x1) + (x2 + x3));
    const float r = __builtin_amdgcn_rcpf(S);
    a = make_float4(x0 * r, x1 * r, x2 * r, x3 * r);
    return S;
}

template <int LPS>
__device__ __forceinline__ float fwd_init(float4& a, const float e[4]) {
    const float S = group_sum<LPS>((e[0] + e[1]) + (e[2] + e[3]));
    const float r = __builtin_amdgcn_rcpf(S);
    a = make_float4(e[0] * r, e[1] * r, e[2] * r, e[3] * r);
    return S;
}

// Backward: from the vector at marker n (emission e, offset emax) to n - 1.
//   b_{n-1}[i] = phi + e^{Bl + emax} (t_o W + u w_i),  w = e * v, W = sum w
// New scale Bl' = max(Bl + emax + log W, -999); with g = exp(Bl + emax +
// log W - Bl') and f = exp(-999 - Bl'), v' = f + g t_o + (g u / W) w.
template <int LPS>
__device__ __forceinline__ void bwd_step(float4& v, double& Bl, const float e[4], float emax,
                                         float t_o, float u) {
    const float w0 = e[0] * v.x, w1 = e[1] * v.y, w2 = e[2] * v.z, w3 = e[3] * v.w;
    const float W = group_sum<LPS>((w0 + w1) + (w2 + w3));
    const double nat = Bl + double(emax) + double(__logf(W));
    const double arg = -999.0 - nat;
    const float x = __expf(float(-fabs(arg)));
    const bool floor_dominates = arg > 0.0;
    const float f = floor_dominates ? 1.0f : x;
    const float g = floor_dominates ? x : 1.0f;
    Bl = floor_dominates ? -999.0 : nat;
    const float gr = g * u * __builtin_amdgcn_rcpf(W);
    const float base = fmaf(g, t_o, f);
    v = make_float4(fmaf(gr, w0, base), fmaf(gr, w1, base), fmaf(gr, w2, base),
                    fmaf(gr, w3, base));
}

// ---------------------------------------------------------------------------
// Chain pass.
// ---------------------------------------------------------------------------
template <int LPS>
__global__ __launch_bounds__(kFbThreads) void fb_chain_kernel(FbParams p) {
    __shared__ float4 eps_l[kMaxPalette * 8];
    __shared__ float emax_l[kMaxPalette * 8];
    load_tables(p, eps_l, emax_l);
    __syncthreads();

    const uint32_t gtid = blockIdx.x * kFbThreads + threadIdx.x;
    const uint32_t grp = gtid / LPS, r = gtid % LPS;
    const bool fwd = grp < p.Bpad;
    const uint32_t s = fwd ? grp : grp - p.Bpad;
    if (s >= p.B) return;  // whole lane groups leave together

    const uint8_t* codes = p.codes + size_t(s) * p.pitch;
    const uint32_t* sa32 = reinterpret_cast<const uint32_t*>(p.sa);
    const uint32_t lp4 = p.Lp / 4;

    if (fwd) {
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
        double ll = 0.0;
        for (uint32_t c = 0; c < p.nblk; ++c) {
            const uint32_t m0 = c * kBlock;
            const uint4 cw = *reinterpret_cast<const uint4*>(codes + m0);
#pragma unroll
            for (uint32_t t = 0; t < kBlock; ++t) {
                const uint32_t m = m0 + t;
                if (m < p.M) {
                    const uint32_t code = byte_of((&cw.x)[t >> 2], t & 3);
                    const uint32_t row = emis_row(p, code, m);
                    float e[4];
                    lane_emission(eps_l[row], sa32[size_t(m) * lp4 + r], e);
                    float S;
                    if (m == 0) {
                        S = fwd_init<LPS>(a, e);
                    } else {
                        const float2 tu = step_tu(p, m);
                        S = fwd_step<LPS>(a, e, tu.x, tu.y);
                    }
                    ll += double(emax_l[row]) + double(__logf(S));
                    if (t == 0)
                        reinterpret_cast<float4*>(p.ck_alpha)[(size_t(s) * p.nblk + c) * lp4 + r] = a;
                }
            }
        }
        if (r == 0 && p.loglik) p.loglik[s] = ll;
    } else {
        float4 v = make_float4(1.f, 1.f, 1.f, 1.f);
        double Bl = 0.0;
        const size_t last = size_t(s) * p.nblk + (p.nblk - 1);
        reinterpret_cast<float4*>(p.ck_beta)[last * lp4 + r] = v;
        if (r == 0) p.ck_bl[last] = Bl;
        for (int c = int(p.nblk) - 1; c >= 0; --c) {
            const uint32_t m0 = uint32_t(c) * kBlock;
            const uint4 cw = *reinterpret_cast<const uint4*>(codes + m0);
#pragma unroll
            for (int t = int(kBlock) - 1; t >= 0; --t) {
                const uint32_t n = m0 + uint32_t(t);
                if (n < p.M && n > 0) {
                    const uint32_t code = byte_of((&cw.x)[t >> 2], t & 3);
                    const uint32_t row = emis_row(p, code, n);
                    float e[4];
                    lane_emission(eps_l[row], sa32[size_t(n) * lp4 + r], e);
                    const float2 tu = step_tu(p, n);
                    bwd_step<LPS>(v, Bl, e, emax_l[row], tu.x, tu.y);
                    if (t == 0) {  // v is now the vector at marker m0 - 1, last of block c - 1
                        const size_t ck = size_t(s) * p.nblk + uint32_t(c - 1);
                        reinterpret_cast<float4*>(p.ck_beta)[ck * lp4 + r] = v;
                        if (r == 0) p.ck_bl[ck] = Bl;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Block pass.
// ---------------------------------------------------------------------------
template <int LPS>
__global__ __launch_bounds__(kFbThreads) void fb_block_kernel(FbParams p) {
    constexpr uint32_t G = kFbThreads / LPS;  // lane groups per workgroup
    constexpr uint32_t LP = 4 * LPS;
    __shared__ float4 eps_l[kMaxPalette * 8];
    __shared__ float emax_l[kMaxPalette * 8];
    __shared__ float pbuf[G][kBlock][LP];
    __shared__ uint4 rng[G][kBlock / 4];
    load_tables(p, eps_l, emax_l);
    __syncthreads();

    const uint32_t gtid = blockIdx.x * kFbThreads + threadIdx.x;
    const uint32_t grp = gtid / LPS, r = gtid % LPS, gl = threadIdx.x / LPS;
    const bool active = grp < p.B * p.nblk;
    const uint32_t s = active ? grp / p.nblk : 0u;
    const uint32_t c = active ? grp % p.nblk : 0u;
    const uint32_t m0 = c * kBlock;
    const uint32_t nm = active ? min(kBlock, p.M - m0) : 0u;
    const uint32_t lp4 = p.Lp / 4;
    const uint32_t* sa32 = reinterpret_cast<const uint32_t*>(p.sa);

    if (active) {
        const uint4 cw = *reinterpret_cast<const uint4*>(p.codes + size_t(s) * p.pitch + m0);
        float4 A[kBlock];
        A[0] = reinterpret_cast<const float4*>(p.ck_alpha)[(size_t(s) * p.nblk + c) * lp4 + r];
#pragma unroll
        for (uint32_t t = 1; t < kBlock; ++t) {
            if (t < nm) {
                const uint32_t m = m0 + t;
                const uint32_t code = byte_of((&cw.x)[t >> 2], t & 3);
                const uint32_t row = emis_row(p, code, m);
                float e[4];
                lane_emission(eps_l[row], sa32[size_t(m) * lp4 + r], e);
                const float2 tu = step_tu(p, m);
                A[t] = A[t - 1];
                fwd_step<LPS>(A[t], e, tu.x, tu.y);
            }
        }
        const size_t ck = size_t(s) * p.nblk + c;
        float4 v = reinterpret_cast<const float4*>(p.ck_beta)[ck * lp4 + r];
        double Bl = p.ck_bl[ck];
#pragma unroll
        for (int t = int(kBlock) - 1; t >= 0; --t) {
            if (uint32_t(t) < nm) {
                const float4 q = make_float4(A[t].x * v.x, A[t].y * v.y, A[t].z * v.z, A[t].w * v.w);
                const float Q = group_sum<LPS>((q.x + q.y) + (q.z + q.w));
                const float rq = 1.0f / Q;
                *reinterpret_cast<float4*>(&pbuf[gl][t][4 * r]) =
                    make_float4(q.x * rq, q.y * rq, q.z * rq, q.w * rq);
                if (t > 0) {
                    const uint32_t n = m0 + uint32_t(t);
                    const uint32_t code = byte_of((&cw.x)[t >> 2], t & 3);
                    const uint32_t row = emis_row(p, code, n);
                    float e[4];
                    lane_emission(eps_l[row], sa32[size_t(n) * lp4 + r], e);
                    const float2 tu = step_tu(p, n);
                    bwd_step<LPS>(v, Bl, e, emax_l[row], tu.x, tu.y);
                }
            }
        }
        const uint2 key = make_uint2(uint32_t(p.seed), uint32_t(p.seed >> 32));
        for (uint32_t j = r; j < kBlock / 4; j += LPS)
            rng[gl][j] = philox4x32_10(make_uint4((m0 >> 2) + j, p.sample_base + s, p.rng_stream, 0u), key);
    }
    __syncthreads();
    if (!active) return;

    const bool argmax = (p.flags & SI_FB_ARGMAX) != 0;
    const uint32_t L = p.L;
    for (uint32_t t = r; t < nm; t += LPS) {
        const uint32_t m = m0 + t;
        const float u = u24((&rng[gl][t >> 2].x)[t & 3]);
        const uint8_t* sam = p.sa + size_t(m) * p.Lp;
        float cs = 0.0f, best = -1.0f, ap0 = 0.0f, ap1 = 0.0f;
        uint32_t kdraw = 0xffffffffu, kbest = 0;
        for (uint32_t k = 0; k < L; ++k) {
            const float pk = pbuf[gl][t][k];
            cs += pk;
            if (kdraw == 0xffffffffu && u < cs) kdraw = k;
            if (pk > best) {
                best = pk;
                kbest = k;
            }
            const uint32_t b = sam[k];
            ap0 += (b & 1u) ? pk : 0.0f;
            ap1 += (b & 2u) ? pk : 0.0f;
        }
        if (kdraw == 0xffffffffu) kdraw = 0;
        const uint32_t ks = argmax ? kbest : kdraw;
        const size_t o = size_t(s) * p.M + m;
        if (p.ds) p.ds[o] = ap0 + ap1;
        if (p.ap) reinterpret_cast<float2*>(p.ap)[o] = make_float2(ap0, ap1);
        if (p.hap) {
            const uint32_t b = sam[ks];
            p.hap[size_t(s) * 2 * p.M + m] = uint8_t((b & 4u) ? 255u : (b & 1u));
            p.hap[size_t(s) * 2 * p.M + p.M + m] = uint8_t((b & 8u) ? 255u : ((b >> 1) & 1u));
        }
        if (p.prob)
            for (uint32_t k = 0; k < L; ++k) p.prob[o * L + k] = pbuf[gl][t][k];
    }
}

size_t fb_workspace_bytes_one(uint32_t B, uint32_t M, uint32_t Lp) {
    const size_t nblk = (size_t(M) + kBlock - 1) / kBlock;
    const size_t vec = size_t(B) * nblk * Lp * sizeof(float);
    const size_t vec_al = (vec + 255) & ~size_t(255);
    const size_t bl = size_t(B) * nblk * sizeof(double);
    return 2 * vec_al + ((bl + 255) & ~size_t(255));
}

template <int LPS>
static int launch_fb_lps(const FbParams& p, uint32_t passes, hipStream_t stream) {
    const uint32_t gpw = kWave / LPS;  // groups per wavefront
    FbParams q = p;
    q.Bpad = (p.B + gpw - 1) / gpw * gpw;
    const size_t chain_threads = size_t(q.Bpad + p.B) * LPS;
    const size_t block_threads = size_t(p.B) * p.nblk * LPS;
    const size_t g1 = (chain_threads + kFbThreads - 1) / kFbThreads;
    const size_t g2 = (block_threads + kFbThreads - 1) / kFbThreads;
    if (g1 > 0x7fffffffu || g2 > 0x7fffffffu) return SI_ERR_UNSUPPORTED;
    if (passes & SI_FB_PASS_CHAIN) {
        hipLaunchKernelGGL(fb_chain_kernel<LPS>, dim3(uint32_t(g1)), dim3(kFbThreads), 0, stream,
                           q);
        if (hipGetLastError() != hipSuccess) return SI_ERR_DEVICE;
    }
    if (passes & SI_FB_PASS_BLOCK) {
        hipLaunchKernelGGL(fb_block_kernel<LPS>, dim3(uint32_t(g2)), dim3(kFbThreads), 0, stream,
                           q);
        if (hipGetLastError() != hipSuccess) return SI_ERR_DEVICE;
    }
    return 0;
}

int launch_fb(const FbParams& p, uint32_t passes, hipStream_t stream) {
    if (p.B == 0 || p.M == 0) return 0;
    switch (p.Lp) {
        case 4: return launch_fb_lps<1>(p, passes, stream);
        case 8: return launch_fb_lps<2>(p, passes, stream);
        case 16: return launch_fb_lps<4>(p, passes, stream);
        case 32: return launch_fb_lps<8>(p, passes, stream);
        case 64: return launch_fb_lps<16>(p, passes, stream);
        default: return SI_ERR_UNSUPPORTED;
    }
}

}  // namespace si
