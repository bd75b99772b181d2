// PBWT prefix arrays for gfx950: a chunked three-phase construction.
//
// Reference (PBWTBuilder::build_marker, src/pbwt/pbwt_index.cpp:153-208): the
// order after marker m is the stable partition of the order after m - 1 by
// the allele at m (allele == 0 first, every nonzero allele after); row m is
// the order after marker m, and every divergence in row m equals m.  So row m
// sorts the haplotypes by (x_m, x_{m-1}, ..., x_0) lexicographically, ties by
// haplotype index.
//
// The markers are cut into chunks of C, and the rows come out of three phases:
//   A  one workgroup per chunk, all chunks at once: the chunk's order at its
//      end when swept from the identity order, and each haplotype's group id
//      (the dense rank of its allele string over the chunk).  For H <= 32768
//      this is an LSD radix sort of per-chunk key bytes (4-bit passes); above,
//      a sweep carrying per-position group-start flags.
//   B  the true order at every chunk end: for H <= 32768 a log-depth prefix
//      scan of rank compositions (one launch per level over every chunk);
//      above, one workgroup serial over the chunk boundaries.
//   C  one workgroup per chunk: re-sweep the chunk from its true starting
//      order (identity for chunk 0, else the previous chunk's last row),
//      writing every row of the chunk but its last.
// The constant divergence rows go out beside A and B on a second stream (or
// inside A or C where that stream is not available).
//
// Inside a workgroup (16 wavefronts) wave w owns the positions
// [w*S, (w+1)*S), S = 64*K; in round j < K lane l holds position
// w*S + 64*j + l, so every load and store of a round is one contiguous
// 256-byte run.  A stable partition needs, per element, the count of zeros
// before it: a ballot and mbcnt within the round, a running count across
// rounds and one LDS exchange of the 16 per-wave totals across waves.
// Haplotype ids: 16 bits in the LDS forms (H <= 49152); the global-memory
// form (49152 < H <= 2^19) keeps 31 bits beside a flag bit, and 64-bit
// element words where an id travels with its group id.
#include "si_device.hpp"
#include "si_internal.hpp"
#include <atomic>
#include <type_traits>

namespace si {

constexpr int kPW = 16;                           // wavefronts per workgroup
constexpr int kPT = kPW * kWave;                  // 1024 threads
constexpr uint32_t kPbwtMaxW32 = kPbwtMaxH / 32;  // column words staged in LDS

struct PbwtArgs {
    const uint32_t* cols32;  // [M][2W] the bit-packed columns as 32-bit words
    uint32_t H, M, W2, C, nchunk, S;
    uint32_t* prefix;        // [M][H]
    uint32_t* div;           // [M][H] or null
    uint32_t* xa;            // [nchunk][2][H] phase A exchange
    uint32_t* gid;           // [nchunk][H] group id of each haplotype at the chunk end
    uint32_t* ngroups;       // [nchunk]
    uint32_t* xb;            // [2][16 * 64 * K] 64-bit words: the boundary composition's exchange
    uint8_t* tp;             // [nchunk][NP][H] key byte p of each haplotype (radix digit planes)
    uint8_t* th;             // [nchunk][H][TS] the same bytes, haplotype-major (key comparison)
    uint32_t NP, TS;         // key bytes per chunk (C / 8), row stride of th (NP rounded to 16)
    uint32_t rows_div;       // phase C also writes the divergence rows of its chunk
    uint16_t* ord0;          // [nchunk][H] each chunk's order sorted by its key (radix form)
    uint16_t* ord1;          // [nchunk][H] the scan's second order buffer
    uint32_t* gcnt;          // [2][nchunk] group counts of the scan's two levels in flight
};

// Divergence row m (every entry m), streaming stores by the whole workgroup.
__device__ __forceinline__ void div_row(const PbwtArgs& a, uint32_t m) {
    uint32_t* drow = a.div + size_t(m) * a.H;
    for (uint32_t i = threadIdx.x; i < a.H; i += blockDim.x) __builtin_nontemporal_store(m, drow + i);
}

__device__ __forceinline__ uint32_t mbcnt64(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

// Inclusive prefix sum over the 64 lanes of a wavefront: shifts within each
// 16-lane row, then row 15's and row 31's last lanes broadcast to the rows
// after them (DPP, no LDS round trips).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    x += __builtin_amdgcn_update_dpp(0u, x, 0x111, 0xf, 0xf, true);   // row_shr:1
    x += __builtin_amdgcn_update_dpp(0u, x, 0x112, 0xf, 0xf, true);   // row_shr:2
    x += __builtin_amdgcn_update_dpp(0u, x, 0x114, 0xf, 0xf, true);   // row_shr:4
    x += __builtin_amdgcn_update_dpp(0u, x, 0x118, 0xf, 0xf, true);   // row_shr:8
    x += __builtin_amdgcn_update_dpp(0u, x, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
    x += __builtin_amdgcn_update_dpp(0u, x, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
    return x;
}

__device__ __forceinline__ uint32_t allele_bit(const uint32_t* col, uint32_t h) {
    return (col[h >> 5] >> (h & 31u)) & 1u;
}

// Number of positions wave v owns that lie inside [0, H).
__device__ __forceinline__ uint32_t wave_valid(const PbwtArgs& a, uint32_t v) {
    const uint32_t b = v * a.S;
    return a.H > b ? min(a.S, a.H - b) : 0u;
}

// Column m into LDS (first column of a sweep).
__device__ __forceinline__ void load_column(const PbwtArgs& a, uint32_t m, uint32_t* col) {
    const uint32_t n = (a.H + 31u) / 32u;
    for (uint32_t i = threadIdx.x; i < n; i += kPT) col[i] = a.cols32[size_t(m) * a.W2 + i];
}

// The next column is fetched into registers at the start of a step and stored
// to the other LDS buffer at its end (n <= 2 * kPT words).
// The global-memory form's columns: up to kPbwtMaxW32 words, kColWide a thread.
constexpr uint32_t kColWide = kPbwtMaxW32 / kPT;
struct ColPrefetchWide {
    uint32_t w[kColWide];
    __device__ __forceinline__ void fetch(const PbwtArgs& a, uint32_t m) {
        const uint32_t n = (a.H + 31u) / 32u, t = threadIdx.x;
#pragma unroll
        for (uint32_t k = 0; k < kColWide; ++k)
            if (t + k * kPT < n) w[k] = a.cols32[size_t(m) * a.W2 + t + k * kPT];
    }
    __device__ __forceinline__ void store(const PbwtArgs& a, uint32_t* col) const {
        const uint32_t n = (a.H + 31u) / 32u, t = threadIdx.x;
#pragma unroll
        for (uint32_t k = 0; k < kColWide; ++k)
            if (t + k * kPT < n) col[t + k * kPT] = w[k];
    }
};
struct ColPrefetch {
    uint32_t w0 = 0, w1 = 0;
    __device__ __forceinline__ void fetch(const PbwtArgs& a, uint32_t m) {
        const uint32_t n = (a.H + 31u) / 32u, t = threadIdx.x;
        if (t < n) w0 = a.cols32[size_t(m) * a.W2 + t];
        if (t + kPT < n) w1 = a.cols32[size_t(m) * a.W2 + t + kPT];
    }
    __device__ __forceinline__ void store(const PbwtArgs& a, uint32_t* col) const {
        const uint32_t n = (a.H + 31u) / 32u, t = threadIdx.x;
        if (t < n) col[t] = w0;
        if (t + kPT < n) col[t + kPT] = w1;
    }
};

// ---------------------------------------------------------------------------
// One stable-partition step of a workgroup, in two passes over the wave's
// rounds with a barrier between them: the first counts (zeros per wave, and
// for group tracking the state of the group open at the wave's end), the
// second recomputes each round (the element loads hit L1/L2) and scatters.
// No per-element state is held across the barrier, so registers do not grow
// with H.
// ---------------------------------------------------------------------------
struct WaveCtx {
    uint32_t t, lane, w, base, nvalid, last;
    uint64_t lt, le;
    __device__ __forceinline__ WaveCtx(const PbwtArgs& a) {
        t = threadIdx.x;
        lane = t & 63u;
        w = t >> 6;
        base = w * a.S;
        nvalid = wave_valid(a, w);
        last = a.H - 1u;
        lt = (1ull << lane) - 1ull;
        le = lt | (1ull << lane);
    }
    __device__ __forceinline__ uint32_t pos(uint32_t j) const { return min(base + j * 64u + lane, last); }
    __device__ __forceinline__ bool valid(uint32_t j) const { return j * 64u + lane < nvalid; }
};

// Zeros before this wave and in all waves, from the per-wave counts in LDS.
__device__ __forceinline__ void zero_offsets(const uint32_t* wz, uint32_t w, uint32_t& zbefore,
                                             uint32_t& ztot) {
    zbefore = 0;
    ztot = 0;
#pragma unroll
    for (int v = 0; v < kPW; ++v) {
        const uint32_t z = wz[v];
        ztot += z;
        zbefore += (uint32_t(v) < w) ? z : 0u;
    }
}

// ---------------------------------------------------------------------------
// Phase C: rows [c*C, min(M, c*C + C) - 1) of chunk c.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kPT) void pbwt_rows_kernel(PbwtArgs a) {
    __shared__ uint32_t col[2][kPbwtMaxW32];
    __shared__ uint32_t wz[kPW];
    const uint32_t c = blockIdx.x;
    const uint32_t m_first = c * a.C, m_end = min(a.M, m_first + a.C);
    if (m_end - m_first < 2) return;
    const WaveCtx q(a);
    const uint32_t K = a.S / 64u;

    load_column(a, m_first, col[0]);
    __syncthreads();
    const uint32_t* prev = (c == 0) ? nullptr : a.prefix + size_t(m_first - 1) * a.H;
    for (uint32_t m = m_first; m + 1 < m_end; ++m) {
        const uint32_t* cb = col[(m - m_first) & 1u];
        const bool pre = m + 2 < m_end;
        ColPrefetchWide nx;
        if (pre) nx.fetch(a, m + 1);

        uint32_t zsum = 0;
#pragma unroll 8
        for (uint32_t j = 0; j < K; ++j) {
            const uint32_t p = q.pos(j);
            const uint32_t h = prev ? prev[p] : p;
            const uint64_t Z = __ballot(q.valid(j) && allele_bit(cb, h) == 0u);
            zsum += uint32_t(__popcll(Z));
        }
        if (q.lane == 0) wz[q.w] = zsum;
        __syncthreads();
        uint32_t zbefore, ztot;
        zero_offsets(wz, q.w, zbefore, ztot);
        uint32_t* row = a.prefix + size_t(m) * a.H;
        uint32_t zrun = 0;
#pragma unroll 8
        for (uint32_t j = 0; j < K; ++j) {
            const uint32_t p = q.pos(j);
            const uint32_t h = prev ? prev[p] : p;
            const bool valid = q.valid(j);
            const uint32_t x = allele_bit(cb, h);
            const uint64_t Z = __ballot(valid && x == 0u);
            const uint32_t zb = zrun + mbcnt64(Z);
            zrun += uint32_t(__popcll(Z));
            const uint32_t e = j * 64u + q.lane;
            const uint32_t dst = x ? ztot + (q.base - zbefore) + (e - zb) : zbefore + zb;
            if (valid) row[dst] = h;
        }
        if (pre) nx.store(a, col[((m - m_first) & 1u) ^ 1u]);
        __syncthreads();
        prev = row;
    }
}

// ---------------------------------------------------------------------------
// Phase A: chunk-local groups (and chunk 0's true final order, and the
// divergence rows of the chunk).  Element word: haplotype | 1 << 31 where a
// group starts.
// ---------------------------------------------------------------------------
constexpr uint32_t kIdMask = 0x7fffffffu;

// One round of group tracking.  Z/O: lanes with allele 0 / 1, Bm: group
// starts.  Returns, for this lane, whether it is the first of its allele in
// its group as far as this wave sees, and whether its group began before
// this wave ("open"); advances the wave-uniform running state.
struct GroupRun {
    uint32_t zs = 0, os = 0;  // same-allele counts since the wave's last group start
    bool hs = false;          // a group started in this wave
    __device__ __forceinline__ void round(const WaveCtx& q, uint64_t Z, uint64_t O, uint64_t Bm,
                                          uint32_t x, bool& first_local, bool& open) {
        const uint64_t sle = Bm & q.le;
        const bool in_round = sle != 0;
        const uint32_t s = in_round ? 63u - uint32_t(__clzll(sle)) : 0u;
        const uint64_t seg = in_round ? (q.lt & ~((1ull << s) - 1ull)) : q.lt;  // [s, lane)
        const uint64_t same = (x ? O : Z) & seg;
        first_local = (same == 0) && (in_round || (x ? os : zs) == 0);
        open = !in_round && !hs;
        if (Bm) {
            const uint32_t sb = 63u - uint32_t(__clzll(Bm));
            zs = uint32_t(__popcll(Z >> sb));
            os = uint32_t(__popcll(O >> sb));
            hs = true;
        } else {
            zs += uint32_t(__popcll(Z));
            os += uint32_t(__popcll(O));
        }
    }
};

__global__ __launch_bounds__(kPT) void pbwt_groups_kernel(PbwtArgs a) {
    __shared__ uint32_t col[2][kPbwtMaxW32];
    __shared__ uint32_t s_z[kPW], s_hs[kPW], s_zs[kPW], s_os[kPW];
    const uint32_t c = blockIdx.x;
    const uint32_t m_first = c * a.C, m_end = min(a.M, m_first + a.C);
    const WaveCtx q(a);
    const uint32_t K = a.S / 64u;
    uint32_t* X0 = a.xa + size_t(c) * 2u * a.H;
    uint32_t* X1 = X0 + a.H;

    load_column(a, m_first, col[0]);
    __syncthreads();
    for (uint32_t m = m_first; m < m_end; ++m) {
        const uint32_t k = m - m_first;
        const uint32_t* cb = col[k & 1u];
        const bool pre = m + 1 < m_end;
        ColPrefetchWide nx;
        if (pre) nx.fetch(a, m + 1);
        const uint32_t* Xin = (k & 1u) ? X0 : X1;
        uint32_t* Xout = (k & 1u) ? X1 : X0;
        auto elem = [&](uint32_t j) -> uint32_t {
            const uint32_t p = q.pos(j);
            return k > 0 ? Xin[p] : (p | (p == 0 ? 0x80000000u : 0u));  // identity: one group
        };

        uint32_t zsum = 0;
        GroupRun g;
#pragma unroll 4
        for (uint32_t j = 0; j < K; ++j) {
            const uint32_t v = elem(j);
            const bool valid = q.valid(j);
            const uint32_t x = allele_bit(cb, v & kIdMask);
            const uint64_t Z = __ballot(valid && x == 0u);
            const uint64_t O = __ballot(valid && x != 0u);
            const uint64_t Bm = __ballot(valid && (v >> 31) != 0u);
            bool fl, op;
            g.round(q, Z, O, Bm, x, fl, op);
            zsum += uint32_t(__popcll(Z));
        }
        if (q.lane == 0) {
            s_z[q.w] = zsum;
            s_hs[q.w] = g.hs ? 1u : 0u;
            s_zs[q.w] = g.zs;
            s_os[q.w] = g.os;
        }
        __syncthreads();
        // zeros before this wave, all zeros, and the same-allele counts of the
        // group still open at this wave's start
        uint32_t zbefore = 0, ztot = 0, cz = 0, co = 0;
#pragma unroll
        for (int v = 0; v < kPW; ++v) {
            const uint32_t z = s_z[v];
            ztot += z;
            if (uint32_t(v) < q.w) {
                zbefore += z;
                if (s_hs[v]) {
                    cz = s_zs[v];
                    co = s_os[v];
                } else {
                    cz += z;
                    co += wave_valid(a, uint32_t(v)) - z;
                }
            }
        }
        uint32_t zrun = 0;
        GroupRun g2;
#pragma unroll 4
        for (uint32_t j = 0; j < K; ++j) {
            const uint32_t v = elem(j);
            const bool valid = q.valid(j);
            const uint32_t x = allele_bit(cb, v & kIdMask);
            const uint64_t Z = __ballot(valid && x == 0u);
            const uint64_t O = __ballot(valid && x != 0u);
            const uint64_t Bm = __ballot(valid && (v >> 31) != 0u);
            bool fl, op;
            g2.round(q, Z, O, Bm, x, fl, op);
            const uint32_t zb = zrun + mbcnt64(Z);
            zrun += uint32_t(__popcll(Z));
            const uint32_t first = (fl && (!op || (x ? co : cz) == 0u)) ? 1u : 0u;
            const uint32_t e = j * 64u + q.lane;
            const uint32_t dst = x ? ztot + (q.base - zbefore) + (e - zb) : zbefore + zb;
            if (valid) Xout[dst] = (v & kIdMask) | (first << 31);
        }
        if (a.div) {
            uint32_t* drow = a.div + size_t(m) * a.H;
            for (uint32_t i = q.t; i < a.H; i += kPT) drow[i] = m;
        }
        if (pre) nx.store(a, col[(k & 1u) ^ 1u]);
        __syncthreads();
    }

    // group ids: dense rank of the group in the final order
    const uint32_t* Xf = ((m_end - 1 - m_first) & 1u) ? X1 : X0;
    uint32_t bsum = 0;
#pragma unroll 8
    for (uint32_t j = 0; j < K; ++j) {
        const uint64_t Bm = __ballot(q.valid(j) && (Xf[q.pos(j)] >> 31) != 0u);
        bsum += uint32_t(__popcll(Bm));
    }
    if (q.lane == 0) s_z[q.w] = bsum;
    __syncthreads();
    uint32_t bbefore, G;
    zero_offsets(s_z, q.w, bbefore, G);
    uint32_t* gid = a.gid + size_t(c) * a.H;
    uint32_t* lastrow = a.prefix + size_t(m_end - 1) * a.H;
    uint32_t brun = 0;
#pragma unroll 8
    for (uint32_t j = 0; j < K; ++j) {
        const uint32_t v = Xf[q.pos(j)];
        const bool valid = q.valid(j);
        const uint64_t Bm = __ballot(valid && (v >> 31) != 0u);
        const uint32_t inc = brun + uint32_t(__popcll(Bm & q.le));
        brun += uint32_t(__popcll(Bm));
        if (valid) {
            gid[v & kIdMask] = bbefore + inc - 1u;
            if (c == 0) lastrow[q.base + j * 64u + q.lane] = v & kIdMask;
        }
    }
    if (q.t == 0) a.ngroups[c] = G;
}

// ---------------------------------------------------------------------------
// Phase B: chunk-boundary orders, one workgroup, boundaries in sequence.
// Element word (64-bit): haplotype | group id << 32; LSD radix over 8-bit
// digits of the id.  Lanes of a round with the same digit agree on all 8 digit-bit
// ballots; the wave's running count per digit lives in LDS and is advanced by
// the highest lane of each digit.  The exchange buffers hold every position a
// wave owns (16 * S >= H), so loads and stores need no validity branch; the
// rounds run in batches of kBatch with the batch's loads issued together.
// ---------------------------------------------------------------------------
constexpr uint32_t kBatch = 8;

__device__ __forceinline__ uint64_t digit_match(uint32_t d, bool valid) {
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = ((d >> b) & 1u) != 0u;
        const uint64_t Bb = __ballot(valid && bit);
        match &= bit ? Bb : ~Bb;
    }
    return match;
}

__global__ __launch_bounds__(kPT) void pbwt_compose_kernel(PbwtArgs a) {
    __shared__ uint32_t cnt[kPW][256];  // per-wave digit counts, then running offsets
    __shared__ uint32_t dbase[256];     // exclusive prefix of the digit totals
    __shared__ uint32_t dsum[4];
    const WaveCtx q(a);
    const uint32_t K = a.S / 64u;
    const uint32_t npos = kPW * a.S;    // padded buffer length
    uint64_t* const xb = reinterpret_cast<uint64_t*>(a.xb);
    const uint32_t own = q.base + q.lane;

    for (uint32_t c = 1; c < a.nchunk; ++c) {
        const uint32_t* prev = a.prefix + size_t(c * a.C - 1u) * a.H;
        uint32_t* out = a.prefix + size_t(min(a.M, c * a.C + a.C) - 1u) * a.H;
        const uint32_t* gid = a.gid + size_t(c) * a.H;
        const uint32_t G = a.ngroups[c];
        const uint32_t nbits = G > 1u ? 32u - uint32_t(__clz(G - 1u)) : 0u;
        const uint32_t npass = (nbits + 7u) / 8u;
        if (npass == 0) {  // a single group: the order is unchanged
            for (uint32_t i = q.t; i < a.H; i += kPT) out[i] = prev[i];
            __syncthreads();
            continue;
        }
        // element words, gathered once into the buffer pass 0 reads
        {
            uint64_t* X0 = xb + npos;
            for (uint32_t j0 = 0; j0 < K; j0 += kBatch) {
                uint32_t h[kBatch];
                uint64_t e[kBatch];
#pragma unroll
                for (uint32_t u = 0; u < kBatch; ++u) h[u] = prev[min(own + (j0 + u) * 64u, q.last)];
#pragma unroll
                for (uint32_t u = 0; u < kBatch; ++u) e[u] = uint64_t(h[u]) | (uint64_t(gid[h[u]]) << 32);
#pragma unroll
                for (uint32_t u = 0; u < kBatch; ++u)
                    if (j0 + u < K) X0[own + (j0 + u) * 64u] = e[u];
            }
            __syncthreads();
        }
        for (uint32_t pass = 0; pass < npass; ++pass) {
            const uint32_t sh = 32u + 8u * pass;
            const uint64_t* X = xb + ((pass + 1u) & 1u) * npos;  // this pass's input
            const bool last = pass + 1 == npass;
            uint64_t* Y = xb + (pass & 1u) * npos;
            for (uint32_t i = q.t; i < uint32_t(kPW) * 256u; i += kPT) (&cnt[0][0])[i] = 0u;
            __syncthreads();
            for (uint32_t j0 = 0; j0 < K; j0 += kBatch) {
                uint64_t v[kBatch];
#pragma unroll
                for (uint32_t u = 0; u < kBatch; ++u) v[u] = X[own + min(j0 + u, K - 1u) * 64u];
#pragma unroll
                for (uint32_t u = 0; u < kBatch; ++u) {
                    if (j0 + u < K) {
                        const bool valid = q.valid(j0 + u);
                        const uint32_t d = uint32_t(v[u] >> sh) & 255u;
                        const uint64_t match = digit_match(d, valid);
                        if (valid && (match >> q.lane) == 1ull)
                            cnt[q.w][d] += uint32_t(__popcll(match));
                    }
                }
            }
            __syncthreads();
            uint32_t tot = 0, incl = 0;
            if (q.t < 256u) {
                for (int w = 0; w < kPW; ++w) {
                    const uint32_t x = cnt[w][q.t];
                    cnt[w][q.t] = tot;
                    tot += x;
                }
                incl = tot;
#pragma unroll
                for (int o = 1; o < kWave; o <<= 1) {
                    const uint32_t n = __shfl_up(incl, o, kWave);
                    if (q.lane >= uint32_t(o)) incl += n;
                }
                if (q.lane == 63u) dsum[q.w] = incl;
            }
            __syncthreads();
            if (q.t < 256u) {
                uint32_t off = 0;
                for (uint32_t w = 0; w < q.w; ++w) off += dsum[w];
                dbase[q.t] = off + incl - tot;
            }
            __syncthreads();
            for (uint32_t j0 = 0; j0 < K; j0 += kBatch) {
                uint64_t v[kBatch];
#pragma unroll
                for (uint32_t u = 0; u < kBatch; ++u) v[u] = X[own + min(j0 + u, K - 1u) * 64u];
#pragma unroll
                for (uint32_t u = 0; u < kBatch; ++u) {
                    if (j0 + u < K) {
                        const bool valid = q.valid(j0 + u);
                        const uint32_t d = uint32_t(v[u] >> sh) & 255u;
                        const uint64_t match = digit_match(d, valid);
                        const uint32_t c0 = cnt[q.w][d];
                        const uint32_t dst = dbase[d] + c0 + mbcnt64(match);
                        if (valid) {
                            if (last) out[dst] = uint32_t(v[u]);
                            else Y[dst] = v[u];
                            if ((match >> q.lane) == 1ull) cnt[q.w][d] = c0 + uint32_t(__popcll(match));
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// LDS-resident forms (H <= 49152).  The current order lives in LDS as 16-bit
// haplotype ids (and, for the group sweep, a bit per position marking group
// starts); each lane holds its K elements in registers across the step's
// barrier, so a step touches global memory only for the next column and the
// rows it writes.  NPOS = 16 * 64 * K positions are allocated; those at or
// beyond H hold 0 and are never written.
// ---------------------------------------------------------------------------
// Keeps the scheduler from hoisting later rounds' ballots (each round holds
// several 64-bit masks; hoisting every round's would overflow the scalar
// registers).
__device__ __forceinline__ void round_end() { __builtin_amdgcn_sched_barrier(0); }

// The same value, opaque to the optimizer: per-round addresses derived from it
// inside a loop are recomputed there (an immediate offset per round) instead
// of being hoisted out of the loop as K live 64-bit registers.
__device__ __forceinline__ uint32_t opaque(uint32_t x) {
    asm volatile("" : "+v"(x));
    return x;
}

template <int K>
struct LdsGeom {
    static constexpr uint32_t NPOS = uint32_t(kPW) * 64u * uint32_t(K);
    static constexpr uint32_t NW32 = NPOS / 32u;
};

// Phase C with the order in LDS (H <= 49152).  Positions past H hold padding
// ids (the position itself) whose allele bits read as 1, so every round of
// every wave is whole: the padding sorts after every haplotype and keeps the
// positions past H.  A step keeps each round's zero ballot in scalar
// registers and each lane's haplotype and allele in vector registers, so the
// scatter needs only the zero count before the lane (mbcnt) and a select
// between the zero and one destinations of its round.
__device__ __forceinline__ void store_col_padded(uint32_t* col, uint32_t i, uint32_t w, uint32_t nw,
                                                 uint32_t padm) {
    col[i] = i + 1u == nw ? (w | padm) : w;
}

// The 32 allele bits of a column word as 32 bytes (0 or 1), for the byte
// columns of the row re-sweep: each nibble spreads to the low bits of four
// bytes through one multiply (its shifted copies land on disjoint bits).
__device__ __forceinline__ void store_col_bytes(uint8_t* colb, uint32_t i, uint32_t w) {
    uint32_t d[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) d[k] = (__umul24((w >> (4 * k)) & 15u, 0x204081u)) & 0x01010101u;
    uint4* p = reinterpret_cast<uint4*>(colb + 32u * i);
    p[0] = make_uint4(d[0], d[1], d[2], d[3]);
    p[1] = make_uint4(d[4], d[5], d[6], d[7]);
}

template <int K>
struct RowsGeom {
    // K <= 32: the column as one byte per position (a haplotype's allele is
    // one LDS byte load at its id); above, bits (the order and two byte
    // columns would not fit in LDS).  K <= 24: the order as 32-bit ids (no
    // widening of each loaded id, one address for the LDS and row stores).
    static constexpr bool BYTES = K <= 32;
    static constexpr bool WIDE = K <= 24;
    using Ord = typename std::conditional<WIDE, uint32_t, uint16_t>::type;
    static constexpr uint32_t NPOS = LdsGeom<K>::NPOS;
    static constexpr uint32_t COLW = BYTES ? NPOS / 4u : LdsGeom<K>::NW32;  // words per column buffer
};

// One step of phase C: the order after marker m (stable partition of the
// order in LDS by the allele of column buffer PAR), written to LDS and to
// prefix row m, with the next column (prefetched) stored to the other buffer.
template <int K, int PAR>
__device__ __forceinline__ void rows_step(const PbwtArgs& a, uint32_t m, bool pre, typename RowsGeom<K>::Ord* ord,
                                          uint32_t (*col)[RowsGeom<K>::COLW], uint32_t* wz, uint32_t nw,
                                          uint32_t padm, bool divs) {
    using R = RowsGeom<K>;
    constexpr int NZ = (K + 31) / 32;  // words of per-round zero bits (bit columns)
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    const uint32_t base = w * uint32_t(64 * K);
    const uint32_t* cb = col[PAR];
    ColPrefetch nx;
    if (pre) nx.fetch(a, m + 1);

    uint32_t hv[K], xs[K];
#pragma unroll
    for (int j = 0; j < K; ++j) hv[j] = ord[base + uint32_t(j) * 64u + lane];
    // every round's allele load issued before the first is used
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if constexpr (R::BYTES) xs[j] = reinterpret_cast<const uint8_t*>(cb)[hv[j]];
        else xs[j] = (cb[hv[j] >> 5] >> (hv[j] & 31u)) & 1u;
    }
    if constexpr (R::BYTES) {
        __builtin_amdgcn_sched_group_barrier(0x0100, K, 0);  // the order loads
        __builtin_amdgcn_sched_group_barrier(0x0100, K, 0);  // the allele loads
    }
    uint64_t Zs[K];
    uint32_t zbits[NZ], zsum = 0;
#pragma unroll
    for (int k = 0; k < NZ; ++k) zbits[k] = 0u;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        Zs[j] = __ballot(xs[j] == 0u);
        if constexpr (!R::BYTES) zbits[j >> 5] |= (xs[j] ^ 1u) << (j & 31);
        zsum += uint32_t(__popcll(Zs[j]));
    }
    if (lane == 0) wz[w] = zsum;
    __syncthreads();
    uint32_t zbefore, ztot;
    zero_offsets(wz, w, zbefore, ztot);
    zbefore = __builtin_amdgcn_readfirstlane(zbefore);
    ztot = __builtin_amdgcn_readfirstlane(ztot);
    uint32_t* const rowp = a.prefix + size_t(m) * a.H;
    // stores past H (padding) fall outside the row's buffer range and are dropped
    const __amdgpu_buffer_rsrc_t rrow = __builtin_amdgcn_make_buffer_rsrc(rowp, 0, int(a.H * 4u), 0x00020000);
    uint32_t zrun = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const uint32_t mb = mbcnt64(Zs[j]);                            // zeros before the lane
        const uint32_t z0 = zbefore + zrun;                           // the round's first zero
        const uint32_t o0 = ztot + (base - zbefore) + uint32_t(j) * 64u - zrun;  // and one
        uint32_t dst;
        if constexpr (R::BYTES) {
            // the ballot itself as the lane condition (no per-lane bit test)
            dst = __builtin_amdgcn_inverse_ballot_w64(Zs[j]) ? z0 + mb : o0 + (lane - mb);
        } else {
            dst = ((zbits[j >> 5] >> (j & 31)) & 1u) ? z0 + mb : o0 + (lane - mb);
        }
        ord[dst] = typename R::Ord(hv[j]);
        if constexpr (R::BYTES) {
            __builtin_amdgcn_raw_buffer_store_b32(hv[j], rrow, int(dst * 4u), 0, 0);
        } else {
            if (dst < a.H) rowp[dst] = hv[j];
        }
        zrun += uint32_t(__popcll(Zs[j]));
    }
    if (divs) div_row(a, m);
    if (pre) {
        uint32_t* nb = col[PAR ^ 1];
        if constexpr (R::BYTES) {
            if (t < nw) store_col_bytes(reinterpret_cast<uint8_t*>(nb), t, t + 1u == nw ? (nx.w0 | padm) : nx.w0);
        } else {
            const uint32_t i0 = t, i1 = t + kPT;
            if (i0 < nw) store_col_padded(nb, i0, nx.w0, nw, padm);
            if (i1 < nw) store_col_padded(nb, i1, nx.w1, nw, padm);
        }
    }
    __syncthreads();
}

template <int K>
__global__ __launch_bounds__(kPT) void pbwt_rows_lds_kernel(PbwtArgs a) {
    using R = RowsGeom<K>;
    // the column buffers first, so the allele loads reach them by immediate offsets
    __shared__ __align__(16) struct {
        uint32_t col[2][R::COLW];
        typename R::Ord ord[R::NPOS];
        uint32_t wz[kPW];
    } sm;
    auto& col = sm.col;
    auto& ord = sm.ord;
    auto& wz = sm.wz;
    const uint32_t c = blockIdx.x;
    const uint32_t m_first = c * a.C, m_end = min(a.M, m_first + a.C);
    const bool divs = a.rows_div && a.div;
    if (divs) div_row(a, m_end - 1u);  // every divergence row of the chunk, its last one too
    if (m_end - m_first < 2) return;
    const uint32_t t = threadIdx.x;
    const uint32_t nw = (a.H + 31u) / 32u;
    const uint32_t padm = (a.H & 31u) ? ~((1u << (a.H & 31u)) - 1u) : 0u;

    const uint32_t* start = (c == 0) ? nullptr : a.prefix + size_t(m_first - 1) * a.H;
    for (uint32_t i = t; i < R::NPOS; i += kPT) ord[i] = typename R::Ord(i < a.H ? (start ? start[i] : i) : i);
    if constexpr (R::BYTES) {
        // padding positions read as allele 1 in both buffers; the last word's
        // padding bits are set by the padded column store
        for (uint32_t i = nw * 8u + t; i < R::COLW; i += kPT) {
            col[0][i] = 0x01010101u;
            col[1][i] = 0x01010101u;
        }
        for (uint32_t i = t; i < nw; i += kPT)
            store_col_bytes(reinterpret_cast<uint8_t*>(col[0]), i,
                            i + 1u == nw ? (a.cols32[size_t(m_first) * a.W2 + i] | padm)
                                         : a.cols32[size_t(m_first) * a.W2 + i]);
    } else {
        for (uint32_t i = nw + t; i < R::COLW; i += kPT) {
            col[0][i] = ~0u;
            col[1][i] = ~0u;
        }
        for (uint32_t i = t; i < nw; i += kPT)
            store_col_padded(col[0], i, a.cols32[size_t(m_first) * a.W2 + i], nw, padm);
    }
    __syncthreads();
    // steps alternate between the column buffers (compile-time, so the allele
    // loads address them by immediate offsets)
    for (uint32_t m = m_first; m + 1 < m_end; m += 2) {
        rows_step<K, 0>(a, m, m + 2 < m_end, ord, col, wz, nw, padm, divs);
        if (m + 2 < m_end) rows_step<K, 1>(a, m + 1, m + 3 < m_end, ord, col, wz, nw, padm, divs);
    }
}

template <int K>
__global__ __launch_bounds__(kPT) void pbwt_groups_lds_kernel(PbwtArgs a) {
    using G = LdsGeom<K>;
    __shared__ uint16_t ord[G::NPOS];
    __shared__ uint32_t fl[2][G::NW32];  // bit p: position p starts a group
    __shared__ uint32_t col[2][G::NW32];
    __shared__ uint32_t s_z[kPW], s_hs[kPW], s_zs[kPW], s_os[kPW];
    const uint32_t c = blockIdx.x;
    const uint32_t m_first = c * a.C, m_end = min(a.M, m_first + a.C);
    const WaveCtx q(a);

    for (uint32_t i = q.t; i < G::NPOS; i += kPT) ord[i] = uint16_t(i < a.H ? i : 0u);
    for (uint32_t i = q.t; i < G::NW32; i += kPT) fl[0][i] = (i == 0) ? 1u : 0u;  // one group
    load_column(a, m_first, col[0]);
    __syncthreads();
    for (uint32_t m = m_first; m < m_end; ++m) {
        const uint32_t k = m - m_first, cur = k & 1u, nxt = cur ^ 1u;
        const uint32_t* cb = col[cur];
        const bool pre = m + 1 < m_end;
        ColPrefetch nx;
        if (pre) nx.fetch(a, m + 1);
        for (uint32_t i = q.t; i < G::NW32; i += kPT) fl[nxt][i] = 0u;

        uint32_t pk[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const uint32_t p = q.base + uint32_t(j) * 64u + q.lane;
            pk[j] = ord[p] | (((fl[cur][p >> 5] >> (p & 31u)) & 1u) << 31);
        }
        uint32_t zsum = 0;
        GroupRun g;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const bool valid = q.valid(uint32_t(j));
            const uint32_t v = pk[j];
            const uint32_t x = allele_bit(cb, v & 0xffffu) | uint32_t(!valid);
            const uint64_t Z = __ballot(x == 0u);
            const uint64_t O = __ballot(valid && x != 0u);
            const uint64_t Bm = __ballot(valid && (v >> 31) != 0u);
            bool fl_, op;
            g.round(q, Z, O, Bm, x, fl_, op);
            pk[j] = (v & 0xffffu) | ((zsum + mbcnt64(Z)) << 16) | (uint32_t(fl_) << 29) |
                    (uint32_t(op) << 30) | (x << 31);
            zsum += uint32_t(__popcll(Z));
            if (K > 12) round_end();
        }
        if (q.lane == 0) {
            s_z[q.w] = zsum;
            s_hs[q.w] = g.hs ? 1u : 0u;
            s_zs[q.w] = g.zs;
            s_os[q.w] = g.os;
        }
        __syncthreads();
        uint32_t zbefore = 0, ztot = 0, cz = 0, co = 0;
#pragma unroll
        for (int v = 0; v < kPW; ++v) {
            const uint32_t z = s_z[v];
            ztot += z;
            if (uint32_t(v) < q.w) {
                zbefore += z;
                if (s_hs[v]) {
                    cz = s_zs[v];
                    co = s_os[v];
                } else {
                    cz += z;
                    co += wave_valid(a, uint32_t(v)) - z;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const uint32_t e = uint32_t(j) * 64u + q.lane;
            const uint32_t in = pk[j];
            const uint32_t x = in >> 31, open = (in >> 30) & 1u, f = (in >> 29) & 1u;
            const uint32_t zb = (in >> 16) & 0x1fffu;
            const bool first = f && (!open || (x ? co : cz) == 0u);
            const uint32_t dst = x ? ztot + (q.base - zbefore) + (e - zb) : zbefore + zb;
            if (e < q.nvalid) {
                ord[dst] = uint16_t(in & 0xffffu);
                if (first) atomicOr(&fl[nxt][dst >> 5], 1u << (dst & 31u));
            }
        }
        if (a.div) {
            uint32_t* drow = a.div + size_t(m) * a.H;
            for (uint32_t i = q.t; i < a.H; i += kPT) __builtin_nontemporal_store(m, drow + i);
        }
        if (pre) nx.store(a, col[nxt]);
        __syncthreads();
    }

    // group ids: dense rank of the group in the final order
    const uint32_t* ff = fl[(m_end - m_first) & 1u];
    uint32_t pk[K];  // haplotype | inclusive count of group starts in the wave << 16
    uint32_t bsum = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const uint32_t p = q.base + uint32_t(j) * 64u + q.lane;
        const uint64_t Bm = __ballot(q.valid(uint32_t(j)) && ((ff[p >> 5] >> (p & 31u)) & 1u));
        pk[j] = uint32_t(ord[p]) | ((bsum + uint32_t(__popcll(Bm & q.le))) << 16);
        bsum += uint32_t(__popcll(Bm));
    }
    if (q.lane == 0) s_z[q.w] = bsum;
    __syncthreads();
    uint32_t bbefore, Gn;
    zero_offsets(s_z, q.w, bbefore, Gn);
    uint32_t* gid = a.gid + size_t(c) * a.H;
    uint32_t* lastrow = a.prefix + size_t(m_end - 1) * a.H;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const uint32_t e = uint32_t(j) * 64u + q.lane;
        if (e < q.nvalid) {
            gid[pk[j] & 0xffffu] = bbefore + (pk[j] >> 16) - 1u;
            if (c == 0) lastrow[q.base + e] = pk[j] & 0xffffu;
        }
    }
    if (q.t == 0) a.ngroups[c] = Gn;
}

// ---------------------------------------------------------------------------
// Radix form of phases A and B (H <= 32768).
//
// A chunk's key of haplotype h is its alleles over the chunk, last marker
// most significant, as NP = C/8 bytes (byte p = markers 8p..8p+7 of the
// chunk, bit k = marker 8p+k).  Sorting the haplotypes by that key, ties by
// index, is the chunk's sweep from the identity order, so phase A is an LSD
// radix sort over the key (2·NP stable 4-bit counting passes instead of C
// stable 1-bit partitions) followed by an exact comparison of adjacent keys
// for the group starts.  Phase B's stable sort by group id uses the same
// pass.  The order stays in registers, thread t holding positions
// [t*K, t*K + K) (two 16-bit ids per register); a pass counts each thread's
// digits in its own LDS counters (so the count is stable without any
// cross-lane matching), scans the digit-major counter array, scatters
// through an LDS exchange buffer and reads the new order back (see the pass
// itself for how the counts are kept).
// ---------------------------------------------------------------------------

// Transpose of a 64 x 64 bit block held one 64-bit row per lane (lane i,
// bit j = entry (i, j)): afterwards lane i holds column i.  Six stages, each
// swapping the off-diagonal sub-blocks of one size between lanes i and i ^ s
// (row index bit s against column index bit s).
__device__ __forceinline__ void wave_transpose64(uint32_t& lo, uint32_t& hi, uint32_t lane) {
    {   // s = 32: lanes below 32 take the partner's low word as their high word
        const bool up = lane >= 32u;
        const uint32_t r = __shfl_xor(up ? lo : hi, 32, kWave);
        if (up) lo = r;
        else hi = r;
    }
#pragma unroll
    for (int k = 4; k >= 0; --k) {
        const uint32_t sh = 1u << k;
        // the columns of each 2s-wide group kept by this lane (low half for a
        // lane with row bit s clear, high half otherwise)
        const uint32_t m = k == 4 ? 0x0000ffffu : k == 3 ? 0x00ff00ffu : k == 2 ? 0x0f0f0f0fu
                         : k == 1 ? 0x33333333u : 0x55555555u;
        const bool up = (lane & sh) != 0u;
        const uint32_t keep = up ? ~m : m;
        const uint32_t ylo = __shfl_xor(lo, int(sh), kWave), yhi = __shfl_xor(hi, int(sh), kWave);
        lo = (lo & keep) | ((up ? (ylo >> sh) : (ylo << sh)) & ~keep);
        hi = (hi & keep) | ((up ? (yhi >> sh) : (yhi << sh)) & ~keep);
    }
}

// Key bytes of a chunk: workgroup (chunk c, 16 column words = 1024
// haplotypes), wave w the 64 haplotypes of word b0 + w.  The chunk's columns
// are staged in LDS 512 markers at a time with coalesced loads (rows padded
// to 17 words, so a lane's row read meets few bank conflicts); a wave then
// transposes each 64-marker block of its word in registers, lane i ending
// with haplotype b*64 + i's 8 key bytes, stored as one 8-byte word of its key
// row and one byte of each of 8 key planes; markers past M count as 0.
constexpr uint32_t kKeyBlock = 512;  // markers staged per round (64 key bytes)
__global__ __launch_bounds__(kPT) void pbwt_keys_kernel(PbwtArgs a) {
    constexpr uint32_t kRow = kPW + 1;  // padded LDS row, in 64-bit words
    __shared__ uint64_t colw[kKeyBlock * kRow];
    const uint32_t c = blockIdx.x;
    const uint32_t t = threadIdx.x, lane = t & 63u;
    const uint32_t w = __builtin_amdgcn_readfirstlane(t >> 6);
    const uint32_t W = a.W2 / 2u, b0 = blockIdx.y * kPW, b = b0 + w;
    const uint64_t* cols = reinterpret_cast<const uint64_t*>(a.cols32);
    const uint32_t h = b * 64u + lane;
    const bool valid = b < W && h < a.H;
    const uint32_t m0 = c * a.C, len = min(a.M, m0 + a.C) - m0;
    uint8_t* tp = a.tp + size_t(c) * a.NP * a.H + (valid ? h : 0u);
    uint8_t* th = a.th + (size_t(c) * a.H + (valid ? h : 0u)) * a.TS;
    for (uint32_t mb = 0; mb < a.TS * 8u; mb += kKeyBlock) {
        __syncthreads();
        for (uint32_t i = t; i < kKeyBlock * kPW; i += kPT) {
            const uint32_t mr = mb + i / kPW, bb = b0 + i % kPW;
            colw[(i / kPW) * kRow + i % kPW] = (mr < len && bb < W) ? cols[size_t(m0 + mr) * W + bb] : 0ull;
        }
        __syncthreads();
        for (uint32_t sb = 0; sb < kKeyBlock / 64u && mb / 8u + 8u * sb < a.TS; ++sb) {
            const uint64_t x = colw[(64u * sb + lane) * kRow + w];  // marker mb + 64 sb + lane
            uint32_t lo = uint32_t(x), hi = uint32_t(x >> 32);
            wave_transpose64(lo, hi, lane);
            const uint32_t p0 = mb / 8u + 8u * sb;  // first key byte of the block
            if (valid) {
                *reinterpret_cast<uint2*>(th + p0) = make_uint2(lo, hi);
#pragma unroll
                for (uint32_t q = 0; q < 8; ++q)
                    if (p0 + q < a.NP) tp[size_t(p0 + q) * a.H] = uint8_t((q < 4 ? lo : hi) >> (8u * (q & 3u)));
            }
        }
    }
}

// The order of a radix kernel: thread t holds positions [t*K, t*K + K),
// two 16-bit haplotype ids per register.
template <int K>
struct BOrd {
    uint32_t v[(K + 1) / 2];
    __device__ __forceinline__ uint32_t get(int j) const { return (v[j >> 1] >> (16 * (j & 1))) & 0xffffu; }
    __device__ __forceinline__ void set(int j, uint32_t x) {
        const uint32_t sh = 16u * uint32_t(j & 1);
        v[j >> 1] = (v[j >> 1] & ~(0xffffu << sh)) | (x << sh);
    }
};

// LDS of a radix pass: per-thread counters of the 16 digits (digit-major,
// [16][kPT] 16-bit) and the exchange buffer (which also carries the 16
// per-wave sums of the scan: it is idle between a pass's read-back and the
// next pass's scatter).
struct R4Lds {
    uint16_t* cnt;
    uint16_t* X;
};

// One stable counting-sort pass of the workgroup's elements by the 4-bit
// digit(h).  Every thread holds K elements: positions past H hold padding
// ids (the position itself) whose digits read as 15, so they sort after
// every haplotype, keep the positions past H and no element needs a range
// check.  The first barrier publishes LDS data the digit function reads,
// written before the call.  If row is set, the new order of positions
// [0, H) is also written there (coalesced, from the exchange buffer).
//
// Each thread counts its own elements' digits in registers: 16 four-bit
// counters in a 64-bit word per group of at most 15 elements (so no counter
// carries into the next), and keeps the digits as packed nibbles.  The
// counts go to LDS once (digit-major), one scan turns them into each
// (digit, thread) pair's first destination, and the scatter adds the
// element's rank among the thread's earlier elements of its digit, read from
// the same counters: no LDS read-modify-write chains and no second digit
// lookup.  Destinations are read in batches before their exchange writes.
template <int K, class Digit>
__device__ __forceinline__ void radix4_pass(BOrd<K>& e, uint32_t H, const R4Lds& L, Digit digit,
                                            uint32_t* row) {
    constexpr int GS = 15;                  // elements per counter group
    constexpr int NA = (K + GS - 1) / GS;   // counter groups
    constexpr int ND = (K + 7) / 8;         // words of packed digits
    constexpr int BT = K > 24 ? 4 : 8;      // scatter batch (fewer registers for long orders)
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    const uint32_t p0 = t * uint32_t(K);
    uint16_t* const cnt = L.cnt;
    uint16_t* const X = L.X;
    uint4* c4 = reinterpret_cast<uint4*>(cnt);  // 2048 uint4
    __syncthreads();
    uint32_t dg[ND];
    uint64_t acc[NA];
#pragma unroll
    for (int k = 0; k < ND; ++k) dg[k] = 0u;
#pragma unroll
    for (int k = 0; k < NA; ++k) acc[k] = 0u;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const uint32_t d = digit(e.get(j));
        dg[j >> 3] |= d << (4 * (j & 7));
        acc[j / GS] += uint64_t(1) << (4u * d);
    }
#pragma unroll
    for (int d = 0; d < 16; ++d) {
        uint32_t n = 0;
#pragma unroll
        for (int k = 0; k < NA; ++k) n += uint32_t(acc[k] >> (4 * d)) & 15u;
        cnt[d * kPT + t] = uint16_t(n);
    }
    __syncthreads();
    // exclusive scan of the digit-major counters: thread t owns entries
    // [16t, 16t + 16) = digit w, threads 16*lane .. 16*lane + 15
    uint32_t* wsum = reinterpret_cast<uint32_t*>(X);
    const uint4 A = c4[2 * t], B = c4[2 * t + 1];
    const uint32_t wd[8] = {A.x, A.y, A.z, A.w, B.x, B.y, B.z, B.w};
    uint32_t ex[16];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        ex[i] = s;
        s += (wd[i >> 1] >> (16 * (i & 1))) & 0xffffu;
    }
    const uint32_t incl = wave_incl_scan(s);
    if (lane == 63u) wsum[w] = incl;
    __syncthreads();
    uint32_t base = incl - s;
#pragma unroll
    for (int v = 0; v < kPW; ++v) {
        const uint32_t x = wsum[v];
        base += uint32_t(v) < w ? x : 0u;
    }
    uint32_t o[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) o[i] = (base + ex[2 * i]) | ((base + ex[2 * i + 1]) << 16);
    c4[2 * t] = make_uint4(o[0], o[1], o[2], o[3]);
    c4[2 * t + 1] = make_uint4(o[4], o[5], o[6], o[7]);
    __syncthreads();
    uint64_t run = 0;  // digits of the current group's earlier elements
#pragma unroll
    for (int j0 = 0; j0 < K; j0 += BT) {
        uint32_t dst[BT];
#pragma unroll
        for (int i = 0; i < BT && j0 + i < K; ++i) {
            const int j = j0 + i;
            const uint32_t d = (dg[j >> 3] >> (4 * (j & 7))) & 15u;
            dst[i] = cnt[d * kPT + t];
        }
#pragma unroll
        for (int i = 0; i < BT && j0 + i < K; ++i) {
            const int j = j0 + i;
            if (j % GS == 0) run = 0u;
            const uint32_t sh = 4u * ((dg[j >> 3] >> (4 * (j & 7))) & 15u);
            uint32_t rk = uint32_t(run >> sh) & 15u;
#pragma unroll
            for (int k = 0; k < j / GS; ++k) rk += uint32_t(acc[k] >> sh) & 15u;
            run += uint64_t(1) << sh;
            X[dst[i] + rk] = uint16_t(e.get(j));
        }
    }
    __syncthreads();
    if (row)
        for (uint32_t i = t; i < H; i += kPT) row[i] = X[i];
#pragma unroll
    for (int j = 0; j < K; ++j) e.set(j, X[p0 + uint32_t(j)]);
}

// Phase A, radix form: one workgroup per chunk.
template <int K>
__global__ __launch_bounds__(kPT) void pbwt_chunk_sort_kernel(PbwtArgs a) {
    using G = LdsGeom<K>;
    __shared__ __align__(16) uint16_t X[G::NPOS];
    __shared__ __align__(16) uint16_t cnt[16 * kPT];
    __shared__ __align__(4) uint8_t dig[G::NPOS];  // the current key byte of each haplotype
    __shared__ uint32_t s_b[kPW];
    const R4Lds L{cnt, X};
    const uint32_t c = blockIdx.x, t = threadIdx.x, lane = t & 63u, w = t >> 6;
    const uint32_t m_first = c * a.C, len = min(a.M, m_first + a.C) - m_first;
    const uint32_t np = (len + 7u) / 8u;
    const uint32_t H = a.H, p0 = t * uint32_t(K);
    BOrd<K> e;
#pragma unroll
    for (int j = 0; j < K; ++j) e.set(j, p0 + uint32_t(j));

    const uint8_t* tpc = a.tp + size_t(c) * a.NP * H;
    const uint32_t nw = (H + 3u) / 4u;  // 32-bit words of a plane
    constexpr uint32_t kPF = (G::NPOS / 4u + kPT - 1) / kPT;
    auto plane_word = [&](uint32_t p, uint32_t i) -> uint32_t {
        // four key bytes; a plane is H bytes, so an unaligned or last word is
        // assembled bytewise, bytes past H reading as 0xff (padding digits)
        const uint8_t* src = tpc + size_t(p) * H + 4u * i;
        if (4u * i + 4u <= H && ((reinterpret_cast<uintptr_t>(src) & 3u) == 0u))
            return *reinterpret_cast<const uint32_t*>(src);
        uint32_t v = 0;
        for (uint32_t k = 0; k < 4u; ++k) v |= (4u * i + k < H ? uint32_t(src[k]) : 0xffu) << (8u * k);
        return v;
    };
    // padding ids' key bytes (words the plane loads below never write)
    for (uint32_t i = nw + t; i < G::NPOS / 4u; i += kPT) reinterpret_cast<uint32_t*>(dig)[i] = ~0u;
    uint32_t pf[kPF];
#pragma unroll
    for (uint32_t k = 0; k < kPF; ++k) {
        const uint32_t i = t + k * kPT;
        pf[k] = i < nw ? plane_word(0, i) : 0u;
    }
    for (uint32_t p = 0; p < np; ++p) {
#pragma unroll
        for (uint32_t k = 0; k < kPF; ++k) {
            const uint32_t i = t + k * kPT;
            if (i < nw) reinterpret_cast<uint32_t*>(dig)[i] = pf[k];
        }
        if (p + 1 < np) {
#pragma unroll
            for (uint32_t k = 0; k < kPF; ++k) {
                const uint32_t i = t + k * kPT;
                pf[k] = i < nw ? plane_word(p + 1, i) : 0u;
            }
        }
        // low nibble, then high nibble (a pass reads its digits before its
        // second barrier, so the buffer is free once the second pass returns)
        radix4_pass<K>(e, H, L, [&](uint32_t h) { return uint32_t(dig[h]) & 15u; }, nullptr);
        radix4_pass<K>(e, H, L, [&](uint32_t h) { return uint32_t(dig[h]) >> 4; }, nullptr);
    }
    // X holds the sorted order by position (also for np == 0, written here)
    if (np == 0) {
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (p0 + uint32_t(j) < H) X[p0 + uint32_t(j)] = uint16_t(e.get(j));
    }
    __syncthreads();
    if (c == 0) {
        uint32_t* lastrow = a.prefix + size_t(m_first + len - 1u) * H;
        for (uint32_t i = t; i < H; i += kPT) lastrow[i] = X[i];
    }
    if (a.ord0) {  // the chunk's sorted order, for the composition scan
        uint16_t* o = a.ord0 + size_t(c) * H;
        for (uint32_t i = t; i < H; i += kPT) o[i] = X[i];
    }
    // group starts: a key differs from its predecessor's (exact, all bytes)
    const uint8_t* thc = a.th + size_t(c) * H * a.TS;
    uint32_t flags = 0, cntf = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const uint32_t pos = p0 + uint32_t(j);
        if (pos >= H) continue;
        bool start = pos == 0u;
        if (!start) {
            const uint4* ka = reinterpret_cast<const uint4*>(thc + size_t(e.get(j)) * a.TS);
            const uint4* kb = reinterpret_cast<const uint4*>(thc + size_t(X[pos - 1u]) * a.TS);
            for (uint32_t u = 0; u < a.TS / 16u && !start; ++u) {
                const uint4 x = ka[u], y = kb[u];
                start = (x.x != y.x) || (x.y != y.y) || (x.z != y.z) || (x.w != y.w);
            }
        }
        flags |= uint32_t(start) << j;
        cntf += uint32_t(start);
    }
    // dense rank of each group in position order: thread counts, block scan
    const uint32_t incl = wave_incl_scan(cntf);
    if (lane == 63u) s_b[w] = incl;
    __syncthreads();
    uint32_t rk = incl - cntf, Gn = 0;
#pragma unroll
    for (int v = 0; v < kPW; ++v) {
        rk += uint32_t(v) < w ? s_b[v] : 0u;
        Gn += s_b[v];
    }
    uint32_t* gid = a.gid + size_t(c) * H;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (p0 + uint32_t(j) >= H) continue;
        rk += (flags >> j) & 1u;
        gid[e.get(j)] = rk - 1u;
    }
    if (t == 0) a.ngroups[c] = Gn;
}

// Radix passes (4 bits each) needed for ids below Gn.
__device__ __forceinline__ uint32_t id_passes(uint32_t Gn) {
    uint32_t n = 0;
    while (Gn > 1u && n < 4u && ((Gn - 1u) >> (4u * n)) != 0u) ++n;
    return n;
}

// Boundary row of chunk c (its last marker).
__device__ __forceinline__ uint32_t* chunk_end_row(const PbwtArgs& a, uint32_t c) {
    return a.prefix + size_t(min(a.M, c * a.C + a.C) - 1u) * a.H;
}

// Load a row into the blocked order.
template <int K>
__device__ __forceinline__ void load_order(BOrd<K>& e, const uint32_t* row, uint32_t H) {
    const uint32_t p0 = threadIdx.x * uint32_t(K);
#pragma unroll
    for (int j = 0; j < K; ++j) e.set(j, p0 + uint32_t(j) < H ? row[p0 + uint32_t(j)] : 0u);
}

// Sort the blocked order e stably by ids (global [H], below Gn), through the
// LDS id array gl; the result also goes to row.
template <int K>
__device__ __forceinline__ void sort_by_ids(BOrd<K>& e, const uint32_t* ids, uint32_t Gn, uint32_t H,
                                            const R4Lds& L, uint16_t* gl, uint32_t* row) {
    // ids of the padding past H read as 0xffff (digit 15 in every pass)
    for (uint32_t i = threadIdx.x; i < LdsGeom<K>::NPOS; i += kPT) gl[i] = i < H ? uint16_t(ids[i]) : uint16_t(0xffffu);
    const uint32_t np = id_passes(Gn);
    for (uint32_t pass = 0; pass < np; ++pass) {
        const uint32_t sh = 4u * pass;
        radix4_pass<K>(e, H, L, [&](uint32_t h) { return (uint32_t(gl[h]) >> sh) & 15u; },
                       pass + 1 == np ? row : nullptr);
    }
    if (np == 0) {  // one group: the order is unchanged
        __syncthreads();
        const uint32_t p0 = threadIdx.x * uint32_t(K);
        if (row) {
#pragma unroll
            for (int j = 0; j < K; ++j)
                if (p0 + uint32_t(j) < H) row[p0 + uint32_t(j)] = e.get(j);
        }
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (p0 + uint32_t(j) < H) L.X[p0 + uint32_t(j)] = uint16_t(e.get(j));
        __syncthreads();
    }
}


// ---------------------------------------------------------------------------
// Phase B as a prefix scan over the chunks (radix form).  With R_k[c] the
// dense rank of each haplotype's key over chunks [c - 2^k + 1, c] (clamped
// at 0) and O_k[c] the haplotypes sorted by (R_k[c], index), level k with
// d = 2^k composes, for every chunk c >= d at once,
//   O_{k+1}[c] = O_k[c - d] stably sorted by R_k[c],
//   R_{k+1}[c] = dense rank of (R_k[c], R_k[c - d]) along O_{k+1}[c],
// since the later chunks' key is the more significant; chunks c < d are
// complete and are copied.  After ceil(log2 n) levels every chunk's range
// starts at marker 0, and O[c] is the true order at its last marker, which
// the level that completes it (d <= c < 2d) writes to that prefix row.  The
// serial depth is log2(n) sorts, each level one launch of n workgroups.
// ---------------------------------------------------------------------------
template <int K>
__global__ __launch_bounds__(kPT) void pbwt_scan_kernel(PbwtArgs a, uint32_t d, const uint32_t* Rin,
                                                        uint32_t* Rout, const uint16_t* Oin, uint16_t* Oout,
                                                        const uint32_t* Gin, uint32_t* Gout) {
    using G = LdsGeom<K>;
    // K <= 20: the earlier chunk's ranks staged in LDS (coalesced, beside the
    // sort) and the new ranks gathered there before one coalesced store,
    // instead of random global loads and stores; larger orders do not fit
    constexpr bool STAGE = K <= 20;
    __shared__ __align__(16) uint16_t X[G::NPOS];
    __shared__ __align__(16) uint16_t cnt[16 * kPT];
    __shared__ __align__(4) uint16_t gl[G::NPOS];
    __shared__ __align__(4) uint16_t lo16[STAGE ? G::NPOS : 2];
    const R4Lds L{cnt, X};
    uint32_t* s_b = reinterpret_cast<uint32_t*>(cnt);  // free after the sort
    const uint32_t c = blockIdx.x, t = threadIdx.x, lane = t & 63u, w = t >> 6, H = a.H;
    const uint32_t p0 = t * uint32_t(K);
    if (c < d) {  // complete: carried to the next level's buffers
        for (uint32_t i = t; i < H; i += kPT) {
            Rout[size_t(c) * H + i] = Rin[size_t(c) * H + i];
            Oout[size_t(c) * H + i] = Oin[size_t(c) * H + i];
        }
        if (t == 0) Gout[c] = Gin[c];
        return;
    }
    BOrd<K> e;
    const uint16_t* lo_ord = Oin + size_t(c - d) * H;
    const uint32_t* lo = Rin + size_t(c - d) * H;
#pragma unroll
    for (int j = 0; j < K; ++j)
        e.set(j, p0 + uint32_t(j) < H ? uint32_t(lo_ord[p0 + uint32_t(j)]) : p0 + uint32_t(j));  // padding ids past H
    if constexpr (STAGE)
        for (uint32_t i = t; i < H; i += kPT) lo16[i] = uint16_t(lo[i]);  // published by the sort's barriers
    const bool done = c < 2u * d;
    sort_by_ids<K>(e, Rin + size_t(c) * H, Gin[c], H, L, gl, done ? chunk_end_row(a, c) : nullptr);
    // X: the new order by position; gl: R_k[c] of each haplotype
    uint32_t flags = 0, cntf = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const uint32_t pos = p0 + uint32_t(j);
        if (pos >= H) continue;
        const uint32_t h = e.get(j);
        bool start = pos == 0u;
        if (!start) {
            const uint32_t hp = X[pos - 1u];
            if constexpr (STAGE) start = gl[h] != gl[hp] || lo16[h] != lo16[hp];
            else start = gl[h] != gl[hp] || lo[h] != lo[hp];
        }
        flags |= uint32_t(start) << j;
        cntf += uint32_t(start);
    }
    const uint32_t incl = wave_incl_scan(cntf);
    if (lane == 63u) s_b[w] = incl;
    __syncthreads();
    uint32_t rk = incl - cntf, Gn = 0;
#pragma unroll
    for (int v = 0; v < kPW; ++v) {
        rk += uint32_t(v) < w ? s_b[v] : 0u;
        Gn += s_b[v];
    }
    uint32_t* out = Rout + size_t(c) * H;
    uint16_t* oo = Oout + size_t(c) * H;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const uint32_t pos = p0 + uint32_t(j);
        if (pos >= H) continue;
        rk += (flags >> j) & 1u;
        if constexpr (STAGE) gl[e.get(j)] = uint16_t(rk - 1u);  // every read of gl ended at the scan's barrier
        else out[e.get(j)] = rk - 1u;
        oo[pos] = uint16_t(e.get(j));
    }
    if constexpr (STAGE) {
        __syncthreads();
        for (uint32_t i = t; i < H; i += kPT) out[i] = gl[i];
    }
    if (t == 0) Gout[c] = Gn;
}

// ---------------------------------------------------------------------------
// Divergence rows alone (row m = m everywhere), for the radix form, which
// runs it on a second stream beside the sort and composition launches: those
// leave most of the HBM bandwidth idle.  Workgroups stride over rows; 16-byte
// stores when rows are 16-byte aligned.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pbwt_div_fill_kernel(uint32_t* __restrict__ div, uint32_t H,
                                                            uint32_t M) {
    const bool vec = (H & 3u) == 0u;
    for (uint32_t m = blockIdx.x; m < M; m += gridDim.x) {
        uint32_t* row = div + size_t(m) * H;
        if (vec) {
            typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
            const u32x4 v = {m, m, m, m};
            u32x4* r4 = reinterpret_cast<u32x4*>(row);
            for (uint32_t i = threadIdx.x; i < H / 4u; i += blockDim.x) __builtin_nontemporal_store(v, r4 + i);
        } else {
            for (uint32_t i = threadIdx.x; i < H; i += blockDim.x) __builtin_nontemporal_store(m, row + i);
        }
    }
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
namespace {

// A second stream per (host thread, device) for work that runs beside a
// call's main launches; created on first use with the default priority.
struct SideStreams {
    hipStream_t s[64] = {};
    ~SideStreams() {
        for (hipStream_t x : s)
            if (x) (void)hipStreamDestroy(x);
    }
};

hipStream_t side_stream() {
    thread_local SideStreams t;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    if (!t.s[dev] && hipStreamCreateWithFlags(&t.s[dev], hipStreamNonBlocking) != hipSuccess)
        t.s[dev] = nullptr;
    return t.s[dev];
}

// Runs the divergence fill on the side stream, forked from and joined back
// into `st` with events (so it also works under graph capture).  Returns
// false if the fork could not be set up (the caller then fills in line).
// Compute units of the current device (memoized per device).
uint32_t device_cus() {
    static std::atomic<uint32_t> cache[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 256;
    if (dev < 64) {
        const uint32_t c = cache[dev].load(std::memory_order_relaxed);
        if (c) return c;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    if (dev < 64) cache[dev].store(uint32_t(cus), std::memory_order_relaxed);
    return uint32_t(cus);
}

bool div_fill_beside(const PbwtArgs& a, hipStream_t st, hipEvent_t* join) {
    hipStream_t side = side_stream();
    hipEvent_t fork = nullptr;
    if (!side || hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess) return false;
    if (hipEventCreateWithFlags(join, hipEventDisableTiming) != hipSuccess) {
        (void)hipEventDestroy(fork);
        return false;
    }
    bool ok = hipEventRecord(fork, st) == hipSuccess && hipStreamWaitEvent(side, fork, 0) == hipSuccess;
    if (ok) {
        // an eighth of the CUs, the ones the chunk launches leave free (the
        // default chunk length keeps them to seven eighths): enough stores in
        // flight to finish during the sort and composition, while those
        // launches still find room (a fill over every CU makes them wait)
        const uint32_t cus = device_cus();
        const uint32_t gmax = cus >= 16 ? cus / 8u : 2u;
        const uint32_t grid = a.M < gmax ? a.M : gmax;
        hipLaunchKernelGGL(pbwt_div_fill_kernel, dim3(grid), dim3(256), 0, side, a.div, a.H, a.M);
        ok = hipGetLastError() == hipSuccess && hipEventRecord(*join, side) == hipSuccess;
    }
    (void)hipEventDestroy(fork);
    if (!ok) {
        (void)hipEventDestroy(*join);
        *join = nullptr;
    }
    return ok;
}

// Rounds per wave K (16 * 64 * K >= H): the smallest instantiated value.
constexpr uint32_t kRounds[] = {1, 2, 4, 6, 8, 12, 16, 20, 24, 32, 40, 48, 64};

uint32_t pbwt_rounds(uint32_t H) {
    const uint32_t need = (H + kPT - 1) / kPT;
    for (uint32_t k : kRounds)
        if (k >= need) return k;
    return 0;
}
// Rounds of the global-memory form, which takes any K (H above 65536 too).
uint32_t pbwt_rounds_any(uint32_t H) {
    const uint32_t k = pbwt_rounds(H);
    return k ? k : (H + kPT - 1) / kPT;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// H <= 49152: phase A and B in the partition form (A also writes the
// divergence rows), phase C with the order in LDS.
template <int K>
void launch_lds(const PbwtArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(pbwt_groups_lds_kernel<K>, dim3(a.nchunk), dim3(kPT), 0, st, a);
    if (a.nchunk > 1) hipLaunchKernelGGL(pbwt_compose_kernel, dim3(1), dim3(kPT), 0, st, a);
    hipLaunchKernelGGL(pbwt_rows_lds_kernel<K>, dim3(a.nchunk), dim3(kPT), 0, st, a);
}

// H <= 32768: key bytes, phase A and B in the radix form, divergence rows on a
// second stream, phase C with the order in LDS.
constexpr uint32_t kRadixMaxK = 32;
template <int K>
void launch_radix(PbwtArgs a, hipStream_t st) {
    const uint32_t W = a.W2 / 2u;
    hipEvent_t join = nullptr;
    hipLaunchKernelGGL(pbwt_keys_kernel, dim3(a.nchunk, (W + kPW - 1) / kPW), dim3(kPT), 0, st, a);
    // the divergence rows are written beside the sort and composition, from
    // the end of the key launch on (fill in the row re-sweep if that fails)
    if (a.div) a.rows_div = div_fill_beside(a, st, &join) ? 0u : 1u;
    hipLaunchKernelGGL(pbwt_chunk_sort_kernel<K>, dim3(a.nchunk), dim3(kPT), 0, st, a);
    // the composition scan: rank and order buffers alternate between levels
    // (level 0 reads the chunk sort's ids, orders and group counts)
    const size_t nh = size_t(a.nchunk) * a.H;
    const uint32_t* Rin = a.gid;
    const uint16_t* Oin = a.ord0;
    const uint32_t* Gin = a.ngroups;
    uint32_t* Rb[2] = {a.xa, a.xa + nh};
    uint16_t* Ob[2] = {a.ord1, a.ord0};
    uint32_t* Gb[2] = {a.gcnt, a.gcnt + a.nchunk};
    uint32_t lv = 0;
    for (uint32_t d = 1; d < a.nchunk; d *= 2, ++lv) {
        hipLaunchKernelGGL(pbwt_scan_kernel<K>, dim3(a.nchunk), dim3(kPT), 0, st, a, d, Rin, Rb[lv & 1u], Oin,
                           Ob[lv & 1u], Gin, Gb[lv & 1u]);
        Rin = Rb[lv & 1u];
        Oin = Ob[lv & 1u];
        Gin = Gb[lv & 1u];
    }
    hipLaunchKernelGGL(pbwt_rows_lds_kernel<K>, dim3(a.nchunk), dim3(kPT), 0, st, a);
    if (join) {
        (void)hipStreamWaitEvent(st, join, 0);
        (void)hipEventDestroy(join);
    }
}

size_t key_bytes(uint32_t H, uint32_t M, uint32_t C) {
    const size_t n = (size_t(M) + C - 1) / C;
    const size_t np = (C + 7u) / 8u, ts = (np + 15u) & ~size_t(15);
    return align256(n * np * H) + align256(n * H * ts);
}

}  // namespace

uint32_t pbwt_default_chunk(uint32_t H, uint32_t M) {
    const uint32_t K = pbwt_rounds(H);
    if (K == 0 || K > kRadixMaxK) return 512;  // partition form: its composition walks the chunks in turn
    // radix form: one round of chunk workgroups (one per CU) on at most seven
    // eighths of the CUs, the rest left to the divergence fill; every phase-A
    // and phase-C step is serial within a chunk, so chunks are as short as
    // that allows, rounded up to a multiple of 128 markers beyond 64 (a
    // chunk's key is then a whole number of 16-byte words)
    const uint32_t cus = device_cus();
    const uint32_t slots = cus >= 16 ? cus - cus / 8u : 1u;
    uint32_t c = (M + slots - 1u) / slots;
    if (c > 64u) c = (c + 127u) / 128u * 128u;
    return c < 16u ? 16u : c;
}

size_t pbwt_workspace_bytes(uint32_t H, uint32_t M, uint32_t C) {
    if (H == 0 || M == 0 || C == 0) return 0;
    const size_t n = (size_t(M) + C - 1) / C, h = H;
    const uint32_t K = pbwt_rounds_any(H);
    const size_t npos = size_t(kPW) * 64u * K;  // padded exchange length (64-bit words)
    const size_t base = align256(n * 2 * h * 4) + align256(n * h * 4) + align256(n * 4) +
                        align256(2 * npos * 8);
    return base + (K != 0 && K <= kRadixMaxK ? key_bytes(H, M, C) + align256(2 * n * h * 2) + align256(2 * n * 4)
                                             : 0);
}

int launch_pbwt_sweep(const uint64_t* cols, uint32_t H, uint32_t M, uint32_t* prefix,
                      uint32_t* div, uint32_t C, void* ws, size_t ws_bytes, hipStream_t stream) {
    if (H == 0 || M == 0) return 0;
    if (H > kPbwtMaxH || C == 0) return SI_ERR_UNSUPPORTED;
    if (ws_bytes < pbwt_workspace_bytes(H, M, C)) return SI_ERR_WORKSPACE;
    const uint32_t K = pbwt_rounds_any(H);
    PbwtArgs a{};
    a.cols32 = reinterpret_cast<const uint32_t*>(cols);
    a.H = H;
    a.M = M;
    a.W2 = 2 * ((H + 63) / 64);
    a.C = C;
    a.nchunk = (M + C - 1) / C;
    a.S = 64 * K;
    a.prefix = prefix;
    a.div = div;
    const size_t n = a.nchunk, h = H;
    char* p = static_cast<char*>(ws);
    a.xa = reinterpret_cast<uint32_t*>(p);
    p += align256(n * 2 * h * 4);
    a.gid = reinterpret_cast<uint32_t*>(p);
    p += align256(n * h * 4);
    a.ngroups = reinterpret_cast<uint32_t*>(p);
    p += align256(n * 4);
    a.xb = reinterpret_cast<uint32_t*>(p);  // 64-bit element words in the boundary composition
    p += align256(2 * size_t(kPW) * 64u * K * 8);
    a.NP = (C + 7u) / 8u;
    a.TS = (a.NP + 15u) & ~15u;
    if (K <= kRadixMaxK) {
        a.tp = reinterpret_cast<uint8_t*>(p);
        p += align256(n * a.NP * h);
        a.th = reinterpret_cast<uint8_t*>(p);
        p += align256(n * h * a.TS);
        a.ord0 = reinterpret_cast<uint16_t*>(p);
        a.ord1 = a.ord0 + n * h;
        p += align256(2 * n * h * 2);
        a.gcnt = reinterpret_cast<uint32_t*>(p);
        a.rows_div = 1;
    }
    switch (K) {
        case 1: launch_radix<1>(a, stream); break;
        case 2: launch_radix<2>(a, stream); break;
        case 4: launch_radix<4>(a, stream); break;
        case 6: launch_radix<6>(a, stream); break;
        case 8: launch_radix<8>(a, stream); break;
        case 12: launch_radix<12>(a, stream); break;
        case 16: launch_radix<16>(a, stream); break;
        case 20: launch_radix<20>(a, stream); break;
        case 24: launch_radix<24>(a, stream); break;
        case 32: launch_radix<32>(a, stream); break;
        case 40: launch_lds<40>(a, stream); break;
        case 48: launch_lds<48>(a, stream); break;
        default:  // H > 49152: the order exchanges through global memory (K = ceil(H / 1024))
            hipLaunchKernelGGL(pbwt_groups_kernel, dim3(a.nchunk), dim3(kPT), 0, stream, a);
            if (a.nchunk > 1)
                hipLaunchKernelGGL(pbwt_compose_kernel, dim3(1), dim3(kPT), 0, stream, a);
            hipLaunchKernelGGL(pbwt_rows_kernel, dim3(a.nchunk), dim3(kPT), 0, stream, a);
    }
    return hipGetLastError() == hipSuccess ? 0 : SI_ERR_DEVICE;
}

}  // namespace si
