// PBWT prefix arrays for gfx950: a chunked three-phase construction.
//
// Reference (PBWTBuilder::build_marker, src/pbwt/pbwt_index.cpp:153-208): the
// order after marker m is the stable partition of the order after m - 1 by
// the allele at m (allele == 0 first, every nonzero allele after); row m is
// the order after marker m, and every divergence in row m equals m.  So row m
// sorts the haplotypes by (x_m, x_{m-1}, ..., x_0) lexicographically, ties by
// haplotype index.
//
// The markers are cut into chunks of C.  Three launches:
//   A  one workgroup per chunk, all chunks at once: sweep the chunk from the
//      identity order, carrying per position a flag "starts a group", a group
//      being the haplotypes with equal alleles over the chunk so far.  At the
//      chunk end a haplotype's group id is the dense rank of its chunk key.
//      Chunk 0's final order is the true order at its last marker and goes to
//      that prefix row.  A also streams out the constant divergence rows.
//   B  one workgroup, serial over chunk boundaries: the order at the end of
//      chunk c is the stable sort, by chunk-c group id, of the order at the end
//      of chunk c - 1 (LSD radix over 8-bit digits of the id), written to the
//      chunk's last prefix row.
//   C  one workgroup per chunk: re-sweep the chunk from its true starting
//      order (identity for chunk 0, else the previous chunk's last row),
//      writing every row of the chunk but its last.
//
// Inside a workgroup (16 wavefronts) wave w owns the positions
// [w*S, (w+1)*S), S = 64*K; in round j < K lane l holds position
// w*S + 64*j + l, so every load and store of a round is one contiguous
// 256-byte run.  A stable partition needs, per element, the count of zeros
// before it: a ballot and mbcnt within the round, a running count across
// rounds and one LDS exchange of the 16 per-wave totals across waves.
// H <= 65536, so a haplotype fits in 16 bits of an element word.
#include "si_device.hpp"
#include "si_internal.hpp"

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
    uint32_t* xb;            // [2][H] phase B exchange
};

__device__ __forceinline__ uint32_t mbcnt64(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
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
        ColPrefetch nx;
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
        ColPrefetch nx;
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
            const uint32_t x = allele_bit(cb, v & 0xffffu);
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
            const uint32_t x = allele_bit(cb, v & 0xffffu);
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
            if (valid) Xout[dst] = (v & 0xffffu) | (first << 31);
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
            gid[v & 0xffffu] = bbefore + inc - 1u;
            if (c == 0) lastrow[q.base + j * 64u + q.lane] = v & 0xffffu;
        }
    }
    if (q.t == 0) a.ngroups[c] = G;
}

// ---------------------------------------------------------------------------
// Phase B: chunk-boundary orders, one workgroup, boundaries in sequence.
// Element word: haplotype | group id << 16; LSD radix over 8-bit digits of
// the id.  Lanes of a round with the same digit agree on all 8 digit-bit
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
            uint32_t* X0 = a.xb + npos;
            for (uint32_t j0 = 0; j0 < K; j0 += kBatch) {
                uint32_t h[kBatch];
#pragma unroll
                for (uint32_t u = 0; u < kBatch; ++u) h[u] = prev[min(own + (j0 + u) * 64u, q.last)];
#pragma unroll
                for (uint32_t u = 0; u < kBatch; ++u) h[u] |= gid[h[u] & 0xffffu] << 16;
#pragma unroll
                for (uint32_t u = 0; u < kBatch; ++u)
                    if (j0 + u < K) X0[own + (j0 + u) * 64u] = h[u];
            }
            __syncthreads();
        }
        for (uint32_t pass = 0; pass < npass; ++pass) {
            const uint32_t sh = 16u + 8u * pass;
            const uint32_t* X = a.xb + ((pass + 1u) & 1u) * npos;  // this pass's input
            const bool last = pass + 1 == npass;
            uint32_t* Y = a.xb + (pass & 1u) * npos;
            for (uint32_t i = q.t; i < uint32_t(kPW) * 256u; i += kPT) (&cnt[0][0])[i] = 0u;
            __syncthreads();
            for (uint32_t j0 = 0; j0 < K; j0 += kBatch) {
                uint32_t v[kBatch];
#pragma unroll
                for (uint32_t u = 0; u < kBatch; ++u) v[u] = X[own + min(j0 + u, K - 1u) * 64u];
#pragma unroll
                for (uint32_t u = 0; u < kBatch; ++u) {
                    if (j0 + u < K) {
                        const bool valid = q.valid(j0 + u);
                        const uint32_t d = (v[u] >> sh) & 255u;
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
                uint32_t v[kBatch];
#pragma unroll
                for (uint32_t u = 0; u < kBatch; ++u) v[u] = X[own + min(j0 + u, K - 1u) * 64u];
#pragma unroll
                for (uint32_t u = 0; u < kBatch; ++u) {
                    if (j0 + u < K) {
                        const bool valid = q.valid(j0 + u);
                        const uint32_t d = (v[u] >> sh) & 255u;
                        const uint64_t match = digit_match(d, valid);
                        const uint32_t c0 = cnt[q.w][d];
                        const uint32_t dst = dbase[d] + c0 + mbcnt64(match);
                        if (valid) {
                            if (last) out[dst] = v[u] & 0xffffu;
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

template <int K>
__global__ __launch_bounds__(kPT) void pbwt_rows_lds_kernel(PbwtArgs a) {
    using G = LdsGeom<K>;
    __shared__ uint16_t ord[G::NPOS];
    __shared__ uint32_t col[2][G::NW32];
    __shared__ uint32_t wz[kPW];
    const uint32_t c = blockIdx.x;
    const uint32_t m_first = c * a.C, m_end = min(a.M, m_first + a.C);
    if (m_end - m_first < 2) return;
    const WaveCtx q(a);

    const uint32_t* start = (c == 0) ? nullptr : a.prefix + size_t(m_first - 1) * a.H;
    for (uint32_t i = q.t; i < G::NPOS; i += kPT)
        ord[i] = uint16_t(i < a.H ? (start ? start[i] : i) : 0u);
    load_column(a, m_first, col[0]);
    __syncthreads();
    for (uint32_t m = m_first; m + 1 < m_end; ++m) {
        const uint32_t* cb = col[(m - m_first) & 1u];
        const bool pre = m + 2 < m_end;
        ColPrefetch nx;
        if (pre) nx.fetch(a, m + 1);

        uint32_t pk[K];
#pragma unroll
        for (int j = 0; j < K; ++j) pk[j] = ord[q.base + uint32_t(j) * 64u + q.lane];
        uint32_t xb[K];
#pragma unroll
        for (int j = 0; j < K; ++j) xb[j] = allele_bit(cb, pk[j]);
        uint32_t zsum = 0;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const uint32_t x = xb[j] | uint32_t(!q.valid(uint32_t(j)));
            const uint64_t Z = __ballot(x == 0u);
            pk[j] |= ((zsum + mbcnt64(Z)) << 16) | (x << 31);
            zsum += uint32_t(__popcll(Z));
        }
        if (q.lane == 0) wz[q.w] = zsum;
        __syncthreads();
        uint32_t zbefore, ztot;
        zero_offsets(wz, q.w, zbefore, ztot);
        uint32_t* row = a.prefix + size_t(m) * a.H;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const uint32_t e = uint32_t(j) * 64u + q.lane;
            const uint32_t v = pk[j], zb = (v >> 16) & 0x1fffu, h = v & 0xffffu;
            const uint32_t dst = (v >> 31) ? ztot + (q.base - zbefore) + (e - zb) : zbefore + zb;
            if (e < q.nvalid) {
                ord[dst] = uint16_t(h);
                row[dst] = h;
            }
        }
        if (pre) nx.store(a, col[((m - m_first) & 1u) ^ 1u]);
        __syncthreads();
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
// Host side.
// ---------------------------------------------------------------------------
namespace {

// Rounds per wave K (16 * 64 * K >= H): the smallest instantiated value.
constexpr uint32_t kRounds[] = {1, 2, 4, 6, 8, 12, 16, 20, 24, 32, 40, 48, 64};

uint32_t pbwt_rounds(uint32_t H) {
    const uint32_t need = (H + kPT - 1) / kPT;
    for (uint32_t k : kRounds)
        if (k >= need) return k;
    return 0;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

template <int K>
void launch_lds(const PbwtArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(pbwt_groups_lds_kernel<K>, dim3(a.nchunk), dim3(kPT), 0, st, a);
    if (a.nchunk > 1) hipLaunchKernelGGL(pbwt_compose_kernel, dim3(1), dim3(kPT), 0, st, a);
    hipLaunchKernelGGL(pbwt_rows_lds_kernel<K>, dim3(a.nchunk), dim3(kPT), 0, st, a);
}

}  // namespace

uint32_t pbwt_default_chunk(uint32_t H, uint32_t M) {
    (void)H;
    (void)M;
    return 512;
}

size_t pbwt_workspace_bytes(uint32_t H, uint32_t M, uint32_t C) {
    if (H == 0 || M == 0 || C == 0) return 0;
    const size_t n = (size_t(M) + C - 1) / C, h = H;
    const size_t npos = size_t(kPW) * 64u * pbwt_rounds(H);  // padded exchange length
    return align256(n * 2 * h * 4) + align256(n * h * 4) + align256(n * 4) + align256(2 * npos * 4);
}

int launch_pbwt_sweep(const uint64_t* cols, uint32_t H, uint32_t M, uint32_t* prefix,
                      uint32_t* div, uint32_t C, void* ws, size_t ws_bytes, hipStream_t stream) {
    if (H == 0 || M == 0) return 0;
    if (H > kPbwtMaxH || C == 0) return SI_ERR_UNSUPPORTED;
    if (ws_bytes < pbwt_workspace_bytes(H, M, C)) return SI_ERR_WORKSPACE;
    const uint32_t K = pbwt_rounds(H);
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
    a.xb = reinterpret_cast<uint32_t*>(p);
    switch (K) {
        case 1: launch_lds<1>(a, stream); break;
        case 2: launch_lds<2>(a, stream); break;
        case 4: launch_lds<4>(a, stream); break;
        case 6: launch_lds<6>(a, stream); break;
        case 8: launch_lds<8>(a, stream); break;
        case 12: launch_lds<12>(a, stream); break;
        case 16: launch_lds<16>(a, stream); break;
        case 20: launch_lds<20>(a, stream); break;
        case 24: launch_lds<24>(a, stream); break;
        case 32: launch_lds<32>(a, stream); break;
        case 40: launch_lds<40>(a, stream); break;
        case 48: launch_lds<48>(a, stream); break;
        default:  // H > 49152: the order exchanges through global memory
            hipLaunchKernelGGL(pbwt_groups_kernel, dim3(a.nchunk), dim3(kPT), 0, stream, a);
            if (a.nchunk > 1)
                hipLaunchKernelGGL(pbwt_compose_kernel, dim3(1), dim3(kPT), 0, stream, a);
            hipLaunchKernelGGL(pbwt_rows_kernel, dim3(a.nchunk), dim3(kPT), 0, stream, a);
    }
    return hipGetLastError() == hipSuccess ? 0 : SI_ERR_DEVICE;
}

}  // namespace si
