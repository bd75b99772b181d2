// Host side of the forward-backward kernels (fb_kernels.hpp, instantiated per
// lane shape in fb_inst.hip): the kernel registry, workspace layout, job
// views, variants and the plain and batched launch sequences.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "fb_kernels.hpp"

namespace si {

// ------------------------------------------------------------- registry --
namespace {
std::mutex& reg_mutex() {
    static std::mutex m;
    return m;
}
std::vector<FbImpl>& registry() {
    static std::vector<FbImpl> r;
    return r;
}
}  // namespace

void fb_register(const FbImpl& impl) {
    std::lock_guard<std::mutex> lk(reg_mutex());
    registry().push_back(impl);
}

const FbImpl* fb_impl(uint32_t Lp, uint32_t G) {
    std::lock_guard<std::mutex> lk(reg_mutex());
    for (const FbImpl& f : registry())
        if (f.Lp == Lp && f.G == G) return &f;
    return nullptr;
}

// ------------------------------------------------------------- variants --
// Default: four states per lane, 16-marker checkpoint blocks.
FbVariant fb_default_variant(uint32_t Lp) { return FbVariant{Lp / 4, Lp / 4, 16}; }

// Chain pass: 8, 4 or 2 states per lane; block pass: 4 or 2 (its LDS rows
// and sampling phase are laid out for those).  Only registered shapes count.
uint32_t fb_candidates(uint32_t Lp, int pass, uint32_t* lanes, uint32_t max) {
    uint32_t n = 0;
    const uint32_t Rs[3] = {8, 4, 2};
    for (uint32_t R : Rs) {
        if (pass == 1 && R > 4) continue;
        if (R > Lp || Lp / R > 64) continue;
        const FbImpl* f = fb_impl(Lp, Lp / R);
        if (!f || (pass == 1 && !f->has_block)) continue;
        if (n < max) lanes[n] = Lp / R;
        ++n;
    }
    return n;
}

bool fb_variant_valid(uint32_t Lp, const FbVariant& v) {
    if (v.cb != 16) return false;  // 16-marker checkpoint blocks only
    const FbImpl* c = fb_impl(Lp, v.gc);
    const FbImpl* b = fb_impl(Lp, v.gb);
    return c && b && b->has_block;
}

// ------------------------------------------------------------ workspace --
namespace {
size_t align256(size_t n) { return (n + 255) & ~size_t(255); }
}  // namespace

// Checkpoint vectors (forward, backward), backward scales, with windows one
// forward vector per window, the typed-marker mask and list and, for banded
// jobs, the bands of the checkpoint vectors, each region 256-byte aligned.
// Sized at the smallest checkpoint block, so one workspace serves every
// variant.
size_t fb_workspace_bytes_one(uint32_t B, uint32_t M, uint32_t Lp, uint32_t W, bool band) {
    const size_t nblk = (size_t(M) + kMinCkptBlock - 1) / kMinCkptBlock;
    const size_t vec = align256(size_t(B) * nblk * Lp * sizeof(float));
    const size_t bl = align256(size_t(B) * nblk * sizeof(double));
    const size_t nwin = W ? (size_t(M) + W - 1) / W : 0;
    const size_t mask = align256((size_t(M) + 15) / 16 * sizeof(uint32_t));
    const size_t list = align256((size_t(M) + 1) * sizeof(uint32_t));
    const size_t bands = band ? 2 * align256(size_t(B) * nblk * Lp) : 0;
    return 2 * vec + bl + align256(size_t(B) * nwin * Lp * sizeof(float)) + mask + list + bands;
}

void fb_bind_workspace(JobView& v, uint32_t Lp, char* ws) {
    const size_t nblk = (size_t(v.M) + kMinCkptBlock - 1) / kMinCkptBlock;
    const size_t vec = align256(size_t(v.B) * nblk * Lp * sizeof(float));
    const size_t bl = align256(size_t(v.B) * nblk * sizeof(double));
    const size_t win = align256(size_t(v.B) * v.nwin * Lp * sizeof(float));
    v.ck_alpha = reinterpret_cast<float*>(ws);
    v.ck_beta = reinterpret_cast<float*>(ws + vec);
    v.ck_bl = reinterpret_cast<double*>(ws + 2 * vec);
    v.ck_win = v.nwin ? reinterpret_cast<float*>(ws + 2 * vec + bl) : nullptr;
    const size_t mask = align256((size_t(v.M) + 15) / 16 * sizeof(uint32_t));
    v.tmask = reinterpret_cast<uint32_t*>(ws + 2 * vec + bl + win);
    v.tcount = reinterpret_cast<uint32_t*>(ws + 2 * vec + bl + win + mask);
    v.tlist = v.tcount + 1;
    if ((v.flags & SI_FB_BANDED) && (v.flags & SI_FB_LOGSPACE) && v.W == 0u) {
        const size_t list = align256((size_t(v.M) + 1) * sizeof(uint32_t));
        const size_t bnd = align256(size_t(v.B) * nblk * Lp);
        v.ck_aband = reinterpret_cast<uint8_t*>(ws + 2 * vec + bl + win + mask + list);
        v.ck_bband = v.ck_aband + bnd;
    } else {
        v.ck_aband = v.ck_bband = nullptr;
    }
}

// ------------------------------------------------------ typed-marker mask --
// Before a fast job's chain pass: its mask of typed markers per load block,
// then the list of those markers.
__global__ __launch_bounds__(kMaskThreads) void fb_tmask_kernel(JobView J, FbTables tb, bool log) {
    tmask_tile(J.codes, J.B, J.M, J.pitch, tb.ucode, untyped_row_ok(tb, log), blockIdx.x * kMaskBlocks, J.tmask);
}

__global__ __launch_bounds__(kListThreads) void fb_tlist_kernel(JobView J) {
    tlist_build(J.tmask, J.M, J.tlist, J.tcount);
}

__global__ __launch_bounds__(kMaskThreads) void fb_tmask_batch_kernel(const JobView* jobs,
                                                                     const uint64_t* off, uint32_t n,
                                                                     FbTables tb, bool log) {
    const uint32_t j = find_job_uniform(off, n, blockIdx.x);
    const JobView* J = jobs + j;
    tmask_tile(J->codes, J->B, J->M, J->pitch, tb.ucode, untyped_row_ok(tb, log),
               uint32_t(blockIdx.x - off[j]) * kMaskBlocks, J->tmask);
}

__global__ __launch_bounds__(kListThreads) void fb_tlist_batch_kernel(const JobView* jobs) {
    const JobView* J = jobs + blockIdx.x;
    tlist_build(J->tmask, J->M, J->tlist, J->tcount);
}

URun fb_urun(const si_fb_job& j) {
    URun q{};
    if (j.flags & SI_FB_PER_MARKER) return q;
    const double to = j.step.t_o, u = j.step.u, L = double(j.n_states);
    const double S = L * to + u;
    const bool log = (j.flags & SI_FB_LOGSPACE) != 0;
    if (!(to > 0.0 && u >= 0.0 && S > 0.0 && S < 1.0 && (!log || to >= 0x1p-30))) return q;
    const double c = u / S;
    q.c1 = float(c);
    q.c2 = float(to / S);
    q.rS = float(1.0 / S);
    q.rL = float(1.0 / L);
    q.lc1 = float(std::log2(c));
    q.lnc1 = float(std::log(c));
    q.lnS = float(std::log(S));
    q.gS = float(1.0 / (1.0 - S));
    q.lnSd = std::log(S);
    q.ok = 1u;
    return q;
}

JobView fb_job_view(const si_fb_job& j, uint32_t cb) {
    JobView v{};
    v.B = j.n_samples;
    v.M = j.n_markers;
    v.L = j.n_states;
    v.flags = j.flags;
    v.pitch = j.code_pitch;
    v.rng_stream = j.rng_stream;
    v.sample_base = j.sample_base;
    v.nblk = (j.n_markers + cb - 1) / cb;
    v.W = fb_window(j);
    v.nwin = v.W ? (j.n_markers + v.W - 1) / v.W : 0u;
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
    v.ur = fb_urun(j);
    return v;
}

// Whether a launch runs the untyped-step kernels: a fast job (untyped code
// in the tables, constant transitions, no windows) at 16-marker checkpoint
// blocks; otherwise the exact kernels.
static bool band_launch(const JobView& v) {
    return (v.flags & SI_FB_BANDED) && (v.flags & SI_FB_LOGSPACE) && v.W == 0u;
}
static bool fast_launch(const JobView& v, const FbTables& t, const FbVariant& var) {
    return v.ur.ok != 0u && v.W == 0u && t.ucode < t.P && t.ucode < 256u && var.cb == 16u && !band_launch(v);
}

// ----------------------------------------------------------- plain form --
int launch_fb(const JobView& v, const FbTables& t, uint32_t Lp, uint32_t passes,
              const FbVariant& var, hipStream_t stream) {
    if (v.B == 0 || v.M == 0) return 0;
    if (!fb_variant_valid(Lp, var)) return SI_ERR_UNSUPPORTED;
    const bool log = (v.flags & SI_FB_LOGSPACE) != 0;
    const bool fast = fast_launch(v, t, var);
    const bool band = band_launch(v);
    if ((passes & SI_FB_PASS_CHAIN) && fast) {
        hipLaunchKernelGGL(fb_tmask_kernel, dim3(uint32_t(fb_tmask_groups(v.M))), dim3(kMaskThreads), 0,
                           stream, v, t, log);
        hipLaunchKernelGGL(fb_tlist_kernel, dim3(1), dim3(kListThreads), 0, stream, v);
        if (hipGetLastError() != hipSuccess) return SI_ERR_DEVICE;
    }
    if (passes & SI_FB_PASS_CHAIN) {
        const FbImpl* f = fb_impl(Lp, var.gc);
        const uint32_t gpw = kWave / var.gc;  // groups per wavefront
        const uint32_t Bpad = (v.B + gpw - 1) / gpw * gpw;
        const size_t threads = size_t(Bpad + v.B) * var.gc;
        const size_t grid = (threads + kFbThreads - 1) / kFbThreads;
        if (grid > 0x7fffffffu) return SI_ERR_UNSUPPORTED;
        FbLaunch L{kFbPlainChain, log, var.cb, fast, band, &v, Bpad, nullptr, t, uint32_t(grid), stream};
        const int rc = f->launch(L);
        if (rc) return rc;
    }
    if (passes & SI_FB_PASS_BLOCK) {
        const FbImpl* f = fb_impl(Lp, var.gb);
        const size_t threads = size_t(v.B) * v.nblk * var.gb;
        const size_t grid = (threads + kFbThreads - 1) / kFbThreads;
        if (grid > 0x7fffffffu) return SI_ERR_UNSUPPORTED;
        FbLaunch L{kFbPlainBlock, log, var.cb, fast, band, &v, 0, nullptr, t, uint32_t(grid), stream};
        const int rc = f->launch(L);
        if (rc) return rc;
    }
    return 0;
}

// --------------------------------------------------------- batched form --
__global__ void zero_words_kernel(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

static bool fast_class(const BatchClass& c, const FbTables& t, const FbVariant& var) {
    return (c.flags & kFbFastClass) && !(c.flags & SI_FB_BANDED) && t.ucode < t.P && t.ucode < 256u &&
           var.cb == 16u;
}

// Workgroups of a batched kernel resident per CU, per device; the runtime's
// occupancy query is kept, since a serving loop plans a batch per launch.
static int batch_occupancy(int dev, uint32_t Lp, uint32_t lanes, int form, bool log, uint32_t cb, uint32_t P,
                           bool fast, bool band) {
    static std::mutex mu;
    static std::map<std::tuple<int, uint32_t, uint32_t, int, bool, uint32_t, uint32_t, bool, bool>, int> memo;
    const auto key = std::make_tuple(dev, Lp, lanes, form, log, cb, P, fast, band);
    {
        std::lock_guard<std::mutex> g(mu);
        const auto it = memo.find(key);
        if (it != memo.end()) return it->second;
    }
    const FbImpl* f = fb_impl(Lp, lanes);
    const int n = f ? f->occupancy(form, log, cb, P, fast, band) : 1;
    std::lock_guard<std::mutex> g(mu);
    memo[key] = n;
    return n;
}

void fb_batch_grid(BatchClass& c, const FbVariant& var, uint32_t P) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    static std::mutex cu_mu;
    static std::map<int, int> cu_count;
    int cus = 0;
    {
        std::lock_guard<std::mutex> g(cu_mu);
        auto it = cu_count.find(dev);
        if (it == cu_count.end()) {
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
            cu_count[dev] = cus;
        } else {
            cus = it->second;
        }
    }
    const uint64_t ncu = uint64_t(cus > 0 ? cus : 1);
    const bool log = (c.flags & SI_FB_LOGSPACE) != 0;
    const bool band = log && (c.flags & SI_FB_BANDED);
    const bool fast = (c.flags & kFbFastClass) && var.cb == 16u && !band;
    const uint64_t occ_c = uint64_t(batch_occupancy(dev, c.Lp, var.gc, kFbBatchChain, log, var.cb, P, fast, band));
    const uint64_t occ_b = uint64_t(batch_occupancy(dev, c.Lp, var.gb, kFbBatchBlock, log, var.cb, P, fast, band));
    const uint64_t gpc = kFbThreads / var.gc, gpb = kFbThreads / var.gb;  // groups per workgroup
    const uint64_t need_c = (c.chain_items + gpc - 1) / gpc, need_b = (c.block_items + gpb - 1) / gpb;
    c.grid_chain = uint32_t(std::max<uint64_t>(1, std::min(need_c, ncu * occ_c)));
    c.grid_block = uint32_t(std::max<uint64_t>(1, std::min(need_b, ncu * occ_b)));
}

int launch_fb_batch(const BatchClass& c, const FbTables& t, uint32_t passes, const FbVariant& var,
                    hipStream_t stream) {
    if (!fb_variant_valid(c.Lp, var)) return SI_ERR_UNSUPPORTED;
    const bool log = (c.flags & SI_FB_LOGSPACE) != 0;
    const bool fast = fast_class(c, t, var);
    const bool band = log && (c.flags & SI_FB_BANDED);
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, stream,
                       reinterpret_cast<uint32_t*>(c.counters), 4u);
    if (hipGetLastError() != hipSuccess) return SI_ERR_DEVICE;
    if ((passes & SI_FB_PASS_CHAIN) && c.chain_items && fast) {
        hipLaunchKernelGGL(fb_tmask_batch_kernel, dim3(uint32_t(c.mask_groups)), dim3(kMaskThreads), 0,
                           stream, c.jobs, c.mask_off, c.n_jobs, t, log);
        hipLaunchKernelGGL(fb_tlist_batch_kernel, dim3(c.n_jobs), dim3(kListThreads), 0, stream, c.jobs);
        if (hipGetLastError() != hipSuccess) return SI_ERR_DEVICE;
    }
    if ((passes & SI_FB_PASS_CHAIN) && c.chain_items) {
        FbLaunch L{kFbBatchChain, log, var.cb, fast, band, nullptr, 0, &c, t, c.grid_chain, stream};
        const int rc = fb_impl(c.Lp, var.gc)->launch(L);
        if (rc) return rc;
    }
    if ((passes & SI_FB_PASS_BLOCK) && c.block_items) {
        FbLaunch L{kFbBatchBlock, log, var.cb, fast, band, nullptr, 0, &c, t, c.grid_block, stream};
        const int rc = fb_impl(c.Lp, var.gb)->launch(L);
        if (rc) return rc;
    }
    return 0;
}

}  // namespace si
