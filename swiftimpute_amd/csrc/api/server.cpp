// Serving form (include/swiftimpute/server.h): resident panel + continuous
// batching of imputation requests over several streams.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../../include/swiftimpute/server.h"
#include "../../../include/swiftimpute/synth.h"
#include "../../../include/swiftimpute_hip.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

using Clock = std::chrono::steady_clock;

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// fn(lo, hi) over [0, n) in up to 8 host threads when the work is large
// (requests of up to ~1e9 genotypes move gigabytes through host memory, which
// one thread copies at a fraction of the memory bandwidth).
template <class F>
void host_parallel(size_t n, size_t bytes, F&& fn) {
    const size_t nt = std::min<size_t>({8, n, bytes / (32u << 20) + 1});
    if (nt <= 1) {
        fn(size_t(0), n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (size_t t = 1; t < nt; ++t) th.emplace_back([&, t] { fn(n * t / nt, n * (t + 1) / nt); });
    fn(size_t(0), n / nt);
    for (auto& x : th) x.join();
}

// A fixed set of host threads for the request path's byte work (copying and
// scanning request codes): run(n, fn) calls fn(k) for k in [0, n) on the
// pool's threads and the caller, and returns when every call has finished
// and every thread is idle again.  One job at a time (callers queue on
// run_mu); no thread is started per request.
class WorkPool {
  public:
    explicit WorkPool(unsigned n) {
        for (unsigned i = 0; i < n; ++i) th_.emplace_back([this] { loop(); });
    }
    ~WorkPool() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    size_t threads() const { return th_.size() + 1; }
    void run(size_t n, const std::function<void(size_t)>& fn) {
        if (n == 0) return;
        std::lock_guard<std::mutex> one(run_mu_);
        {
            std::lock_guard<std::mutex> g(mu_);
            job_ = &fn;
            n_ = n;
            next_.store(0);
            remaining_.store(n);
            ++gen_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> g(mu_);
        done_.wait(g, [&] { return remaining_.load() == 0 && active_ == 0; });
        job_ = nullptr;
        n_ = 0;
    }

  private:
    void work() {
        for (;;) {
            const size_t k = next_.fetch_add(1);
            if (k >= n_) return;
            (*job_)(k);
            if (remaining_.fetch_sub(1) == 1) {
                std::lock_guard<std::mutex> g(mu_);
                done_.notify_all();
            }
        }
    }
    void loop() {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> g(mu_);
        for (;;) {
            cv_.wait(g, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            if (!job_) continue;
            ++active_;
            g.unlock();
            work();
            g.lock();
            if (--active_ == 0) done_.notify_all();
        }
    }
    std::vector<std::thread> th_;
    std::mutex run_mu_, mu_;
    std::condition_variable cv_, done_;
    const std::function<void(size_t)>* job_ = nullptr;
    size_t n_ = 0;
    std::atomic<size_t> next_{0}, remaining_{0};
    uint64_t gen_ = 0;
    unsigned active_ = 0;
    bool stop_ = false;
};

// Page-locked host memory (hipHostMalloc, torch's pin_memory): results are
// copied from the device into it directly, without the staging copy.
bool host_pinned(const void* p) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // pageable memory: clear the error
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

void par_memcpy(void* dst, const void* src, size_t bytes) {
    constexpr size_t kPiece = 4u << 20;
    const size_t pieces = (bytes + kPiece - 1) / kPiece;
    host_parallel(pieces, bytes, [&](size_t lo, size_t hi) {
        const size_t a = lo * kPiece, b = std::min(bytes, hi * kPiece);
        if (b > a) std::memcpy(static_cast<char*>(dst) + a, static_cast<const char*>(src) + a, b - a);
    });
}

// First-fit allocator over one device buffer (host-side bookkeeping only).
class Arena {
  public:
    void init(char* base, size_t bytes) {
        base_ = base;
        size_ = bytes;
        free_.clear();
        free_[0] = bytes;
        used_ = 0;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = align256(std::max<size_t>(bytes, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes, start = it->first + bytes;
                free_.erase(it);
                if (rest) free_[start] = rest;
                used_ += bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = align256(std::max<size_t>(bytes, 1));
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    char* ptr(size_t off) const { return base_ + off; }
    size_t used() const { return used_; }
    size_t size() const { return size_; }

  private:
    char* base_ = nullptr;
    size_t size_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;
};

struct Request {
    uint64_t id = 0;
    uint32_t m0 = 0, M = 0, B = 0, pitch = 0;
    uint64_t seed = 0;
    // codes [B][pitch]: in the server's page-locked pool (uploaded from there)
    // or, when the pool is full, in codes_heap (uploaded through slot staging)
    uint8_t* codes = nullptr;
    std::unique_ptr<uint8_t[]> codes_heap;
    size_t codes_bytes = 0, pool_off = 0;
    bool in_pool = false;
    uint64_t untyped = 0;
    Clock::time_point arrival, done_at;
    size_t off_codes = 0, off_hap = 0, off_ds = 0, own_bytes = 0;
    // region re-sweep: the request's own state table, prefix rows and sweep workspace
    size_t off_sa = 0, off_mmask = 0, off_prefix = 0, off_pws = 0, pws_bytes = 0;
    std::vector<uint8_t> hap;
    std::vector<float> ds;
    uint8_t* dst_hap = nullptr;  // caller buffers (si_server_submit_into), else hap/ds above
    float* dst_ds = nullptr;
    bool into = false;
    bool hap_pinned = false, ds_pinned = false;  // caller buffers in page-locked memory
    bool done = false;
    int status = 0;
};

enum class SlotState { Free, Running, Copying };

struct Slot {
    hipStream_t stream = nullptr;
    // stream marks: before the code upload, before the forward-backward,
    // before the result copies, after them
    hipEvent_t t0 = nullptr, k0 = nullptr, k1 = nullptr, done = nullptr;
    SlotState state = SlotState::Free;
    std::vector<std::shared_ptr<Request>> reqs;
    si_fb_batch* batch = nullptr;  // kept across launches (si_fb_batch_reset)
    size_t ws_off = 0, ws_bytes = 0;
    char* pinned = nullptr;  // staging for codes in (pool misses) and results out
    size_t pinned_bytes = 0;
    size_t out_base = 0;     // where the results start in `pinned`
};

}  // namespace

struct si_server {
    si_server_cfg cfg{};
    uint32_t H = 0, M = 0, Lp = 0;
    int device = 0;
    // resident
    uint64_t* d_cols = nullptr;
    uint8_t* d_sa = nullptr;
    uint8_t* d_mmask = nullptr;
    si_step_coef* d_steps = nullptr;  // [M], entry m = step into m
    std::vector<uint32_t> perm;       // selection permutation (min(L, H) entries)
    std::vector<si_step_coef> steps;  // host copy
    std::vector<uint32_t> tiny_switch_prefix;  // intervals below the linear range, prefix count
    float* d_eps = nullptr;
    float* d_emax = nullptr;
    float* d_palette = nullptr;
    char* d_arena = nullptr;
    Arena arena;
    // page-locked pool for request codes: submit writes them there once and
    // the upload reads them from there (no second host copy, no page faults
    // of fresh pageable memory in the request path)
    char* h_codes = nullptr;
    Arena codes_pool;
    std::unique_ptr<WorkPool> pool;  // host threads copying and scanning codes in submit
    // scheduling
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> reqs;
    std::vector<Slot> slots;
    std::thread worker;
    // host-side result copies run on their own threads, so the scheduler keeps
    // launching while finished batches are copied out of pinned staging
    std::vector<std::thread> copiers;
    std::deque<Slot*> copy_queue;
    std::condition_variable cv_copy;
    bool stop = false;
    bool copiers_stop = false;
    uint64_t next_id = 1;
    si_server_stats stats{};
    std::string worker_error;
    // Testing aid: when nonzero, the launch of this batch (1-based count)
    // raises an error on the host, as a failing launch would
    // (SI_SERVER_FAIL_AT_BATCH, read at creation).  No device work is affected.
    uint64_t fail_at_batch = 0;
    // requests whose results the device copied straight into page-locked
    // caller buffers (counted outside the lock)
    std::atomic<uint64_t> n_direct{0};

    ~si_server() {
        for (auto& sl : slots) {
            if (sl.batch) si_fb_batch_destroy(sl.batch);
            if (sl.pinned) (void)hipHostFree(sl.pinned);
            for (hipEvent_t e : {sl.t0, sl.k0, sl.k1, sl.done})
                if (e) (void)hipEventDestroy(e);
            if (sl.stream) (void)hipStreamDestroy(sl.stream);
        }
        if (h_codes) (void)hipHostFree(h_codes);
        for (void* p : {static_cast<void*>(d_cols), static_cast<void*>(d_sa), static_cast<void*>(d_mmask),
                        static_cast<void*>(d_steps), static_cast<void*>(d_eps), static_cast<void*>(d_emax),
                        static_cast<void*>(d_palette), static_cast<void*>(d_arena)})
            if (p) (void)hipFree(p);
    }

    bool uses_logspace(uint32_t m0, uint32_t m1) const {
        // steps used: compat -> interval m0 only; per-marker -> intervals m0 .. m1-2
        if (m1 - m0 < 2) return false;
        const uint32_t a = m0, b = cfg.per_marker ? m1 - 1 : m0 + 1;
        return tiny_switch_prefix[b] - tiny_switch_prefix[a] > 0;
    }

    // Device footprint of a request apart from its checkpoint workspace:
    // codes, haplotypes, dosages and, with region re-sweeps, its state table,
    // prefix rows and sweep workspace (offsets set here, relative to the start).
    size_t own_bytes(Request& r) const {
        size_t o = align256(size_t(r.B) * r.pitch);
        r.off_hap = o;
        o += align256(size_t(r.B) * 2 * r.M);
        r.off_ds = o;
        o += align256(size_t(r.B) * r.M * sizeof(float));
        if (cfg.region_resweep) {
            r.off_sa = o;
            o += align256(size_t(r.M) * Lp);
            r.off_mmask = o;
            o += align256(r.M);
            r.off_prefix = o;
            o += align256(size_t(r.M) * H * 4);
            r.off_pws = o;
            r.pws_bytes = si_pbwt_workspace_bytes(H, r.M);
            o += align256(r.pws_bytes);
        }
        return o;
    }
    // A request's job as far as its shape goes: sizes, numeric form, window
    // (what its workspace depends on).
    si_fb_job job_shape(const Request& r) const {
        si_fb_job j{};
        j.n_samples = r.B;
        j.n_markers = r.M;
        j.n_states = cfg.n_states;
        // the log form is chosen for switch probabilities below 1e-30, which
        // the banded kernels carry exactly against the -999 floor
        j.flags = (cfg.per_marker ? uint32_t(SI_FB_PER_MARKER) : 0u) |
                  (uses_logspace(r.m0, r.m0 + r.M) ? uint32_t(SI_FB_LOGSPACE | SI_FB_BANDED) : 0u);
        j.window = cfg.window_size;
        return j;
    }
    size_t ws_bytes_of(const Request& r) const {
        const si_fb_job j = job_shape(r);
        return align256(si_fb_workspace_bytes(&j, 1));
    }

    si_fb_job job_of(const Request& r) const {
        si_fb_job j = job_shape(r);
        j.code_pitch = r.pitch;
        j.n_palette = 4;
        j.flags |= SI_FB_UNTYPED;  // code 3 of the hard-call palette: (0, 0, 0)
        j.untyped_code = 3;
        j.seed = r.seed;
        j.step = (r.M >= 2) ? steps[r.m0 + 1] : si_step_coef{1.0f, 0.0f, 0.0f, -INFINITY};
        j.steps = d_steps + r.m0;
        j.sa = cfg.region_resweep ? reinterpret_cast<const uint8_t*>(arena.ptr(r.off_codes + r.off_sa))
                                  : d_sa + size_t(r.m0) * Lp;
        j.mmask = cfg.region_resweep ? reinterpret_cast<const uint8_t*>(arena.ptr(r.off_codes + r.off_mmask))
                                     : d_mmask + r.m0;
        j.codes = reinterpret_cast<const uint8_t*>(arena.ptr(r.off_codes));
        j.eps = d_eps;
        j.emax = d_emax;
        j.palette = d_palette;
        j.hap = reinterpret_cast<uint8_t*>(arena.ptr(r.off_codes + r.off_hap));
        j.ds = reinterpret_cast<float*>(arena.ptr(r.off_codes + r.off_ds));
        return j;
    }

    // The batch of a slot finished on the device: its device memory is free
    // again, and its results wait in the slot's pinned staging for a copier.
    void retire(Slot& sl) {
        for (auto& r : sl.reqs) {
            arena.release(r->off_codes, r->own_bytes);
            if (r->in_pool) codes_pool.release(r->pool_off, r->codes_bytes);
            r->in_pool = false;
            r->codes = nullptr;
            r->codes_heap.reset();
        }
        arena.release(sl.ws_off, sl.ws_bytes);
        float a = 0.0f, b = 0.0f, c = 0.0f;
        if (hipEventElapsedTime(&a, sl.t0, sl.k0) == hipSuccess) stats.h2d_s += 1e-3 * a;
        if (hipEventElapsedTime(&b, sl.k0, sl.k1) == hipSuccess) stats.kernel_s += 1e-3 * b;
        if (hipEventElapsedTime(&c, sl.k1, sl.done) == hipSuccess) stats.d2h_s += 1e-3 * c;
        sl.state = SlotState::Copying;
        copy_queue.push_back(&sl);
        cv_copy.notify_one();
    }

    // Copier thread: results from pinned staging to their destinations
    // (outside the lock), then the requests complete and the slot frees.
    void copy_loop() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv_copy.wait(lk, [&] { return !copy_queue.empty() || copiers_stop; });
            if (copy_queue.empty()) return;
            Slot* sl = copy_queue.front();
            copy_queue.pop_front();
            lk.unlock();
            const auto c0 = Clock::now();
            // the staged parts (destinations not in page-locked memory), in
            // the order launch() laid them out
            size_t o = sl->out_base;
            for (auto& r : sl->reqs) {
                const size_t hb = size_t(r->B) * 2 * r->M, db = size_t(r->B) * r->M * sizeof(float);
                const char* sh = r->hap_pinned ? nullptr : sl->pinned + o;
                if (!r->hap_pinned) o += hb;
                const char* sd = r->ds_pinned ? nullptr : sl->pinned + o;
                if (!r->ds_pinned) o += db;
                if (r->into) {
                    if (r->dst_hap && sh) par_memcpy(r->dst_hap, sh, hb);
                    if (r->dst_ds && sd) par_memcpy(r->dst_ds, sd, db);
                } else {
                    r->hap.resize(hb);
                    r->ds.resize(db / sizeof(float));
                    par_memcpy(r->hap.data(), sh, hb);
                    par_memcpy(r->ds.data(), sd, db);
                }
            }
            const auto now = Clock::now();
            lk.lock();
            stats.copy_host_s += std::chrono::duration<double>(now - c0).count();
            for (auto& r : sl->reqs) {
                r->done = true;
                r->done_at = now;
                stats.requests++;
                stats.genotypes += uint64_t(r->B) * r->M;
                stats.untyped_genotypes += r->untyped;
            }
            sl->reqs.clear();
            sl->state = SlotState::Free;
            cv_done.notify_all();
            cv_work.notify_one();
        }
    }

    // Admits queued requests (arrival order) into a free slot; false if none.
    // Called with the lock held; drops it while staging and enqueueing.
    bool launch(Slot& sl, std::unique_lock<std::mutex>& lk) {
        std::vector<std::shared_ptr<Request>> take;
        size_t ws_total = 0;
        while (!queue.empty() && take.size() < cfg.max_batch) {
            auto& r = queue.front();
            const size_t own = own_bytes(*r), ws = ws_bytes_of(*r);
            size_t off = 0;
            // the batch workspace is allocated after selection; keep room for it
            if (arena.size() - arena.used() < own + ws_total + ws + 4096 || !arena.alloc(own, &off)) {
                if (take.empty()) stats.deferred_admissions++;
                break;
            }
            r->own_bytes = own;
            r->off_codes = off;
            ws_total += ws;
            take.push_back(r);
            queue.pop_front();
        }
        if (take.empty()) return false;
        size_t ws_off = 0;
        if (!arena.alloc(ws_total, &ws_off)) {  // fragmentation: put them back
            for (auto it = take.rbegin(); it != take.rend(); ++it) {
                arena.release((*it)->off_codes, (*it)->own_bytes);
                queue.push_front(*it);
            }
            stats.deferred_admissions++;
            return false;
        }
        sl.state = SlotState::Running;
        sl.reqs = std::move(take);
        sl.ws_off = ws_off;
        sl.ws_bytes = ws_total;
        stats.batches++;
        stats.max_batch_requests = std::max<uint64_t>(stats.max_batch_requests, sl.reqs.size());
        stats.peak_bytes = std::max<uint64_t>(stats.peak_bytes, arena.used());
        const auto& take_ref = sl.reqs;
        std::vector<si_fb_job> jobs;
        for (auto& r : take_ref) jobs.push_back(job_of(*r));
        lk.unlock();  // the slot is ours; staging and launch need no shared state
        const auto h0 = Clock::now();
        // page-locked staging holds the codes of pool misses and the results
        // bound for pageable destinations; the rest moves by DMA directly
        size_t in_bytes = 0, out_bytes = 0, stage_in = 0, stage_out = 0;
        for (auto& r : take_ref) {
            const size_t hb = size_t(r->B) * 2 * r->M, db = size_t(r->B) * r->M * sizeof(float);
            in_bytes += r->codes_bytes;
            if (!r->in_pool) stage_in += r->codes_bytes;
            out_bytes += hb + db;
            stage_out += (r->hap_pinned ? 0 : hb) + (r->ds_pinned ? 0 : db);
        }
        struct Relock {
            si_server* s;
            std::unique_lock<std::mutex>& l;
            Clock::time_point h0;
            size_t in, out;
            ~Relock() {
                const double dt = std::chrono::duration<double>(Clock::now() - h0).count();
                l.lock();
                s->stats.launch_host_s += dt;
                s->stats.h2d_bytes += in;
                s->stats.d2h_bytes += out;
            }
        } relock{this, lk, h0, in_bytes, out_bytes};
        if (sl.pinned_bytes < stage_in + stage_out) {
            if (sl.pinned) (void)hipHostFree(sl.pinned);
            sl.pinned = nullptr;
            sl.pinned_bytes = std::max<size_t>(stage_in + stage_out, sl.pinned_bytes * 2);
            if (hipHostMalloc(reinterpret_cast<void**>(&sl.pinned), sl.pinned_bytes, 0) != hipSuccess)
                throw std::runtime_error("hipHostMalloc failed");
        }
        if (hipEventRecord(sl.t0, sl.stream) != hipSuccess) throw std::runtime_error("event record failed");
        size_t o = 0;
        for (auto& r : take_ref) {
            const uint8_t* src = r->codes;
            if (!r->in_pool) {
                par_memcpy(sl.pinned + o, r->codes, r->codes_bytes);
                src = reinterpret_cast<const uint8_t*>(sl.pinned + o);
                o += r->codes_bytes;
            }
            if (hipMemcpyAsync(arena.ptr(r->off_codes), src, r->codes_bytes, hipMemcpyHostToDevice,
                               sl.stream) != hipSuccess)
                throw std::runtime_error("codes upload failed");
        }
        sl.out_base = o;
        if (cfg.region_resweep) {
            // each request's region swept from the identity order at its first marker
            const size_t W = (H + 63) / 64;
            for (auto& r : take_ref) {
                const uint64_t* cols = d_cols + size_t(r->m0) * W;
                auto* prefix = reinterpret_cast<uint32_t*>(arena.ptr(r->off_codes + r->off_prefix));
                if (si_pbwt_sweep(cols, H, r->M, prefix, nullptr, arena.ptr(r->off_codes + r->off_pws),
                                  r->pws_bytes, sl.stream) != SI_OK ||
                    si_state_table(prefix, cols, nullptr, H, r->M, cfg.n_states, 0, perm.data(), nullptr,
                                   reinterpret_cast<uint8_t*>(arena.ptr(r->off_codes + r->off_sa)),
                                   reinterpret_cast<uint8_t*>(arena.ptr(r->off_codes + r->off_mmask)),
                                   sl.stream) != SI_OK)
                    throw std::runtime_error(std::string("region sweep failed: ") + si_last_error());
            }
        }
        if (fail_at_batch && stats.batches == fail_at_batch)
            throw std::runtime_error("injected launch failure (SI_SERVER_FAIL_AT_BATCH)");
        if (hipEventRecord(sl.k0, sl.stream) != hipSuccess) throw std::runtime_error("event record failed");
        if (cfg.plain) {
            // one job per launch, each with its own part of the slot's workspace
            size_t wo = 0;
            for (size_t k = 0; k < jobs.size(); ++k) {
                const size_t wb = ws_bytes_of(*take_ref[k]);
                if (si_fb_run(&jobs[k], 1, arena.ptr(ws_off + wo), wb, sl.stream) != SI_OK)
                    throw std::runtime_error(std::string("plain launch failed: ") + si_last_error());
                wo += wb;
            }
        } else {
            // the slot's batch is planned again in place: its device tables and
            // staging are reused, so a launch allocates nothing once the
            // batches have reached their working size
            const int rc = sl.batch ? si_fb_batch_reset(sl.batch, jobs.data(), uint32_t(jobs.size()))
                                    : si_fb_batch_create(jobs.data(), uint32_t(jobs.size()), &sl.batch);
            if (rc != SI_OK) throw std::runtime_error(std::string("batch launch failed: ") + si_last_error());
            // Default lanes per sample, not the autotuner: a batch of a
            // cohort-size bucket not seen before would otherwise time every
            // candidate inside its requests' latency, and on these shapes the
            // tuner keeps the defaults anyway.  Outputs do not depend on the
            // variant.
            for (uint32_t k = 0; k < si_fb_batch_n_classes(sl.batch); ++k) {
                uint32_t lp = 0, flags = 0;
                int tuned = 0;
                si_fb_variant v{};
                if (si_fb_batch_class_info(sl.batch, k, &lp, &flags, &v, &tuned) != SI_OK ||
                    si_fb_default_variant(lp, &v) != SI_OK || si_fb_batch_set_variant(sl.batch, k, &v) != SI_OK)
                    throw std::runtime_error(std::string("batch variant failed: ") + si_last_error());
            }
            if (si_fb_batch_run(sl.batch, arena.ptr(ws_off), ws_total, 0, sl.stream) != SI_OK)
                throw std::runtime_error(std::string("batch launch failed: ") + si_last_error());
        }
        if (hipEventRecord(sl.k1, sl.stream) != hipSuccess) throw std::runtime_error("event record failed");
        for (auto& r : take_ref) {
            const size_t hb = size_t(r->B) * 2 * r->M, db = size_t(r->B) * r->M * sizeof(float);
            if (r->hap_pinned || r->ds_pinned) n_direct.fetch_add(1, std::memory_order_relaxed);
            // into the staging, or straight into page-locked caller buffers
            void* dh = r->hap_pinned ? static_cast<void*>(r->dst_hap) : static_cast<void*>(sl.pinned + o);
            (void)hipMemcpyAsync(dh, arena.ptr(r->off_codes + r->off_hap), hb, hipMemcpyDeviceToHost, sl.stream);
            if (!r->hap_pinned) o += hb;
            void* dd = r->ds_pinned ? static_cast<void*>(r->dst_ds) : static_cast<void*>(sl.pinned + o);
            (void)hipMemcpyAsync(dd, arena.ptr(r->off_codes + r->off_ds), db, hipMemcpyDeviceToHost, sl.stream);
            if (!r->ds_pinned) o += db;
        }
        if (hipEventRecord(sl.done, sl.stream) != hipSuccess) throw std::runtime_error("event record failed");
        return true;
    }

    void run() {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        try {
            for (;;) {
                bool inflight = false, copying = false;
                for (auto& sl : slots)
                    if (sl.state == SlotState::Running) {
                        if (hipEventQuery(sl.done) == hipSuccess) retire(sl);
                        else inflight = true;
                    }
                for (auto& sl : slots) copying |= sl.state == SlotState::Copying;
                bool launched = false;
                // a partial batch launches once its oldest request has waited
                // max_queue_delay_us (or at shutdown)
                const bool ready =
                    !queue.empty() &&
                    (stop || queue.size() >= cfg.max_batch ||
                     Clock::now() - queue.front()->arrival >=
                         std::chrono::microseconds(cfg.max_queue_delay_us));
                for (auto& sl : slots)
                    if (ready && sl.state == SlotState::Free && !queue.empty() && launch(sl, lk)) {
                        launched = inflight = true;
                    }
                if (stop && queue.empty() && !inflight && !copying) break;
                if (launched) continue;
                if (inflight) {
                    lk.unlock();
                    std::this_thread::sleep_for(std::chrono::microseconds(50));
                    lk.lock();
                } else {
                    cv_work.wait_for(lk, std::chrono::milliseconds(5));
                }
            }
        } catch (const std::exception& e) {
            worker_error = e.what();
            // Batches still on the device may be copying results straight into
            // page-locked caller buffers: wait for their streams before any of
            // their requests is reported, so that no copy lands after
            // si_server_wait has returned.
            for (auto& sl : slots)
                if (sl.state == SlotState::Running) (void)hipStreamSynchronize(sl.stream);
            // Requests of batches already handed to a copier complete normally.
            std::vector<const Request*> copying;
            for (auto& sl : slots)
                if (sl.state == SlotState::Copying)
                    for (auto& r : sl.reqs) copying.push_back(r.get());
            for (auto& kv : reqs)
                if (!kv.second->done &&
                    std::find(copying.begin(), copying.end(), kv.second.get()) == copying.end()) {
                    kv.second->done = true;
                    kv.second->status = SI_ERR_DEVICE;
                }
            cv_done.notify_all();
        }
    }
};

extern "C" {

const char* si_server_last_error(void) { return g_err.c_str(); }

void si_server_defaults(si_server_cfg* c) {
    c->n_states = 8;
    c->ne = 10000;
    c->rho_rate = 4.0f;
    c->per_marker = 0;
    c->max_streams = 4;
    c->max_batch = 256;
    c->hbm_budget_bytes = 32ull << 30;
    c->device_id = 0;
    c->max_queue_delay_us = 0;
    c->selection = 0;
    c->region_resweep = 0;
    c->plain = 0;
    c->staging_bytes = 4ull << 30;
    c->window_size = 0;
}

int si_server_create(const si_server_cfg* cfg, const uint64_t* cols, uint32_t H, uint32_t M,
                     const uint64_t* pos, const double* cM, si_server** out) {
    if (!cfg || !cols || !pos || !out || H == 0 || M == 0)
        return fail(SI_ERR_ARG, "si_server_create: bad arguments");
    if (cfg->n_states == 0 || cfg->n_states > 256) return fail(SI_ERR_ARG, "n_states must be in 1..256");
    if (cfg->selection > 1) return fail(SI_ERR_ARG, "selection must be 0 (CPU path) or 1 (GPU kernels)");
    if (cfg->region_resweep > 1) return fail(SI_ERR_ARG, "region_resweep must be 0 or 1");
    if (cfg->plain > 1) return fail(SI_ERR_ARG, "plain must be 0 or 1");
    if (cfg->window_size != 0 && cfg->window_size < 32)
        return fail(SI_ERR_ARG, "window_size must be 0 or at least 32 markers");
    auto s = std::make_unique<si_server>();
    s->cfg = *cfg;
    s->cfg.max_streams = std::max(1u, cfg->max_streams);
    s->cfg.max_batch = std::max(1u, cfg->max_batch);
    s->H = H;
    s->M = M;
    s->Lp = si_padded_states(cfg->n_states);
    s->device = cfg->device_id < 0 ? 0 : cfg->device_id;
    if (hipSetDevice(s->device) != hipSuccess) return fail(SI_ERR_DEVICE, "hipSetDevice failed");
    const size_t W = (H + 63) / 64;
    const uint32_t L = cfg->n_states;
    // index build: PBWT sweep into a temporary prefix array, then the state table
    uint32_t* prefix = nullptr;
    hipStream_t st = nullptr;
    if (hipStreamCreate(&st) != hipSuccess ||
        hipMalloc(&s->d_cols, M * W * 8) != hipSuccess ||
        hipMalloc(&s->d_sa, size_t(M) * s->Lp) != hipSuccess ||
        hipMalloc(&s->d_mmask, M) != hipSuccess ||
        hipMalloc(&prefix, size_t(M) * H * 4) != hipSuccess ||
        hipMemcpy(s->d_cols, cols, M * W * 8, hipMemcpyHostToDevice) != hipSuccess) {
        if (prefix) (void)hipFree(prefix);
        if (st) (void)hipStreamDestroy(st);
        return fail(SI_ERR_DEVICE, "si_server_create: device allocation failed");
    }
    std::vector<uint32_t>& perm = s->perm;
    perm.resize(std::min(L, H));
    int rc = si_pbwt_sweep(s->d_cols, H, M, prefix, nullptr, nullptr, 0, st);
    if (rc == SI_OK)
        rc = cfg->selection == 1 ? si_select_perm_gpu_compat(H, L, perm.data())
                                 : si_select_perm(H, L, perm.data());
    if (rc == SI_OK)
        rc = si_state_table(prefix, s->d_cols, nullptr, H, M, L, 0, perm.data(), nullptr, s->d_sa,
                            s->d_mmask, st);
    if (rc == SI_OK && hipStreamSynchronize(st) != hipSuccess) rc = SI_ERR_DEVICE;
    (void)hipFree(prefix);
    (void)hipStreamDestroy(st);
    if (rc != SI_OK) return fail(rc, std::string("index build failed: ") + si_last_error());
    // transitions over the whole panel
    std::vector<double> g(M);
    si_genetic_map(pos, cM, M, g.data());
    std::vector<float> Td(M > 1 ? M - 1 : 1), To(Td.size());
    si_transition_constants(g.data(), M, cfg->ne, cfg->rho_rate, Td.data(), To.data());
    std::vector<si_step_coef> co(Td.size());
    si_step_coefs(Td.data(), To.data(), M > 1 ? M - 1 : 0, co.data());
    s->steps.assign(M, si_step_coef{1.0f, 0.0f, 0.0f, -INFINITY});
    s->tiny_switch_prefix.assign(M + 1, 0);
    for (uint32_t m = 1; m < M; ++m) s->steps[m] = co[m - 1];
    // tiny_switch_prefix[t+1] counts intervals 0..t whose switch is below the linear range
    for (uint32_t t = 0; t < M; ++t)
        s->tiny_switch_prefix[t + 1] =
            s->tiny_switch_prefix[t] + ((t + 1 < M && !(s->steps[t + 1].t_o >= 1e-30f)) ? 1u : 0u);
    float pal[12], eps[4 * 32], emax[32];
    si_hardcall_palette(pal);
    si_emission_table(pal, 4, eps, emax);
    if (hipMalloc(&s->d_steps, M * sizeof(si_step_coef)) != hipSuccess ||
        hipMalloc(&s->d_eps, sizeof(eps)) != hipSuccess || hipMalloc(&s->d_emax, sizeof(emax)) != hipSuccess ||
        hipMalloc(&s->d_palette, sizeof(pal)) != hipSuccess ||
        hipMemcpy(s->d_steps, s->steps.data(), M * sizeof(si_step_coef), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->d_eps, eps, sizeof(eps), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->d_emax, emax, sizeof(emax), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->d_palette, pal, sizeof(pal), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&s->d_arena, cfg->hbm_budget_bytes) != hipSuccess)
        return fail(SI_ERR_DEVICE, "si_server_create: device allocation failed");
    s->arena.init(s->d_arena, cfg->hbm_budget_bytes);
    if (cfg->staging_bytes) {
        if (hipHostMalloc(reinterpret_cast<void**>(&s->h_codes), cfg->staging_bytes, 0) != hipSuccess) {
            s->h_codes = nullptr;
            return fail(SI_ERR_DEVICE, "si_server_create: page-locked staging allocation failed");
        }
        s->codes_pool.init(s->h_codes, cfg->staging_bytes);
    }
    s->slots.resize(s->cfg.max_streams);
    for (auto& sl : s->slots)
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&sl.t0) != hipSuccess || hipEventCreate(&sl.k0) != hipSuccess ||
            hipEventCreate(&sl.k1) != hipSuccess || hipEventCreate(&sl.done) != hipSuccess)
            return fail(SI_ERR_DEVICE, "si_server_create: stream creation failed");
    if (const char* f = std::getenv("SI_SERVER_FAIL_AT_BATCH")) s->fail_at_batch = std::strtoull(f, nullptr, 10);
    {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        s->pool = std::make_unique<WorkPool>(std::min(hw, 16u) - 1u);
    }
    si_server* raw = s.release();
    raw->worker = std::thread([raw] { raw->run(); });
    for (uint32_t i = 0; i < raw->cfg.max_streams; ++i)
        raw->copiers.emplace_back([raw] { raw->copy_loop(); });
    *out = raw;
    return SI_OK;
}

int si_server_submit_into(si_server* s, uint32_t m0, uint32_t m1, uint32_t B,
                          const uint8_t* codes, uint64_t seed, uint8_t* hap, float* ds,
                          uint64_t* id) {
    if (!s || !id || m1 <= m0 || m1 > s->M || B == 0 || !codes)
        return fail(SI_ERR_ARG, "si_server_submit: bad arguments");
    auto r = std::make_shared<Request>();
    r->arrival = Clock::now();
    r->m0 = m0;
    r->M = m1 - m0;
    r->B = B;
    r->pitch = (r->M + 15) / 16 * 16;
    r->seed = seed;
    r->into = hap != nullptr || ds != nullptr;
    r->dst_hap = hap;
    r->dst_ds = ds;
    r->hap_pinned = hap && host_pinned(hap);
    r->ds_pinned = ds && host_pinned(ds);
    const size_t need = s->own_bytes(*r) + s->ws_bytes_of(*r) + 8192;
    if (need > s->cfg.hbm_budget_bytes)
        return fail(SI_ERR_ARG, "si_server_submit: request needs " + std::to_string(need) +
                                    " bytes, above the server's HBM budget");
    r->codes_bytes = size_t(B) * r->pitch;
    if (s->h_codes) {
        std::lock_guard<std::mutex> g(s->mu);
        r->in_pool = s->codes_pool.alloc(r->codes_bytes, &r->pool_off);
    }
    if (r->in_pool) {
        r->codes = reinterpret_cast<uint8_t*>(s->codes_pool.ptr(r->pool_off));
    } else {
        r->codes_heap.reset(new uint8_t[r->codes_bytes]);
        r->codes = r->codes_heap.get();
    }
    auto unreserve = [&] {
        if (!r->in_pool) return;
        std::lock_guard<std::mutex> g(s->mu);
        s->codes_pool.release(r->pool_off, r->codes_bytes);
        r->in_pool = false;
    };
    // rows copied into the padded layout while checking the codes and counting
    // the untyped ones (branch-free, so the scan vectorizes), in tasks of
    // about 1 MiB of codes over the server's host threads
    const size_t total = size_t(B) * r->M;
    const size_t ntask = std::min<size_t>(B, std::max<size_t>(1, total >> 20));
    std::vector<uint64_t> part_untyped(ntask, 0);
    std::vector<uint8_t> part_max(ntask, 0);
    const std::function<void(size_t)> task = [&](size_t me) {
        const size_t lo = size_t(B) * me / ntask, hi = size_t(B) * (me + 1) / ntask;
        uint64_t un = 0;
        uint8_t mx = 0;
        for (size_t b = lo; b < hi; ++b) {
            const uint8_t* src = codes + b * r->M;
            uint8_t* dst = r->codes + b * r->pitch;
            std::memcpy(dst, src, r->M);
            std::memset(dst + r->M, 3, r->pitch - r->M);
            uint32_t row_un = 0;
            uint8_t row_mx = 0;
            for (uint32_t m = 0; m < r->M; ++m) {
                row_mx = std::max(row_mx, src[m]);
                row_un += src[m] == 3;
            }
            un += row_un;
            mx = std::max(mx, row_mx);
        }
        part_untyped[me] = un;
        part_max[me] = mx;
    };
    if (ntask == 1 || !s->pool) {
        for (size_t k = 0; k < ntask; ++k) task(k);
    } else {
        s->pool->run(ntask, task);
    }
    for (size_t t = 0; t < part_untyped.size(); ++t) {
        if (part_max[t] > 3) {
            unreserve();
            return fail(SI_ERR_ARG, "si_server_submit: code outside 0..3");
        }
        r->untyped += part_untyped[t];
    }
    std::lock_guard<std::mutex> g(s->mu);
    if (!s->worker_error.empty()) {
        if (r->in_pool) s->codes_pool.release(r->pool_off, r->codes_bytes);
        r->in_pool = false;
        return fail(SI_ERR_DEVICE, s->worker_error);
    }
    r->id = s->next_id++;
    s->reqs[r->id] = r;
    s->queue.push_back(r);
    *id = r->id;
    s->cv_work.notify_one();
    return SI_OK;
}

int si_server_submit(si_server* s, uint32_t m0, uint32_t m1, uint32_t B, const uint8_t* codes,
                     uint64_t seed, uint64_t* id) {
    return si_server_submit_into(s, m0, m1, B, codes, seed, nullptr, nullptr, id);
}

int si_server_wait(si_server* s, uint64_t id, uint8_t* hap, float* ds, double* latency_s) {
    if (!s) return fail(SI_ERR_ARG, "si_server_wait: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(SI_ERR_ARG, "si_server_wait: unknown request");
    auto r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    s->reqs.erase(it);
    if (r->status != 0) return fail(r->status, "request failed: " + s->worker_error);
    if (!r->into) {
        if (hap) std::memcpy(hap, r->hap.data(), r->hap.size());
        if (ds) std::memcpy(ds, r->ds.data(), r->ds.size() * sizeof(float));
    }
    if (latency_s) *latency_s = std::chrono::duration<double>(r->done_at - r->arrival).count();
    return SI_OK;
}

int si_server_get_stats(si_server* s, si_server_stats* out) {
    if (!s || !out) return fail(SI_ERR_ARG, "si_server_get_stats: bad arguments");
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->stats;
    out->direct_copies = s->n_direct.load(std::memory_order_relaxed);
    return SI_OK;
}

int si_server_destroy(si_server* s) {
    if (!s) return SI_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->copiers_stop = true;
    }
    s->cv_copy.notify_all();
    for (auto& th : s->copiers) th.join();
    delete s;
    return SI_OK;
}

}  // extern "C"
