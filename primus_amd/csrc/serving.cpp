// serving.cpp -- continuous-batching request loop over the batched transforms.
//
// A dispatcher thread turns the pending queue into batches: it admits
// requests in arrival order while their device footprint fits the HBM budget
// and a stream slot is free, groups them by kind and ring degree (one launch
// per poly-multiply ring degree whatever the limb counts, one persistent
// launch for all RLWE encryptions, one for all external products), and
// enqueues the batch as three stages: input copies on the server's
// host-to-device stream, kernels on the slot's stream, result copies on the
// device-to-host stream, handed over by events.  A completion thread retires
// batches in order, copies staged results to the callers' buffers and frees
// the budget.  Requests that arrive while batches are on the GPU join the
// next batch as soon as a slot and memory free up; one batch's compute
// overlaps the next batch's inputs and the previous batch's results.  Tables
// and keys are built once and stay resident.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <random>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/primus_hip.h"
#include "capi_internal.hpp"
#include "copy_launch.hpp"

namespace {

using Clock = std::chrono::steady_clock;

constexpr uint64_t kAlign = 256;

uint64_t round_up(uint64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

struct Ticket {
    ph_serve_request req;
    Clock::time_point arrival, done_at;
    int status = -1;  // -1 pending, else a PH_* code
    uint64_t bytes = 0;
    bool direct = false;  // every buffer page-locked: copied without staging
};

struct Part {          // one request's place in a batch
    uint64_t ticket;
    uint64_t in0, in1, out, small;  // arena offsets (bytes)
    uint64_t out_bytes;
};

struct Batch {
    std::vector<Part> parts;
    uint64_t arena_bytes = 0, budget_bytes = 0;
    int slot = -1;
    hipEvent_t done = nullptr;
    int status = PH_OK;
};

struct Slot {
    hipStream_t stream = nullptr;  // the batch's compute, allocation and completion
    // stage hand-offs: arena allocated -> inputs copied -> compute done ->
    // results copied (a slot holds one batch at a time, so they are reused)
    hipEvent_t ev_alloc = nullptr, ev_in = nullptr, ev_comp = nullptr, ev_out = nullptr;
    void *pinned = nullptr;
    uint64_t pinned_cap = 0;
    ph::CopySpan *spans = nullptr;  // page-locked: read by the copy kernels
    uint64_t spans_cap = 0;
    void *items = nullptr;  // page-locked: the RLWE launch's item list
    uint64_t items_cap = 0;
    uint8_t *maps = nullptr;  // page-locked: the products' row -> limb map
    uint64_t maps_cap = 0;
    bool busy = false;
};

struct CopyJob {
    void *dst;
    const void *src;
    uint64_t bytes;
};

// Host-side staging copies of one thread (the dispatcher's inputs or the
// completer's outputs) spread over a few helper threads: a single thread's
// memcpy bandwidth, not PCIe, otherwise caps the request rate.  run() returns
// when every job is done; small batches are copied by the caller alone.  One
// caller at a time.
class CopyPool {
  public:
    explicit CopyPool(unsigned helpers) {
        for (unsigned i = 0; i < helpers; ++i) threads_.emplace_back([this] { worker(); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread &t : threads_) t.join();
    }
    void run(const std::vector<CopyJob> &jobs) {
        uint64_t total = 0;
        for (const CopyJob &j : jobs) total += j.bytes;
        if (threads_.empty() || total < (4ull << 20)) {
            for (const CopyJob &j : jobs) std::memcpy(j.dst, j.src, j.bytes);
            return;
        }
        // pieces of at most 2 MiB so that the helpers share the work evenly
        std::vector<CopyJob> pieces;
        for (const CopyJob &j : jobs)
            for (uint64_t o = 0; o < j.bytes; o += 2ull << 20)
                pieces.push_back({(char *)j.dst + o, (const char *)j.src + o, std::min<uint64_t>(2ull << 20, j.bytes - o)});
        std::unique_lock<std::mutex> lk(mu_);
        // a helper still leaving the previous run reads pieces_: let it go first
        cv_done_.wait(lk, [&] { return active_ == 0; });
        pieces_.swap(pieces);
        next_.store(0);
        done_ = 0;
        ++generation_;
        ++active_;  // the caller works too
        lk.unlock();
        cv_.notify_all();
        const size_t finished = drain();
        lk.lock();
        done_ += finished;
        --active_;
        cv_done_.wait(lk, [&] { return done_ == pieces_.size() && active_ == 0; });
    }

  private:
    size_t drain() {
        size_t n = 0;
        for (size_t i = next_.fetch_add(1); i < pieces_.size(); i = next_.fetch_add(1), ++n)
            std::memcpy(pieces_[i].dst, pieces_[i].src, pieces_[i].bytes);
        return n;
    }
    void worker() {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(mu_);
        while (true) {
            cv_.wait(lk, [&] { return stop_ || generation_ != seen; });
            if (stop_) return;
            seen = generation_;
            ++active_;
            lk.unlock();
            const size_t finished = drain();
            lk.lock();
            done_ += finished;
            --active_;
            cv_done_.notify_all();
        }
    }
    std::vector<std::thread> threads_;
    std::vector<CopyJob> pieces_;  // replaced only under mu_ with no helper active
    std::atomic<size_t> next_{0};
    size_t done_ = 0, active_ = 0;
    uint64_t generation_ = 0;
    bool stop_ = false;
    std::mutex mu_;
    std::condition_variable cv_, cv_done_;
};

}  // namespace

struct ph_server {
    ph_server_config cfg{};
    std::vector<uint64_t> moduli;
    std::vector<ph_ntt_table *> tables;                 // [(log_n - min) * max_limbs + i]
    std::map<std::pair<uint32_t, uint32_t>, ph_dcrt_table *> dcrt;  // (log_n, limbs) -> view
    std::map<std::pair<uint32_t, uint32_t>, ph_dcrt_table *> single;  // (log_n, limb) -> one-limb view
    ph_rlwe_ctx *rlwe = nullptr;
    uint32_t rlwe_min = 0, rlwe_max = 0, rlwe_limbs = 0;
    // batch arenas: a pool of the server's own that keeps up to the HBM budget
    // mapped between batches instead of returning it at every synchronisation
    hipMemPool_t pool = nullptr;
    // every batch's input copies run on h2d and its result copies on d2h, in
    // batch order: each direction of the host link carries one copy at a time
    // at its full rate, and the two directions overlap each other and the
    // slots' compute
    hipStream_t h2d = nullptr, d2h = nullptr;
    std::vector<Slot> slots;

    std::mutex mu;
    std::condition_variable cv_pending, cv_done, cv_slot;
    std::deque<uint64_t> pending;
    // a deque: submissions append while the dispatcher and completer hold
    // references to earlier tickets outside the lock
    std::deque<Ticket> tickets;
    std::deque<Batch> inflight;
    uint64_t inflight_bytes = 0;
    // batches taken off the pending queue whose launch is being enqueued
    // (outside the lock) and that are not in `inflight` yet
    uint32_t launching = 0;
    bool stopping = false;
    bool dispatcher_done = false;  // set by the dispatcher as it exits
    // host-forced failures: the next fault_batches batches are completed with
    // PH_ERR_DEVICE without being launched
    uint32_t fault_batches = 0;
    bool paused = false;  // the dispatcher takes no new batch while set
    std::thread dispatcher, completer;

    ph_server_stats stats{};
    Clock::time_point started;
    // staging copies: inputs (dispatcher thread), outputs (completion thread)
    CopyPool copy_in{4}, copy_out{4};
};

namespace {

// True when [p, p + bytes) starts and ends in page-locked host memory mapped
// at the same address on the device, which kernels read and write directly.
bool page_locked(const void *p, uint64_t bytes) {
    if (!p || bytes == 0) return false;
    for (const char *q : {(const char *)p, (const char *)p + bytes - 1}) {
        hipPointerAttribute_t at{};
        if (hipPointerGetAttributes(&at, q) != hipSuccess) {
            (void)hipGetLastError();  // pageable memory reports an error: clear it
            return false;
        }
        if (at.type != hipMemoryTypeHost || at.devicePointer != (const void *)q) return false;
    }
    return true;
}

// Whether every buffer of a (validated) request is page-locked.
bool request_direct(const ph_serve_request &r) {
    const uint64_t n = 1ull << r.log_n;
    switch (r.kind) {
        case PH_REQ_POLY_MUL:
            return page_locked(r.in0, r.limbs * n * 8) && page_locked(r.in1, r.limbs * n * 8) &&
                   page_locked(r.out, r.limbs * n * 8);
        case PH_REQ_RLWE_ENCRYPT:
            return page_locked(r.in0, r.limbs * n * 8) && page_locked(r.s, n) && page_locked(r.e, n) &&
                   page_locked(r.m, n) && page_locked(r.out, r.limbs * n * 8);
        case PH_REQ_EXT_PRODUCT: return page_locked(r.in0, 2 * n * 8) && page_locked(r.out, 2 * n * 8);
    }
    return false;
}

// Device bytes a request holds while in flight (inputs, outputs, scratch).
uint64_t request_bytes(const ph_serve_request &r) {
    const uint64_t n = 1ull << r.log_n;
    switch (r.kind) {
        case PH_REQ_POLY_MUL: return 3 * round_up(r.limbs * n * 8);
        case PH_REQ_RLWE_ENCRYPT: return 2 * round_up(r.limbs * n * 8) + 3 * round_up(n) + 64ull * r.limbs;
        case PH_REQ_EXT_PRODUCT: return 2 * round_up(2 * n * 8);
    }
    return 0;
}

int validate(const ph_server *s, const ph_serve_request &r) {
    switch (r.kind) {
        case PH_REQ_POLY_MUL:
            if (r.log_n < s->cfg.min_log_n || r.log_n > s->cfg.max_log_n || r.limbs == 0 ||
                r.limbs > s->cfg.max_limbs || !r.in0 || !r.in1 || !r.out)
                return PH_ERR_INVALID;
            return PH_OK;
        case PH_REQ_RLWE_ENCRYPT:
            if (!s->rlwe || r.log_n < s->rlwe_min || r.log_n > s->rlwe_max || r.limbs == 0 || r.limbs > s->rlwe_limbs ||
                !r.in0 || !r.s || !r.e || !r.m || !r.out)
                return PH_ERR_INVALID;
            return PH_OK;
        case PH_REQ_EXT_PRODUCT:
            if (!s->cfg.ext_key || !r.in0 || !r.out || r.log_n != ph_ext_key_log_n(s->cfg.ext_key))
                return PH_ERR_INVALID;
            return PH_OK;
    }
    return PH_ERR_INVALID;
}

// Device scratch of an RLWE launch whose largest ring degree is 2^log_n: one
// polynomial of that degree for each of the kernel's resident workgroups (two
// per compute unit) when log_n > 12, none otherwise.
uint64_t rlwe_scratch_bytes(uint32_t log_n) {
    if (log_n <= 12) return 0;
    static const int cus = [] {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess) n = 256;
        return std::max(n, 1);
    }();
    return (uint64_t)cus * 2 * (1ull << log_n) * 8;
}

int ensure_spans(Slot &sl, uint64_t count) {
    if (sl.spans_cap >= count) return PH_OK;
    if (sl.spans) (void)hipHostFree(sl.spans);
    sl.spans = nullptr;
    sl.spans_cap = 0;
    const uint64_t cap = std::max<uint64_t>(count, 4096);
    if (hipHostMalloc((void **)&sl.spans, cap * sizeof(ph::CopySpan), hipHostMallocDefault) != hipSuccess)
        return PH_ERR_NO_MEMORY;
    sl.spans_cap = cap;
    return PH_OK;
}

// Appends the spans of one copy, split at kCopySpanMax bytes.
void add_spans(std::vector<ph::CopySpan> &v, uint64_t src, uint64_t dst, uint64_t bytes) {
    for (uint64_t o = 0; o < bytes; o += ph::kCopySpanMax)
        v.push_back({src + o, dst + o, std::min<uint64_t>(ph::kCopySpanMax, bytes - o)});
}

int ensure_items(Slot &sl, uint64_t bytes) {
    if (bytes <= sl.items_cap) return PH_OK;
    if (sl.items) (void)hipHostFree(sl.items);
    sl.items = nullptr;
    sl.items_cap = 0;
    const uint64_t cap = std::max<uint64_t>(bytes, 64 << 10);
    if (hipHostMalloc(&sl.items, cap, hipHostMallocDefault) != hipSuccess) return PH_ERR_NO_MEMORY;
    sl.items_cap = cap;
    return PH_OK;
}

int ensure_maps(Slot &sl, uint64_t bytes) {
    if (bytes <= sl.maps_cap) return PH_OK;
    if (sl.maps) (void)hipHostFree(sl.maps);
    sl.maps = nullptr;
    sl.maps_cap = 0;
    const uint64_t cap = std::max<uint64_t>(bytes, 16 << 10);
    if (hipHostMalloc((void **)&sl.maps, cap, hipHostMallocDefault) != hipSuccess) return PH_ERR_NO_MEMORY;
    sl.maps_cap = cap;
    return PH_OK;
}

int ensure_pinned(Slot &sl, uint64_t bytes) {
    if (sl.pinned_cap >= bytes) return PH_OK;
    if (sl.pinned) (void)hipHostFree(sl.pinned);
    sl.pinned = nullptr;
    sl.pinned_cap = 0;
    uint64_t cap = std::max<uint64_t>(bytes, 64ull << 20);
    if (hipHostMalloc(&sl.pinned, cap, hipHostMallocDefault) != hipSuccess) return PH_ERR_NO_MEMORY;
    sl.pinned_cap = cap;
    return PH_OK;
}

// Enqueues one batch on its slot's stream.  Called by the dispatcher without
// the lock held; the batch's tickets are not touched by anyone else meanwhile.
//
// Arena layout (staging mirrors it byte for byte): per poly-multiply ring
// degree an a block and a b block of limb-polynomial rows (each request's
// limbs consecutive; the product replaces a) and the two-pass scratch, one
// byte per row naming its limb (all degrees together), the RLWE requests'
// a_hat / b_hat / s / e / m regions with shared per-request offsets and the
// launch's item list and scratch, and the external products' input and output
// blocks.  Each group is one launch: the products of one ring degree share it
// whatever their limb counts.
int launch_batch(ph_server *s, Batch &b, const std::vector<Ticket *> &tk) {
    Slot &sl = s->slots[b.slot];
    hipStream_t st = sl.stream;
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        uint64_t o = off;
        off += round_up(bytes);
        return o;
    };
    // poly-multiply groups, one per ring degree
    std::map<uint32_t, std::vector<size_t>> groups;
    std::vector<size_t> rlwe, ext;
    for (size_t i = 0; i < tk.size(); ++i) {
        const ph_serve_request &r = tk[i]->req;
        if (r.kind == PH_REQ_POLY_MUL) groups[r.log_n].push_back(i);
        else if (r.kind == PH_REQ_RLWE_ENCRYPT) rlwe.push_back(i);
        else ext.push_back(i);
    }
    struct GroupAt { uint64_t a, b, scratch, rows, row0; ph_dcrt_table *t; uint32_t log_n; };
    std::vector<GroupAt> gat;
    std::vector<uint8_t> limb_of;  // every group's rows, in group order
    for (auto &g : groups) {
        const uint32_t log_n = g.first;
        uint64_t rows = 0;
        for (size_t i : g.second) rows += tk[i]->req.limbs;
        ph_dcrt_table *t = s->dcrt[{log_n, s->cfg.max_limbs}];
        const uint64_t row_bytes = 8ull << log_n;
        GroupAt x{take(rows * row_bytes), take(rows * row_bytes), take(ph::negacyclic_mul_scratch_bytes(t, rows)), rows,
                  limb_of.size(), t, log_n};
        uint64_t r0 = 0;
        for (size_t i : g.second) {
            Part &p = b.parts[i];
            const uint32_t limbs = tk[i]->req.limbs;
            p.in0 = x.a + r0 * row_bytes;
            p.in1 = x.b + r0 * row_bytes;
            p.out = p.in0;
            p.out_bytes = limbs * row_bytes;
            for (uint32_t j = 0; j < limbs; ++j) limb_of.push_back((uint8_t)j);
            r0 += limbs;
        }
        gat.push_back(x);
    }
    const uint64_t lmap = take(limb_of.size());
    // RLWE regions
    uint64_t a_words = 0, small = 0;
    std::vector<ph_rlwe_request> rq;
    for (size_t i : rlwe) {
        const ph_serve_request &r = tk[i]->req;
        ph_rlwe_request x;
        x.log_n = r.log_n;
        x.limbs = r.limbs;
        x.a_offset = a_words;
        x.small_offset = small;
        rq.push_back(x);
        a_words += (uint64_t)r.limbs << r.log_n;
        small += 1ull << r.log_n;
    }
    const uint64_t ra = take(a_words * 8), rout = take(a_words * 8), rs = take(small), re = take(small), rm = take(small);
    // the RLWE launch's item list and scratch, staged in the arena: no
    // allocation or host synchronisation inside the launch
    const ph::RlweStageSize rstage = rlwe.empty() ? ph::RlweStageSize{0, 0}
                                                  : ph::rlwe_stage_size(s->rlwe, rq.data(), (uint32_t)rq.size());
    if (!rlwe.empty() && rstage.item_bytes == 0) return PH_ERR_INVALID;
    const uint64_t ritems = take(rstage.item_bytes), rscratch = take(rstage.scratch_bytes);
    for (size_t k = 0; k < rlwe.size(); ++k) {
        Part &p = b.parts[rlwe[k]];
        p.in0 = ra + rq[k].a_offset * 8;
        p.out = rout + rq[k].a_offset * 8;
        p.small = rq[k].small_offset;
        p.out_bytes = ((uint64_t)rq[k].limbs << rq[k].log_n) * 8;
    }
    // external products (all with the server key's ring degree)
    uint64_t ext_row = 0;
    if (!ext.empty()) ext_row = 2ull * (1ull << tk[ext[0]]->req.log_n) * 8;
    const uint64_t xin = take(ext_row * ext.size()), xout = take(ext_row * ext.size());
    for (size_t k = 0; k < ext.size(); ++k) {
        Part &p = b.parts[ext[k]];
        p.in0 = xin + k * ext_row;
        p.out = xout + k * ext_row;
        p.out_bytes = ext_row;
    }
    b.arena_bytes = off;

    // Inputs: page-locked requests are copied to the arena directly; the
    // others through the slot's staging mirror of the arena (one copy, issued
    // first so that the direct copies land over it).
    size_t staged = 0;
    for (const Ticket *t : tk) staged += t->direct ? 0 : 1;
    if (staged && ensure_pinned(sl, off) != PH_OK) return PH_ERR_NO_MEMORY;
    char *host = (char *)sl.pinned;
    std::vector<CopyJob> jobs;  // staging copies
    std::vector<std::tuple<uint64_t, const void *, uint64_t>> direct;  // (arena offset, source, bytes)
    for (size_t i = 0; i < tk.size(); ++i) {
        const ph_serve_request &r = tk[i]->req;
        const Part &p = b.parts[i];
        const uint64_t n = 1ull << r.log_n;
        auto add = [&](uint64_t at, const void *src, uint64_t bytes) {
            if (tk[i]->direct) direct.emplace_back(at, src, bytes);
            else jobs.push_back({host + at, src, bytes});
        };
        add(p.in0, r.in0, p.out_bytes);
        if (r.kind == PH_REQ_POLY_MUL) add(p.in1, r.in1, p.out_bytes);
        if (r.kind == PH_REQ_RLWE_ENCRYPT) {
            add(rs + p.small, r.s, n);
            add(re + p.small, r.e, n);
            add(rm + p.small, r.m, n);
        }
    }
    s->copy_in.run(jobs);
    void *arena = nullptr;
    if (!rlwe.empty() && ensure_items(sl, rstage.item_bytes) != PH_OK) return PH_ERR_NO_MEMORY;
    if (!limb_of.empty() && !s->cfg.per_poly_launches && ensure_maps(sl, limb_of.size()) != PH_OK)
        return PH_ERR_NO_MEMORY;
    if (hipMallocFromPoolAsync(&arena, std::max<uint64_t>(off, kAlign), s->pool, st) != hipSuccess)
        return PH_ERR_NO_MEMORY;
    char *dev = (char *)arena;
    int rc = PH_OK;
    const hipStream_t sin = s->h2d, sout = s->d2h;
    if (hipEventRecord(sl.ev_alloc, st) != hipSuccess || hipStreamWaitEvent(sin, sl.ev_alloc, 0) != hipSuccess)
        rc = PH_ERR_DEVICE;
    // staged inputs: the arena's input regions only (not the results, the
    // scratch or the item list), each one copy
    if (staged) {
        std::vector<std::pair<uint64_t, uint64_t>> in_regions;  // (offset, bytes)
        for (const GroupAt &g : gat) in_regions.push_back({g.a, g.scratch - g.a});  // a and b blocks
        if (!rlwe.empty()) {
            in_regions.push_back({ra, rout - ra});     // a_hat
            in_regions.push_back({rs, ritems - rs});   // s, e, m
        }
        if (!ext.empty()) in_regions.push_back({xin, xout - xin});
        for (const auto &rg : in_regions) {
            if (rc != PH_OK || rg.second == 0) continue;
            if (hipMemcpyAsync(dev + rg.first, host + rg.first, rg.second, hipMemcpyHostToDevice, sin) != hipSuccess)
                rc = PH_ERR_DEVICE;
        }
    }
    // page-locked requests: every copy of the batch, in and out, as spans of
    // two copy-kernel launches (inputs now, results after the compute)
    std::vector<ph::CopySpan> spans_in, spans_out;
    for (const auto &[at, src, bytes] : direct) add_spans(spans_in, (uint64_t)(uintptr_t)src, (uint64_t)(uintptr_t)(dev + at), bytes);
    for (size_t i = 0; i < tk.size(); ++i)
        if (tk[i]->direct)
            add_spans(spans_out, (uint64_t)(uintptr_t)(dev + b.parts[i].out), (uint64_t)(uintptr_t)tk[i]->req.out,
                      b.parts[i].out_bytes);
    if (rc == PH_OK && (!spans_in.empty() || !spans_out.empty())) {
        if (ensure_spans(sl, spans_in.size() + spans_out.size()) != PH_OK) rc = PH_ERR_NO_MEMORY;
        else {
            std::memcpy(sl.spans, spans_in.data(), spans_in.size() * sizeof(ph::CopySpan));
            std::memcpy(sl.spans + spans_in.size(), spans_out.data(), spans_out.size() * sizeof(ph::CopySpan));
            if (ph::launch_copy_spans(sl.spans, (uint32_t)spans_in.size(), ph::kCopyGroupsIn, sin) != hipSuccess)
                rc = PH_ERR_DEVICE;
        }
    }
    // the row -> limb map, after the staging copy of the whole arena
    if (rc == PH_OK && !limb_of.empty() && !s->cfg.per_poly_launches) {
        std::memcpy(sl.maps, limb_of.data(), limb_of.size());
        if (hipMemcpyAsync(dev + lmap, sl.maps, limb_of.size(), hipMemcpyHostToDevice, sin) != hipSuccess)
            rc = PH_ERR_DEVICE;
    }
    if (rc == PH_OK && (hipEventRecord(sl.ev_in, sin) != hipSuccess || hipStreamWaitEvent(st, sl.ev_in, 0) != hipSuccess))
        rc = PH_ERR_DEVICE;
    for (const GroupAt &g : gat) {
        if (rc != PH_OK) break;
        const uint64_t row_bytes = 8ull << g.log_n;
        if (s->cfg.per_poly_launches) {
            for (uint64_t k = 0; k < g.rows && rc == PH_OK; ++k) {
                rc = ph_dcrt_negacyclic_mul_device(s->single[{g.log_n, (uint32_t)limb_of[g.row0 + k]}],
                                                   (uint64_t *)(dev + g.a + k * row_bytes),
                                                   (const uint64_t *)(dev + g.b + k * row_bytes), 1, st);
                ++s->stats.launches;
            }
            continue;
        }
        rc = ph::negacyclic_mul_rows_staged(g.t, (uint64_t *)(dev + g.a), (const uint64_t *)(dev + g.b),
                                            (const uint8_t *)(dev + lmap + g.row0), g.rows, dev + g.scratch, st);
        ++s->stats.launches;
    }
    if (rc == PH_OK && !rlwe.empty()) {
        rc = ph::rlwe_encrypt_staged(s->rlwe, rq.data(), (uint32_t)rq.size(), (const uint64_t *)(dev + ra),
                                     (const int8_t *)(dev + rs), (const int8_t *)(dev + re),
                                     (const uint8_t *)(dev + rm), (uint64_t *)(dev + rout), st, sl.items,
                                     dev + ritems, dev + rscratch, rstage);
        ++s->stats.launches;
    }
    if (rc == PH_OK && !ext.empty()) {
        rc = ph_ext_product_device(s->cfg.ext_key, (const uint64_t *)(dev + xin), (uint64_t *)(dev + xout), ext.size(),
                                   0, st);
        ++s->stats.launches;
    }
    // Results: page-locked requests straight into the caller's buffer; the
    // others into the staging area (one copy per contiguous result region --
    // each product group's a blocks, the RLWE b_hat rows, the external
    // products' outputs -- when no request of the batch is direct), for the
    // completion thread to hand over.  Then the event the completer waits on.
    std::vector<std::pair<uint64_t, uint64_t>> regions;  // (offset, bytes)
    if (staged == tk.size()) {
        for (const GroupAt &g : gat) regions.push_back({g.a, g.b - g.a});
        if (!rlwe.empty()) regions.push_back({rout, a_words * 8});
        if (!ext.empty()) regions.push_back({xout, ext_row * ext.size()});
    } else {
        for (size_t i = 0; i < tk.size(); ++i)
            if (!tk[i]->direct) regions.push_back({b.parts[i].out, b.parts[i].out_bytes});
    }
    if (rc == PH_OK &&
        (hipEventRecord(sl.ev_comp, st) != hipSuccess || hipStreamWaitEvent(sout, sl.ev_comp, 0) != hipSuccess))
        rc = PH_ERR_DEVICE;
    for (const auto &rg : regions) {
        if (rc != PH_OK || rg.second == 0) continue;
        if (hipMemcpyAsync(host + rg.first, dev + rg.first, rg.second, hipMemcpyDeviceToHost, sout) != hipSuccess)
            rc = PH_ERR_DEVICE;
    }
    if (rc == PH_OK && !spans_out.empty() &&
        ph::launch_copy_spans(sl.spans + spans_in.size(), (uint32_t)spans_out.size(), ph::kCopyGroupsOut, sout) !=
            hipSuccess)
        rc = PH_ERR_DEVICE;
    // the arena goes back (and the batch completes) once the results are out;
    // on an error the slot stream still waits for whatever the copy streams
    // were given
    if (hipEventRecord(sl.ev_out, sout) != hipSuccess || hipStreamWaitEvent(st, sl.ev_out, 0) != hipSuccess)
        rc = rc == PH_OK ? PH_ERR_DEVICE : rc;
    if (hipEventRecord(sl.ev_in, sin) != hipSuccess || hipStreamWaitEvent(st, sl.ev_in, 0) != hipSuccess)
        rc = rc == PH_OK ? PH_ERR_DEVICE : rc;
    (void)hipFreeAsync(arena, st);
    if (hipEventRecord(b.done, st) != hipSuccess && rc == PH_OK) rc = PH_ERR_DEVICE;
    return rc;
}

void dispatcher_loop(ph_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        s->cv_pending.wait(lk, [&] { return s->stopping || (!s->pending.empty() && !s->paused); });
        if (s->pending.empty() && s->stopping) {
            s->dispatcher_done = true;
            s->cv_done.notify_all();  // the completer exits once in-flight work is retired
            return;
        }
        // a free slot
        int slot = -1;
        const auto w0 = Clock::now();
        s->cv_slot.wait(lk, [&] {
            for (size_t i = 0; i < s->slots.size(); ++i)
                if (!s->slots[i].busy) return true;
            return false;
        });
        s->stats.slot_wait_s += std::chrono::duration<double>(Clock::now() - w0).count();
        for (size_t i = 0; i < s->slots.size(); ++i)
            if (!s->slots[i].busy) {
                slot = (int)i;
                break;
            }
        // admit in arrival order while the budget allows
        Batch b;
        std::vector<Ticket *> tk;
        uint32_t rlwe_max_ln = 0;  // largest RLWE ring degree admitted to this batch
        while (!s->pending.empty() && tk.size() < s->cfg.max_batch) {
            Ticket &t = s->tickets[s->pending.front()];
            uint64_t need = t.bytes;
            uint32_t ln = rlwe_max_ln;
            if (t.req.kind == PH_REQ_RLWE_ENCRYPT && t.req.log_n > ln) {
                ln = t.req.log_n;
                need += rlwe_scratch_bytes(ln) - rlwe_scratch_bytes(rlwe_max_ln);
            }
            // every request fits the budget alone (checked at submission), so
            // an empty device always admits the first one
            if (s->inflight_bytes + b.budget_bytes + need > s->cfg.hbm_budget_bytes) {
                if (!tk.empty() || !s->inflight.empty()) break;  // wait for memory to free up
            }
            rlwe_max_ln = ln;
            b.budget_bytes += need;
            Part p{};
            p.ticket = s->pending.front();
            b.parts.push_back(p);
            tk.push_back(&t);
            s->pending.pop_front();
        }
        if (tk.empty()) {
            s->cv_done.wait(lk);  // memory is held by in-flight batches
            continue;
        }
        b.slot = slot;
        s->slots[slot].busy = true;
        s->inflight_bytes += b.budget_bytes;
        s->stats.peak_inflight_bytes = std::max(s->stats.peak_inflight_bytes, s->inflight_bytes);
        (void)hipEventCreateWithFlags(&b.done, hipEventDisableTiming);
        const bool fault = s->fault_batches > 0;
        if (fault) --s->fault_batches;
        ++s->launching;
        lk.unlock();
        const auto d0 = Clock::now();
        // a forced fault enqueues nothing: the unrecorded event is complete
        // and the batch retires with the error
        b.status = fault ? PH_ERR_DEVICE : launch_batch(s, b, tk);
        const double took = std::chrono::duration<double>(Clock::now() - d0).count();
        lk.lock();
        s->stats.dispatch_s += took;
        --s->launching;
        s->inflight.push_back(std::move(b));
        ++s->stats.batches;
        s->cv_done.notify_all();  // wakes the completer
    }
}

void completer_loop(ph_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        // exits only after the dispatcher has: a batch it is still enqueueing
        // is in neither queue, and its results must still be retired
        s->cv_done.wait(lk, [&] { return !s->inflight.empty() || s->dispatcher_done; });
        if (s->inflight.empty()) {
            if (s->dispatcher_done) return;
            continue;
        }
        Batch &b = s->inflight.front();
        hipEvent_t ev = b.done;
        Slot &sl = s->slots[b.slot];
        // the callers' buffers, read while the ticket list cannot grow
        std::vector<CopyJob> jobs;
        for (const Part &p : b.parts) {
            const Ticket &t = s->tickets[p.ticket];
            if (!t.direct) jobs.push_back({t.req.out, (char *)sl.pinned + p.out, p.out_bytes});
        }
        lk.unlock();
        int st = b.status;
        const auto e0 = Clock::now();
        if (hipEventSynchronize(ev) != hipSuccess && st == PH_OK) st = PH_ERR_DEVICE;
        const double waited = std::chrono::duration<double>(Clock::now() - e0).count();
        // results from the pinned staging area to the callers' buffers
        if (st == PH_OK) s->copy_out.run(jobs);
        const auto now = Clock::now();
        lk.lock();
        s->stats.device_wait_s += waited;
        for (const Part &p : b.parts) {
            Ticket &t = s->tickets[p.ticket];
            t.done_at = now;
            t.status = st;
            ++s->stats.completed;
        }
        s->inflight_bytes -= b.budget_bytes;
        sl.busy = false;
        (void)hipEventDestroy(ev);
        s->inflight.pop_front();
        s->cv_done.notify_all();
        s->cv_slot.notify_all();
    }
}

void destroy(ph_server *s) {
    if (!s) return;
    for (auto &kv : s->dcrt) ph_dcrt_table_free(kv.second);
    for (auto &kv : s->single) ph_dcrt_table_free(kv.second);
    for (ph_ntt_table *t : s->tables) ph_ntt_table_free(t);
    if (s->rlwe) ph_rlwe_ctx_free(s->rlwe);
    for (Slot &sl : s->slots) {
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        for (hipEvent_t e : {sl.ev_alloc, sl.ev_in, sl.ev_comp, sl.ev_out})
            if (e) (void)hipEventDestroy(e);
        if (sl.pinned) (void)hipHostFree(sl.pinned);
        if (sl.spans) (void)hipHostFree(sl.spans);
        if (sl.items) (void)hipHostFree(sl.items);
        if (sl.maps) (void)hipHostFree(sl.maps);
    }
    if (s->h2d) (void)hipStreamDestroy(s->h2d);
    if (s->d2h) (void)hipStreamDestroy(s->d2h);
    if (s->pool) (void)hipMemPoolDestroy(s->pool);
    delete s;
}

}  // namespace

extern "C" {

int ph_server_new(const ph_server_config *cfg, ph_server **out) {
    if (!cfg || !out || !cfg->moduli || cfg->max_limbs == 0 || cfg->min_log_n > cfg->max_log_n ||
        cfg->streams == 0 || cfg->max_batch == 0)
        return PH_ERR_INVALID;
    *out = nullptr;
    ph_server *s = new (std::nothrow) ph_server();
    if (!s) return PH_ERR_NO_MEMORY;
    s->cfg = *cfg;
    const uint32_t degrees = cfg->max_log_n - cfg->min_log_n + 1;
    s->moduli.assign(cfg->moduli, cfg->moduli + (size_t)degrees * cfg->max_limbs);
    s->cfg.moduli = s->moduli.data();
    int rc = PH_OK;
    for (uint32_t d = 0; d < degrees && rc == PH_OK; ++d)
        for (uint32_t i = 0; i < cfg->max_limbs && rc == PH_OK; ++i) {
            ph_ntt_table *t = nullptr;
            rc = ph_ntt_table_new(cfg->min_log_n + d, s->moduli[(size_t)d * cfg->max_limbs + i], &t);
            if (rc == PH_OK) s->tables.push_back(t);
        }
    for (uint32_t d = 0; d < degrees && rc == PH_OK; ++d)
        for (uint32_t L = 1; L <= cfg->max_limbs && rc == PH_OK; ++L) {
            ph_dcrt_table *v = nullptr;
            rc = ph_dcrt_table_view(&s->tables[(size_t)d * cfg->max_limbs], L, &v);
            if (rc == PH_OK) s->dcrt[{cfg->min_log_n + d, L}] = v;
        }
    for (uint32_t d = 0; d < degrees && rc == PH_OK && cfg->per_poly_launches; ++d)
        for (uint32_t i = 0; i < cfg->max_limbs && rc == PH_OK; ++i) {
            ph_dcrt_table *v = nullptr;
            rc = ph_dcrt_table_view(&s->tables[(size_t)d * cfg->max_limbs + i], 1, &v);
            if (rc == PH_OK) s->single[{cfg->min_log_n + d, i}] = v;
        }
    // RLWE tables over the part of the grid its kernel covers (2^10 .. 2^16)
    s->rlwe_min = std::max(cfg->min_log_n, 10u);
    s->rlwe_max = std::min(cfg->max_log_n, 16u);
    s->rlwe_limbs = cfg->max_limbs;
    if (rc == PH_OK && s->rlwe_min <= s->rlwe_max) {
        std::vector<uint64_t> sub;
        for (uint32_t ln = s->rlwe_min; ln <= s->rlwe_max; ++ln)
            for (uint32_t i = 0; i < s->rlwe_limbs; ++i)
                sub.push_back(s->moduli[(size_t)(ln - cfg->min_log_n) * cfg->max_limbs + i]);
        rc = ph_rlwe_ctx_new(s->rlwe_min, s->rlwe_max, s->rlwe_limbs, sub.data(), &s->rlwe);
    }
    if (rc == PH_OK) {
        int dev = 0;
        hipMemPoolProps props{};
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        if (hipGetDevice(&dev) != hipSuccess) rc = PH_ERR_DEVICE;
        props.location.id = dev;
        if (rc == PH_OK && hipMemPoolCreate(&s->pool, &props) != hipSuccess) {
            s->pool = nullptr;
            rc = PH_ERR_DEVICE;
        }
        uint64_t keep = cfg->hbm_budget_bytes;
        if (rc == PH_OK && hipMemPoolSetAttribute(s->pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess)
            rc = PH_ERR_DEVICE;
        // memory freed on one slot's stream is reused by another only once the
        // free has completed, never by making that stream wait for it (which
        // would serialise the slots)
        int no = 0;
        if (rc == PH_OK && hipMemPoolSetAttribute(s->pool, hipMemPoolReuseAllowInternalDependencies, &no) != hipSuccess)
            rc = PH_ERR_DEVICE;
    }
    s->slots.resize(cfg->streams);
    for (Slot &sl : s->slots) {
        if (rc == PH_OK && hipStreamCreate(&sl.stream) != hipSuccess) rc = PH_ERR_DEVICE;
        for (hipEvent_t *e : {&sl.ev_alloc, &sl.ev_in, &sl.ev_comp, &sl.ev_out})
            if (rc == PH_OK && hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) rc = PH_ERR_DEVICE;
    }
    if (rc == PH_OK && (hipStreamCreate(&s->h2d) != hipSuccess || hipStreamCreate(&s->d2h) != hipSuccess))
        rc = PH_ERR_DEVICE;
    if (rc != PH_OK) {
        destroy(s);
        return rc;
    }
    s->started = Clock::now();
    s->dispatcher = std::thread(dispatcher_loop, s);
    s->completer = std::thread(completer_loop, s);
    *out = s;
    return PH_OK;
}

int ph_server_submit(ph_server *s, const ph_serve_request *req, uint64_t *ticket) {
    if (!s || !req) return PH_ERR_INVALID;
    const auto t0 = Clock::now();
    int rc = validate(s, *req);
    if (rc != PH_OK) return rc;
    const uint64_t bytes = request_bytes(*req);
    const uint64_t scratch = req->kind == PH_REQ_RLWE_ENCRYPT ? rlwe_scratch_bytes(req->log_n) : 0;
    if (bytes + scratch > s->cfg.hbm_budget_bytes) return PH_ERR_NO_MEMORY;  // can never be admitted
    const bool direct = request_direct(*req);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return PH_ERR_INVALID;
    Ticket t;
    t.req = *req;
    t.direct = direct;
    t.arrival = Clock::now();
    t.bytes = bytes;
    s->tickets.push_back(t);
    const uint64_t id = s->tickets.size() - 1;
    s->pending.push_back(id);
    if (ticket) *ticket = id;
    s->stats.submit_s += std::chrono::duration<double>(Clock::now() - t0).count();
    s->cv_pending.notify_one();
    return PH_OK;
}

int ph_server_wait(ph_server *s, uint64_t ticket, int *status) {
    if (!s) return PH_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    if (ticket >= s->tickets.size()) return PH_ERR_INVALID;
    s->cv_done.wait(lk, [&] { return s->tickets[ticket].status >= 0; });
    if (status) *status = s->tickets[ticket].status;
    return PH_OK;
}

int ph_server_latency(ph_server *s, uint64_t ticket, double *seconds) {
    if (!s || !seconds) return PH_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (ticket >= s->tickets.size() || s->tickets[ticket].status < 0) return PH_ERR_INVALID;
    const Ticket &t = s->tickets[ticket];
    *seconds = std::chrono::duration<double>(t.done_at - t.arrival).count();
    return PH_OK;
}

int ph_server_drain(ph_server *s) {
    if (!s) return PH_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->pending.empty() && s->launching == 0 && s->inflight.empty(); });
    return PH_OK;
}

int ph_server_cancel(ph_server *s, uint64_t ticket) {
    if (!s) return PH_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (ticket >= s->tickets.size()) return PH_ERR_INVALID;
    auto it = std::find(s->pending.begin(), s->pending.end(), ticket);
    if (it == s->pending.end()) return PH_ERR_INVALID;  // already in a batch, or finished
    s->pending.erase(it);
    Ticket &t = s->tickets[ticket];
    t.status = PH_ERR_CANCELLED;
    t.done_at = Clock::now();
    ++s->stats.cancelled;
    s->cv_done.notify_all();
    return PH_OK;
}

int ph_server_pause(ph_server *s, int paused) {
    if (!s) return PH_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->paused = paused != 0;
    }
    s->cv_pending.notify_all();
    return PH_OK;
}

int ph_server_inject_fault(ph_server *s, uint32_t batches) {
    if (!s) return PH_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    s->fault_batches = batches;
    return PH_OK;
}

int ph_server_stats_get(ph_server *s, ph_server_stats *out) {
    if (!s || !out) return PH_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    return PH_OK;
}

void ph_server_free(ph_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        s->paused = false;
    }
    s->cv_pending.notify_all();
    s->cv_done.notify_all();
    s->cv_slot.notify_all();
    // the dispatcher launches what is still pending and exits; the completer
    // retires every in-flight batch after that, so all tickets complete
    if (s->dispatcher.joinable()) s->dispatcher.join();
    s->cv_done.notify_all();
    if (s->completer.joinable()) s->completer.join();
    destroy(s);
}

int ph_server_run_poisson(ph_server *s, const ph_serve_request *reqs, size_t n, double rate_per_s, uint64_t seed,
                          double *latency_s, double *elapsed_s) {
    if (!s || (!reqs && n) || rate_per_s <= 0) return PH_ERR_INVALID;
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(rate_per_s);
    std::vector<uint64_t> ids(n);
    const auto t0 = Clock::now();
    double at = 0;
    for (size_t i = 0; i < n; ++i) {
        at += gap(rng);
        const auto due = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(at));
        // sleep most of the gap, spin the last stretch for sub-100us arrival accuracy
        auto now = Clock::now();
        if (due - now > std::chrono::microseconds(200)) std::this_thread::sleep_until(due - std::chrono::microseconds(100));
        while (Clock::now() < due) {
        }
        int rc = ph_server_submit(s, &reqs[i], &ids[i]);
        if (rc != PH_OK) return rc;
    }
    int rc = ph_server_drain(s);
    const auto t1 = Clock::now();
    if (elapsed_s) *elapsed_s = std::chrono::duration<double>(t1 - t0).count();
    for (size_t i = 0; i < n && rc == PH_OK; ++i) {
        int st = 0;
        ph_server_wait(s, ids[i], &st);
        if (st != PH_OK) rc = st;
        if (latency_s) ph_server_latency(s, ids[i], &latency_s[i]);
    }
    return rc;
}

}  // extern "C"
