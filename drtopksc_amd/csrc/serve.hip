// serve.hip -- continuous-batching top-k server over a device-resident key store
// (drtopk_server_*), its trace replayer and the one-at-a-time baseline.
//
// The reference has no serving layer: every driver handles one vector per process
// (baseline+filter+beta+shuffle/main.cu:56-323) and the MultiGPU driver loops vectors
// (MultiGPU/main.cu:366-424).  Here max_inflight worker threads each own one HIP stream
// (default priority) and keep one launch on it.  A worker whose stream is free takes the
// queue: a large request becomes a single-vector drtopk_select call, otherwise every queued
// small request goes into one drtopk_select_batched launch.  Launches are prepared under the
// server lock and issued outside it, so the host cost of launching overlaps across workers.  All scratch and outputs come from one device arena of
// budget_bytes (first fit, 256-byte granules); a launch that does not fit waits for running
// launches to finish, which is the admission rule.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/drtopk.h"
#include "common.hpp"
#include "host_util.hpp"

namespace drtopk {
namespace {

u64 now_ns() {
    return (u64)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// First-fit allocator over [0, size) of one device allocation.
class Arena {
   public:
    void init(u64 size) {
        size_ = size;
        free_.clear();
        free_[0] = size;
    }
    bool alloc(u64 bytes, u64* off) {
        bytes = align_up(std::max<u64>(bytes, 1), 256);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const u64 rest = it->second - bytes, at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used_ += bytes;
            return true;
        }
        return false;
    }
    void release(u64 off, u64 bytes) {
        bytes = align_up(std::max<u64>(bytes, 1), 256);
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
    u64 largest() const {  // the largest free block
        u64 m = 0;
        for (const auto& f : free_) m = std::max(m, f.second);
        return m;
    }
    u64 size() const { return size_; }
    u64 used() const { return used_; }

   private:
    std::map<u64, u64> free_;
    u64 size_ = 0, used_ = 0;
};

enum ReqState { QUEUED = 0, RUNNING = 1, DONE = 2 };

struct OutBlock {  // outputs of one launch; freed when every request of it is released
    u64 off = 0, bytes = 0;
    int refs = 0;
};

struct Req {
    drtopk_query q;
    int state = QUEUED;
    drtopk_status status = DRTOPK_OK;
    u64 t_submit = 0, t_launch = 0, t_done = 0;
    u32 kth = 0;
    u64 keys_off = 0, idx_off = 0;  // byte offsets of its outputs in the arena
    u64 block = 0;                  // OutBlock id
    u32 launch_rows = 0, solo = 0;
    u64 need = 0;  // arena bytes of its outputs and scratch when it runs as a single-vector call
    bool released = false;
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    std::vector<u64> tickets;
    u64 scratch_off = 0, scratch_bytes = 0;  // workspace (+ device plan) of the launch
    unsigned char* h_plan = nullptr;          // pinned staging of the plan
    size_t h_plan_cap = 0;
    u32* h_kth = nullptr;                     // pinned k-th keys, copied back by the launch
    size_t h_kth_cap = 0;
    u32* h_err = nullptr;                     // pinned device-side error bits of the launch
    drtopk_query solo_q{};                    // the request of a single-vector launch
    std::vector<void*> retired;               // outgrown pinned buffers, freed by destroy
};

}  // namespace
}  // namespace drtopk

struct drtopk_server {
    drtopk_server_config cfg;
    drtopk_tuning tuning;
    bool has_tuning = false;
    unsigned char* d_arena = nullptr;
    drtopk::Arena arena;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> queue;
    std::unordered_map<uint64_t, drtopk::Req> reqs;
    std::unordered_map<uint64_t, drtopk::OutBlock> blocks;
    uint64_t next_ticket = 1, next_block = 1;
    std::vector<drtopk::Slot> slots;
    std::vector<std::thread> workers;
    bool stop = false;
    int device = 0;
};

namespace drtopk {
namespace {

bool is_solo(const drtopk_server* s, const drtopk_query& q) { return q.n >= s->cfg.solo_n || q.k > s->cfg.solo_k; }

// Parameters of a single-vector launch (the server's large requests and the one-at-a-time
// baseline): the reference configuration with the measured (alpha, beta) table, which returns
// the same outputs as the reference policy.
void solo_params(drtopk_params* p) {
    drtopk_params_default(p);
    p->alpha = DRTOPK_ALPHA_TUNED;
}

u64 solo_ws(const drtopk_query& q) {
    drtopk_params p;
    solo_params(&p);
    size_t b = 0;
    if (drtopk_workspace_size(DRTOPK_ALGO_RADIX, (drtopk_dtype)q.dtype, q.n, q.k, &p, &b) != DRTOPK_OK) return ~0ull;
    return b;
}
u64 out_bytes(u64 rows, u64 ksum) { return align_up(4 * rows, 256) + align_up(4 * ksum, 256) + 8 * ksum; }

// Grow a pinned host buffer (only while its slot is idle).  The outgrown buffer is kept until
// destroy: hipHostFree synchronises the whole device, and this runs under the server lock, so
// freeing here would hold every worker until all running launches had finished.  Sizes double,
// so the kept buffers add up to less than the current one.
template <class T>
bool ensure_pinned(T** p, size_t* cap, size_t bytes, std::vector<void*>& retired) {
    if (*cap >= bytes) return true;
    T* np = nullptr;
    const size_t nb = std::max<size_t>(bytes, 4096) * 2;
    if (hipHostMalloc((void**)&np, nb, hipHostMallocDefault) != hipSuccess) return false;
    if (*p) retired.push_back(*p);
    *p = np;
    *cap = nb;
    return true;
}

void finish_request(drtopk_server* s, u64 t, drtopk_status st, u64 done_at, u32 kth) {
    Req& r = s->reqs[t];
    r.state = DONE;
    r.status = st;
    r.t_done = done_at;
    r.kth = kth;
}

// A launch prepared under the lock and issued outside it.
struct Prep {
    std::vector<u64> take;
    std::vector<drtopk_row> rows;
    bool solo = false;
    u64 ksum = 0, out_off = 0, out_bytes = 0, scratch_off = 0, scratch = 0;
    size_t plan_bytes = 0, ws_bytes = 0;
};

// Queue entries a batch looks at past its first row: bounds the work done under the lock.
constexpr size_t BATCH_SCAN = 16384;

// Pick the next launch from the queue and reserve its arena space.  Lock held.  Returns false
// when the queue is empty or nothing fits the arena now.
//
// A batch takes the small requests of the queue in arrival order, also those behind queued large
// requests, which keeps launches few when large and small requests interleave.  A client that
// waits and releases in submission order holds the outputs of such a batch until it has waited
// for the large requests the batch passed, so the batch is taken only if the arena's largest free
// block still holds the outputs and scratch of each of them afterwards; otherwise only the small
// requests ahead of the first large one go.  Every batch that passes a request leaves it room,
// and whatever else holds arena space then is scratch of running launches or outputs the client
// releases before waiting for it, so the request fits once it reaches the head of the queue.
bool prepare(drtopk_server* s, Slot& sl, Prep& pr) {
    if (s->queue.empty()) return false;
    pr = Prep();
    const u64 head = s->queue.front();
    pr.solo = is_solo(s, s->reqs[head].q);
    std::vector<u64> passed;  // per taken request: the largest need of the large requests before it
    size_t in_order = 1;      // taken requests ahead of the first large one
    if (pr.solo) {
        pr.take.push_back(head);
    } else {
        u64 need = 0;
        size_t seen = 0;
        for (u64 t : s->queue) {
            if (pr.take.size() >= s->cfg.max_batch_rows || seen++ > BATCH_SCAN) break;
            const Req& r = s->reqs[t];
            if (is_solo(s, r.q)) {
                need = std::max(need, r.need);
                continue;
            }
            pr.take.push_back(t);
            passed.push_back(need);
            if (need == 0) in_order = pr.take.size();
        }
    }
    for (;;) {  // shrink the batch until outputs + scratch fit the arena
        const u64 rows = pr.take.size();
        pr.ksum = 0;
        for (u64 t : pr.take) pr.ksum += s->reqs[t].q.k;
        pr.out_bytes = out_bytes(rows, pr.ksum);
        if (pr.solo) {
            pr.scratch = solo_ws(s->reqs[pr.take[0]].q);
        } else {
            pr.rows.resize(rows);
            u64 oo = 0;
            for (u64 i = 0; i < rows; ++i) {
                const drtopk_query& q = s->reqs[pr.take[i]].q;
                pr.rows[i] = drtopk_row{q.offset, q.n, q.k, oo, q.dtype, 0};
                oo += q.k;
            }
            if (drtopk_batch_plan(pr.rows.data(), rows, s->cfg.store_len, pr.ksum, s->has_tuning ? &s->tuning : nullptr,
                                  nullptr, &pr.plan_bytes, &pr.ws_bytes) != DRTOPK_OK)
                pr.scratch = ~0ull;
            else
                pr.scratch = align_up(pr.ws_bytes, 256) + align_up(pr.plan_bytes, 256);
        }
        if (!ensure_pinned(&sl.h_kth, &sl.h_kth_cap, 4 * rows, sl.retired) ||
            (!pr.solo && !ensure_pinned(&sl.h_plan, &sl.h_plan_cap, pr.plan_bytes, sl.retired)))
            return false;
        bool ok = pr.scratch != ~0ull && s->arena.alloc(pr.out_bytes, &pr.out_off);
        if (ok && !s->arena.alloc(pr.scratch, &pr.scratch_off)) {
            s->arena.release(pr.out_off, pr.out_bytes);
            ok = false;
        }
        const bool passes = !pr.solo && passed[rows - 1] > 0;
        if (ok && passes && s->arena.largest() < passed[rows - 1]) {
            s->arena.release(pr.scratch_off, pr.scratch);
            s->arena.release(pr.out_off, pr.out_bytes);
            pr.take.resize(in_order);  // the passed requests would no longer fit
            continue;
        }
        if (ok) break;
        if (rows == 1) return false;  // wait for running launches to free the arena
        pr.take.resize(passes && rows / 2 < in_order ? in_order : rows / 2);
    }
    const u64 rows = pr.take.size();
    const u64 keys_off = pr.out_off + align_up(4 * rows, 256), idx_off = keys_off + align_up(4 * pr.ksum, 256);
    const u64 blk = s->next_block++;
    s->blocks[blk] = OutBlock{pr.out_off, pr.out_bytes, (int)rows};
    u64 kacc = 0;
    for (u64 i = 0; i < rows; ++i) {
        Req& r = s->reqs[pr.take[i]];
        r.state = RUNNING;
        r.block = blk;
        r.keys_off = keys_off + 4 * kacc;
        r.idx_off = idx_off + 8 * kacc;
        r.launch_rows = (u32)rows;
        r.solo = pr.solo ? 1u : 0u;
        kacc += r.q.k;
    }
    std::deque<u64> rest;  // the taken requests are a subsequence of the queue
    size_t j = 0;
    for (u64 t : s->queue) {
        if (j < pr.take.size() && pr.take[j] == t) ++j;
        else rest.push_back(t);
    }
    s->queue.swap(rest);
    sl.busy = true;
    sl.tickets = pr.take;
    sl.scratch_off = pr.scratch_off;
    sl.scratch_bytes = pr.scratch;
    return true;
}

// Issue a prepared launch on the slot's stream.  No lock held.
drtopk_status issue(drtopk_server* s, Slot& sl, Prep& pr) {
    unsigned char* A = s->d_arena;
    const u64 rows = pr.take.size();
    const u64 kth_off = pr.out_off, keys_off = pr.out_off + align_up(4 * rows, 256),
              idx_off = keys_off + align_up(4 * pr.ksum, 256);
    drtopk_status st = DRTOPK_OK;
    if (pr.solo) {
        drtopk_params p;
        solo_params(&p);
        const drtopk_query& q = sl.solo_q;
        st = drtopk_select(DRTOPK_ALGO_RADIX, (drtopk_dtype)q.dtype, (const u32*)s->cfg.d_store + q.offset, q.n, q.k,
                           &p, A + pr.scratch_off, pr.scratch, A + kth_off, A + keys_off, (int64_t*)(A + idx_off),
                           nullptr, sl.st);
    } else {
        size_t pb = pr.plan_bytes, wb = 0;
        st = drtopk_batch_plan(pr.rows.data(), rows, s->cfg.store_len, pr.ksum, s->has_tuning ? &s->tuning : nullptr,
                               sl.h_plan, &pb, &wb);
        unsigned char* d_plan = A + pr.scratch_off + align_up(pr.ws_bytes, 256);
        if (st == DRTOPK_OK &&
            hipMemcpyAsync(d_plan, sl.h_plan, pr.plan_bytes, hipMemcpyHostToDevice, sl.st) != hipSuccess)
            st = DRTOPK_ERR_HIP;
        if (st == DRTOPK_OK)
            st = drtopk_select_batched(sl.h_plan, d_plan, s->cfg.d_store, A + pr.scratch_off, pr.ws_bytes, A + kth_off,
                                       A + keys_off, (int64_t*)(A + idx_off), sl.st);
    }
    if (st == DRTOPK_OK && hipMemcpyAsync(sl.h_kth, A + kth_off, 4 * rows, hipMemcpyDeviceToHost, sl.st) != hipSuccess)
        st = DRTOPK_ERR_HIP;
    // the launch's device-side error bits: both workspaces start with the same head
    if (st == DRTOPK_OK && hipMemcpyAsync(sl.h_err, A + pr.scratch_off + offsetof(WsHead, error), sizeof(u32),
                                          hipMemcpyDeviceToHost, sl.st) != hipSuccess)
        st = DRTOPK_ERR_HIP;
    if (st == DRTOPK_OK && hipEventRecord(sl.ev, sl.st) != hipSuccess) st = DRTOPK_ERR_HIP;
    return st;
}

// Retire the slot's launch if it finished.  Lock held.
bool retire(drtopk_server* s, Slot& sl, drtopk_status failed = DRTOPK_OK) {
    if (!sl.busy) return false;
    drtopk_status st = failed;
    if (st == DRTOPK_OK) {
        const hipError_t q = hipEventQuery(sl.ev);
        if (q == hipErrorNotReady) return false;
        if (q != hipSuccess) st = DRTOPK_ERR_HIP;
        else if (*sl.h_err) st = DRTOPK_ERR_INTERNAL;  // a device-side check failed
    }
    const u64 t = now_ns();
    for (size_t i = 0; i < sl.tickets.size(); ++i)
        finish_request(s, sl.tickets[i], st, t, st == DRTOPK_OK ? sl.h_kth[i] : 0u);
    s->arena.release(sl.scratch_off, sl.scratch_bytes);
    sl.busy = false;
    sl.tickets.clear();
    s->cv_done.notify_all();
    s->cv_work.notify_one();  // arena space is free again (the retiring worker prepares next)
    return true;
}

bool any_busy(const drtopk_server* s) {
    for (const Slot& x : s->slots)
        if (x.busy) return true;
    return false;
}

// One worker per slot: prepares under the lock, issues outside it, so the host cost of
// launching overlaps across slots.
void worker_main(drtopk_server* s, size_t si) {
    (void)hipSetDevice(s->device);
    Slot& sl = s->slots[si];
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        if (sl.busy) {
            // only this worker retires its slot: wait for the launch without the lock, so that
            // polling workers do not keep the submitting client (and each other) off it
            lk.unlock();
            while (hipEventQuery(sl.ev) == hipErrorNotReady)
                std::this_thread::yield();  // a sleep would add tens of microseconds of timer slack
            lk.lock();
            retire(s, sl);
            continue;
        }
        if (s->stop && s->queue.empty()) break;
        Prep pr;
        if (!prepare(s, sl, pr)) {
            if (s->stop && !s->queue.empty() && !any_busy(s)) {
                // shutting down, nothing running and the queue head cannot fit: only a client
                // release could free the arena, and the client is destroying the server
                const u64 t = now_ns();
                for (u64 q : s->queue) finish_request(s, q, DRTOPK_ERR_BUDGET, t, 0u);
                s->queue.clear();
                s->cv_done.notify_all();
                continue;
            }
            s->cv_work.wait_for(lk, std::chrono::milliseconds(1));
            continue;
        }
        if (pr.solo) sl.solo_q = s->reqs[pr.take[0]].q;
        // wakeups are chained, one idle worker at a time: waking every worker on each submit
        // keeps the submitting thread off the lock
        if (!s->queue.empty()) s->cv_work.notify_one();
        const u64 tl = now_ns();
        for (u64 t : pr.take) s->reqs[t].t_launch = tl;
        lk.unlock();
        const drtopk_status st = issue(s, sl, pr);
        if (st != DRTOPK_OK) (void)hipStreamSynchronize(sl.st);
        lk.lock();
        if (st != DRTOPK_OK) retire(s, sl, st);
    }
}

drtopk_status fill_result(drtopk_server* s, const Req& r, drtopk_result* res) {
    std::memset(res, 0, sizeof(*res));
    res->status = r.status;
    res->kth_bits = r.kth;
    res->d_topk_keys = s->d_arena + r.keys_off;
    res->d_topk_idx = (const int64_t*)(s->d_arena + r.idx_off);
    res->t_submit_ns = r.t_submit;
    res->t_launch_ns = r.t_launch;
    res->t_done_ns = r.t_done;
    res->launch_rows = r.launch_rows;
    res->solo = r.solo;
    return r.status;
}

drtopk_status check_query(const drtopk_server_config& c, const drtopk_query* q) {
    if (!q) return fail(DRTOPK_ERR_INVALID_ARG, "query is NULL");
    if (q->dtype != DRTOPK_U32 && q->dtype != DRTOPK_F32) return fail(DRTOPK_ERR_INVALID_ARG, "bad dtype %d", q->dtype);
    if (q->reserved) return fail(DRTOPK_ERR_INVALID_ARG, "reserved must be 0");
    if (q->n == 0 || q->n > (1ull << 31) - 1) return fail(DRTOPK_ERR_INVALID_ARG, "n = %llu not in [1, 2^31-1]", (unsigned long long)q->n);
    if (q->k == 0 || q->k > q->n) return fail(DRTOPK_ERR_INVALID_ARG, "k = %llu not in [1, n]", (unsigned long long)q->k);
    if (q->offset % 4) return fail(DRTOPK_ERR_INVALID_ARG, "offset must be a multiple of 4 keys");
    if (q->offset > c.store_len || q->n > c.store_len - q->offset)
        return fail(DRTOPK_ERR_INVALID_ARG, "query outside the key store");
    return DRTOPK_OK;
}

}  // namespace
}  // namespace drtopk

using namespace drtopk;

extern "C" {

void drtopk_server_config_default(drtopk_server_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->max_inflight = 8;
    c->max_batch_rows = 4096;
    c->solo_n = 1ull << 25;
    c->solo_k = 4096;
}

drtopk_status drtopk_server_create(const drtopk_server_config* cfg, drtopk_server** out) {
    clear_error();
    if (!cfg || !out) return fail(DRTOPK_ERR_INVALID_ARG, "cfg and out must not be NULL");
    for (int i = 0; i < 4; ++i)
        if (cfg->reserved[i]) return fail(DRTOPK_ERR_INVALID_ARG, "reserved must be zero");
    if (!cfg->d_store || (uintptr_t)cfg->d_store % 16) return fail(DRTOPK_ERR_INVALID_ARG, "d_store must be a 16-byte aligned device pointer");
    std::unique_ptr<drtopk_server> s(new drtopk_server());
    s->cfg = *cfg;
    if (!s->cfg.max_inflight) s->cfg.max_inflight = 8;
    if (s->cfg.max_inflight > 16) return fail(DRTOPK_ERR_INVALID_ARG, "max_inflight %u > 16", s->cfg.max_inflight);
    if (!s->cfg.max_batch_rows) s->cfg.max_batch_rows = 4096;
    if (!s->cfg.solo_n) s->cfg.solo_n = 1ull << 25;
    if (!s->cfg.solo_k) s->cfg.solo_k = 4096;
    if (cfg->tuning) {
        s->tuning = *cfg->tuning;
        s->has_tuning = true;
    }
    HIPCHK(hipGetDevice(&s->device));
    if (!s->cfg.budget_bytes) {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        s->cfg.budget_bytes = std::min<u64>(64ull << 30, (u64)(fr * 0.9));
    }
    s->cfg.budget_bytes = align_up(s->cfg.budget_bytes, 256);
    if (hipMalloc((void**)&s->d_arena, s->cfg.budget_bytes) != hipSuccess)
        return fail(DRTOPK_ERR_BUDGET, "cannot allocate a %llu-byte arena", (unsigned long long)s->cfg.budget_bytes);
    s->arena.init(s->cfg.budget_bytes);
    s->slots.resize(s->cfg.max_inflight);
    for (Slot& sl : s->slots) {
        HIPCHK(hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
        HIPCHK(hipHostMalloc((void**)&sl.h_err, 256, hipHostMallocDefault));
        *sl.h_err = 0;
    }
    drtopk_server* raw = s.get();
    for (size_t i = 0; i < raw->slots.size(); ++i) raw->workers.emplace_back(worker_main, raw, i);
    *out = s.release();
    return DRTOPK_OK;
}

drtopk_status drtopk_server_submit(drtopk_server* s, const drtopk_query* q, uint64_t* ticket) {
    clear_error();
    if (!s || !ticket) return fail(DRTOPK_ERR_INVALID_ARG, "server and ticket must not be NULL");
    drtopk_status st = check_query(s->cfg, q);
    if (st != DRTOPK_OK) return st;
    // a request that could never fit the arena on its own is refused
    const u64 ob = align_up(out_bytes(1, q->k), 256);
    u64 need;
    if (is_solo(s, *q)) {
        need = solo_ws(*q);
    } else {
        drtopk_row r{q->offset, q->n, q->k, 0, q->dtype, 0};
        size_t pb = 0, wb = 0;
        st = drtopk_batch_plan(&r, 1, s->cfg.store_len, q->k, s->has_tuning ? &s->tuning : nullptr, nullptr, &pb, &wb);
        if (st != DRTOPK_OK) return st;
        need = align_up(wb, 256) + align_up(pb, 256);
    }
    if (need == ~0ull || align_up(need, 256) + ob > s->cfg.budget_bytes)
        return fail(DRTOPK_ERR_BUDGET, "request needs %llu bytes, the arena holds %llu",
                    (unsigned long long)(need + ob), (unsigned long long)s->cfg.budget_bytes);
    std::lock_guard<std::mutex> lk(s->mu);
    const u64 t = s->next_ticket++;
    Req r;
    r.q = *q;
    r.need = align_up(need, 256) + ob;
    r.t_submit = now_ns();
    s->reqs.emplace(t, r);
    s->queue.push_back(t);
    *ticket = t;
    s->cv_work.notify_one();
    return DRTOPK_OK;
}

drtopk_status drtopk_server_wait(drtopk_server* s, uint64_t ticket, int64_t timeout_ns, drtopk_result* res) {
    clear_error();
    if (!s || !res) return fail(DRTOPK_ERR_INVALID_ARG, "server and res must not be NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(DRTOPK_ERR_INVALID_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    auto done = [&] { return s->reqs[ticket].state == DONE; };
    if (timeout_ns < 0) s->cv_done.wait(lk, done);
    else if (!s->cv_done.wait_for(lk, std::chrono::nanoseconds(timeout_ns), done))
        return fail(DRTOPK_ERR_BUDGET, "request %llu not done within the timeout", (unsigned long long)ticket);
    const drtopk_status st = fill_result(s, s->reqs[ticket], res);
    if (st != DRTOPK_OK) return fail(st, "request %llu failed on the device", (unsigned long long)ticket);
    return DRTOPK_OK;
}

drtopk_status drtopk_server_release(drtopk_server* s, uint64_t ticket) {
    clear_error();
    if (!s) return fail(DRTOPK_ERR_INVALID_ARG, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(DRTOPK_ERR_INVALID_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    if (it->second.state != DONE) return fail(DRTOPK_ERR_INVALID_ARG, "request %llu is not done", (unsigned long long)ticket);
    const u64 blk = it->second.block;
    s->reqs.erase(it);
    auto b = s->blocks.find(blk);
    if (b != s->blocks.end() && --b->second.refs == 0) {
        s->arena.release(b->second.off, b->second.bytes);
        s->blocks.erase(b);
        s->cv_work.notify_one();
    }
    return DRTOPK_OK;
}

drtopk_status drtopk_server_destroy(drtopk_server* s) {
    clear_error();
    if (!s) return DRTOPK_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    for (auto& w : s->workers)
        if (w.joinable()) w.join();
    for (Slot& sl : s->slots) {
        if (sl.st) (void)hipStreamSynchronize(sl.st);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.st) (void)hipStreamDestroy(sl.st);
        if (sl.h_plan) (void)hipHostFree(sl.h_plan);
        if (sl.h_kth) (void)hipHostFree(sl.h_kth);
        if (sl.h_err) (void)hipHostFree(sl.h_err);
        for (void* b : sl.retired) (void)hipHostFree(b);
    }
    if (s->d_arena) (void)hipFree(s->d_arena);
    delete s;
    return DRTOPK_OK;
}

drtopk_status drtopk_server_replay(drtopk_server* s, const drtopk_query* q, const uint64_t* arrival_ns, uint64_t count,
                                   drtopk_result* results, uint64_t* t0_ns) {
    clear_error();
    if (!s || !q || !arrival_ns || !results) return fail(DRTOPK_ERR_INVALID_ARG, "NULL argument");
    std::vector<uint64_t> tk(count);
    const u64 t0 = now_ns();
    if (t0_ns) *t0_ns = t0;
    for (u64 i = 0; i < count; ++i) {
        const u64 at = t0 + arrival_ns[i];
        while (now_ns() < at) {
        }
        const drtopk_status st = drtopk_server_submit(s, &q[i], &tk[i]);
        if (st != DRTOPK_OK) return st;
    }
    drtopk_status first = DRTOPK_OK;
    for (u64 i = 0; i < count; ++i) {
        const drtopk_status st = drtopk_server_wait(s, tk[i], -1, &results[i]);
        if (st != DRTOPK_OK && first == DRTOPK_OK) first = st;
        results[i].d_topk_keys = nullptr;
        results[i].d_topk_idx = nullptr;
        (void)drtopk_server_release(s, tk[i]);
    }
    return first;
}

drtopk_status drtopk_replay_serial(const void* d_store, uint64_t store_len, const drtopk_query* q,
                                   const uint64_t* arrival_ns, uint64_t count, drtopk_result* results,
                                   uint64_t* t0_ns) {
    clear_error();
    if (!d_store || !q || !arrival_ns || !results) return fail(DRTOPK_ERR_INVALID_ARG, "NULL argument");
    drtopk_server_config c;
    drtopk_server_config_default(&c);
    c.d_store = d_store;
    c.store_len = store_len;
    drtopk_params p;
    solo_params(&p);
    u64 ws = 256, kmax = 1;
    for (u64 i = 0; i < count; ++i) {
        drtopk_status st = check_query(c, &q[i]);
        if (st != DRTOPK_OK) return st;
        size_t b = 0;
        st = drtopk_workspace_size(DRTOPK_ALGO_RADIX, (drtopk_dtype)q[i].dtype, q[i].n, q[i].k, &p, &b);
        if (st != DRTOPK_OK) return st;
        ws = std::max<u64>(ws, b);
        kmax = std::max<u64>(kmax, q[i].k);
    }
    // everything is allocated before the clock starts
    unsigned char* d = nullptr;
    const u64 o_k = align_up(ws, 256), o_keys = o_k + 256, o_idx = o_keys + align_up(4 * kmax, 256);
    HIPCHK(hipMalloc((void**)&d, o_idx + 8 * kmax));
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        (void)hipFree(d);
        return fail(DRTOPK_ERR_HIP, "stream creation failed");
    }
    u32* h_kth = nullptr;
    if (hipHostMalloc((void**)&h_kth, 256, hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(st);
        (void)hipFree(d);
        return fail(DRTOPK_ERR_HIP, "pinned allocation failed");
    }
    drtopk_status first = DRTOPK_OK;
    const u64 t0 = now_ns();
    if (t0_ns) *t0_ns = t0;
    for (u64 i = 0; i < count; ++i) {
        const u64 at = t0 + arrival_ns[i];
        while (now_ns() < at) {
        }
        drtopk_result& r = results[i];
        std::memset(&r, 0, sizeof(r));
        r.t_submit_ns = at;  // the request's arrival: waiting behind earlier requests counts
        r.t_launch_ns = now_ns();
        drtopk_status s = drtopk_select(DRTOPK_ALGO_RADIX, (drtopk_dtype)q[i].dtype, (const u32*)d_store + q[i].offset,
                                        q[i].n, q[i].k, &p, d, ws, d + o_k, d + o_keys, (int64_t*)(d + o_idx), nullptr, st);
        if (s == DRTOPK_OK && hipMemcpyAsync(h_kth, d + o_k, 4, hipMemcpyDeviceToHost, st) != hipSuccess) s = DRTOPK_ERR_HIP;
        if (hipStreamSynchronize(st) != hipSuccess && s == DRTOPK_OK) s = DRTOPK_ERR_HIP;
        r.t_done_ns = now_ns();
        r.status = s;
        r.kth_bits = h_kth[0];
        r.launch_rows = 1;
        r.solo = 1;
        if (s != DRTOPK_OK && first == DRTOPK_OK) first = s;
    }
    (void)hipHostFree(h_kth);
    (void)hipStreamDestroy(st);
    (void)hipFree(d);
    return first;
}

}  // extern "C"
