// Batched and continuous proving on one GPU (the scheduler half of liblabinius_gpu).
//
// Worker threads each drive one proof at a time through the whole-proof path (lbn_prove) on a
// context of their own, so every proof in flight has its own HIP streams: one proof's kernels run
// while another's host phases (transcript, challenge sampling, the switch round trips) run on
// other cores, and kernels of different proofs share the device. Contexts are cached per table
// size and admitted against an HBM budget: a job takes an idle context of its size, else a new
// one if its footprint fits, else idle contexts of other sizes are evicted (least recently used
// first); a job that fits nowhere waits until a running proof releases memory. Admission is in
// arrival order with bounded overtaking: a later job may start on a free context or free memory
// while the first waits, but the first cannot be overtaken more than MAX_BYPASS times.
//
// With continuous batching (lbn_sched_set_batch) a worker that wakes takes every job it can admit
// at that moment, up to the limit, and proves them together with lbn_prove_batch (on the contexts'
// own streams): under load the queue drains in batches, whose switch, rows, challenges and folds
// run as one launch each.
//
// Each worker owns a pair of HIP streams and runs its proofs on them, whatever context they use:
// HIP maps streams onto its hardware queues in creation order, so the workers' streams (created
// together) sit on distinct queues, while the contexts' own streams (created one per context, any
// number of them) would put two proofs in flight on one queue, one after the other, by chance.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "lbn_internal.h"

namespace {

constexpr int MAX_BYPASS = 64;

double now_us() {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Slot {
    lbn_ctx *ctx = nullptr;
    int log_vars = 0;
    size_t bytes = 0;
    bool busy = false;
    bool ready = false;  // created and modulus set
    uint64_t last_used = 0;
};

struct Entry {
    lbn_job *job;
    double submitted;
    int bypassed = 0;
};

}  // namespace

struct lbn_sched {
    int device = 0;
    uint64_t q_lo = 0, q_hi = 0;
    size_t budget = 0;
    std::mutex m;
    std::condition_variable wake, done;
    std::deque<Entry> queue;
    std::vector<Slot *> slots;
    std::vector<std::thread> workers;
    std::vector<hipStream_t> streams;  // two per worker: main, copy (created together at the first job)
    int nworkers = 0;

    // Under m.
    bool make_streams() {
        if (!streams.empty()) return true;
        if (hipSetDevice(device) != hipSuccess) return false;
        for (int i = 0; i < 2 * nworkers; i++) {
            hipStream_t st = nullptr;
            if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
                for (hipStream_t x : streams) hipStreamDestroy(x);
                streams.clear();
                return false;
            }
            streams.push_back(st);
        }
        return true;
    }
    size_t used = 0;
    size_t pending = 0;
    uint64_t tick = 0;
    bool stop = false;
    int max_batch = 1, batch_threads = 8;  // continuous batching (lbn_sched_set_batch)
    lbn_sched_stats stats{};

    // Under m: pick a job and a slot for it, or return false. Reserves memory for a new slot and
    // collects evicted contexts in `evict` (destroyed by the caller outside the lock).
    bool admit(Entry &out, Slot *&slot, std::vector<lbn_ctx *> &evict) {
        for (size_t qi = 0; qi < queue.size(); qi++) {
            Entry &e = queue[qi];
            const bool front = qi == 0;
            if (!front && queue.front().bypassed >= MAX_BYPASS) break;
            const int k = e.job->log_vars;
            Slot *idle = nullptr;
            for (Slot *s : slots)
                if (!s->busy && s->ready && s->log_vars == k && (!idle || s->last_used > idle->last_used)) idle = s;
            if (!idle) {
                const size_t need = lbn_ctx_footprint(k, q_lo, q_hi);
                if (used + need > budget) {
                    if (!front) continue;
                    // the first job may evict idle contexts of other sizes, least recently used first
                    std::vector<Slot *> cand;
                    for (Slot *s : slots)
                        if (!s->busy && s->ready) cand.push_back(s);
                    std::sort(cand.begin(), cand.end(), [](Slot *a, Slot *b) { return a->last_used < b->last_used; });
                    size_t freed = 0;
                    size_t take = 0;
                    while (take < cand.size() && used - freed + need > budget) freed += cand[take++]->bytes;
                    if (used - freed + need > budget) return false;  // wait for running proofs
                    for (size_t i = 0; i < take; i++) {
                        Slot *s = cand[i];
                        evict.push_back(s->ctx);
                        used -= s->bytes;
                        slots.erase(std::find(slots.begin(), slots.end(), s));
                        delete s;
                        stats.evictions++;
                    }
                }
                idle = new Slot();
                idle->log_vars = k;
                idle->bytes = need;
                used += need;
                stats.peak_bytes = std::max<uint64_t>(stats.peak_bytes, used);
                slots.push_back(idle);
            }
            idle->busy = true;
            idle->last_used = ++tick;
            slot = idle;
            out = e;
            for (size_t i = 0; i < qi; i++) queue[i].bypassed++;
            queue.erase(queue.begin() + (long)qi);
            return true;
        }
        return false;
    }

    // Outside m: the slot's context created with the modulus if it is new; LBN_OK or the error.
    int ready_slot(Slot *slot, int log_vars) {
        if (slot->ready) return LBN_OK;
        lbn_ctx *ctx = nullptr;
        int rc = lbn_create(device, log_vars, &ctx);
        if (rc == LBN_OK) rc = lbn_set_modulus(ctx, q_lo, q_hi);
        if (rc == LBN_OK) {
            slot->ctx = ctx;
            slot->ready = true;
            std::lock_guard<std::mutex> lk(m);
            stats.contexts_created++;
        } else if (ctx) {
            lbn_destroy(ctx);
        }
        return rc;
    }

    // Under m: a finished job's slot back (or its reservation, if the context was never made).
    void release(Entry &e, Slot *slot, int rc, double start, double end) {
        lbn_job *job = e.job;
        job->queue_us = start - e.submitted;
        job->run_us = end - start;
        job->status = rc;
        if (slot->ready) {
            slot->busy = false;
        } else {
            used -= slot->bytes;
            slots.erase(std::find(slots.begin(), slots.end(), slot));
            delete slot;
        }
        stats.completed++;
        if (rc != LBN_OK) stats.failed++;
        pending--;
    }

    // Several admitted jobs as one lbn_prove_batch on their contexts (own streams).
    void run_batch(std::vector<Entry> &es, std::vector<Slot *> &ss, int threads) {
        const double start = now_us();
        const int n = (int)es.size();
        std::vector<int> rc(n, LBN_OK);
        std::vector<lbn_batch_item> items;
        std::vector<int> idx;
        for (int i = 0; i < n; i++) {
            lbn_job *job = es[i].job;
            rc[i] = ready_slot(ss[i], job->log_vars);
            if (rc[i] == LBN_OK && job->table)
                rc[i] = lbn_load_table(ss[i]->ctx, job->table, ((size_t)1 << job->log_vars) / 8);
            if (rc[i] != LBN_OK) continue;
            lbn_batch_item it{};
            it.ctx = ss[i]->ctx;
            memcpy(it.seed, job->seed, 32);
            it.gen_table = job->table ? 0 : 1;
            it.proof = job->proof;
            it.cap = job->cap;
            it.status = 1;  // a call rejected as a whole leaves it so
            items.push_back(it);
            idx.push_back(i);
        }
        if (!items.empty()) {
            const int whole = lbn_prove_batch(items.data(), (int)items.size(), threads);
            for (size_t k = 0; k < items.size(); k++) {
                lbn_job *job = es[idx[k]].job;
                rc[idx[k]] = items[k].status == 1 ? (whole != LBN_OK ? whole : LBN_E_STATE) : items[k].status;
                job->len = items[k].len;
                job->profile = items[k].profile;
            }
        }
        const double end = now_us();
        std::lock_guard<std::mutex> lk(m);
        for (int i = 0; i < n; i++) release(es[i], ss[i], rc[i], start, end);
        done.notify_all();
        wake.notify_all();
    }

    void run_job(Entry &e, Slot *slot, int w) {
        lbn_job *job = e.job;
        const double start = now_us();
        int rc = ready_slot(slot, job->log_vars);
        if (rc == LBN_OK) {
            std::lock_guard<std::mutex> lk(m);
            if (!make_streams()) rc = LBN_E_DEVICE;
        }
        if (rc == LBN_OK) {
            lbn_ctx *ctx = slot->ctx;
            lbn::dev_use_streams(ctx, streams[2 * w], streams[2 * w + 1]);
            rc = job->table ? lbn_load_table(ctx, job->table, ((size_t)1 << job->log_vars) / 8)
                            : lbn_generate_table(ctx, job->seed);
            if (rc == LBN_OK) rc = lbn_prove(ctx, job->seed, job->proof, job->cap, &job->len, &job->profile);
            lbn::dev_use_streams(ctx, nullptr, nullptr);
        }
        const double end = now_us();
        std::lock_guard<std::mutex> lk(m);
        release(e, slot, rc, start, end);
        done.notify_all();
        wake.notify_all();
    }

    void loop(int w) {
        for (;;) {
            Entry e{};
            Slot *slot = nullptr;
            std::vector<lbn_ctx *> evict;
            std::vector<Entry> es;
            std::vector<Slot *> ss;
            int threads = 1;
            {
                std::unique_lock<std::mutex> lk(m);
                wake.wait(lk, [&] { return stop || (!queue.empty() && admit(e, slot, evict)); });
                if (!slot) return;  // stop with nothing admitted
                // continuous batching: every other job admissible now joins this one
                if (max_batch > 1) {
                    es.push_back(e), ss.push_back(slot);
                    Entry e2{};
                    Slot *s2 = nullptr;
                    while ((int)es.size() < max_batch && !queue.empty() && admit(e2, s2, evict))
                        es.push_back(e2), ss.push_back(s2);
                    threads = batch_threads;
                }
            }
            for (lbn_ctx *c : evict) lbn_destroy(c);
            if (es.size() > 1)
                run_batch(es, ss, threads);
            else
                run_job(e, slot, w);
        }
    }
};

extern "C" int lbn_sched_create(int device, uint64_t q_lo, uint64_t q_hi, size_t hbm_budget, int workers,
                                lbn_sched **out) {
    if (!out || workers < 1 || workers > 64 || hbm_budget == 0) return LBN_E_ARG;
    *out = nullptr;
    if (lbn_ctx_footprint(31, q_lo, q_hi) == 0 && lbn_ctx_footprint(20, q_lo, q_hi) == 0) return LBN_E_MODULUS;
    auto *s = new lbn_sched();
    s->device = device;
    s->q_lo = q_lo;
    s->q_hi = q_hi;
    s->budget = hbm_budget;
    s->nworkers = workers;
    for (int i = 0; i < workers; i++) s->workers.emplace_back([s, i] { s->loop(i); });
    *out = s;
    return LBN_OK;
}

extern "C" int lbn_sched_submit(lbn_sched *s, lbn_job *job) {
    if (!s || !job || !job->proof) return LBN_E_ARG;
    const size_t need = lbn_ctx_footprint(job->log_vars, s->q_lo, s->q_hi);
    if (need == 0) return LBN_E_ARG;
    std::lock_guard<std::mutex> lk(s->m);
    if (s->stop) return LBN_E_STATE;
    if (need > s->budget) return LBN_E_RANGE;  // could never be admitted
    job->status = 1;
    job->len = 0;
    job->queue_us = job->run_us = 0;
    s->queue.push_back(Entry{job, now_us(), 0});
    s->pending++;
    s->stats.submitted++;
    s->wake.notify_one();
    return LBN_OK;
}

extern "C" int lbn_sched_wait(lbn_sched *s, const lbn_job *job) {
    if (!s || !job) return LBN_E_ARG;
    std::unique_lock<std::mutex> lk(s->m);
    s->done.wait(lk, [&] { return job->status != 1; });
    return job->status;
}

extern "C" int lbn_sched_drain(lbn_sched *s) {
    if (!s) return LBN_E_ARG;
    std::unique_lock<std::mutex> lk(s->m);
    s->done.wait(lk, [&] { return s->pending == 0; });
    return LBN_OK;
}

extern "C" int lbn_sched_get_stats(lbn_sched *s, lbn_sched_stats *out) {
    if (!s || !out) return LBN_E_ARG;
    std::lock_guard<std::mutex> lk(s->m);
    *out = s->stats;
    out->used_bytes = s->used;
    return LBN_OK;
}

extern "C" void lbn_sched_destroy(lbn_sched *s) {
    if (!s) return;
    lbn_sched_drain(s);
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
    }
    s->wake.notify_all();
    for (auto &t : s->workers) t.join();
    for (Slot *slot : s->slots) {
        if (slot->ctx) lbn_destroy(slot->ctx);
        delete slot;
    }
    for (hipStream_t x : s->streams) hipStreamDestroy(x);
    delete s;
}

extern "C" int lbn_sched_set_batch(lbn_sched *s, int max_batch, int threads) {
    if (!s || max_batch < 1 || max_batch > lbn::SWB_MAX || threads < 1 || threads > 64) return LBN_E_ARG;
    std::lock_guard<std::mutex> lk(s->m);
    s->max_batch = max_batch;
    s->batch_threads = threads;
    return LBN_OK;
}

extern "C" int lbn_sched_prewarm(lbn_sched *s, int log_vars, int count) {
    if (!s || count < 0 || count > 64) return LBN_E_ARG;
    const size_t need = lbn_ctx_footprint(log_vars, s->q_lo, s->q_hi);
    if (need == 0) return LBN_E_ARG;
    std::vector<Slot *> mine;
    {
        std::lock_guard<std::mutex> lk(s->m);
        if (s->stop) return LBN_E_STATE;
        int have = 0;
        for (Slot *x : s->slots)
            if (x->log_vars == log_vars) have++;
        for (; have < count && s->used + need <= s->budget; have++) {
            Slot *x = new Slot();
            x->log_vars = log_vars;
            x->bytes = need;
            x->busy = true;  // reserved: no job takes it before it is ready
            s->used += need;
            s->stats.peak_bytes = std::max<uint64_t>(s->stats.peak_bytes, s->used);
            s->slots.push_back(x);
            mine.push_back(x);
        }
        if (have < count) {
            for (Slot *x : mine) {
                s->used -= x->bytes;
                s->slots.erase(std::find(s->slots.begin(), s->slots.end(), x));
                delete x;
            }
            return LBN_E_RANGE;  // the budget holds fewer
        }
    }
    int rc = LBN_OK;
    for (Slot *x : mine) {
        lbn_ctx *ctx = nullptr;
        int r = lbn_create(s->device, log_vars, &ctx);
        if (r == LBN_OK) r = lbn_set_modulus(ctx, s->q_lo, s->q_hi);
        std::lock_guard<std::mutex> lk(s->m);
        if (r == LBN_OK) {
            x->ctx = ctx;
            x->ready = true;
            x->busy = false;
            x->last_used = ++s->tick;
            s->stats.contexts_created++;
        } else {
            if (ctx) lbn_destroy(ctx);
            s->used -= x->bytes;
            s->slots.erase(std::find(s->slots.begin(), s->slots.end(), x));
            delete x;
            rc = r;
        }
    }
    s->wake.notify_all();
    return rc;
}
