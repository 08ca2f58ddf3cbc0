// Serving loop: continuous batching of SCF requests on one GPU (declared in include/hfgpu.h).
//
// Lockstep mode (the default, lockstep = k >= 1): `slots` contexts hold the admitted requests; `prep_threads`
// threads prepare each admitted request on its slot's own stream (molecule set-up, one-electron integrals,
// S^-1/2, the in-core store), and k batch threads, each with its own stream, step their prepared requests
// together: each phase of an SCF iteration is one launch over all the requests of a batch (scf_step_phases),
// so a batch of B requests costs the launches and host round trips of one, and the batches fill each other's
// round trips on the device.  A request joins the smallest batch at its next iteration boundary and leaves it
// when its SCF ends.
//
// Slot mode (lockstep = 0): `slots` worker threads each own a context (stream, side streams, device memory)
// and run one request's SCF alone at a time, the slots' kernels overlapping on the device.
//
// Both modes: requests wait in a FIFO queue; the head request is admitted to a free slot once its estimated HBM
// footprint fits in the budget together with what the other slots hold (no overtaking, so a large request is
// not starved); a request larger than the whole budget is rejected at submission.  Admission happens whenever a
// slot or memory frees up, so requests that arrive while others run join them.  Results go to a completion queue
// read by hf_serve_poll.  All times are steady-clock milliseconds since the server was created.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <memory>

#include "scf_run.h"

namespace {

struct Request {
    uint64_t id = 0, tag = 0;
    std::string basis;
    std::vector<int> Z;
    std::vector<double> xyz;
    int unrestricted = 0;
    int nbf = 0;
    double footprint = 0;
    double store_bytes = 0;   // the in-core store's share of the footprint (0: none)
    double arrival_ms = 0;
};

}  // namespace

struct hf_server {
    hf_serve_options opt;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request> queue;
    std::deque<hf_serve_result> done;
    std::vector<std::thread> workers;
    std::vector<hf_ctx*> ctxs;
    // Device bytes charged to each slot against the budget: while it runs a request, the larger of that
    // request's footprint and what the slot already held; while idle, what its context keeps (SCF matrices,
    // J/K scratch, an in-core store) for the next request.
    std::vector<double> charged;
    std::vector<char> busy;
    // kernel statistics of every finished request, summed over the slots (hf_serve_kernel_stats)
    std::vector<hf_kernel_stat> stats;
    void add_stats(hf_ctx* ctx) {   // caller holds mu; ctx belongs to the calling worker
        hf_kernel_stat buf[128];
        const int k = hf_kernel_stats(ctx, buf, 128);
        for (int i = 0; i < k; ++i) {
            hf_kernel_stat* t = nullptr;
            for (hf_kernel_stat& x : stats)
                if (x.la == buf[i].la && x.lb == buf[i].lb && x.lc == buf[i].lc && x.ld == buf[i].ld &&
                    x.fp32 == buf[i].fp32)
                    t = &x;
            if (!t) {
                stats.push_back(buf[i]);
                continue;
            }
            t->launches += buf[i].launches;
            t->quartets += buf[i].quartets;
            t->prim_quartets += buf[i].prim_quartets;
            t->eris += buf[i].eris;
            t->flops += buf[i].flops;
            t->ms += buf[i].ms;
        }
        (void)hf_kernel_stats_reset(ctx);
    }
    double budget = 0, keep_store = 0;
    uint64_t next_id = 1;
    int pending = 0;
    bool stop = false;

    // ---- lockstep mode ----
    std::vector<std::unique_ptr<ScfRun>> runs;   // per slot: the request's SCF session
    std::vector<hf_serve_result> slot_res;       // per slot: the result being filled
    std::vector<hipEvent_t> slot_ev;             // per slot: end of its last step on the batch stream
    std::vector<char> slot_ev_pending;
    std::deque<int> ready;                       // prepared slots waiting to join the batch
    int preparing = 0;                           // slots being prepared
    std::condition_variable cv_batch;
    // one stream, step scratch and active count per batch thread (lockstep = the number of batch threads); a
    // prepared request joins the batch with the fewest requests
    std::vector<hipStream_t> bstreams;
    std::vector<std::unique_ptr<BatchScratch>> bss;
    std::vector<int> nactive;

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    double charged_total() const {
        double t = 0;
        for (double c : charged) t += c;
        return t;
    }
    // The head request fits on slot w: the other slots' charges plus this slot's during the request.
    bool fits(int w, double fp) const { return charged_total() - charged[w] + std::max(charged[w], fp) <= budget; }
    // No idle slot can take the head request because idle slots hold memory: those slots release their stores.
    bool head_blocked() const {
        if (queue.empty()) return false;
        for (size_t v = 0; v < charged.size(); ++v)
            if (!busy[v] && fits((int)v, queue.front().footprint)) return false;
        return true;
    }

    void run(int w) {
        hf_ctx* ctx = ctxs[w];
        for (;;) {
            Request rq;
            {
                std::unique_lock<std::mutex> lock(mu);
                bool release = false;
                cv_work.wait(lock, [&] {
                    if (stop) return true;
                    if (!queue.empty() && fits(w, queue.front().footprint)) return true;
                    release = head_blocked() && (ctx->d_G.n + ctx->d_G32.n > 0 || ctx->scf_ws);
                    return release;
                });
                if (stop) return;
                if (release) {
                    // free this idle slot's in-core store (then its SCF matrices) so that the queue head can
                    // be admitted somewhere
                    lock.unlock();
                    if (ctx->d_G.n + ctx->d_G32.n > 0)
                        hf_release_incore(ctx, 0);
                    else
                        hf_scf_workspace_free(ctx);
                    const double held = hf_ctx_held_bytes(ctx);
                    lock.lock();
                    charged[w] = held;
                    cv_work.notify_all();
                    continue;
                }
                rq = std::move(queue.front());
                queue.pop_front();
                charged[w] = std::max(charged[w], rq.footprint);
                busy[w] = 1;
            }
            hf_serve_result r{};
            r.id = rq.id;
            r.tag = rq.tag;
            r.nbf = rq.nbf;
            r.footprint = rq.footprint;
            r.arrival_ms = rq.arrival_ms;
            r.admit_ms = now_ms();
            int rc = hf_set_molecule(ctx, rq.basis.c_str(), (int)rq.Z.size(), rq.Z.data(), rq.xyz.data());
            if (rc == HF_OK) {
                hf_scf_params p = opt.scf;
                p.unrestricted = rq.unrestricted;
                p.terminate = nullptr;
                // the in-core store only within the allowance the admission counted (never from the
                // device's free memory, which the other slots are using at the same time)
                if (p.incore) p.incore_max_bytes = rq.store_bytes > 0 ? rq.store_bytes : -1.0;
                hf_scf_result sr{};
                rc = hf_scf(ctx, &p, &sr, nullptr);
                // the slot keeps its store for its next request up to keep_store bytes (freeing synchronises
                // the whole device, which stalls every other slot); what it keeps stays charged
                hf_release_incore(ctx, (size_t)keep_store);
                r.energy = sr.energy;
                r.iterations = sr.iterations;
                r.converged = sr.converged;
            }
            r.status = rc;
            const double held = hf_ctx_held_bytes(ctx);
            r.done_ms = now_ms();
            {
                std::lock_guard<std::mutex> lock(mu);
                charged[w] = held;
                busy[w] = 0;
                add_stats(ctx);
                done.push_back(r);
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }

    // A request leaves its slot: result to the completion queue, the slot's memory charged as what it keeps.
    void complete(int w, hf_serve_result& r) {
        r.done_ms = now_ms();
        const double held = hf_ctx_held_bytes(ctxs[w]);
        {
            std::lock_guard<std::mutex> lock(mu);
            charged[w] = held;
            busy[w] = 0;
            add_stats(ctxs[w]);
            done.push_back(r);
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }

    // Preparation thread: admits the queue head to a free slot whose memory fits it (or frees an idle slot's
    // memory when only that blocks the head), installs the molecule and runs Init on the slot's stream, then
    // hands the slot to the batch thread.
    void prep() {
        const char* tenv = std::getenv("HFGPU_SERVE_TRACE");
        const bool prep_trace = tenv && tenv[0] == '2';
        for (;;) {
            int w = -1;
            bool release = false;
            Request rq;
            {
                std::unique_lock<std::mutex> lock(mu);
                cv_work.wait(lock, [&] {
                    if (stop) return true;
                    if (queue.empty()) return false;
                    for (size_t v = 0; v < charged.size(); ++v)
                        if (!busy[v] && fits((int)v, queue.front().footprint)) {
                            w = (int)v;
                            return true;
                        }
                    if (head_blocked())
                        for (size_t v = 0; v < charged.size(); ++v)
                            if (!busy[v] && (ctxs[v]->d_G.n + ctxs[v]->d_G32.n > 0 || ctxs[v]->scf_ws)) {
                                w = (int)v;
                                release = true;
                                return true;
                            }
                    return false;
                });
                if (stop) return;
                busy[w] = 1;
                if (!release) {
                    rq = std::move(queue.front());
                    queue.pop_front();
                    charged[w] = std::max(charged[w], rq.footprint);
                    ++preparing;
                }
            }
            hf_ctx* ctx = ctxs[w];
            if (slot_ev_pending[w]) {   // the slot's previous request may still have work queued on the batch stream
                (void)hipEventSynchronize(slot_ev[w]);
                slot_ev_pending[w] = 0;
            }
            if (release) {
                runs[w].reset();
                if (ctx->d_G.n + ctx->d_G32.n > 0)
                    hf_release_incore(ctx, 0);
                else
                    hf_scf_workspace_free(ctx);
                const double held = hf_ctx_held_bytes(ctx);
                {
                    std::lock_guard<std::mutex> lock(mu);
                    charged[w] = held;
                    busy[w] = 0;
                }
                cv_work.notify_all();
                continue;
            }
            // a store above what a slot keeps, from the slot's previous request
            hf_release_incore(ctx, (size_t)keep_store);
            hf_serve_result& r = slot_res[w];
            r = hf_serve_result{};
            r.id = rq.id;
            r.tag = rq.tag;
            r.nbf = rq.nbf;
            r.footprint = rq.footprint;
            r.arrival_ms = rq.arrival_ms;
            r.admit_ms = now_ms();
            runs[w].reset();
            int rc = hf_set_molecule(ctx, rq.basis.c_str(), (int)rq.Z.size(), rq.Z.data(), rq.xyz.data());
            if (rc == HF_OK) {
                hf_scf_params p = opt.scf;
                p.unrestricted = rq.unrestricted;
                p.terminate = nullptr;
                // the in-core store only within the allowance the admission counted
                if (p.incore) p.incore_max_bytes = rq.store_bytes > 0 ? rq.store_bytes : -1.0;
                runs[w].reset(new (std::nothrow) ScfRun(ctx, p));
                if (!runs[w]) rc = HF_E_OOM;
                else {
                    runs[w]->orbitals = false;   // a served result is the energy and the convergence verdict
                    rc = runs[w]->init();
                }
                if (prep_trace)
                    std::fprintf(stderr, "[serve] prep slot %d n=%d: %.1f ms (admitted %.0f ms after arrival)\n", w,
                                 rq.nbf, now_ms() - r.admit_ms, r.admit_ms - r.arrival_ms);
            }
            if (rc != HF_OK) {
                r.status = rc;
                runs[w].reset();
                {
                    std::lock_guard<std::mutex> lock(mu);
                    --preparing;
                }
                cv_batch.notify_one();
                complete(w, r);
                continue;
            }
            {
                std::lock_guard<std::mutex> lock(mu);
                --preparing;
                ready.push_back(w);
            }
            cv_batch.notify_one();
        }
    }

    // Batch thread: steps every prepared request together until its SCF ends.  HFGPU_SERVE_TRACE=1 prints, about
    // once a second, the batch steps taken, the mean batch size and the time spent in steps and in waiting.
    void batch_loop(int b) {
        hipStream_t bstream = bstreams[b];
        BatchScratch& bs = *bss[b];
        std::vector<int> active;
        std::vector<ScfRun*> ms;
        const char* tenv = std::getenv("HFGPU_SERVE_TRACE");
        const bool trace = tenv && (tenv[0] == '1' || tenv[0] == '2');
        double t_last = now_ms(), t_step = 0, t_wait = 0;
        long steps = 0, members = 0;
        for (;;) {
            if (trace && now_ms() - t_last > 1000.0) {
                int q = 0, prep_n = 0, rd = 0;
                {
                    std::lock_guard<std::mutex> lock(mu);
                    q = (int)queue.size();
                    prep_n = preparing;
                    rd = (int)ready.size();
                }
                std::fprintf(stderr, "[serve] t=%.0f ms: %ld steps, mean batch %.1f, step %.0f ms, idle %.0f ms; "
                             "queue %d, preparing %d, ready %d\n", now_ms(), steps, steps ? (double)members / steps : 0.0,
                             t_step, t_wait, q, prep_n, rd);
                t_last = now_ms();
                steps = members = 0;
                t_step = t_wait = 0;
            }
            const double tw0 = now_ms();
            std::vector<int> joining;
            {
                std::unique_lock<std::mutex> lock(mu);
                if (active.empty())
                    cv_batch.wait_for(lock, std::chrono::milliseconds(trace ? 200 : 100000),
                                      [&] { return !ready.empty() || (stop && preparing == 0); });
                if (active.empty() && ready.empty() && stop && preparing == 0) return;
                // take ready requests while this batch is not larger than the smallest other one
                while (!ready.empty()) {
                    int others = 1 << 30;
                    for (size_t c = 0; c < nactive.size(); ++c)
                        if ((int)c != b) others = std::min(others, nactive[c]);
                    if (nactive.size() > 1 && nactive[b] > others && !active.empty()) break;
                    joining.push_back(ready.front());
                    ready.pop_front();
                    ++nactive[b];
                }
                if (!ready.empty()) cv_batch.notify_one();
            }
            for (int w : joining) {
                ScfRun* m = runs[w].get();
                (void)hipStreamWaitEvent(bstream, ctxs[w]->scf_ws->ready_ev, 0);
                m->begin();
                if (m->done()) {
                    finish_slot(w, HF_OK, b);
                    continue;
                }
                active.push_back(w);
            }
            t_wait += now_ms() - tw0;
            if (active.empty()) continue;
            const double ts0 = now_ms();
            ms.clear();
            for (int w : active) {
                ScfRun* m = runs[w].get();
                m->step_iter = m->next_iter();
                ms.push_back(m);
            }
            const int rc = scf_step_phases(ms, bstream, bs);
            ++steps;
            members += (long)ms.size();
            t_step += now_ms() - ts0;
            std::vector<int> keep;
            for (int w : active) {
                ScfRun* m = runs[w].get();
                int st = rc != HF_OK ? rc : m->status;
                if (st == HF_OK) st = m->after_step();
                if (st != HF_OK || m->done()) {
                    finish_slot(w, st, b);
                    continue;
                }
                keep.push_back(w);
            }
            active.swap(keep);
        }
    }

    void finish_slot(int w, int st, int b) {
        ScfRun* m = runs[w].get();
        hf_serve_result& r = slot_res[w];
        r.energy = m->cur;
        r.iterations = m->iter + (m->converged ? 1 : 0);
        r.converged = m->converged ? 1 : 0;
        r.status = st != HF_OK ? st : m->cancelled ? HF_E_CANCELLED : m->converged ? HF_OK : HF_E_NOT_CONVERGED;
        if (hipEventRecord(slot_ev[w], bstreams[b]) == hipSuccess) slot_ev_pending[w] = 1;
        {
            std::lock_guard<std::mutex> lock(mu);
            --nactive[b];
        }
        complete(w, r);
    }
};

extern "C" {

void hf_serve_default_options(hf_serve_options* o) {
    o->slots = 32;
    o->hbm_budget = 0;
    o->lockstep = 2;
    o->prep_threads = 8;
    hf_default_options(&o->ctx);
    hf_scf_defaults(&o->scf);
}

double hf_serve_footprint(int nbf, int unrestricted) {
    // SCF state on the device (S, h, S^-1/2, Fock, densities, DIIS ring of 6 per spin, SP2 and J/K
    // work matrices: about 56 n x n FP64 matrices for RHF, 64 for UHF) plus pair and task data.
    const double n2 = (double)nbf * nbf;
    return 8.0 * n2 * (unrestricted ? 64 : 56) + 32.0 * (1 << 20);
}

int hf_serve_create(const hf_serve_options* o, hf_server** out) {
    if (!o || !out || o->slots < 1) return HF_E_ARG;
    *out = nullptr;
    hf_server* s = new hf_server();
    s->opt = *o;
    s->t0 = std::chrono::steady_clock::now();
    s->budget = o->hbm_budget;
    if (s->budget <= 0) {
        size_t free_b = 0, total_b = 0;
        if (hipSetDevice(o->ctx.device) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
            delete s;
            return HF_E_HIP;
        }
        s->budget = 0.8 * (double)free_b;
    }
    for (int w = 0; w < o->slots; ++w) {
        hf_ctx* c = nullptr;
        const int rc = hf_create(&o->ctx, &c);
        if (rc != HF_OK) {
            for (hf_ctx* x : s->ctxs) hf_destroy(x);
            delete s;
            return rc;
        }
        s->ctxs.push_back(c);
    }
    s->charged.assign(o->slots, 0.0);
    s->busy.assign(o->slots, 0);
    s->keep_store = std::min(16.0 * (1 << 30), s->budget / o->slots);
    if (o->lockstep) {
        s->runs.resize(o->slots);
        s->slot_res.resize(o->slots);
        s->slot_ev.assign(o->slots, nullptr);
        s->slot_ev_pending.assign(o->slots, 0);
        const int nb = std::max(1, std::min(o->lockstep, o->slots));
        bool ok = true;
        for (int b = 0; b < nb && ok; ++b) {
            hipStream_t st = nullptr;
            ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
            if (!ok) break;
            s->bstreams.push_back(st);
            s->bss.emplace_back(new BatchScratch());
            ok = s->bss.back()->reserve(s->ctxs[0], 2 * o->slots) == HF_OK;
        }
        s->nactive.assign(nb, 0);
        for (int w = 0; w < o->slots && ok; ++w)
            ok = hipEventCreateWithFlags(&s->slot_ev[w], hipEventDisableTiming) == hipSuccess;
        // Each slot's in-core store is allocated now at the size a slot keeps between requests (budget / slots, at
        // most 16 GiB; FP64 store and, in the mixed tier, its FP32 copy): growing stores while requests run means
        // hipFree and hipMalloc of gigabytes, which synchronise the device and stall every batch.  The bytes are
        // charged to the slots, as retained stores are; a request whose store is larger than this still gets one
        // (the slot's store is reallocated), and idle slots release theirs when the queue head needs the memory.
        if (ok && o->scf.incore) {
            const double per = o->ctx.precision == HF_PREC_MIXED ? 12.0 : 8.0;
            const size_t len = (size_t)(s->keep_store / per);
            for (int w = 0; w < o->slots && ok && len > 0; ++w) {
                hf_ctx* c = s->ctxs[w];
                ok = c->d_G.alloc_exact(len) == hipSuccess &&
                     (o->ctx.precision != HF_PREC_MIXED || c->d_G32.alloc_exact(len) == hipSuccess);
                s->charged[w] = hf_ctx_held_bytes(c);
            }
        }
        if (!ok) {
            for (hipEvent_t e : s->slot_ev)
                if (e) (void)hipEventDestroy(e);
            s->bss.clear();
            for (hipStream_t st : s->bstreams) (void)hipStreamDestroy(st);
            for (hf_ctx* x : s->ctxs) hf_destroy(x);
            delete s;
            return HF_E_HIP;
        }
        const int np = std::max(1, o->prep_threads);
        for (int t = 0; t < np; ++t) s->workers.emplace_back([s] { s->prep(); });
        for (int b = 0; b < nb; ++b) s->workers.emplace_back([s, b] { s->batch_loop(b); });
    } else {
        for (int w = 0; w < o->slots; ++w) s->workers.emplace_back([s, w] { s->run(w); });
    }
    *out = s;
    return HF_OK;
}

int hf_serve_submit(hf_server* s, const char* basis_file, int natoms, const int* Z, const double* xyz,
                    int unrestricted, uint64_t tag, uint64_t* id) {
    if (!s || !basis_file || natoms <= 0 || !Z || !xyz) return HF_E_ARG;
    Request rq;
    rq.basis = basis_file;
    rq.Z.assign(Z, Z + natoms);
    rq.xyz.assign(xyz, xyz + 3 * (size_t)natoms);
    rq.unrestricted = unrestricted ? 1 : 0;
    rq.tag = tag;
    hfhost::Molecule m;
    std::string err;
    if (!hfhost::load_molecule(basis_file, natoms, Z, xyz, m, err)) return HF_E_IO;
    rq.nbf = m.nbf;
    rq.footprint = hf_serve_footprint(m.nbf, rq.unrestricted);
    if (s->opt.scf.incore && m.nbf <= 512) {   // the in-core ERI store (and its FP32 copy in the mixed tier)
        rq.store_bytes = (double)hf_pair_matrix_len(m.nbf) * (s->opt.ctx.precision == HF_PREC_MIXED ? 12.0 : 8.0);
        if (s->opt.scf.incore_max_bytes > 0 && rq.store_bytes > s->opt.scf.incore_max_bytes) rq.store_bytes = 0;
        rq.footprint += rq.store_bytes;
    }
    if (rq.footprint > s->budget) return HF_E_OOM;   // could never be admitted: rejected, never queued
    {
        std::lock_guard<std::mutex> lock(s->mu);
        rq.id = s->next_id++;
        rq.arrival_ms = s->now_ms();
        if (id) *id = rq.id;
        s->queue.push_back(std::move(rq));
        ++s->pending;
    }
    s->cv_work.notify_one();
    return HF_OK;
}

int hf_serve_poll(hf_server* s, hf_serve_result* out, int max, int wait_ms) {
    if (!s || !out || max <= 0) return 0;
    std::unique_lock<std::mutex> lock(s->mu);
    if (s->done.empty() && wait_ms > 0)
        s->cv_done.wait_for(lock, std::chrono::milliseconds(wait_ms), [&] { return !s->done.empty(); });
    int k = 0;
    while (k < max && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
        --s->pending;
    }
    return k;
}

int hf_serve_kernel_stats(hf_server* s, hf_kernel_stat* out, int max, int reset) {
    if (!s || !out || max <= 0) return 0;
    std::lock_guard<std::mutex> lock(s->mu);
    int k = 0;
    for (const hf_kernel_stat& x : s->stats)
        if (k < max) out[k++] = x;
    if (reset) s->stats.clear();
    return k;
}

int hf_serve_pending(hf_server* s) {
    if (!s) return 0;
    std::lock_guard<std::mutex> lock(s->mu);
    return s->pending;
}

int hf_serve_destroy(hf_server* s) {
    if (!s) return HF_E_ARG;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
        s->queue.clear();
    }
    s->cv_work.notify_all();
    s->cv_batch.notify_all();
    for (std::thread& t : s->workers) t.join();
    s->runs.clear();
    for (hipEvent_t e : s->slot_ev)
        if (e) (void)hipEventDestroy(e);
    s->bss.clear();
    for (hipStream_t st : s->bstreams) (void)hipStreamDestroy(st);
    for (hf_ctx* c : s->ctxs) hf_destroy(c);
    delete s;
    return HF_OK;
}

}  // extern "C"
