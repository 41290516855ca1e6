// serve.cpp -- the solve server of nqes.h: continuous batching of whole solves.
//
// One solve is the reference's run() loop (rbm.cc:1173-1197) over nqes_sr_step.
// The server keeps a FIFO of requests and max_slots slots.  Each slot is a worker
// thread with its own context (own HIP stream, own device buffers), created once and
// kept: a solve re-initialises the problem on it, and its buffers, which only grow
// (to the largest solve the slot has run), are reused by the next solve, so a slot in
// steady state makes no device allocation.  A slot is accounted at the bytes its
// context holds: the largest need (nqes_required_bytes plus a fixed per-solve
// allowance) of the solves it has run.
// Admission happens under the server's lock whenever a request arrives or a solve
// ends.  A request goes to an idle slot whose buffers already hold it (the smallest
// such slot), else to the idle slot that needs the least growth, if that growth fits
// the budget minus what all slots hold; if it does not, idle slots whose buffers are
// larger than the request are trimmed (their contexts re-created empty by their own
// threads) and the request waits.  A request that does not fit blocks the ones behind
// it unless they fit as well; each pass of that kind is counted, and after
// 2*max_slots passes the blocked request is waited for.  Concurrent solves share the
// device only through the hardware queues, and a solve's result depends only on its
// request, not on the slot or on what the slot ran before.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <complex>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nqes.h"
#include "srloop.hpp"

namespace {

using clk = std::chrono::steady_clock;

// device bytes a context needs beyond nqes_required_bytes: stream, events, J,
// tables, moments and the solver's vectors (generous; checked against the
// measured free memory in the tests)
constexpr int64_t kPerSolveBytes = 64ll << 20;

struct Job {
    uint64_t id = 0;
    nqes_request req{};
    std::vector<double> J, W;
    int64_t bytes = 0;
    int passed = 0;  // times a later request was admitted ahead of this one
    bool done = false;
    bool claimed = false;  // a waiter has taken it (one waiter per id)
    std::atomic<int> progress{0};  // SR steps completed so far (read by nqes_server_progress)
    std::atomic<bool> cancel{false};  // nqes_server_cancel: stop before the next SR step
    nqes_result res{};
    std::string err;  // message of a failed solve
};

struct Slot {
    nqes_ctx *ctx = nullptr;
    int64_t reserved = 0;  // bytes its context holds (largest need it has served)
    Job *job = nullptr;    // assigned, not finished
    bool trim = false;     // re-create the context empty (idle slot, memory short)
};

double ms_since(clk::time_point t0) { return std::chrono::duration<double, std::milli>(clk::now() - t0).count(); }

} // namespace

struct nqes_server {
    int device = 0;
    int max_slots = 1;
    int64_t budget = 0;
    clk::time_point t0;
    std::string err;

    std::mutex mu;
    std::condition_variable cv_slot;   // workers: a job or a trim for their slot, or stopping
    std::condition_variable cv_done;   // waiters: some solve finished
    std::deque<Job *> queue;           // submitted, not admitted (arrival order)
    std::map<uint64_t, std::unique_ptr<Job>> jobs;
    std::vector<Slot> slots;
    uint64_t next_id = 1;
    bool stopping = false;
    std::vector<std::thread> workers;

    int64_t reserved_total() const
    {
        int64_t t = 0;
        for (const Slot &s : slots) t += s.reserved;
        return t;
    }

    // under mu: the idle slot for a request of `need` bytes (-1: none fits now)
    int pick_slot(int64_t need)
    {
        int best_fit = -1, least_growth = -1;
        for (int i = 0; i < (int)slots.size(); i++) {
            const Slot &s = slots[i];
            if (s.job || s.trim) continue;
            if (s.reserved >= need) {
                if (best_fit < 0 || s.reserved < slots[best_fit].reserved) best_fit = i;
            } else if (least_growth < 0 || s.reserved > slots[least_growth].reserved) {
                least_growth = i;
            }
        }
        if (best_fit >= 0) return best_fit;
        if (least_growth >= 0 && need - slots[least_growth].reserved <= budget - reserved_total()) return least_growth;
        return -1;
    }

    // under mu
    void admit()
    {
        Job *blocked = nullptr;
        for (auto it = queue.begin(); it != queue.end();) {
            Job *j = *it;
            const int si = (blocked && blocked->passed >= 2 * max_slots) ? -1 : pick_slot(j->bytes);
            if (si >= 0) {
                if (blocked) blocked->passed++;
                Slot &s = slots[si];
                if (s.reserved < j->bytes) s.reserved = j->bytes;
                s.job = j;
                j->res.t_admit_ms = ms_since(t0);
                it = queue.erase(it);
                cv_slot.notify_all();
            } else {
                if (!blocked) blocked = j;
                ++it;
            }
        }
        // the first waiting request cannot grow any idle slot into the budget: trim idle
        // slots that hold more than it needs (their threads free the buffers)
        if (blocked) {
            bool any_idle = false;
            for (Slot &s : slots)
                if (!s.job && !s.trim) any_idle = true;
            if (any_idle && pick_slot(blocked->bytes) < 0) {
                for (Slot &s : slots)
                    if (!s.job && !s.trim && s.reserved > 0) s.trim = true;
                cv_slot.notify_all();
            }
        }
    }

    void run(Slot &slot, Job *j)
    {
        const nqes_request &q = j->req;
        nqes_result &r = j->res;
        r.K = q.K;
        if (!slot.ctx) slot.ctx = nqes_create(device);
        nqes_ctx *ctx = slot.ctx;
        if (!ctx) {
            r.status = NQES_EDEVICE;
            j->err = "cannot create a context on the device";
            return;
        }
        int rc = nqes_problem_init(ctx, q.N, q.M, q.K, q.h, j->J.data(), q.model, q.precision);
        srloop::Schedule sched{q.lrate_init,  q.lrate_decay,  q.lrate_cutoff,
                               q.reglam_init, q.reglam_decay, q.reglam_cutoff};
        srloop::Convergence conv(q.K);
        std::vector<std::complex<double>> emat(q.K * q.K);
        for (int it = 0; rc == NQES_OK && it < q.niters; it++) {
            if (j->cancel.load(std::memory_order_relaxed)) {
                rc = NQES_ECANCELED;
                break;
            }
            nqes_stats st{};
            rc = nqes_sr_step(ctx, j->W.data(), q.seed + (uint64_t)it * q.C, q.C, q.m, q.B, sched.step_lrate(),
                              sched.step_reglam(), q.rtol, q.maxit, reinterpret_cast<double *>(emat.data()), &st);
            if (rc != NQES_OK) break;
            r.t_device_ms += st.t_total_ms;
            r.steps = it + 1;
            j->progress.store(it + 1, std::memory_order_relaxed);
            for (auto &e : emat) e /= (double)q.N;
            sched.advance();
            std::vector<double> ev = srloop::sorted_real_eigs(emat, q.K);
            for (int k = 0; k < q.K; k++) r.energies[k] = ev[k];
            if (conv.update(it, ev, q.tol)) {
                r.converged = 1;
                break;
            }
        }
        r.status = rc;
        if (rc == NQES_ECANCELED) {
            j->err = "solve cancelled";
        } else if (rc != NQES_OK) {
            j->err = nqes_last_error(ctx);
            // a failed step may leave the context mid-solve; start the next one fresh
            nqes_destroy(ctx);
            slot.ctx = nullptr;
        }
    }

    void worker(int si)
    {
        for (;;) {
            Job *j = nullptr;
            bool trim = false;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_slot.wait(lk, [&] {
                    return slots[si].job || slots[si].trim || (stopping && queue.empty() && all_idle());
                });
                if (slots[si].job) {
                    j = slots[si].job;
                } else if (slots[si].trim) {
                    trim = true;
                } else {
                    break;  // stopping
                }
            }
            if (trim) {
                if (slots[si].ctx) nqes_destroy(slots[si].ctx);
                std::lock_guard<std::mutex> lk(mu);
                slots[si].ctx = nullptr;
                slots[si].reserved = 0;
                slots[si].trim = false;
                admit();
                continue;
            }
            run(slots[si], j);
            {
                std::lock_guard<std::mutex> lk(mu);
                j->res.t_done_ms = ms_since(t0);
                j->done = true;
                slots[si].job = nullptr;
                if (!slots[si].ctx) slots[si].reserved = 0;  // failed solve: context dropped
                admit();
                if (stopping && queue.empty() && all_idle()) cv_slot.notify_all();
            }
            cv_done.notify_all();
        }
        if (slots[si].ctx) nqes_destroy(slots[si].ctx);
        slots[si].ctx = nullptr;
    }

    bool all_idle() const
    {
        for (const Slot &s : slots)
            if (s.job) return false;
        return true;
    }
};

extern "C" {

nqes_server *nqes_server_create(int device, int max_slots, int64_t budget)
{
    int n = 0;
    if (max_slots < 1 || max_slots > 64) return nullptr;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    if (budget <= 0) {
        size_t freeb = 0, totalb = 0;
        if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) return nullptr;
        budget = (int64_t)(0.9 * (double)freeb);
    }
    auto *s = new nqes_server();
    s->device = device;
    s->max_slots = max_slots;
    s->budget = budget;
    s->t0 = clk::now();
    s->slots.resize(max_slots);
    for (int i = 0; i < max_slots; i++)
        s->workers.emplace_back([s, device, i] {
            (void)hipSetDevice(device);
            s->worker(i);
        });
    return s;
}

void nqes_server_destroy(nqes_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_slot.notify_all();
    for (auto &t : s->workers) t.join();
    delete s;
}

const char *nqes_server_last_error(const nqes_server *s) { return s ? s->err.c_str() : "null server"; }

int nqes_server_submit(nqes_server *s, const nqes_request *q, uint64_t *id)
{
    if (!s || !q || !id) return NQES_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto bad = [&](const char *msg) {
        s->err = msg;
        return NQES_EINVAL;
    };
    if (s->stopping) return bad("server is stopping");
    if (q->N < 1 || q->N > 256 || q->M < 1 || q->M > 512 || q->K < 1 || q->K > 4) return bad("N, M or K out of range");
    if (q->M % q->N != 0) return bad("M must be a multiple of N (rbm.cc:148)");
    if (!q->J || !q->W0) return bad("J and W0 are required");
    if (q->C < 1 || q->m < 1 || q->B < 0 || q->niters < 0) return bad("C, m, B or niters out of range");
    if (q->model != NQES_MODEL_TRAPPEDION && q->model != NQES_MODEL_HS && q->model != NQES_MODEL_XXZ)
        return bad("unsupported model");
    if (q->precision != NQES_PREC_FP64 && q->precision != NQES_PREC_MIXED) return bad("unsupported precision");
    int64_t need = nqes_required_bytes(q->N, q->M, q->K, q->C, q->m, q->precision);
    if (need < 0) return bad("bad sizes");
    need += kPerSolveBytes;
    if (need > s->budget) {
        s->err = "request needs " + std::to_string(need) + " device bytes, budget is " + std::to_string(s->budget);
        return NQES_ENOMEM;
    }
    auto j = std::make_unique<Job>();
    j->id = s->next_id++;
    j->req = *q;
    j->J.assign(q->J, q->J + (size_t)q->N * q->N);
    const size_t P = (size_t)q->K * (q->N + q->M + (size_t)q->N * q->M);
    j->W.assign(q->W0, q->W0 + 2 * P);
    j->req.J = nullptr;
    j->req.W0 = nullptr;
    j->bytes = need;
    j->res.bytes = need;
    j->res.t_submit_ms = ms_since(s->t0);
    *id = j->id;
    s->queue.push_back(j.get());
    s->jobs.emplace(j->id, std::move(j));
    s->admit();
    return NQES_OK;
}

int nqes_server_poll(nqes_server *s, uint64_t id, int *done)
{
    if (!s || !done) return NQES_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) {
        s->err = "unknown request id";
        return NQES_EINVAL;
    }
    *done = it->second->done ? 1 : 0;
    return NQES_OK;
}

int nqes_server_progress(nqes_server *s, uint64_t id, int *steps)
{
    if (!s || !steps) return NQES_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) {
        s->err = "unknown request id";
        return NQES_EINVAL;
    }
    *steps = it->second->progress.load(std::memory_order_relaxed);
    return NQES_OK;
}

int nqes_server_cancel(nqes_server *s, uint64_t id)
{
    if (!s) return NQES_EINVAL;
    bool notify = false;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) {
            s->err = "unknown request id";
            return NQES_EINVAL;
        }
        Job *j = it->second.get();
        if (j->done) return NQES_OK;
        j->cancel.store(true, std::memory_order_relaxed);
        for (auto qi = s->queue.begin(); qi != s->queue.end(); ++qi)
            if (*qi == j) {  // still queued: finished now, never admitted
                s->queue.erase(qi);
                j->res.K = j->req.K;
                j->res.status = NQES_ECANCELED;
                j->err = "solve cancelled";
                j->res.t_admit_ms = j->res.t_done_ms = ms_since(s->t0);
                j->done = true;
                notify = true;
                break;
            }
        if (notify) s->admit();  // a blocked request may fit now
    }
    if (notify) s->cv_done.notify_all();
    return NQES_OK;
}

int nqes_server_wait(nqes_server *s, uint64_t id, double *W, nqes_result *res)
{
    if (!s) return NQES_EINVAL;
    std::unique_ptr<Job> j;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) {
            s->err = "unknown request id";
            return NQES_EINVAL;
        }
        Job *jp = it->second.get();
        if (jp->claimed) {  // each id has one waiter; a second one is refused
            s->err = "request id already being waited for";
            return NQES_EINVAL;
        }
        jp->claimed = true;
        s->cv_done.wait(lk, [&] { return jp->done; });
        auto it2 = s->jobs.find(id);  // looked up again after the wait
        j = std::move(it2->second);
        s->jobs.erase(it2);
        if (j->res.status != NQES_OK) s->err = j->err;
    }
    if (W) std::memcpy(W, j->W.data(), j->W.size() * sizeof(double));
    if (res) *res = j->res;
    return j->res.status;
}

} // extern "C"
