// Serving engine: continuous batching of independent regressions on one GPU.
//
// Requests enter a FIFO queue. A request is admitted when a slot is free and its HBM footprint
// fits the budget next to the requests already on the device (the head waits otherwise, unless
// nothing is running). Each slot is a worker thread with its own absorber stream, so the
// kernels of the requests in flight interleave on the device and a request that arrives while
// others run starts as soon as a slot and memory free up.
//
// Packing: a slot that admits a small unweighted OLS request (iid or hetero variance) also
// admits the consecutive queued requests of that kind, up to `pack` of them, and solves them as
// ONE absorber: their fixed effects on disjoint level ranges, one structure build, one multi-RHS
// PCG launch sequence over all their columns (ols_fit_packed). Each packed result equals its
// solve alone within the request's tolerance; with pack = 1 every request runs alone, and a
// result then does not depend on the mix at all (every reduction is in a fixed order).
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "fh_fit.hpp"
#include "fh_internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    uint64_t ticket;
    fh_request req;
    uint64_t bytes;
    Clock::time_point submitted, started, finished;
    fh_response resp{};
    bool done = false;
};

// Packable: small unweighted OLS with iid or hetero variance (the packed solve's range).
constexpr size_t PACK_ROWS = 200'000;      // largest request packed
constexpr size_t PACK_TOTAL = 2'000'000;   // rows of one packed solve
bool packable(const fh_request& r) {
    if (!(r.model == FH_MODEL_OLS && (r.vcov_kind == FH_VCOV_IID || r.vcov_kind == FH_VCOV_HETERO) && r.K <= 15 &&
          r.nterms <= 8 && r.n <= PACK_ROWS && !r.cluster))
        return false;
    for (size_t q = 0; q < r.nterms; q++)  // intercept-only terms: their cell structures pack
        if (r.terms[q].intercept != 1 || r.terms[q].nslopes != 0) return false;
    return true;
}

uint64_t footprint(const fh_request& r) {
    // inputs (codes, y, X, cluster), the structure (pi, cells, level transposes), the row
    // block and its residual/solver copies, weights and scratch, with a fixed margin
    const double n = (double)r.n, K = (double)r.K, q = (double)r.nterms;
    const double per_row = 4.0 * q + 8.0 + 8.0 * K + 4.0 + 48.0 + 16.0 * q + 3.0 * 8.0 * (K + 1.0) +
                           (r.model == FH_MODEL_PPML ? 8.0 * (6.0 + K) : 0.0);
    return (uint64_t)(n * per_row) + (32ull << 20);
}

}  // namespace

struct fh_server {
    fh_server_config cfg{};
    int device = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job*> queue;
    std::unordered_map<uint64_t, Job*> jobs;
    uint64_t next_ticket = 1, inflight_bytes = 0, running = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    Clock::time_point created = Clock::now();

    void run_job(Job* j) {
        fh_response& out = j->resp;
        const fh_request& r = j->req;
        fh_solver_config sc;
        fh_config_default(&sc);
        sc.precision = cfg.precision;
        fh_absorber* ab = nullptr;
        std::vector<fh_term> terms(r.terms, r.terms + r.nterms);
        int rc = fh_absorber_create(terms.data(), r.nterms, nullptr, r.n, &sc, nullptr, &ab);
        fh_fit_stats st{};
        if (rc == FH_OK) {
            fh_solve_opts o{r.tol, r.maxiter, 0, -1};
            if (r.model == FH_MODEL_PPML)
                rc = fh_ppml_fit(ab, r.y, r.X, r.K, nullptr, r.vcov_kind, r.cluster, r.nested, &r.ssc, r.tol, r.maxiter,
                                 &o, r.beta, r.vcov, r.kept, &st, nullptr);
            else
                rc = fh_ols_fit(ab, r.y, r.X, r.K, r.vcov_kind, r.cluster, r.nested, &r.ssc, &o, r.beta, r.vcov,
                                r.kept, &st);
        }
        if (ab) fh_absorber_destroy(ab);
        out.rc = rc;
        out.packed = 1;
        out.iterations = st.iterations;
        out.pcg_iterations = st.pcg_iterations;
    }

    // One packed solve of several OLS requests.
    void run_batch(const std::vector<Job*>& js) {
        std::vector<fh::PackedOls> reqs(js.size());
        std::vector<fh::PackedOls*> ptrs;
        for (size_t i = 0; i < js.size(); i++) {
            const fh_request& r = js[i]->req;
            fh::PackedOls& p = reqs[i];
            p.n = r.n;
            p.K = r.K;
            for (size_t q = 0; q < r.nterms; q++) {
                p.codes.push_back(r.terms[q].codes);
                p.g.push_back(r.terms[q].g);
            }
            p.y = r.y;
            p.X = r.X;
            p.kind = r.vcov_kind;
            p.ssc = r.ssc;
            p.tol = r.tol;
            p.maxiter = r.maxiter;
            p.beta = r.beta;
            p.vcov = r.vcov;
            p.kept = r.kept;
            ptrs.push_back(&p);
        }
        try {
            fh::ols_fit_packed(ptrs, cfg.precision);
        } catch (const fh::Error& e) {
            for (auto& p : reqs) p.rc = e.rc;
        } catch (...) {
            for (auto& p : reqs) p.rc = FH_INTERNAL;
        }
        for (size_t i = 0; i < js.size(); i++) {
            js[i]->resp.rc = reqs[i].rc;
            js[i]->resp.iterations = reqs[i].iterations;
            js[i]->resp.pcg_iterations = reqs[i].iterations;
            js[i]->resp.packed = (int32_t)js.size();
        }
    }

    void worker() {
        (void)hipSetDevice(device);
        for (;;) {
            std::vector<Job*> js;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] {
                    if (stop) return true;
                    if (queue.empty()) return false;
                    const Job* h = queue.front();
                    return running == 0 || inflight_bytes + h->bytes <= cfg.hbm_budget;
                });
                if (stop && queue.empty()) return;
                if (queue.empty()) continue;
                js.push_back(queue.front());
                queue.pop_front();
                uint64_t bytes = js[0]->bytes;
                size_t rows = js[0]->req.n;
                // the consecutive packable requests behind a packable head join its solve
                if (cfg.pack > 1 && packable(js[0]->req)) {
                    while (!queue.empty() && js.size() < (size_t)cfg.pack) {
                        const Job* h = queue.front();
                        if (!packable(h->req) || rows + h->req.n > PACK_TOTAL ||
                            inflight_bytes + bytes + h->bytes > cfg.hbm_budget)
                            break;
                        js.push_back(queue.front());
                        queue.pop_front();
                        bytes += h->bytes;
                        rows += h->req.n;
                    }
                }
                inflight_bytes += bytes;
                running++;
                const auto now = Clock::now();
                for (Job* j : js) j->started = now;
            }
            if (js.size() == 1) run_job(js[0]);
            else run_batch(js);
            {
                std::lock_guard<std::mutex> lk(mu);
                const auto fin = Clock::now();
                for (Job* j : js) {
                    j->finished = fin;
                    j->resp.queued_s = std::chrono::duration<double>(j->started - j->submitted).count();
                    j->resp.service_s = std::chrono::duration<double>(j->finished - j->started).count();
                    j->resp.total_s = std::chrono::duration<double>(j->finished - j->submitted).count();
                    j->resp.start_s = std::chrono::duration<double>(j->started - created).count();
                    j->resp.end_s = std::chrono::duration<double>(j->finished - created).count();
                    j->done = true;
                    inflight_bytes -= j->bytes;
                }
                running--;
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

void fh_server_config_default(fh_server_config* c) {
    c->slots = 8;
    c->hbm_budget = 0;
    c->precision = FH_MIXED;
    c->pack = 32;
}

int fh_server_create(const fh_server_config* cfg, fh_server** out) {
    if (!out) return FH_INVALID;
    *out = nullptr;
    try {
        auto s = std::make_unique<fh_server>();
        if (cfg) s->cfg = *cfg;
        else fh_server_config_default(&s->cfg);
        if (s->cfg.slots < 1 || s->cfg.slots > 64 || s->cfg.pack < 0 || s->cfg.pack > 256) return FH_INVALID;
        if (hipGetDevice(&s->device) != hipSuccess) return FH_INTERNAL;
        if (s->cfg.hbm_budget == 0) {
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) != hipSuccess) return FH_INTERNAL;
            s->cfg.hbm_budget = fr / 4 * 3;
        }
        for (uint32_t i = 0; i < s->cfg.slots; i++) s->workers.emplace_back([p = s.get()] { p->worker(); });
        *out = s.release();
        return FH_OK;
    } catch (...) {
        return FH_INTERNAL;
    }
}

int fh_server_submit(fh_server* s, const fh_request* r, uint64_t* ticket) {
    if (!s || !r || !ticket || !r->terms || !r->y || !r->X || !r->beta || !r->vcov || !r->kept || r->n == 0 ||
        r->K == 0)
        return FH_INVALID;
    auto* j = new Job();
    j->req = *r;
    j->bytes = footprint(*r);
    j->submitted = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        j->ticket = s->next_ticket++;
        s->jobs[j->ticket] = j;
        s->queue.push_back(j);
        *ticket = j->ticket;
    }
    s->cv_work.notify_all();
    return FH_OK;
}

int fh_server_wait(fh_server* s, uint64_t ticket, fh_response* resp) {
    if (!s || !resp) return FH_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return FH_INVALID;
    Job* j = it->second;
    s->cv_done.wait(lk, [&] { return j->done; });
    *resp = j->resp;
    s->jobs.erase(it);
    delete j;
    return FH_OK;
}

void fh_server_destroy(fh_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    for (auto& kv : s->jobs) delete kv.second;
    delete s;
}

}  // extern "C"
