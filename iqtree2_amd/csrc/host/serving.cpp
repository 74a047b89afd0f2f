// Continuous-batching request server over the likelihood engine.
#include "plk_serve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <thread>
#include <vector>

#include "newton.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Req {
    uint64_t ticket = 0;
    plks_request r{};
    // owned copies of the request arrays
    std::vector<uint8_t> states;
    std::vector<double> freq, invar, model;
    std::vector<plk_op> ops;
    std::vector<float> weights;
    size_t footprint = 0;
    int inst = -1;
    enum Phase { TRAVERSE, THETA, NEWTON, FINAL, DONE } phase = TRAVERSE;
    plk::NewtonRun nr;
    double optx = 0.0, lnl_opt = 0.0, lnl_last = 0.0, lnl_orig = 0.0;
    bool need_orig = false;   // divergence guard: lnL at the last evaluated and the old length
    bool need_last = false;   // the last evaluated length differs from optx
    plks_result res{};
    std::vector<double> rell;
};

}  // namespace

struct plks_server {
    plk_context* ctx = nullptr;
    int max_inflight = 256;
    size_t budget = 0, used = 0;
    Clock::time_point t0;
    std::mutex mu;   // queue, finished, live, stop, stats
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> queue;
    std::map<uint64_t, Req*> finished;
    std::set<uint64_t> live;
    std::vector<Req*> inflight;   // scheduler thread only
    uint64_t next_ticket = 1;
    bool stop = false;
    plks_stats stats{};
    std::thread worker;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void complete(plks_server* s, Req* q) {
    q->res.finish_s = s->now();
    std::lock_guard<std::mutex> lk(s->mu);
    s->finished[q->ticket] = q;
    s->stats.completed++;
    s->cv_done.notify_all();
}

void release(plks_server* s, Req* q) {
    if (q->inst >= 0) {
        plk_instance_destroy(s->ctx, q->inst);
        q->inst = -1;
        s->used -= q->footprint;
    }
}

// admit queued requests in arrival order while slots and memory allow
void admit(plks_server* s) {
    for (;;) {
        Req* q = nullptr;
        bool too_big = false;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->queue.empty()) return;
            q = s->queue.front();
            if (q->footprint > s->budget) {
                too_big = true;
            } else if ((int)s->inflight.size() >= s->max_inflight || s->used + q->footprint > s->budget) {
                return;
            }
            s->queue.pop_front();
        }
        q->res.admit_s = s->now();
        if (too_big) {
            q->res.status = PLK_ERR_NOMEM;
            complete(s, q);
            continue;
        }
        const plks_request& r = q->r;
        int st = plk_instance_create(s->ctx, &r.desc, &q->inst);
        if (st == PLK_OK) {
            s->used += q->footprint;
            st = plk_set_patterns(s->ctx, q->inst, r.states, r.freq, r.invar);
        }
        if (st == PLK_OK)
            st = plk_set_model(s->ctx, q->inst, r.eval, r.evec, r.inv_evec, r.tip_eigen, r.cat_rates,
                               r.cat_props);
        if (st != PLK_OK) {
            release(s, q);
            q->res.status = st;
            complete(s, q);
            continue;
        }
        q->phase = r.kind == PLKS_NEWTON ? Req::THETA : Req::TRAVERSE;
        s->inflight.push_back(q);
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.peak_inflight = std::max(s->stats.peak_inflight, (int)s->inflight.size());
        s->stats.peak_bytes = std::max(s->stats.peak_bytes, s->used);
    }
}

void finish_rell(plks_server* s, Req* q) {
    if (q->res.status == PLK_OK && q->r.kind == PLKS_RELL) {
        q->rell.assign(q->r.nboot, 0.0);
        q->res.status = plk_rell(s->ctx, q->inst, q->weights.data(), q->r.nboot, q->rell.data());
        q->res.nboot = q->r.nboot;
    }
    q->phase = Req::DONE;
}

// one scheduler iteration over the resident requests
void step(plks_server* s) {
    plk_context* ctx = s->ctx;
    bool did = false;
    // every request that needs a traversal (tree lnL, or theta for a branch
    // optimisation just admitted): one fused launch
    std::vector<Req*> tv;
    for (Req* q : s->inflight)
        if (q->phase == Req::TRAVERSE || q->phase == Req::THETA) tv.push_back(q);
    if (!tv.empty()) {
        did = true;
        std::vector<plk_batch_item> items(tv.size());
        std::vector<int> modes(tv.size());
        std::vector<double> lnl(tv.size(), 0.0);
        for (size_t k = 0; k < tv.size(); k++) {
            const plks_request& r = tv[k]->r;
            items[k] = plk_batch_item{tv[k]->inst, r.ops, r.nops, r.node, r.dad, r.length};
            modes[k] = tv[k]->phase == Req::THETA ? PLK_BATCH_THETA : PLK_BATCH_LNL;
        }
        const double tb = s->now();
        std::vector<int> status(tv.size(), PLK_OK);
        const int st = plk_batch_evaluate(ctx, items.data(), modes.data(), (int)items.size(), lnl.data());
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stats.traversal_batches++;
        }
        if (st != PLK_OK && st != PLK_ERR_NONFINITE) {
            // a request the batch rejected: find it by giving each its own
            // batch (the same per-item arithmetic as in a shared one)
            for (size_t k = 0; k < tv.size(); k++) {
                const int sk = plk_batch_evaluate(ctx, &items[k], &modes[k], 1, &lnl[k]);
                status[k] = sk == PLK_ERR_NONFINITE ? PLK_OK : sk;
            }
        }
        for (size_t k = 0; k < tv.size(); k++) {
            Req* q = tv[k];
            q->res.status = status[k];
            if (status[k] != PLK_OK) {
                q->phase = Req::DONE;
            } else if (q->phase == Req::THETA) {
                const plks_request& r = q->r;
                q->nr.start(r.length, r.min_len, r.max_len, r.min_len, r.max_nr);
                q->phase = Req::NEWTON;
            } else {
                q->res.lnl = lnl[k];
                finish_rell(s, q);
            }
        }
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.traverse_s += s->now() - tb;
    }
    // derivative rounds: one derivative evaluation for every optimisation in
    // progress per round, and the final lnL evaluations of those whose Newton
    // iteration has ended. Several rounds run back to back (a round is one
    // small batched launch, far cheaper than an iteration with admissions and
    // a traversal batch), so an optimisation takes few scheduler iterations;
    // the rounds stop when none is left, after kDervRounds, or after two when
    // new requests wait to be admitted. Every request's Newton steps are the
    // same whatever the rounds.
    constexpr int kDervRounds = 16;
    for (int round = 0; round < kDervRounds; round++) {
        if (round >= 2) {
            std::lock_guard<std::mutex> lk(s->mu);
            if (!s->queue.empty()) break;
        }
        std::vector<Req*> dv;
        // which: 0 Newton step, 1 lnL at optx, 2 lnL at the old length, 3 lnL at
        // the last evaluated length
        std::vector<int> ids, which;
        std::vector<double> xs;
        for (Req* q : s->inflight) {
            if (q->phase == Req::NEWTON) {
                dv.push_back(q);
                ids.push_back(q->inst);
                xs.push_back(q->nr.x());
                which.push_back(0);
            } else if (q->phase == Req::FINAL) {
                dv.push_back(q);
                ids.push_back(q->inst);
                xs.push_back(q->optx);
                which.push_back(1);
                if (q->need_orig) {
                    dv.push_back(q);
                    ids.push_back(q->inst);
                    xs.push_back(q->r.length);
                    which.push_back(2);
                }
                if (q->need_last) {
                    dv.push_back(q);
                    ids.push_back(q->inst);
                    xs.push_back(q->nr.last_x());
                    which.push_back(3);
                }
            }
        }
        if (!dv.empty()) {
            did = true;
            const size_t n = dv.size();
            const double tb = s->now();
            std::vector<double> df(n), ddf(n), lnl(n);
            const int st = plk_likelihood_derv_batch(ctx, ids.data(), xs.data(), (int)n, df.data(), ddf.data(),
                                                     lnl.data());
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->stats.derv_batches++;
            }
            if (st != PLK_OK && st != PLK_ERR_NONFINITE) {
                for (Req* q : dv) {
                    q->res.status = st;
                    q->phase = Req::DONE;
                }
            } else {
                std::vector<Req*> finals;
                for (size_t k = 0; k < n; k++) {
                    Req* q = dv[k];
                    if (q->phase == Req::DONE) continue;
                    if (which[k] == 0) {
                        // a non-finite derivative (marked by lnl = NaN, df = ddf = 0):
                        // an error of the normal kernel, df = ddf = 0 for the safe
                        // one (tree/phylokernelnew.h:2577-2579, 2629-2632)
                        if (std::isnan(lnl[k]) && !q->r.desc.safe_numeric) {
                            q->res.status = PLK_ERR_NONFINITE;
                            q->phase = Req::DONE;
                            continue;
                        }
                        q->nr.feed(-df[k], -ddf[k]);
                        if (!q->nr.done()) continue;
                        if (q->nr.failed()) {
                            q->res.status = PLK_ERR_NEWTON;
                            q->phase = Req::DONE;
                            continue;
                        }
                        // optimizeOneBranch's divergence guard needs lnL at optx and
                        // at the old length (tree/phylotree.cpp:2616-2625)
                        q->optx = q->nr.result();
                        q->need_orig = q->optx > q->r.max_len * 0.95;
                        q->need_last = q->need_orig && q->nr.last_x() != q->optx;
                        q->phase = Req::FINAL;
                    } else if (which[k] == 1) {
                        q->lnl_opt = lnl[k];
                        finals.push_back(q);
                    } else if (which[k] == 2) {
                        q->lnl_orig = lnl[k];
                    } else {
                        q->lnl_last = lnl[k];
                    }
                }
                for (Req* q : finals) {
                    // lnL values the batch could not form: computeLikelihoodFromBuffer
                    // with its per-pattern fix-up, one request at a time
                    int fst = PLK_OK;
                    auto fixed = [&](double x, double& v) {
                        if (fst == PLK_OK && !std::isfinite(v))
                            fst = plk_likelihood_from_buffer(ctx, q->inst, x, &v, nullptr);
                    };
                    fixed(q->optx, q->lnl_opt);
                    if (q->need_orig) fixed(q->r.length, q->lnl_orig);
                    if (q->need_last) fixed(q->nr.last_x(), q->lnl_last);
                    if (fst != PLK_OK) {
                        q->res.status = fst;
                        q->phase = Req::DONE;
                        continue;
                    }
                    const double guard_lh = q->need_last ? q->lnl_last : q->lnl_opt;
                    if (q->need_orig && q->lnl_orig > guard_lh) {
                        q->res.length = q->r.length;
                        q->res.lnl = q->lnl_orig;
                    } else {
                        q->res.length = q->optx;
                        q->res.lnl = q->lnl_opt;
                    }
                    q->res.n_derv = q->nr.evals();
                    q->res.status = PLK_OK;
                    q->phase = Req::DONE;
                }
            }
            std::lock_guard<std::mutex> lk(s->mu);
            s->stats.derv_s += s->now() - tb;
        }
        if (dv.empty()) break;
    }
    if (did) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.iterations++;
    }
}

void retire(plks_server* s) {
    std::vector<Req*> keep;
    for (Req* q : s->inflight) {
        if (q->phase != Req::DONE) {
            keep.push_back(q);
            continue;
        }
        release(s, q);
        complete(s, q);
    }
    s->inflight.swap(keep);
}

void run(plks_server* s) {
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty() || !s->inflight.empty(); });
            if (s->stop && s->queue.empty() && s->inflight.empty()) return;
        }
        const double t0 = s->now();
        admit(s);
        const double t1 = s->now();
        if (!s->inflight.empty()) step(s);
        const double t2 = s->now();
        retire(s);
        const double t3 = s->now();
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.admit_s += t1 - t0;
        s->stats.retire_s += t3 - t2;
    }
}

template <typename T>
bool copy_into(std::vector<T>& dst, const T* src, size_t n) {
    if (n > 0 && !src) return false;
    dst.assign(src, src + n);
    return true;
}

}  // namespace

extern "C" {

plk_status plks_create(plk_context* ctx, const plks_config* cfg, plks_server** out) {
    if (!ctx || !out) return PLK_ERR_ARG;
    size_t ctx_budget = 0, ctx_used = 0;
    plk_status st = plk_memory(ctx, &ctx_used, &ctx_budget);
    if (st != PLK_OK) return st;
    plks_server* s = new plks_server();
    s->ctx = ctx;
    s->t0 = Clock::now();
    s->max_inflight = cfg && cfg->max_inflight > 0 ? cfg->max_inflight : 256;
    s->budget = cfg && cfg->hbm_budget > 0 ? cfg->hbm_budget : ctx_budget / 4 * 3;
    s->worker = std::thread(run, s);
    *out = s;
    return PLK_OK;
}

plk_status plks_destroy(plks_server* s) {
    if (!s) return PLK_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->worker.join();
    for (auto& kv : s->finished) delete kv.second;
    delete s;
    return PLK_OK;
}

plk_status plks_submit(plks_server* s, const plks_request* r, uint64_t* ticket) {
    if (!s || !r || !ticket) return PLK_ERR_ARG;
    const plk_instance_desc& d = r->desc;
    if (d.ntaxa < 3 || d.npatterns <= 0 || d.nstates < 2 || d.ncat < 1 || d.state_unknown < d.nstates ||
        r->nops < 0 || (r->nops > 0 && !r->ops))
        return PLK_ERR_ARG;
    if (r->kind != PLKS_LNL && r->kind != PLKS_NEWTON && r->kind != PLKS_RELL) return PLK_ERR_ARG;
    if (r->kind == PLKS_NEWTON && !(r->min_len > 0.0 && r->max_len > r->min_len)) return PLK_ERR_ARG;
    if (r->kind == PLKS_RELL && (r->nboot < 1 || !r->weights)) return PLK_ERR_ARG;
    Req* q = new Req();
    const size_t P = d.npatterns, n = d.nstates, C = d.ncat, SU = d.state_unknown;
    bool ok = copy_into(q->states, r->states, (size_t)d.ntaxa * P) && copy_into(q->freq, r->freq, P) &&
              copy_into(q->invar, r->invar, P) && copy_into(q->ops, r->ops, (size_t)r->nops);
    const double* parts[6] = {r->eval, r->evec, r->inv_evec, r->tip_eigen, r->cat_rates, r->cat_props};
    const size_t sizes[6] = {n, n * n, n * n, (SU + 1) * n, C, C};
    size_t off[6], total = 0;
    for (int k = 0; k < 6; k++) {
        off[k] = total;
        total += sizes[k];
        if (!parts[k]) ok = false;
    }
    if (ok) {
        q->model.resize(total);
        for (int k = 0; k < 6; k++) std::copy(parts[k], parts[k] + sizes[k], q->model.begin() + off[k]);
    }
    if (ok && r->kind == PLKS_RELL) ok = copy_into(q->weights, r->weights, (size_t)r->nboot * P);
    if (!ok) {
        delete q;
        return PLK_ERR_ARG;
    }
    q->r = *r;
    q->r.states = q->states.data();
    q->r.freq = q->freq.data();
    q->r.invar = q->invar.data();
    q->r.ops = q->ops.empty() ? nullptr : q->ops.data();
    q->r.eval = q->model.data() + off[0];
    q->r.evec = q->model.data() + off[1];
    q->r.inv_evec = q->model.data() + off[2];
    q->r.tip_eigen = q->model.data() + off[3];
    q->r.cat_rates = q->model.data() + off[4];
    q->r.cat_props = q->model.data() + off[5];
    q->r.weights = q->weights.empty() ? nullptr : q->weights.data();
    // device state plus RELL's temporary weight and output buffers
    q->footprint = plk_instance_footprint(&d);
    if (r->kind == PLKS_RELL) q->footprint += (size_t)r->nboot * (P * sizeof(float) + sizeof(double)) + 512;
    q->res.status = PLK_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        q->ticket = s->next_ticket++;
        q->res.submit_s = s->now();
        s->queue.push_back(q);
        s->live.insert(q->ticket);
        *ticket = q->ticket;
    }
    s->cv_work.notify_all();
    return PLK_OK;
}

plk_status plks_wait(plks_server* s, uint64_t ticket, double timeout_s, plks_result* res,
                     double* rell_out, int* done) {
    if (!s || !done) return PLK_ERR_ARG;
    *done = 0;
    std::unique_lock<std::mutex> lk(s->mu);
    if (!s->live.count(ticket)) return PLK_ERR_ARG;
    auto ready = [&] { return s->finished.count(ticket) > 0; };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, ready);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
        return PLK_OK;
    }
    Req* q = s->finished[ticket];
    s->finished.erase(ticket);
    s->live.erase(ticket);
    lk.unlock();
    if (res) *res = q->res;
    if (rell_out && !q->rell.empty()) std::copy(q->rell.begin(), q->rell.end(), rell_out);
    *done = 1;
    delete q;
    return PLK_OK;
}

plk_status plks_get_stats(plks_server* s, plks_stats* out) {
    if (!s || !out) return PLK_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    return PLK_OK;
}

double plks_clock(plks_server* s) { return s ? s->now() : 0.0; }

}  // extern "C"
