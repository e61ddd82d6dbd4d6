// server.cpp -- continuous-batching request loop over one engine (pdds_server_* of
// include/pdds_engine.h).  No counterpart in the reference, whose PDDSparseGM::Solve runs one
// mesh per MPI job; the per-request work is the same assembly and solve as pdds_assemble /
// pdds_solve / pdds_point_eval.
//
// Threads: prep_threads host workers prepare requests (a solve builds its mesh, Psi^-1 and
// host stencil tables, each build spread over up to 16 host threads of its own, so a few
// workers already occupy the host and keep preparation close to arrival order); one device
// thread runs rounds; a solver thread runs the interface solves.  A round
// admits prepared requests first-fit by device footprint, builds one knot list over every
// admitted request (point paths first, then the next knots of each solve, up to round_paths
// paths), runs it as one assembly over the requests' stencil tables, scatters the per-knot
// results back to their requests, completes the points and hands each solve whose knots are
// all done to the solver thread, which solves (I + G) u = B on a second engine.  Each
// request's paths are keyed by its own Philox stream and every sum has a fixed order, so its
// results do not depend on the other requests of a round.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>

#include "engine_internal.h"

namespace {

enum ReqState { RS_PREP = 0, RS_READY, RS_ADMITTED, RS_DONE };

// device bytes per path of a round: its exit record
constexpr int64_t kRecordBytes = (int64_t)sizeof(ExitRec);
// device bytes per work item of a knot besides its records and partials: the item's descriptor
// and statistics, and (counted per item, an over-estimate) the knot's descriptor, statistics and
// step counters
constexpr int64_t kKnotBytes = (int64_t)(sizeof(KnotDesc) + sizeof(KnotStat) + 2 * sizeof(unsigned long long) +
                                         sizeof(ItemDesc) + sizeof(ItemStat));

struct SrvReq {
    uint64_t ticket = 0;
    pdds_request rq{};
    pdds_result res{};
    ReqState state = RS_PREP;
    double t_submit = 0, t_admit = 0;
    int64_t footprint = 0;            // device bytes while admitted
    bool ran = false;                 // point: its paths have run
    // solve
    std::unique_ptr<pdds_mesh> mesh;
    std::vector<StencilDesc> sten;    // host stencil tables (uploaded on admission)
    std::vector<double> entx, enty, ipsi, entic2;
    std::vector<int8_t> ents;
    std::vector<int32_t> e1, e2;
    DevTables tab;
    std::vector<int> mc;              // Monte Carlo knots in knot order
    size_t next = 0, finished = 0;    // next knot to launch, knots done
    std::vector<double> G, B;
    std::vector<double> u;
};

// work items of a knot of n paths in an assembly with run r (as run_assembly cuts them)
int64_t items_of(const pdds_run &r, int64_t n) {
    const int S = r.sub_batch > 0 ? r.sub_batch : 32;
    const int64_t P = r.item_paths > 0 ? r.item_paths : choose_P(n, S);
    return (n + P - 1) / P;
}

double table_bytes(const SrvReq &r) {
    return (double)r.sten.size() * sizeof(StencilDesc) + 8.0 * (r.entx.size() + r.enty.size() + r.ipsi.size()) +
           (double)r.ents.size() + 4.0 * (r.e1.size() + r.e2.size());
}

}  // namespace

struct pdds_server {
    pdds_engine *e = nullptr;
    pdds_server_config cfg{};
    PdeParams pp{};
    std::mutex mu;
    std::condition_variable cv_prep, cv_work, cv_done;
    std::deque<SrvReq *> prep_q;
    std::deque<SrvReq *> ready;
    std::vector<SrvReq *> admitted;
    std::map<uint64_t, std::unique_ptr<SrvReq>> reqs;
    std::vector<std::thread> preps;
    std::thread dev;
    // interface solves run on a second engine (its own stream and buffers) in their own thread,
    // so that the next round's paths run on the device while a finished request's BiCGSTAB does
    pdds_engine *se = nullptr;
    std::thread solver;
    std::deque<SrvReq *> to_solve;
    std::condition_variable cv_solve;
    bool dev_done = false;
    bool stop = false;
    uint64_t next_ticket = 1;
    int64_t inflight_bytes = 0;
    int64_t n_submitted = 0, n_done = 0, n_rounds = 0, n_steps = 0;
    int64_t n_solver = 0;             // requests handed to the solver thread and not finished
    double busy_ms = 0;
    // G values of a round, kept across rounds (the device thread's only): a round's knots can carry
    // millions of values, and a fresh zeroed vector per round cost tens of milliseconds
    std::vector<double> round_G;
    DevTables dom;                    // the domain's one-stencil table (point requests)
    bool dom_ready = false;
};

namespace {

void finish(pdds_server *s, SrvReq *r, pdds_status st) {
    // caller holds s->mu
    r->res.status = st;
    r->res.latency_ms = now_ms() - r->t_submit;
    r->state = RS_DONE;
    ++s->n_done;
    s->cv_done.notify_all();
}

// host preparation: mesh, Psi^-1 and stencil tables of a solve; nothing for a point
void prepare(pdds_server *s, SrvReq *r) {
    pdds_status st = PDDS_OK;
    if (r->rq.kind == PDDS_REQ_SOLVE) {
        const double *d = s->cfg.domain;
        r->mesh.reset(new pdds_mesh(d[0], d[1], d[2], d[3], r->rq.n_sub, r->rq.n_node, r->rq.fac));
        const Mesh &M = r->mesh->m;
        if (M.any_singular) st = PDDS_ERR_SINGULAR;
        if (!st) {
            mesh_tables(M, s->cfg.run.quirk_last_side, r->sten, r->entx, r->enty, r->ents, r->entic2, r->e1, r->e2, r->ipsi);
            const int n = M.n_knots();
            r->G.assign(std::max<int64_t>(M.row_ptr.back(), 1), 0.0);
            r->B.assign(n, 0.0);
            for (int q = 0; q < n; ++q) {
                if (M.on_boundary[q]) r->B[q] = host_g(&s->cfg.problem, M.x0[q], M.y0[q], s->pp);
                else r->mc.push_back(q);
            }
            r->res.n_knots = n;
            // tables, the records of at most one round of its paths, per knot descriptors and
            // statistics, and its CSR values: G and its variance (16 B per entry), and the
            // reduction partials (2 per entry and work item)
            const int64_t paths = (int64_t)r->mc.size() * r->rq.n_paths;
            const int64_t items_per_knot = items_of(s->cfg.run, r->rq.n_paths);
            r->footprint = (int64_t)table_bytes(*r) + std::min<int64_t>(paths, s->cfg.round_paths) * kRecordBytes +
                           (int64_t)r->mc.size() * items_per_knot * kKnotBytes +
                           (16 + 16 * items_per_knot) * (int64_t)M.row_ptr.back();
        }
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (!st && r->footprint > s->cfg.hbm_budget) st = PDDS_ERR_FULL;
    if (st) { finish(s, r, st); return; }
    r->state = RS_READY;
    s->ready.push_back(r);
    s->cv_work.notify_all();
}

void prep_loop(pdds_server *s) {
    while (true) {
        SrvReq *r;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_prep.wait(lk, [&] { return s->stop || !s->prep_q.empty(); });
            if (s->prep_q.empty()) return;
            r = s->prep_q.front();
            s->prep_q.pop_front();
        }
        prepare(s, r);
    }
}

struct RoundKnot {
    SrvReq *r;
    int knot;      // solve: knot index; point: -1
};

// One round on the device; the engine lock is held.  Returns false if nothing ran.
bool run_round(pdds_server *s) {
    pdds_engine *e = s->e;
    std::vector<SrvReq *> adm;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        // first-fit admission by device footprint, in arrival order
        for (auto it = s->ready.begin(); it != s->ready.end();) {
            SrvReq *r = *it;
            if ((int)s->admitted.size() >= s->cfg.max_inflight) break;
            if (s->inflight_bytes + r->footprint <= s->cfg.hbm_budget) {
                r->state = RS_ADMITTED;
                r->t_admit = now_ms();
                r->res.queue_ms = r->t_admit - r->t_submit;
                s->inflight_bytes += r->footprint;
                s->admitted.push_back(r);
                it = s->ready.erase(it);
            } else {
                ++it;
            }
        }
        adm = s->admitted;
    }
    if (adm.empty()) return false;
    if (hipSetDevice(e->device) != hipSuccess) return false;
    const double t0 = now_ms();
    // device tables of newly admitted solves; the domain table once
    std::vector<SrvReq *> failed;
    for (SrvReq *r : adm)
        if (r->rq.kind == PDDS_REQ_SOLVE && r->tab.sten == nullptr &&
            build_tables(e, r->sten, r->entx, r->enty, r->ents, r->entic2, r->e1, r->e2, r->ipsi, r->tab) != PDDS_OK)
            failed.push_back(r);
    if (!s->dom_ready) {
        StencilDesc d;
        std::memset(&d, 0, sizeof(d));
        for (int i = 0; i < 4; ++i) { d.rect[i] = s->cfg.domain[i]; d.bbox[i] = s->cfg.domain[i]; d.inv_c2[i] = 1.0; }
        d.gmask = 15;
        std::vector<StencilDesc> sd = {d};
        std::vector<double> none;
        std::vector<int8_t> none8;
        std::vector<int32_t> none32;
        if (build_tables(e, sd, none, none, none8, none, none32, none32, none, s->dom) == PDDS_OK) s->dom_ready = true;
    }
    // the round's knots: point paths first, then the next knots of each solve
    std::vector<const DevTables *> tabs = {&s->dom};
    std::vector<AsmKnot> knots;
    std::vector<RoundKnot> who;
    int64_t paths = 0, n_vals = 0;
    const int64_t cap = s->cfg.round_paths;
    for (SrvReq *r : adm) {
        if (r->rq.kind != PDDS_REQ_POINT || r->ran || !s->dom_ready) continue;
        if (!knots.empty() && paths + r->rq.n_paths > cap) continue;
        AsmKnot k{r->rq.x, r->rq.y, 0u, 0, r->rq.n_paths, 0};
        k.has_k2 = true; k.k2 = r->rq.k2; k.has_stream = true; k.stream = r->rq.stream; k.h = r->rq.h;
        knots.push_back(k);
        who.push_back({r, -1});
        paths += r->rq.n_paths;
        r->ran = true;
    }
    for (SrvReq *r : adm) {
        if (r->rq.kind != PDDS_REQ_SOLVE || r->tab.sten == nullptr) continue;
        const Mesh &M = r->mesh->m;
        int ti = -1;
        while (r->next < r->mc.size() && (knots.empty() || paths + r->rq.n_paths <= cap)) {
            if (ti < 0) { ti = (int)tabs.size(); tabs.push_back(&r->tab); }
            const int q = r->mc[r->next++];
            AsmKnot k{M.x0[q], M.y0[q], (uint32_t)q, M.stencil_of[q], r->rq.n_paths, n_vals};
            k.table = ti; k.has_k2 = true; k.k2 = r->rq.k2; k.has_stream = true; k.stream = r->rq.stream; k.h = r->rq.h;
            knots.push_back(k);
            who.push_back({r, q});
            n_vals += M.row_ptr[q + 1] - M.row_ptr[q];
            paths += r->rq.n_paths;
        }
    }
    pdds_status st = PDDS_OK;
    AsmOut o;
    // (G variances are not part of a served result: none are copied back)
    if ((int64_t)s->round_G.size() < std::max<int64_t>(n_vals, 1)) s->round_G.resize(std::max<int64_t>(n_vals, 1));
    std::vector<double> &Gr = s->round_G;
    if (!knots.empty())
        st = run_assembly(e, &s->cfg.problem, &s->cfg.run, s->cfg.domain, tabs, knots, n_vals, Gr.data(), nullptr, o);
    int64_t steps = 0;
    for (size_t i = 0; i < who.size() && !st; ++i) {
        SrvReq *r = who[i].r;
        const KnotStat &ks = o.kstats[i];
        r->res.path_steps += (int64_t)ks.steps;
        steps += (int64_t)ks.steps;
        if (who[i].knot < 0) {
            r->res.u = ks.B;
            r->res.u_var = ks.BB - ks.B * ks.B;
        } else {
            const Mesh &M = r->mesh->m;
            const int q = who[i].knot;
            const int64_t len = M.row_ptr[q + 1] - M.row_ptr[q];
            std::copy(Gr.begin() + knots[i].out_off, Gr.begin() + knots[i].out_off + len, r->G.begin() + M.row_ptr[q]);
            r->B[q] = ks.B;
            r->finished += 1;
        }
    }
    // requests whose paths are all done: points complete here, solves go to the solver thread
    std::vector<SrvReq *> complete, solve;
    for (SrvReq *r : adm) {
        pdds_status rs = st;
        bool done = false;
        if (std::find(failed.begin(), failed.end(), r) != failed.end()) { done = true; rs = PDDS_ERR_NOMEM; }
        else if (st) done = true;
        else if (r->rq.kind == PDDS_REQ_POINT) done = r->ran;
        else if (r->finished == r->mc.size()) solve.push_back(r);
        if (done) {
            r->res.status = rs;
            complete.push_back(r);
        }
    }
    const double dt = now_ms() - t0;
    std::lock_guard<std::mutex> lk(s->mu);
    s->busy_ms += dt;
    s->n_steps += steps;
    ++s->n_rounds;
    auto release = [&](SrvReq *r) {
        r->tab.release();
        r->sten.clear(); r->entx.clear(); r->enty.clear(); r->ipsi.clear(); r->entic2.clear(); r->ents.clear(); r->e1.clear(); r->e2.clear();
        s->inflight_bytes -= r->footprint;
        s->admitted.erase(std::find(s->admitted.begin(), s->admitted.end(), r));
    };
    for (SrvReq *r : complete) {
        release(r);
        finish(s, r, r->res.status);
    }
    for (SrvReq *r : solve) {
        release(r);
        s->to_solve.push_back(r);
        ++s->n_solver;
    }
    if (!solve.empty()) s->cv_solve.notify_all();
    return true;
}

// (I + G) u = B of the solves whose knots are done, on the solver engine, in hand-over order
void solver_loop(pdds_server *s) {
    while (true) {
        SrvReq *r;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_solve.wait(lk, [&] { return !s->to_solve.empty() || s->dev_done; });
            if (s->to_solve.empty()) return;
            r = s->to_solve.front();
            s->to_solve.pop_front();
        }
        const Mesh &M = r->mesh->m;
        const int n = M.n_knots();
        r->u.assign(n, 0.0);
        int32_t it = 0;
        double rr = 0;
        const double tol = r->rq.tol > 0 ? r->rq.tol : 1e-10;
        std::vector<int32_t> cols(M.cols.begin(), M.cols.end());
        pdds_status rs;
        {
            std::lock_guard<std::mutex> lk(s->se->mu);
            rs = solve_system(s->se, n, M.row_ptr.data(), cols.data(), r->G.data(), r->B.data(), r->u.data(), tol, 0,
                              &it, &rr);
        }
        r->res.iters = it;
        r->res.relres = rr;
        std::lock_guard<std::mutex> lk(s->mu);
        --s->n_solver;
        finish(s, r, rs);
        s->cv_work.notify_all();
    }
}

void dev_loop(pdds_server *s) {
    while (true) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                return !s->ready.empty() || !s->admitted.empty() ||
                       (s->stop && s->prep_q.empty() && s->n_done + s->n_solver == s->n_submitted);
            });
            if (s->ready.empty() && s->admitted.empty()) return;
        }
        bool ran;
        {
            std::lock_guard<std::mutex> lk(s->e->mu);
            ran = run_round(s);
        }
        if (!ran) {
            // nothing admissible now (budget taken by requests that cannot run): wait for news
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait_for(lk, std::chrono::milliseconds(1));
        }
    }
}

pdds_status take_result(pdds_server *s, std::map<uint64_t, std::unique_ptr<SrvReq>>::iterator it, pdds_result *res,
                        double *u, int64_t cap) {
    SrvReq *r = it->second.get();
    if (res) *res = r->res;
    if (u && cap > 0)
        for (int64_t i = 0; i < std::min<int64_t>(cap, (int64_t)r->u.size()); ++i) u[i] = r->u[i];
    s->reqs.erase(it);
    // other threads waiting on this ticket wake up, find it gone and return PDDS_ERR_INVALID
    s->cv_done.notify_all();
    return PDDS_OK;
}

}  // namespace

extern "C" {

pdds_status pdds_server_create(pdds_engine *e, const pdds_server_config *cfg, pdds_server **out) {
    if (!e || !cfg || !out) return PDDS_ERR_INVALID;
    *out = nullptr;
    pdds_server_config c = *cfg;
    if (!(c.domain[2] > c.domain[0]) || !(c.domain[3] > c.domain[1])) return fail(e, PDDS_ERR_INVALID, "bad domain");
    if (c.hbm_budget <= 0) c.hbm_budget = (int64_t)16 << 30;
    if (c.max_inflight <= 0) c.max_inflight = 256;
    if (c.prep_threads <= 0) c.prep_threads = 2;
    if (c.round_paths <= 0) c.round_paths = (int64_t)1 << 25;
    std::unique_ptr<pdds_server> s(new pdds_server());
    s->e = e;
    s->cfg = c;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        pdds_status st = check_run(e, &c.run);
        if (!st) st = make_params(e, &c.problem, &c.run, s->pp);
        if (st) return st;
        // a round carries every request's own stream and k2 (per-lane kernels), which the
        // reflecting-side kernels do not take
        if (s->pp.refl) return fail(e, PDDS_ERR_UNSUPPORTED, "serving with reflecting sides");
    }
    pdds_status st = pdds_engine_create(e->device, &s->se);
    if (st) return fail(e, st, "server: solver engine");
    pdds_server *sp = s.get();
    for (int i = 0; i < c.prep_threads; ++i) sp->preps.emplace_back(prep_loop, sp);
    sp->dev = std::thread(dev_loop, sp);
    sp->solver = std::thread(solver_loop, sp);
    *out = s.release();
    return PDDS_OK;
}

void pdds_server_destroy(pdds_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_prep.notify_all();
    for (std::thread &t : s->preps) t.join();
    s->cv_work.notify_all();
    s->dev.join();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->dev_done = true;
    }
    s->cv_solve.notify_all();
    s->solver.join();
    {
        std::lock_guard<std::mutex> lk(s->e->mu);
        s->dom.release();
        for (auto &kv : s->reqs) kv.second->tab.release();
    }
    pdds_engine_destroy(s->se);
    delete s;
}

pdds_status pdds_server_submit(pdds_server *s, const pdds_request *rq, uint64_t *ticket) {
    if (!s || !rq || !ticket) return PDDS_ERR_INVALID;
    if (rq->kind != PDDS_REQ_POINT && rq->kind != PDDS_REQ_SOLVE) return PDDS_ERR_INVALID;
    if (rq->n_paths <= 0 || !(rq->h >= 0.0) || !std::isfinite(rq->h)) return PDDS_ERR_INVALID;
    const double *d = s->cfg.domain;
    std::unique_ptr<SrvReq> r(new SrvReq());
    r->rq = *rq;
    r->res.kind = rq->kind;
    r->t_submit = now_ms();
    if (rq->kind == PDDS_REQ_SOLVE) {
        if (rq->n_sub < 2 || rq->n_node < 3) return PDDS_ERR_INVALID;
    } else {
        if (!(rq->x >= d[0] && rq->x <= d[2] && rq->y >= d[1] && rq->y <= d[3])) return PDDS_ERR_INVALID;
        r->footprint = std::min<int64_t>(rq->n_paths, s->cfg.round_paths) * kRecordBytes +
                       items_of(s->cfg.run, rq->n_paths) * kKnotBytes;
        if (r->footprint > s->cfg.hbm_budget) return PDDS_ERR_FULL;
    }
    SrvReq *rp = r.get();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return PDDS_ERR_INVALID;
    rp->ticket = s->next_ticket++;
    *ticket = rp->ticket;
    s->reqs[rp->ticket] = std::move(r);
    ++s->n_submitted;
    if (rq->kind == PDDS_REQ_POINT) {
        double ex, ey, nx, ny;
        int kd;
        const double dist = rect_dist<double>(d, rq->x, rq->y, ex, ey, nx, ny, kd);
        if (std::fabs(dist) <= 1e-8 || dist > 0.0) {
            // on the boundary: the Dirichlet value, no Monte Carlo
            rp->res.u = host_g(&s->cfg.problem, ex, ey, s->pp);
            rp->res.u_var = 0.0;
            finish(s, rp, PDDS_OK);
            return PDDS_OK;
        }
        rp->state = RS_READY;
        s->ready.push_back(rp);
        s->cv_work.notify_all();
    } else {
        s->prep_q.push_back(rp);
        s->cv_prep.notify_one();
    }
    return PDDS_OK;
}

pdds_status pdds_server_poll(pdds_server *s, uint64_t ticket, pdds_result *res, double *u, int64_t cap) {
    if (!s) return PDDS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return PDDS_ERR_INVALID;
    if (it->second->state != RS_DONE) return PDDS_ERR_NOT_READY;
    return take_result(s, it, res, u, cap);
}

pdds_status pdds_server_wait(pdds_server *s, uint64_t ticket, pdds_result *res, double *u, int64_t cap) {
    if (!s) return PDDS_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    // the ticket is looked up afresh on every wake-up: another thread may have taken the result
    // (poll or wait on the same ticket) and released the request meanwhile
    s->cv_done.wait(lk, [&] {
        auto f = s->reqs.find(ticket);
        return f == s->reqs.end() || f->second->state == RS_DONE;
    });
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return PDDS_ERR_INVALID;
    return take_result(s, it, res, u, cap);
}

pdds_status pdds_server_stats(pdds_server *s, int64_t *submitted, int64_t *done, int64_t *rounds, int64_t *path_steps,
                              double *busy_ms) {
    if (!s) return PDDS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (submitted) *submitted = s->n_submitted;
    if (done) *done = s->n_done;
    if (rounds) *rounds = s->n_rounds;
    if (path_steps) *path_steps = s->n_steps;
    if (busy_ms) *busy_ms = s->busy_ms;
    return PDDS_OK;
}

}  // extern "C"
