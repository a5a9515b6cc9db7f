// wql_serve.hip -- continuous-batching loop over independent rupture
// scenarios (part of libwql; included by wql_api.hip after the C ABI).
//
// A scenario is a full two-block problem with its own context: its state is
// on the device from admission until it is retired.  Admission is first-fit
// by device footprint (wql_estimate_bytes) against the server's HBM budget
// and a cap on concurrently active scenarios.  Every active scenario runs on
// its own stream with its LSRK step captured once into a HIP graph; the
// server advances each one in chunks of steps and tracks chunk completion
// with an event, so scenarios that arrive while others run are admitted as
// soon as memory and a slot free up, and their work overlaps on the device.
//
// Retired contexts are not freed: they wait in a pool keyed by everything
// their allocation and set-up depend on (block shapes, boundary and PML
// descriptors, a hash of the coordinates and material, precision,
// plasticity, spatial operator).  A scenario with the same key takes a pooled context, which
// is cleared on its stream and keeps its metrics, face frames and material,
// so neither admission nor retirement allocates, frees or recomputes
// geometry, and the poll loop never waits for the device except to download
// a retired scenario's fault arrays.  Pooled contexts count against the
// budget; the least recently used ones are freed when a scenario with a new
// key needs the room.  Queued scenarios are taken in arrival order (policy
// 0) or shortest job first with aging (policy 1: the key is nodes x steps /
// (1 + seconds waited), so a large scenario overtakes fresh small ones
// after waiting as many seconds as its work is a multiple of theirs).

#include <algorithm>
#include <chrono>
#include <deque>
#include <memory>
#include <unordered_map>

struct wql_scenario {
    int id = 0;
    int state = 0;  // 0 queued, 1 active, 2 finished (result ready), 3 failed
    wql_problem prob;
    wql_block_desc blocks[2];
    std::vector<double> X[2], M[2];
    int nsteps = 0, steps_done = 0, steps_inflight = 0;
    size_t need = 0;
    uint64_t key = 0;       // pool key: what the context's allocation and set-up depend on
    double work = 0.0;      // nodes x steps
    double t_submit = 0.0;  // host clock at submission (s)
    wql_ctx *ctx = nullptr;
    hipEvent_t ev = nullptr;
    std::vector<double> S, trup, psr;
    double stats[6] = {0, 0, 0, 0, 0, 0};
    int err = 0;
    std::string msg;
};

struct wql_pooled {
    wql_ctx *ctx;
    uint64_t key;
    size_t bytes;
    long last_use;
};

struct wql_server {
    int max_active = 8;
    size_t budget = 0;
    int chunk = 50;
    int next_id = 1;
    int policy = 0;  // 0 arrival order, 1 shortest job first with aging
    size_t used = 0;  // active and pooled contexts
    long clock = 0;
    long created = 0, reused = 0, freed = 0;
    std::deque<int> queue;
    std::vector<int> active;
    std::deque<int> unreported;  // retired (finished or failed) ids not yet returned by poll
    std::vector<wql_pooled> pool;
    std::unordered_map<int, std::unique_ptr<wql_scenario>> all;
    // tuned launch shapes of the main volume kernel by block size, precision
    // and grid kind (wql_server_set_shape), applied at admission
    struct shape { int n[3], prec, cart, variant, zchunk, order, group; };
    std::vector<shape> shapes;
};

namespace {

double host_seconds()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// FNV-1a over 64-bit words
uint64_t fnv(uint64_t h, const void *p, size_t bytes)
{
    const unsigned char *b = (const unsigned char *)p;
    size_t n = bytes / 8;
    for (size_t k = 0; k < n; k++) {
        uint64_t w;
        memcpy(&w, b + 8 * k, 8);
        h = (h ^ w) * 1099511628211ull;
    }
    for (size_t k = 8 * n; k < bytes; k++) h = (h ^ b[k]) * 1099511628211ull;
    return h;
}

uint64_t scenario_key(const wql_scenario *sc)
{
    uint64_t h = 1469598103934665603ull;
    // what a pooled context keeps from its create: precision, plasticity,
    // operator, the keyed fault's per-node tables (recomputed by the reset,
    // but allocated only for a keyed problem)
    const int flags[6] = {sc->prob.precision, sc->prob.plastic, sc->prob.s0_kind, sc->prob.fd_type,
                          sc->prob.fd_type ? sc->prob.fd_order : 0, sc->prob.fault.key};
    h = fnv(h, flags, sizeof flags);
    const double an[4] = {(double)sc->prob.anelastic, sc->prob.an_c, sc->prob.an_weight_exp, sc->prob.an_fref};
    if (sc->prob.anelastic) h = fnv(h, an, sizeof an);
    for (int ib = 0; ib < 2; ib++) {
        const wql_block_desc &b = sc->blocks[ib];
        const int shape[17] = {b.nq, b.nr, b.ns, b.bc[0], b.bc[1], b.bc[2], b.bc[3], b.bc[4], b.bc[5],
                               b.pml[0], b.pml[1], b.pml[2], b.pml[3], b.pml[4], b.pml[5], b.npml, b.cartesian};
        h = fnv(h, shape, sizeof shape);
        h = fnv(h, b.rho_s_p, sizeof b.rho_s_p);
        h = fnv(h, sc->X[ib].data(), sc->X[ib].size() * sizeof(double));
        const int het = sc->M[ib].empty() ? 0 : 1;
        h = fnv(h, &het, sizeof het);
        if (het) h = fnv(h, sc->M[ib].data(), sc->M[ib].size() * sizeof(double));
    }
    return h;
}

// a retired scenario's context goes to the pool (it stays counted in used)
void scenario_release(wql_server *s, wql_scenario *sc, bool keep)
{
    if (sc->ev) {
        hipEventDestroy(sc->ev);
        sc->ev = nullptr;
    }
    if (sc->ctx) {
        if (keep) {
            s->pool.push_back({sc->ctx, sc->key, sc->need, s->clock++});
        } else {
            wql_destroy(sc->ctx);
            s->used -= sc->need;
            s->freed++;
        }
        sc->ctx = nullptr;
    }
    for (int ib = 0; ib < 2; ib++) {
        std::vector<double>().swap(sc->X[ib]);
        std::vector<double>().swap(sc->M[ib]);
    }
}

// free pooled contexts, least recently used first, until `need` more
// bytes fit the budget; false if they cannot
bool pool_make_room(wql_server *s, size_t need)
{
    while (s->used + need > s->budget && !s->pool.empty()) {
        auto it = std::min_element(s->pool.begin(), s->pool.end(),
                                   [](const wql_pooled &a, const wql_pooled &b) { return a.last_use < b.last_use; });
        wql_destroy(it->ctx);
        s->used -= it->bytes;
        s->pool.erase(it);
        s->freed++;
    }
    return s->used + need <= s->budget;
}

int scenario_fail(wql_server *s, wql_scenario *sc, int rc)
{
    sc->err = rc;
    sc->msg = wql_last_error();
    sc->state = 3;
    scenario_release(s, sc, false);  // a failed context is not reused
    return rc;
}

// enqueue the next chunk of steps and the event that marks its end
int scenario_launch(wql_scenario *sc, int chunk)
{
    int n = std::min(chunk, sc->nsteps - sc->steps_done);
    if (int rc = wql_step(sc->ctx, n)) return rc;
    if (hipEventRecord(sc->ev, sc->ctx->stream) != hipSuccess) return fail(WQL_E_HIP, "hipEventRecord");
    sc->steps_inflight = n;
    return 0;
}

int scenario_admit(wql_server *s, wql_scenario *sc)
{
    wql_ctx *ctx = nullptr;
    auto hit = std::find_if(s->pool.begin(), s->pool.end(), [&](const wql_pooled &p) { return p.key == sc->key; });
    if (hit != s->pool.end()) {
        ctx = hit->ctx;
        sc->need = hit->bytes;
        s->pool.erase(hit);
        s->reused++;
        sc->ctx = ctx;
        if (int rc = ctx_reset(ctx, &sc->prob)) return rc;
    } else {
        if (int rc = wql_create(&sc->prob, sc->blocks, &ctx)) return rc;
        // one stream per scenario: the scenarios in flight already overlap
        ctx->overlap = 0;
        sc->ctx = ctx;
        s->used += sc->need;
        s->created++;
    }
    for (int ib = 0; ib < 2; ib++) {  // host copies are no longer needed
        std::vector<double>().swap(sc->X[ib]);
        std::vector<double>().swap(sc->M[ib]);
    }
    for (const auto &e : s->shapes) {
        const wql_block_desc &b = sc->blocks[0];
        if (e.n[0] == b.nq && e.n[1] == b.nr && e.n[2] == b.ns && e.prec == sc->prob.precision &&
            e.cart == (b.cartesian ? 1 : 0)) {
            if (int rc = wql_set_tile(ctx, e.variant, e.zchunk)) return rc;
            if (int rc = wql_set_tile_order(ctx, e.order, e.group)) return rc;
            break;
        }
    }
    if (int rc = wql_init_state(ctx)) return rc;
    if (int rc = wql_use_graph(ctx, 1)) return rc;
    if (hipEventCreateWithFlags(&sc->ev, hipEventDisableTiming) != hipSuccess) return fail(WQL_E_HIP, "hipEventCreate");
    sc->state = 1;
    return scenario_launch(sc, s->chunk);
}

int scenario_finish(wql_server *s, wql_scenario *sc)
{
    if (int rc = wql_sync(sc->ctx)) return rc;
    const long FP = (long)sc->blocks[0].nr * sc->blocks[0].ns;
    sc->S.resize(4 * FP);
    sc->trup.resize(FP);
    sc->psr.resize(FP);
    if (int rc = wql_get_fault(sc->ctx, 0, sc->S.data(), nullptr, sc->trup.data(), nullptr, nullptr, nullptr))
        return rc;
    if (int rc = wql_get_fault_peak(sc->ctx, 0, sc->psr.data())) return rc;
    if (int rc = wql_fault_stats(sc->ctx, 0, sc->stats)) return rc;
    sc->state = 2;
    scenario_release(s, sc, true);
    return 0;
}

}  // namespace

extern "C" {

int wql_server_create(int max_active, size_t hbm_budget, int steps_per_launch, wql_server **out)
{
    if (!out || max_active < 1 || steps_per_launch < 1) return fail(WQL_E_ARG, "bad server argument");
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return fail(WQL_E_HIP, "hipMemGetInfo");
    wql_server *s = new wql_server();
    s->max_active = max_active;
    // default budget: 90% of what is free now
    s->budget = hbm_budget ? std::min(hbm_budget, free_b) : (size_t)(0.9 * (double)free_b);
    s->chunk = steps_per_launch;
    *out = s;
    return 0;
}

int wql_server_set_policy(wql_server *s, int policy)
{
    if (!s || policy < 0 || policy > 1) return fail(WQL_E_ARG, "bad server policy");
    s->policy = policy;
    return 0;
}

int wql_server_set_shape(wql_server *s, const int n[3], int precision, int cartesian, int variant, int zchunk,
                         int order, int group)
{
    if (!s || !n) return fail(WQL_E_ARG, "null argument");
    if (variant < 0 || variant >= wql_tile_variants() || zchunk < 1 || order < 0 || order > 3 || group < 1)
        return fail(WQL_E_ARG, "bad launch shape");
    wql_server::shape e = {{n[0], n[1], n[2]}, precision, cartesian ? 1 : 0, variant, zchunk, order, group};
    for (auto &o : s->shapes)
        if (o.n[0] == n[0] && o.n[1] == n[1] && o.n[2] == n[2] && o.prec == precision && o.cart == e.cart) {
            o = e;
            return 0;
        }
    s->shapes.push_back(e);
    return 0;
}

int wql_server_destroy(wql_server *s)
{
    if (!s) return 0;
    for (auto &kv : s->all) scenario_release(s, kv.second.get(), false);
    for (auto &p : s->pool) wql_destroy(p.ctx);
    delete s;
    return 0;
}

int wql_server_submit(wql_server *s, const wql_problem *prob, const wql_block_desc blocks[2], int nsteps, int *id)
{
    if (!s || !prob || !blocks || !id || nsteps < 1) return fail(WQL_E_ARG, "bad submit argument");
    auto sc = std::make_unique<wql_scenario>();
    sc->id = s->next_id++;
    sc->prob = *prob;
    sc->nsteps = nsteps;
    for (int ib = 0; ib < 2; ib++) {
        sc->blocks[ib] = blocks[ib];
        if (int rc = valid_block(blocks[ib])) return rc;
        const size_t nd = (size_t)blocks[ib].nq * blocks[ib].nr * blocks[ib].ns;
        if (!blocks[ib].X) return fail(WQL_E_ARG, "block coordinates are required");
        sc->X[ib].assign(blocks[ib].X, blocks[ib].X + 3 * nd);
        sc->blocks[ib].X = sc->X[ib].data();
        if (blocks[ib].M) {
            sc->M[ib].assign(blocks[ib].M, blocks[ib].M + 3 * nd);
            sc->blocks[ib].M = sc->M[ib].data();
        }
    }
    sc->need = wql_estimate_bytes(prob, sc->blocks);
    sc->key = scenario_key(sc.get());
    sc->work = (double)nsteps * ((double)blocks[0].nq * blocks[0].nr * blocks[0].ns +
                                 (double)blocks[1].nq * blocks[1].nr * blocks[1].ns);
    sc->t_submit = host_seconds();
    if (sc->need > s->budget) return fail(WQL_E_NOMEM, "scenario needs %zu bytes, budget %zu", sc->need, s->budget);
    *id = sc->id;
    s->queue.push_back(sc->id);
    s->all[sc->id] = std::move(sc);
    return 0;
}

int wql_server_poll(wql_server *s, int *done, int max_done)
{
    if (!s || (max_done > 0 && !done)) return fail(WQL_E_ARG, "bad poll argument");
    int ndone = 0;
    // progress of the active scenarios
    std::vector<int> still;
    for (int id : s->active) {
        wql_scenario *sc = s->all[id].get();
        hipError_t q = hipEventQuery(sc->ev);
        if (q == hipErrorNotReady) {
            still.push_back(id);
            continue;
        }
        if (q != hipSuccess) {
            fail(WQL_E_HIP, "scenario %d: %s", id, hipGetErrorString(q));
            scenario_fail(s, sc, WQL_E_HIP);
        } else {
            sc->steps_done += sc->steps_inflight;
            sc->steps_inflight = 0;
            int rc = sc->steps_done >= sc->nsteps ? scenario_finish(s, sc) : scenario_launch(sc, s->chunk);
            if (rc) scenario_fail(s, sc, rc);
        }
        if (sc->state == 1) still.push_back(id);
        else s->unreported.push_back(id);
    }
    s->active.swap(still);
    // admission: first fit by footprint, in arrival order or by the aged
    // job-size key (stable, so equal keys keep arrival order)
    if (s->policy == 1 && s->queue.size() > 1 && (int)s->active.size() < s->max_active) {
        const double now = host_seconds();
        std::vector<std::pair<double, int>> key;
        key.reserve(s->queue.size());
        for (int id : s->queue) {
            const wql_scenario *q = s->all[id].get();
            key.emplace_back(q->work / (1.0 + std::max(0.0, now - q->t_submit)), id);
        }
        std::stable_sort(key.begin(), key.end(),
                         [](const std::pair<double, int> &x, const std::pair<double, int> &y) { return x.first < y.first; });
        s->queue.clear();
        for (const auto &k : key) s->queue.push_back(k.second);
    }
    for (auto it = s->queue.begin(); it != s->queue.end() && (int)s->active.size() < s->max_active;) {
        wql_scenario *sc = s->all[*it].get();
        const bool pooled = std::any_of(s->pool.begin(), s->pool.end(),
                                        [&](const wql_pooled &p) { return p.key == sc->key; });
        if (!pooled && !pool_make_room(s, sc->need)) {
            ++it;
            continue;
        }
        int rc = scenario_admit(s, sc);
        if (rc) {
            scenario_fail(s, sc, rc);
            s->unreported.push_back(sc->id);
        } else {
            s->active.push_back(sc->id);
        }
        it = s->queue.erase(it);
    }
    // retirements are reported in order, at most max_done per call; the
    // rest wait for the next call
    while (ndone < max_done && !s->unreported.empty()) {
        done[ndone++] = s->unreported.front();
        s->unreported.pop_front();
    }
    return ndone;
}

int wql_server_result(wql_server *s, int id, double *S, double *trup, double *psr, double stats[6])
{
    if (!s) return fail(WQL_E_ARG, "null server");
    auto f = s->all.find(id);
    if (f == s->all.end()) return fail(WQL_E_ARG, "unknown scenario %d", id);
    wql_scenario *sc = f->second.get();
    if (sc->state == 3) {
        int rc = sc->err;
        std::string msg = sc->msg;
        s->all.erase(f);
        return fail(rc, "scenario %d failed: %s", id, msg.c_str());
    }
    if (sc->state != 2) return fail(WQL_E_STATE, "scenario %d is not finished", id);
    if (S) memcpy(S, sc->S.data(), sc->S.size() * sizeof(double));
    if (trup) memcpy(trup, sc->trup.data(), sc->trup.size() * sizeof(double));
    if (psr) memcpy(psr, sc->psr.data(), sc->psr.size() * sizeof(double));
    if (stats) memcpy(stats, sc->stats, sizeof sc->stats);
    s->all.erase(f);
    return 0;
}

int wql_server_pool_stats(const wql_server *s, long out[4])
{
    if (!s || !out) return fail(WQL_E_ARG, "bad argument");
    out[0] = s->created;
    out[1] = s->reused;
    out[2] = s->freed;
    out[3] = (long)s->pool.size();
    return 0;
}

int wql_server_status(const wql_server *s, long out[4])
{
    if (!s || !out) return fail(WQL_E_ARG, "bad argument");
    out[0] = (long)s->queue.size();
    out[1] = (long)s->active.size();
    out[2] = (long)s->used;
    out[3] = (long)s->budget;
    return 0;
}

}  // extern "C"
