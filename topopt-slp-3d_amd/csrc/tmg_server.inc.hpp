// tmg_server.inc.hpp -- continuous-batching server for streams of equilibrium
// solves of uneven sizes (included at the end of tmg_solver.hip: it drives the
// solver through its internal begin / launch / poll steps).
//
// Requests are admitted FIFO while a slot is free and the estimated device
// memory of the request fits the budget. Each slot owns a HIP stream, a
// pinned scalar block, a pinned staging buffer for the request's host vectors,
// a rocBLAS handle and a stream-ordered memory pool that never hands memory to
// another stream, so admitting and retiring requests never synchronises the
// device and no request can be given memory another request's kernels still
// use. Each slot is served by its own worker thread: it builds the admitted
// request's hierarchy (Galerkin refresh, coarse factor, graph capture), checks
// that the device holds the request's densities, then drives the solve with
// graph launches of growing chunks of PCG iterations, waiting on its own
// stream only. The workers' host-side launch work runs in parallel and the
// kernels of different requests overlap on the device. Results are bitwise
// identical to a solve run alone: every kernel is deterministic and a
// request's launches depend only on its own problem.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace {

// headroom of the admission estimate for what the handle does not allocate
// itself (rocSOLVER's workspace): not part of the handle's arena
constexpr size_t SRV_ARENA_SLACK = (size_t)32 << 20;

struct SrvReq {
    int64_t id = 0;
    tmg_problem prob{};
    std::vector<int64_t> freedofs;
    tmg_mg_params mg{};
    int prec = TMG_MIXED;
    std::vector<double> x, f, x0;
    size_t est = 0;
    int slot = -1;
    int state = 0;   // 0 queued, 1 setting up, 2 running, 3 done, < 0 failed (error code)
    tmg_solver *h = nullptr;
    std::vector<double> u;
    int it = 0, rc = 0;
    double relres = 0.0;
    char *arena = nullptr;    // device block of the slot's pool holding the handle's buffers
    size_t arena_bytes = 0;
    double t_arrival = 0, t_admit = 0, t_start = 0, t_done = 0;
    std::string err;
};

// Device bytes a handle allocates (tmg_setup + tmg_update_density + one
// host-vector upload), rounded like dalloc.
size_t estimate_bytes(int nelx, int nely, int nelz, int L, int prec, int64_t nfree) {
    size_t b = 0;
    auto add = [&](size_t n) { b += round256(std::max<size_t>(n, 1)); };
    int64_t ne1 = 1, ne2 = 1, ncoarse = 0;
    int ex = nelx, ey = nely, ez = nelz;
    int64_t ne0 = (int64_t)nelx * nely * nelz, ndof0 = 0, nn0x = nelx + 1, nn0y = nely + 1, nn0z = nelz + 1;
    for (int l = 0; l < L; ++l) {
        const int64_t nn = (int64_t)(ex + 1) * (ey + 1) * (ez + 1), ndof = 3 * nn;
        if (l == 0) ndof0 = ndof;
        add(ndof);                        // mask
        add(8 * ndof);                    // d_free (upper bound: all free)
        add(8 * ndof);                    // d_cidx
        add(8 * ndof); add(8 * ndof); add(8 * ndof);   // V-cycle work vectors (FP64-sized)
        add(8 * ndof);                    // inv64
        if (prec == TMG_MIXED) add(4 * ndof);
        if (l > 0) {
            const size_t ns = (size_t)((nn + 63) / 64) * 64 * NCOEF;
            add(8 * ns);
            if (prec == TMG_MIXED) add(4 * ns);
        }
        if (l == 1) ne1 = (int64_t)ex * ey * ez;
        if (l == 2) ne2 = (int64_t)ex * ey * ez;
        if (l == L - 1) ncoarse = ndof;
        ex /= 2; ey /= 2; ez /= 2;
    }
    add(576 * 8); add(576 * 4);
    add(WHT_MR_NCOEF * 8); add(WHT_MR_NCOEF * 4); add(WHT_MR_NCOEF * 8); add(WHT_MR_NCOEF * 4);   // spectral tables
    add(8 * ne0); add(4 * ne0); add(8 * ne0);
    add(8 * ne0); add(4 * ne0);                                // level-1 children's moduli (upper bound)
    add(8 * ne1 * 576); add(8 * (L > 2 ? ne2 * 576 : 1));
    add(8 * ncoarse * ncoarse); add(8 * ncoarse); add(8 * ncoarse); add(4);
    if (prec == TMG_MIXED) add(4 * ncoarse * ncoarse);         // FP32 copy of the coarse inverse
    // symmetric coarse solve (coarse systems of 16,384 DOFs and more): the
    // run table and the partial products (upper bound: every coarse DOF free,
    // runs of one tile)
    if (ncoarse >= 16384) {
        const int64_t nb = (ncoarse + CS_B - 1) / CS_B;
        const int64_t nruns = nb * (nb + 1) / 2;
        add(sizeof(int4) * (size_t)nruns);
        add(sizeof(int) * (size_t)(nb + 1));
        add(8 * (size_t)((nb * (nb - 1) / 2 + nruns) * CS_B));
    }
    for (int k = 0; k < 6; ++k) add(8 * ndof0);
    // per-block partial sums: z chunks of at least one plane, of either the
    // 256-node blocks of the dense apply or the overlapping tiles of the
    // spectral apply (the larger count)
    int tyl = 4, ntx = 1, nty = 1;
    wht_tiles((int)nn0x, (int)nn0y, tyl, ntx, nty);
    const int64_t pb = std::max<int64_t>((nn0x * nn0y + TPB - 1) / TPB, (int64_t)ntx * nty);
    const int64_t nblocks = std::max<int64_t>(pb * nn0z, 4096);
    add(8 * nblocks); add(8 * nblocks); add(sizeof(PcgScalars));
    add(8 * nfree);                                            // upload staging
    return b + SRV_ARENA_SLACK;                                // rocSOLVER workspace headroom
}

}  // namespace

struct tmg_server {
    size_t budget = 0, used = 0, peak = 0;
    int nslots = 0;
    int peak_running = 0;          // most requests solving at the same time
    int rebuilds = 0;              // operators rebuilt after a failed first coarse factorization
    // per slot: stream, pinned scalar block, pinned staging for the
    // request's host vectors, and its own stream-ordered memory pool (device
    // memory is reused only within a slot, i.e. only on the slot's stream)
    std::vector<hipStream_t> streams;
    std::vector<PcgScalars *> hsc;
    std::vector<double *> stage;
    std::vector<size_t> stage_cap;
    std::vector<hipMemPool_t> pools;
    std::vector<char> slot_busy;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<SrvReq *> queue;
    std::vector<SrvReq *> running;
    std::unordered_map<int64_t, SrvReq *> all;
    int64_t next_id = 1;
    bool stop = false;
    bool trace = false;            // TMG_SERVER_TRACE=1: per-request set-up phase times on stderr
    // one worker thread per slot (set-up, solve and retirement of the slot's
    // requests); each owns a rocBLAS handle, whose stream binding is per handle
    std::vector<std::thread> workers;
    std::vector<rocblas_handle> rbs;
    std::chrono::steady_clock::time_point t0;
    std::string err;
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

void srv_fail(tmg_server *S, SrvReq *r, int rc, const std::string &msg) {
    // called with S->mu held
    r->state = rc < 0 ? rc : TMG_ERR_HIP;
    r->err = msg;
    r->t_done = S->now();
}

void srv_release(tmg_server *S, SrvReq *r) {
    // called with S->mu held; the handle is already destroyed
    if (r->slot >= 0) S->slot_busy[r->slot] = 0;
    r->slot = -1;
    S->used -= r->est;
}

// The densities the handle holds in device memory are the ones the request
// brought (compared through an order-independent checksum of their bits).
int srv_verify_upload(tmg_solver *h, const double *d, const double *hv, int64_t n, const char *what) {
    uint64_t dev = 0;
    TRY(device_checksum(h, d, n, &dev));
    const uint64_t host = host_checksum(hv, n);
    if (dev != host) {
        h->err = std::string("the ") + what + " in device memory differ from the request's (checksum " +
                 std::to_string(dev) + " vs " + std::to_string(host) + ")";
        return TMG_ERR_HIP;
    }
    return 0;
}

// A failed coarse factorisation is examined for what tells a corrupted
// set-up from a singular operator: whether the device still holds the
// request's densities, and whether building the same hierarchy again on the
// slot's stream, with nothing else in flight on it, factorises. If the
// densities are intact and the rebuild factorises, the request goes on with
// the rebuilt hierarchy (every level and the factor are recomputed from the
// densities), the event is printed and counted (tmg_server_stats); otherwise
// the request fails with both messages. Returns the request's status.
int srv_diagnose_not_spd(tmg_server *S, tmg_solver *h, SrvReq *r) {
    std::string first = h->err;
    uint64_t dev = 0;
    const bool readable = device_checksum(h, h->xdens, (int64_t)r->x.size(), &dev) == 0;
    const bool intact = readable && dev == host_checksum(r->x.data(), (int64_t)r->x.size());
    const std::string dens = readable ? (intact ? "intact" : "corrupted") : "unreadable";
    const int rc2 = intact ? refresh(h) : TMG_ERR_NOT_SPD;
    h->err = first + "; densities on the device " + dens + "; rebuilding the same hierarchy " +
             (rc2 == 0 ? "factorised (the first build was corrupted)" : "failed again: " + h->err);
    if (rc2 != 0) return TMG_ERR_NOT_SPD;
    fprintf(stderr, "tmg_server: request %lld: %s\n", (long long)r->id, h->err.c_str());
    std::lock_guard<std::mutex> lk(S->mu);
    ++S->rebuilds;
    return 0;
}

// The request's arena goes back to its slot's pool, ordered after the
// request's work on the slot's stream.
void srv_free_arena(tmg_server *S, SrvReq *r) {
    if (r->arena) (void)hipFreeAsync(r->arena, S->streams[r->slot]);
    r->arena = nullptr;
    r->arena_bytes = 0;
}

// Set-up of an admitted request on its slot: hierarchy, densities check,
// host vectors, PCG start. Returns the handle (nullptr and *rc < 0 on failure).
tmg_solver *srv_setup(tmg_server *S, SrvReq *r, rocblas_handle rb, int *rc, std::string *msg) {
    const int slot = r->slot;
    *rc = 0;
    // pinned staging large enough for the densities and the free vectors
    const size_t need = std::max(r->x.size(), r->f.size());
    if (S->stage_cap[slot] < need) {
        if (S->stage[slot]) (void)hipHostFree(S->stage[slot]);
        S->stage[slot] = nullptr;
        S->stage_cap[slot] = 0;
        const size_t cap = need + need / 4;
        if (hipHostMalloc((void **)&S->stage[slot], cap * sizeof(double), hipHostMallocDefault) != hipSuccess) {
            S->stage[slot] = nullptr;
            *rc = TMG_ERR_OOM;
            *msg = "pinned staging allocation failed";
            return nullptr;
        }
        S->stage_cap[slot] = cap;
    }
    // one block of the slot's pool for all of the handle's buffers (the
    // estimate bounds them; the handle falls back to the pool beyond it)
    r->arena_bytes = r->est > SRV_ARENA_SLACK ? r->est - SRV_ARENA_SLACK : r->est;
    if (hipMallocFromPoolAsync((void **)&r->arena, r->arena_bytes, S->pools[slot], S->streams[slot]) != hipSuccess) {
        r->arena = nullptr;
        r->arena_bytes = 0;
    }
    ExtRes ext;
    ext.st = S->streams[slot];
    ext.rb = rb;
    ext.h_sc = S->hsc[slot];
    ext.pool = S->pools[slot];
    ext.arena = r->arena;
    ext.arena_cap = r->arena_bytes;
    ext.stage = S->stage[slot];
    ext.stage_cap = S->stage_cap[slot];
    r->prob.freedofs = r->freedofs.data();
    tmg_solver *h = nullptr;
    const double tp0 = S->now();
    *rc = setup_impl(&r->prob, &r->mg, r->prec, &ext, &h);
    const double tp1 = S->now();
    if (*rc) {
        *msg = g_setup_err;
        srv_free_arena(S, r);
        return nullptr;
    }
    *rc = tmg_update_density(h, r->x.data());   // each worker's own rocBLAS handle and stream
    const double tp2 = S->now();
    if (!*rc) *rc = srv_verify_upload(h, h->xdens, r->x.data(), (int64_t)r->x.size(), "densities");
    if (*rc == TMG_ERR_NOT_SPD) *rc = srv_diagnose_not_spd(S, h, r);
    if (!*rc && (*rc = upload_free(h, 0, r->x0.data(), h->px)) == 0) *rc = upload_free(h, 0, r->f.data(), h->pf);
    h->have_rhs = true;
    const double tp3 = S->now();
    if (!*rc) *rc = solve_begin(h);
    if (S->trace)
        fprintf(stderr, "tmg_server: request %lld %dx%dx%d slot %d: structure %.2f ms, operator %.2f ms, vectors %.2f ms, "
                "PCG start %.2f ms\n", (long long)r->id, r->prob.nelx, r->prob.nely, r->prob.nelz, slot,
                1e3 * (tp1 - tp0), 1e3 * (tp2 - tp1), 1e3 * (tp3 - tp2), 1e3 * (S->now() - tp3));
    if (*rc) {
        *msg = h->err;
        destroy_all(h);
        delete h;
        srv_free_arena(S, r);
        return nullptr;
    }
    return h;
}

// The PCG of a running request in graph launches of 1, 2, 4, 8, 8, ... chunks
// (a chunk is the captured run of iterations; iterations past convergence
// return at once on the device), with the mixed-precision stall check between
// launches; then the deferred last update and the download of u.
int srv_solve(SrvReq *r, tmg_solver *h) {
    int chunk = 1;
    for (;;) {
        TRY(solve_launch(h, chunk));
        HIPX(hipEventSynchronize(h->ev));
        if (h->h_sc->done) break;
        TRY(solve_check(h));
        chunk = std::min(chunk * 2, 8);
    }
    int rc = solve_result(h, &r->it, &r->relres);
    if (rc >= 0) rc = std::min(rc, solve_finish(h));
    if (rc >= 0) {
        r->u.assign(h->lev[0].nfree, 0.0);
        const int d = download_free(h, 0, h->px, r->u.data());
        if (d < 0) rc = d;
    }
    return rc;
}

// Worker of one slot: admits the queue's head when the budget allows (FIFO),
// sets it up, solves it and retires it.
void srv_worker(tmg_server *S, int slot) {
    (void)hipSetDevice(0);
    rocblas_handle rb = S->rbs[slot];
    for (;;) {
        SrvReq *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv.wait(lk, [&] {
                if (S->stop) return true;
                return !S->queue.empty() && S->used + S->queue.front()->est <= S->budget;
            });
            if (S->stop) return;
            r = S->queue.front();
            S->queue.pop_front();
            r->slot = slot;
            S->slot_busy[slot] = 1;
            S->used += r->est;
            S->peak = std::max(S->peak, S->used);
            r->state = 1;
            r->t_admit = S->now();
        }
        int rc = 0;
        std::string msg;
        tmg_solver *h = srv_setup(S, r, rb, &rc, &msg);
        if (h) {
            std::lock_guard<std::mutex> lk(S->mu);
            r->h = h;
            r->state = 2;
            r->t_start = S->now();
            S->running.push_back(r);
            S->peak_running = std::max(S->peak_running, (int)S->running.size());
        }
        if (h) {
            rc = srv_solve(r, h);
            r->rc = rc;
            if (rc < 0) msg = h->err;
            destroy_all(h);
            delete h;
            srv_free_arena(S, r);
        }
        std::lock_guard<std::mutex> lk(S->mu);
        if (h) {
            r->h = nullptr;
            S->running.erase(std::find(S->running.begin(), S->running.end(), r));
        }
        srv_release(S, r);
        if (rc < 0) srv_fail(S, r, rc, msg);
        else { r->state = 3; r->t_done = S->now(); }
        S->cv.notify_all();
    }
}

}  // namespace

extern "C" {

int tmg_server_create(double budget_bytes, int max_active, tmg_server **out) {
    g_setup_err.clear();
    if (!out || !(budget_bytes > 0) || max_active < 1 || max_active > 64) {
        g_setup_err = "invalid server arguments";
        return TMG_ERR_ARG;
    }
    *out = nullptr;
    tmg_server *S = new tmg_server();
    S->budget = (size_t)budget_bytes;
    S->nslots = max_active;
    S->trace = getenv_flag("TMG_SERVER_TRACE");
    S->t0 = std::chrono::steady_clock::now();
    auto bad = [&](const std::string &m) {
        g_setup_err = m;
        for (hipStream_t st : S->streams) (void)hipStreamDestroy(st);
        for (PcgScalars *p : S->hsc) (void)hipHostFree(p);
        for (rocblas_handle h : S->rbs) rocblas_destroy_handle(h);
        for (hipMemPool_t pl : S->pools) (void)hipMemPoolDestroy(pl);
        delete S;
        return TMG_ERR_HIP;
    };
    for (int i = 0; i < max_active; ++i) {
        hipStream_t st;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return bad("hipStreamCreate failed");
        S->streams.push_back(st);
        PcgScalars *p = nullptr;
        if (hipHostMalloc((void **)&p, sizeof(PcgScalars), hipHostMallocDefault) != hipSuccess) return bad("hipHostMalloc failed");
        S->hsc.push_back(p);
        hipMemPoolProps props;
        memset(&props, 0, sizeof(props));
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = 0;
        hipMemPool_t pl = nullptr;
        if (hipMemPoolCreate(&pl, &props) != hipSuccess) return bad("hipMemPoolCreate failed");
        S->pools.push_back(pl);
        uint64_t keep = UINT64_MAX;  // retain freed memory in the slot's pool
        (void)hipMemPoolSetAttribute(pl, hipMemPoolAttrReleaseThreshold, &keep);
        int off = 0;                 // never hand memory to another stream
        (void)hipMemPoolSetAttribute(pl, hipMemPoolReuseAllowOpportunistic, &off);
        (void)hipMemPoolSetAttribute(pl, hipMemPoolReuseAllowInternalDependencies, &off);
    }
    S->stage.assign(max_active, nullptr);
    S->stage_cap.assign(max_active, 0);
    S->slot_busy.assign(max_active, 0);
    for (int i = 0; i < max_active; ++i) {
        rocblas_handle h = nullptr;
        if (rocblas_create_handle(&h) != rocblas_status_success) return bad("rocblas_create_handle failed");
        S->rbs.push_back(h);
    }
    for (int i = 0; i < max_active; ++i) S->workers.emplace_back(srv_worker, S, i);
    *out = S;
    return TMG_OK;
}

int tmg_server_submit(tmg_server *S, const tmg_problem *prob, const tmg_mg_params *mg, int precision,
                      const double *x, const double *f, const double *x0, double t_arrival, int64_t *id) {
    if (!S || !prob || !mg || !x || !f || !id) return TMG_ERR_ARG;
    if (prob->nelx <= 0 || prob->nely <= 0 || prob->nelz <= 0 || prob->nfree < 0 || (prob->nfree && !prob->freedofs) ||
        mg->ngrids < 2 || mg->ngrids > MAXL) {
        S->err = "invalid problem or multigrid parameters";
        return TMG_ERR_ARG;
    }
    SrvReq *r = new SrvReq();
    r->prob = *prob;
    r->freedofs.assign(prob->freedofs, prob->freedofs + prob->nfree);
    r->mg = *mg;
    r->prec = precision;
    const int64_t ne = (int64_t)prob->nelx * prob->nely * prob->nelz;
    r->x.assign(x, x + ne);
    r->f.assign(f, f + prob->nfree);
    r->x0.assign(prob->nfree, 0.0);
    if (x0) r->x0.assign(x0, x0 + prob->nfree);
    r->est = estimate_bytes(prob->nelx, prob->nely, prob->nelz, mg->ngrids, precision, prob->nfree);
    std::lock_guard<std::mutex> lk(S->mu);
    if (r->est > S->budget) {
        S->err = "request needs " + std::to_string(r->est) + " bytes, more than the whole budget";
        delete r;
        return TMG_ERR_BUDGET;
    }
    r->id = S->next_id++;
    r->t_arrival = t_arrival >= 0 ? t_arrival : S->now();
    S->all[r->id] = r;
    S->queue.push_back(r);
    *id = r->id;
    S->cv.notify_all();
    return TMG_OK;
}

int tmg_server_wait(tmg_server *S, int64_t id, double timeout_s, int *state) {
    if (!S) return TMG_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    auto itr = S->all.find(id);
    if (itr == S->all.end()) return TMG_ERR_ARG;
    SrvReq *r = itr->second;
    auto pred = [&] { return r->state == 3 || r->state < 0; };
    if (timeout_s < 0) S->cv.wait(lk, pred);
    else S->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
    if (state) *state = r->state;
    return pred() ? TMG_OK : 1;
}

int tmg_server_result(tmg_server *S, int64_t id, double *u, int *it, double *relres, double *times) {
    if (!S) return TMG_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto itr = S->all.find(id);
    if (itr == S->all.end()) return TMG_ERR_ARG;
    SrvReq *r = itr->second;
    if (r->state < 0) { S->err = r->err; return r->state; }
    if (r->state != 3) return TMG_ERR_STATE;
    if (u) std::copy(r->u.begin(), r->u.end(), u);
    if (it) *it = r->it;
    if (relres) *relres = r->relres;
    if (times) { times[0] = r->t_arrival; times[1] = r->t_admit; times[2] = r->t_start; times[3] = r->t_done; }
    return r->rc;
}

int tmg_server_release(tmg_server *S, int64_t id) {
    if (!S) return TMG_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto itr = S->all.find(id);
    if (itr == S->all.end()) return TMG_ERR_ARG;
    if (itr->second->state != 3 && itr->second->state >= 0) return TMG_ERR_STATE;
    delete itr->second;
    S->all.erase(itr);
    return TMG_OK;
}

int tmg_server_stats(tmg_server *S, double *stats) {
    if (!S || !stats) return TMG_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    int done = 0;
    for (auto &kv : S->all) done += kv.second->state == 3;
    stats[0] = (double)S->used;
    stats[1] = (double)S->peak;
    stats[2] = (double)S->queue.size();
    stats[3] = (double)S->running.size();
    stats[4] = (double)done;
    stats[5] = S->now();
    stats[6] = (double)S->peak_running;
    stats[7] = (double)S->rebuilds;
    return TMG_OK;
}

double tmg_server_clock(const tmg_server *S) { return S ? S->now() : 0.0; }

int64_t tmg_server_estimate(const tmg_problem *prob, const tmg_mg_params *mg, int precision) {
    if (!prob || !mg) return TMG_ERR_ARG;
    return (int64_t)estimate_bytes(prob->nelx, prob->nely, prob->nelz, mg->ngrids, precision, prob->nfree);
}

const char *tmg_server_last_error(const tmg_server *S) { return S ? S->err.c_str() : g_setup_err.c_str(); }

void tmg_server_destroy(tmg_server *S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv.notify_all();
    for (std::thread &t : S->workers)
        if (t.joinable()) t.join();
    for (SrvReq *r : S->running) {
        if (r->h) { destroy_all(r->h); delete r->h; }
    }
    for (auto &kv : S->all) delete kv.second;
    for (hipStream_t st : S->streams) { (void)hipStreamSynchronize(st); (void)hipStreamDestroy(st); }
    for (PcgScalars *p : S->hsc) (void)hipHostFree(p);
    for (double *p : S->stage)
        if (p) (void)hipHostFree(p);
    for (rocblas_handle h : S->rbs) rocblas_destroy_handle(h);
    for (hipMemPool_t pl : S->pools) (void)hipMemPoolDestroy(pl);
    delete S;
}

}  // extern "C"
