// tmg_server.inc.hpp -- continuous-batching server for streams of equilibrium
// solves of uneven sizes (included at the end of tmg_solver.hip: it drives the
// solver through its internal begin / launch / poll steps).
//
// Requests are admitted FIFO while a slot is free and the estimated device
// memory of the request fits the budget. Each slot owns a HIP stream, a
// pinned scalar block, a pinned staging buffer for the request's host vectors
// and a stream-ordered memory pool that never hands memory to another stream,
// so admitting and retiring requests never synchronises the device and no
// request can be given memory another request's kernels still use. A setup
// thread builds each admitted hierarchy (Galerkin refresh, coarse factor,
// graph capture) and checks that the device holds the request's densities;
// a driver thread keeps every running solve supplied with chunks of PCG
// iterations; kernels of different requests overlap on the device. Results
// are bitwise identical to a solve run alone: every kernel is deterministic
// and a request's launches depend only on its own problem.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace {

constexpr int SRV_SETUP_THREADS = 4;

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
    int chunk = 1;
    bool in_flight = false;
    tmg_solver *h = nullptr;
    std::vector<double> u;
    int it = 0, rc = 0;
    double relres = 0.0;
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
    // symmetric coarse solve (coarse systems of 16,384 DOFs and more): block
    // pairs and per-pair partial products (upper bound: every coarse DOF free)
    if (ncoarse >= 16384) {
        const int64_t nb = (ncoarse + CS_B - 1) / CS_B;
        add(sizeof(int2) * (size_t)(nb * (nb + 1) / 2));
        add(8 * (size_t)(nb * (nb + 1) * CS_B));
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
    return b + (32u << 20);                                    // rocSOLVER workspace headroom
}

}  // namespace

struct tmg_server {
    size_t budget = 0, used = 0, peak = 0;
    int nslots = 0;
    int peak_running = 0;          // most requests solving at the same time
    // per slot: stream, pinned scalar block, pinned staging for the
    // request's host vectors, and its own stream-ordered memory pool (device
    // memory is reused only within a slot, i.e. only on the slot's stream)
    std::vector<hipStream_t> streams;
    std::vector<PcgScalars *> hsc;
    std::vector<double *> stage;
    std::vector<size_t> stage_cap;
    std::vector<hipMemPool_t> pools;
    std::vector<char> slot_busy;
    // the coarse factorisations of concurrent set-ups run one at a time
    std::mutex factor_mu;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<SrvReq *> queue;
    std::vector<SrvReq *> running;
    std::unordered_map<int64_t, SrvReq *> all;
    int64_t next_id = 1;
    bool stop = false;
    // several setup threads build admitted hierarchies concurrently (host work,
    // graph capture and the coarse factorization of one request overlap with
    // the others'); each owns a rocBLAS handle, whose stream binding is
    // per handle
    std::vector<std::thread> setup_thr;
    std::vector<rocblas_handle> rbs;
    std::thread drive_thr;
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

// A failed coarse factorisation is reported with what tells a corrupted
// set-up from a singular operator: whether the device still holds the
// request's densities, and whether building the same hierarchy again on the
// slot's stream, with nothing else in flight on it, factorises.
void srv_diagnose_not_spd(tmg_server *S, tmg_solver *h, SrvReq *r) {
    std::string first = h->err;
    uint64_t dev = 0;
    std::string dens = device_checksum(h, h->xdens, (int64_t)r->x.size(), &dev) == 0
                           ? (dev == host_checksum(r->x.data(), (int64_t)r->x.size()) ? "intact" : "corrupted")
                           : "unreadable";
    int rc2;
    {
        std::lock_guard<std::mutex> fk(S->factor_mu);
        rc2 = refresh(h);
    }
    h->err = first + "; densities on the device " + dens + "; rebuilding the same hierarchy " +
             (rc2 == 0 ? "factorised (the first build was corrupted)" : "failed again: " + h->err);
}

void srv_setup_loop(tmg_server *S, rocblas_handle rb) {
    (void)hipSetDevice(0);
    for (;;) {
        SrvReq *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv.wait(lk, [&] {
                if (S->stop) return true;
                if (S->queue.empty()) return false;
                SrvReq *q = S->queue.front();
                bool slot = false;
                for (int i = 0; i < S->nslots; ++i) slot |= !S->slot_busy[i];
                return slot && S->used + q->est <= S->budget;
            });
            if (S->stop) return;
            r = S->queue.front();
            S->queue.pop_front();
            for (int i = 0; i < S->nslots; ++i)
                if (!S->slot_busy[i]) { r->slot = i; S->slot_busy[i] = 1; break; }
            S->used += r->est;
            S->peak = std::max(S->peak, S->used);
            r->state = 1;
            r->t_admit = S->now();
        }
        const int slot = r->slot;
        int rc = 0;
        std::string msg;
        // pinned staging large enough for the densities and the free vectors
        const size_t need = std::max(r->x.size(), r->f.size());
        if (S->stage_cap[slot] < need) {
            if (S->stage[slot]) (void)hipHostFree(S->stage[slot]);
            S->stage[slot] = nullptr;
            S->stage_cap[slot] = 0;
            const size_t cap = need + need / 4;
            if (hipHostMalloc((void **)&S->stage[slot], cap * sizeof(double), hipHostMallocDefault) != hipSuccess) {
                S->stage[slot] = nullptr;
                rc = TMG_ERR_OOM;
                msg = "pinned staging allocation failed";
            } else {
                S->stage_cap[slot] = cap;
            }
        }
        ExtRes ext;
        ext.st = S->streams[slot];
        ext.rb = rb;
        ext.h_sc = S->hsc[slot];
        ext.pool = S->pools[slot];
        ext.stage = S->stage[slot];
        ext.stage_cap = S->stage_cap[slot];
        r->prob.freedofs = r->freedofs.data();
        tmg_solver *h = nullptr;
        if (!rc) {
            rc = setup_impl(&r->prob, &r->mg, r->prec, &ext, &h);
            if (rc) msg = g_setup_err;
        }
        if (!rc) {
            {
                std::lock_guard<std::mutex> fk(S->factor_mu);
                rc = tmg_update_density(h, r->x.data());
            }
            if (!rc) rc = srv_verify_upload(h, h->xdens, r->x.data(), (int64_t)r->x.size(), "densities");
            if (rc == TMG_ERR_NOT_SPD) srv_diagnose_not_spd(S, h, r);
            if (!rc && (rc = upload_free(h, 0, r->x0.data(), h->px)) == 0) rc = upload_free(h, 0, r->f.data(), h->pf);
            h->have_rhs = true;
            if (!rc) rc = solve_begin(h);
            if (rc) msg = h->err;
        }
        std::lock_guard<std::mutex> lk(S->mu);
        if (rc) {
            if (h) { destroy_all(h); delete h; }
            srv_release(S, r);
            srv_fail(S, r, rc, msg);
        } else {
            r->h = h;
            r->state = 2;
            r->t_start = S->now();
            S->running.push_back(r);
            S->peak_running = std::max(S->peak_running, (int)S->running.size());
        }
        S->cv.notify_all();
    }
}

void srv_drive_loop(tmg_server *S) {
    (void)hipSetDevice(0);
    std::vector<SrvReq *> act;
    for (;;) {
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (S->stop) return;
            act = S->running;
        }
        bool progressed = false;
        std::vector<SrvReq *> finished;
        for (SrvReq *r : act) {
            tmg_solver *h = r->h;
            if (r->in_flight) {
                int p = solve_poll(h);
                if (p == 0) continue;
                r->in_flight = false;
                progressed = true;
                if (p < 0) { r->rc = p; finished.push_back(r); continue; }
                if (h->h_sc->done) {
                    r->rc = solve_result(h, &r->it, &r->relres);
                    finished.push_back(r);
                    continue;
                }
                const int sck = solve_check(h);
                if (sck < 0) { r->rc = sck; finished.push_back(r); continue; }
            }
            int rc = solve_launch(h, r->chunk);
            r->chunk = std::min(r->chunk * 2, 4);
            if (rc < 0) { r->rc = rc; finished.push_back(r); continue; }
            r->in_flight = true;
            progressed = true;
        }
        for (SrvReq *r : finished) {
            tmg_solver *h = r->h;
            std::string msg = h->err;
            int rc = r->rc;
            if (rc >= 0) rc = std::min(rc, solve_finish(h));
            if (rc >= 0) {
                r->u.assign(h->lev[0].nfree, 0.0);
                int d = download_free(h, 0, h->px, r->u.data());
                if (d < 0) { rc = d; msg = h->err; }
            }
            destroy_all(h);
            delete h;
            std::lock_guard<std::mutex> lk(S->mu);
            r->h = nullptr;
            S->running.erase(std::find(S->running.begin(), S->running.end(), r));
            srv_release(S, r);
            if (rc < 0) srv_fail(S, r, rc, msg);
            else { r->state = 3; r->t_done = S->now(); }
            S->cv.notify_all();
        }
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
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
    const int nsetup = std::min(max_active, SRV_SETUP_THREADS);
    for (int i = 0; i < nsetup; ++i) {
        rocblas_handle h = nullptr;
        if (rocblas_create_handle(&h) != rocblas_status_success) return bad("rocblas_create_handle failed");
        S->rbs.push_back(h);
    }
    for (int i = 0; i < nsetup; ++i) S->setup_thr.emplace_back(srv_setup_loop, S, S->rbs[i]);
    S->drive_thr = std::thread(srv_drive_loop, S);
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
    for (std::thread &t : S->setup_thr)
        if (t.joinable()) t.join();
    if (S->drive_thr.joinable()) S->drive_thr.join();
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
