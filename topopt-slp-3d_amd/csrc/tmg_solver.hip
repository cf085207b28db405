// tmg_solver.hip -- host side of libtmg: device-resident GMG hierarchy,
// Galerkin refresh, V-cycle orchestration, device-driven PCG and the C-ABI
// declared in include/tmg.h.
#include "tmg_kernels.hpp"
#include "tmg.h"

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cfloat>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

using namespace tmg;

namespace {

constexpr int MAXL = 12;
constexpr int GRAPH_ITERS = 4;
constexpr int64_t FUSED_LEVEL_MAXNODES = 100000;
thread_local std::string g_setup_err;

struct DevLevel {
    int nelx = 0, nely = 0, nelz = 0, nx = 0, ny = 0, nz = 0;
    int64_t nn = 0, ne = 0, ndof = 0, nfree = 0;
    uint8_t *mask = nullptr;
    double *S64 = nullptr;
    float *S32 = nullptr;
    double *inv64 = nullptr;
    float *inv32 = nullptr;
    void *b = nullptr, *x = nullptr, *t = nullptr;  // V-cycle vectors, preconditioner precision
    std::vector<int64_t> free_full;                 // compact -> full dof (0-based)
    int64_t *d_free = nullptr;                      // device copy of free_full
    int64_t *d_cidx = nullptr;                      // full dof -> compact index or -1
};

// true when the environment variable is set to a value starting with '1'
bool getenv_flag(const char *name) {
    const char *v = getenv(name);
    return v && v[0] == '1';
}

}  // namespace

struct tmg_solver {
    int nelx, nely, nelz;
    double el, eh, ew, E0, nu, penal, emin;
    tmg_mg_params mg;
    int prec = TMG_FP64;
    int L = 0;
    DevLevel lev[MAXL];
    double ke[576];
    int mr = 1;                       // multiresolution factor n (1 = off; n = 2: 8 density sub-elements)
    int nsub = 1;                     // n^3
    double *K0_64 = nullptr;          // Ke, or the n^3 sub-element integrands (576 each)
    float *K0_32 = nullptr;
    double *E64 = nullptr;
    float *E32 = nullptr;
    // level 1 applied matrix-free (the Galerkin operator of an 8-child coarse
    // element has the multiresolution form): moduli of the children, table
    bool l1_wht = false;
    bool l1_maskok = false;           // every fixed fine DOF interpolates only from fixed coarse DOFs
    double *E1_64 = nullptr;
    float *E1_32 = nullptr;
    double *kwg64_d = nullptr;
    float *kwg32_d = nullptr;
    int l1_tyl = 4, l1_ntx = 1, l1_nty = 1, l1_zc32 = 4, l1_zc64 = 4;
    double *xdens = nullptr;
    double *Ahat[2] = {nullptr, nullptr};
    // coarsest grid
    int64_t nc = 0;
    double *Ainv = nullptr, *cb = nullptr, *cx = nullptr;
    // mixed precision: the FP32 cycle's coarse solve reads an FP32 copy of
    // the inverse (half the bytes; the FP64 hierarchy reads Ainv)
    bool coarse_f32 = false;
    float *Ainv32 = nullptr;
    // large coarse systems: solve reading the upper block triangle of Ainv
    // (k_coarse_symv_run/sum); below the threshold its extra passes cost more
    // than the halved matrix read saves
    bool coarse_symv = false;
    int cs_nb = 0;                    // 64-row blocks
    int cs_nruns = 0;                 // runs of consecutive upper tiles of one block row
    int4 *cs_runs = nullptr;          // per run: block row, first and end tile column
    int *cs_rs = nullptr;             // first run of each block row (cs_nb + 1)
    double *cs_part = nullptr;        // column products per upper tile pair, then row products per run
    int *d_info = nullptr;
    // large coarse systems: multifrontal Cholesky over a nested dissection of
    // the coarse node grid instead of the dense inverse (tmg_coarse_nd.inc.hpp)
    bool coarse_nd = false;
    int nd_nfront = 0, nd_maxld = 0, nd_nheight = 0;
    std::vector<NdFront> nd_fr;       // host copy of the front table (height order)
    std::vector<int> nd_hstart;       // first front of each height (nd_nheight + 1)
    std::vector<int> nd_ea_start, nd_fw_start, nd_bw_start;   // work-list ranges per height
    NdFront *nd_fronts = nullptr;
    int64_t *nd_P = nullptr, *nd_B = nullptr, *nd_src = nullptr, *nd_dst = nullptr, *nd_bsrc = nullptr, *nd_tsrc = nullptr;
    int *nd_cmap = nullptr, *nd_info = nullptr;
    int2 *nd_ea = nullptr, *nd_fw = nullptr, *nd_bw = nullptr, *nd_tp = nullptr;
    double *nd_F = nullptr, *nd_Mt = nullptr, *nd_T = nullptr, *nd_y = nullptr, *nd_xc = nullptr;
    int64_t nd_nent = 0, nd_Fsize = 0, nd_ntp = 0;
    // the fronts of one height are factorised on up to 4 side streams, each
    // with its own rocBLAS handle (handles that own their stream only)
    rocblas_handle rb = nullptr;
    // PCG
    double *px = nullptr, *pr = nullptr, *pd = nullptr, *pd2 = nullptr, *pq = nullptr, *pf = nullptr;
    double *partial = nullptr, *partial2 = nullptr;
    int npartial = 0;
    PcgScalars *d_sc = nullptr;
    PcgScalars *h_sc = nullptr;
    int zchunk = 4;
    // PCG iteration captured once as a HIP graph (all pointers are fixed for
    // the handle's lifetime), replayed GRAPH_ITERS iterations per launch
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // mixed precision: a solve whose residual stalls switches its
    // preconditioner to the FP64 hierarchy (the handle keeps FP64 masters)
    int pprec = TMG_FP64;             // preconditioner precision of the running solve
    int vec_prec = -1;                // precision the V-cycle work vectors were last written in
    hipGraph_t graph64 = nullptr;
    hipGraphExec_t graph_exec64 = nullptr;
    bool escalate = true;
    int stall_it = 0;
    double stall_nr = 0.0;
    int escalations = 0;              // solves switched to the FP64 preconditioner
    int escalated_at = 0;             // iteration of the last solve's switch (0: none)
    int stall_window = 64;            // iterations between progress checks
    double stall_factor = 0.5;        // required residual reduction per window
    bool use_graph = true;
    bool ebe_tiled = false;           // LDS-staged fine-level apply (ny <= EBE_TILE_MAXNY)
    int xcd_remap = 1;                // XCD-contiguous block order in the stencil kernels
    int xcd_ebe = 0;                  // the same for the fine-level apply (measured neutral or slower)
    int zchunk64 = 8;                 // z planes per thread of the FP64 fine apply
    bool cube = false;                // cube elements: Ke as ten scalars (KCube)
    bool ebe_prefetch = false;        // cube fine apply loading the next plane one step ahead (measured slower)
    bool fine_stencil = false;        // fine level held as an assembled block stencil (multiresolution)
    KCube<double> kc64{};
    KCube<float> kc32{};
    // spectral element form of the fine apply (k_ebe_wht): coupling
    // coefficients, tile shape (TY = 1 << wht_tyl, TX = TPB / TY) and tiles
    bool wht = false;
    bool wht_pf64 = true;             // spectral apply with the next plane prefetched: FP64 (faster)
    bool wht_pf32 = false;            // and FP32 (slower: the extra registers cost a wave)
    bool ebe_lds = true;              // SRC_X applies of the spectral form with the node plane staged through LDS (k_ebe_lds)
    bool mr_lds = true;               // FP64 multiresolution fine level through k_ebe_lds as well (FP32: no gain)
    bool fin_consumer = true;         // rho/beta and alpha reduced in their consumers' blocks (no k_finalize<1,2>)
    bool prolong4 = true;             // prolongation with one thread per coarse node and fine plane (k_prolong_add4)
    bool restrict_jz = true;          // the restriction also forms the next level's first sweep from zero
    bool debug_fail_first = false;    // TMG_DEBUG_FAIL_FIRST_FACTOR=1: the first coarse factorization reports failure
    int xcd_wht = 1;                  // XCD-contiguous tile order for the spectral apply
    KWht<double> kw64{};
    KWht<float> kw32{};
    KWht<double> kwg64{};             // level-1 (Galerkin) table
    KWht<float> kwg32{};
    double *kw64_d = nullptr;         // the same coefficients on the device
    float *kw32_d = nullptr;
    int wht_tyl = 4, wht_ntx = 1, wht_nty = 1;
    bool fused = false;               // fused fine-level V-cycle / direction update (nu1 = nu2 = 1)
    bool fused_coarse = true;         // fused stencil-level sweeps (nu1 = nu2 = 1)
    int64_t fused_maxnodes = FUSED_LEVEL_MAXNODES;  // largest stencil level that uses them
    int64_t st3_maxnodes = 200000;    // largest stencil level swept in the three-wave split form
    bool fine_r = true;               // fine-level cycle reading the FP64 residual, r's in its post-sweep
    // PCG reductions finished by the producing kernel's last block instead of
    // a k_finalize launch: measured much slower (each block's device-scope
    // fence writes back and invalidates its XCD's L2), so off by default
    bool fin_fused = false;
    // solution update deferred into the next direction update, and the
    // cycle's first Jacobi sweep folded into the residual update
    bool xdefer = true;
    hipStream_t st = nullptr;
    int dev = 0;                      // device current at set-up (the stream's)
    hipEvent_t ev = nullptr;          // completion marker of the last launched chunk
    // externally owned resources (a server slot): device memory comes from a
    // stream-ordered pool, and the stream, rocBLAS handle and pinned scalar
    // block are lent, so nothing here synchronises the whole device
    hipMemPool_t pool = nullptr;
    // or from one lent block (the server's per-request arena): allocations are
    // carved from it in order and a free of the most recent one gives its
    // space back (the temporaries of set-up and of the host-vector transfers
    // are freed in that order); the block itself belongs to the lender
    char *arena = nullptr;
    size_t arena_cap = 0, arena_off = 0;
    double *stage = nullptr;          // pinned host staging lent by a server slot: host vectors pass through it
    size_t stage_cap = 0;             // its capacity in doubles
    bool own_stream = true, own_rb = true, own_hsc = true;
    size_t bytes = 0;                 // device bytes held by this handle now
    size_t peak_bytes = 0;            // and at most, over its lifetime
    std::unordered_map<void *, size_t> sizes;   // rounded size of each live allocation
    std::unordered_map<void *, std::pair<size_t, int>> canary;   // TMG_DEBUG_CANARY: data bytes, allocating line
    bool have_op = false;
    bool have_rhs = false;            // pf holds a right-hand side (a solve was started)
    // kernel timing inside a solve (tmg_time_in_solve): event pairs recorded
    // around each launch of kernel kt_which while the PCG runs eagerly
    int kt_which = -1;
    int kt_n = 0;
    std::vector<hipEvent_t> kt_ev;
    std::string err;
};

namespace {

#define HIPX(expr)                                                                          \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            h->err = std::string(#expr) + ": " + hipGetErrorString(e_);                     \
            return TMG_ERR_HIP;                                                             \
        }                                                                                   \
    } while (0)

#define TRY(expr)                      \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ < 0) return rc_;       \
    } while (0)

inline unsigned nblk(int64_t n) { return (unsigned)((n + TPB - 1) / TPB); }

// kernels timed inside a solve by tmg_time_in_solve (ids of tmg_time_kernel)
constexpr int KT_APPLY = 0;        // FP64 fine apply q = K d with d'q
constexpr int KT_POST = 8;         // fine post-sweep on the FP64 residual
constexpr int KT_MAX_PAIRS = 1024;

// Records the start (end = false) or the end of one launch of kernel id on
// the solver's stream while tmg_time_in_solve runs (never inside a graph
// capture: the timed solve runs eagerly).
inline void kt_mark(tmg_solver *h, int id, bool end) {
    if (h->kt_which != id || h->kt_n >= (int)h->kt_ev.size() / 2) return;
    (void)hipEventRecord(h->kt_ev[2 * h->kt_n + (end ? 1 : 0)], h->st);
    if (end) ++h->kt_n;
}

inline size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// TMG_DEBUG_POISON=1: every new device buffer starts as 0xFF bytes (NaN in
// either float view, -1 as an index), so a read before the first write shows
// up in the results instead of reading whatever the memory held before
bool poison_buffers() {
    static const bool on = getenv_flag("TMG_DEBUG_POISON");
    return on;
}

// TMG_DEBUG_CANARY=1: every device buffer is followed by 4 KiB of 0xA5
// bytes, checked after set-up, each density update and each solve, so that
// a write past the end of a buffer is reported with the source line of the
// buffer's allocation (instead of landing in a neighbouring buffer)
constexpr size_t CANARY_BYTES = 4096;
bool canary_buffers() {
    static const bool on = getenv_flag("TMG_DEBUG_CANARY");
    return on;
}

template <typename T>
int dalloc(tmg_solver *h, T **p, int64_t n, int line = __builtin_LINE()) {
    const size_t bytes = sizeof(T) * (size_t)std::max<int64_t>(n, 1);
    const size_t guard = canary_buffers() ? CANARY_BYTES : 0;
    const size_t total = round256(bytes) + guard;
    hipError_t e = hipSuccess;
    if (h->arena && h->arena_off + total <= h->arena_cap) {
        *p = reinterpret_cast<T *>(h->arena + h->arena_off);
        h->arena_off += total;
    } else {
        e = h->pool ? hipMallocFromPoolAsync((void **)p, total, h->pool, h->st) : hipMalloc((void **)p, total);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        h->err = "device allocation of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e);
        return TMG_ERR_OOM;
    }
    if (poison_buffers() && h->st) (void)hipMemsetAsync((void *)*p, 0xFF, bytes, h->st);
    if (guard) {
        (void)hipMemsetAsync(reinterpret_cast<char *>(*p) + round256(bytes), 0xA5, guard, h->st);
        h->canary[(void *)*p] = std::make_pair(round256(bytes), line);
    }
    h->bytes += total;
    h->peak_bytes = std::max(h->peak_bytes, h->bytes);
    h->sizes[(void *)*p] = total;
    return 0;
}

// reports (and returns TMG_ERR_HIP for) the first buffer whose guard bytes changed
int canary_check(tmg_solver *h, const char *where) {
    if (!canary_buffers() || h->canary.empty()) return 0;
    if (hipStreamSynchronize(h->st) != hipSuccess) return 0;
    std::vector<uint8_t> g(CANARY_BYTES);
    for (const auto &kv : h->canary) {
        // stream-ordered: other threads may be capturing graphs
        if (hipMemcpyAsync(g.data(), reinterpret_cast<char *>(kv.first) + kv.second.first, CANARY_BYTES,
                           hipMemcpyDeviceToHost, h->st) != hipSuccess ||
            hipStreamSynchronize(h->st) != hipSuccess)
            continue;
        size_t bad = 0;
        while (bad < CANARY_BYTES && g[bad] == 0xA5) ++bad;
        if (bad < CANARY_BYTES) {
            h->err = std::string("guard bytes overwritten ") + where + ": buffer of " + std::to_string(kv.second.first) +
                     " bytes allocated at tmg_solver.hip line " + std::to_string(kv.second.second) + ", first changed byte " +
                     std::to_string(bad) + " past its end";
            fprintf(stderr, "tmg: %s\n", h->err.c_str());
            return TMG_ERR_HIP;
        }
    }
    return 0;
}

template <typename T>
void dfree(tmg_solver *h, T *&p) {
    if (p) {
        auto it = h->sizes.find((void *)p);
        const size_t sz = it != h->sizes.end() ? it->second : 0;
        char *c = reinterpret_cast<char *>(p);
        if (h->arena && c >= h->arena && c < h->arena + h->arena_cap) {
            if (sz && c + sz == h->arena + h->arena_off) h->arena_off -= sz;   // the most recent block
        } else if (h->pool) {
            (void)hipFreeAsync((void *)p, h->st);
        } else {
            (void)hipFree((void *)p);
        }
        if (it != h->sizes.end()) {
            h->bytes -= it->second;
            h->sizes.erase(it);
        }
        h->canary.erase((void *)c);
    }
    p = nullptr;
}

#include "tmg_coarse_nd.inc.hpp"

// ---------------------------------------------------------------------------
// operator applies
// ---------------------------------------------------------------------------
template <typename T> const T *K0of(tmg_solver *h);
template <> const double *K0of<double>(tmg_solver *h) { return h->K0_64; }
template <> const float *K0of<float>(tmg_solver *h) { return h->K0_32; }
template <typename T>
const KCube<T> &KCof(tmg_solver *h);
template <> const KCube<double> &KCof<double>(tmg_solver *h) { return h->kc64; }
template <> const KCube<float> &KCof<float>(tmg_solver *h) { return h->kc32; }
template <typename T> const T *Eof(tmg_solver *h);
template <> const double *Eof<double>(tmg_solver *h) { return h->E64; }
template <> const float *Eof<float>(tmg_solver *h) { return h->E32; }
template <typename T> const T *Sof(const DevLevel &l);
template <> const double *Sof<double>(const DevLevel &l) { return l.S64; }
template <> const float *Sof<float>(const DevLevel &l) { return l.S32; }
template <typename T> const T *invof(const DevLevel &l);
template <> const double *invof<double>(const DevLevel &l) { return l.inv64; }
template <> const float *invof<float>(const DevLevel &l) { return l.inv32; }

// Number of blocks an apply of this level launches (partial-sum slots).
// dbl: the FP64 launch (its z chunk differs from the FP32 one).
// Grid of the transfer kernels: in-plane blocks x planes.
inline dim3 plane_grid(int nx, int ny, int nz) { return dim3((unsigned)nblk((int64_t)nx * ny), (unsigned)nz); }

// Coupling table A of the multiresolution form from 8 sub-element matrices K
// (8 x 576, row-major 24x24 each): C_s[c][d][k][l] = 1/64 sum_ab H[k][a]
// K_s[3a+c][3b+d] H[l][b] must equal A[c][d][k][l] H[m][s] with
// m = k ^ l ^ e_c ^ e_d, and A must vanish outside WHT_MR_NZ.
bool wht_mr_table(const double *K, KWht<double> &t64, KWht<float> &t32) {
    const int bitk[3] = {2, 1, 4};
    auto Hm = [](int k, int a) { return (__builtin_popcount(k & a) & 1) ? -1.0 : 1.0; };
    std::vector<double> Cs((size_t)8 * 576);
    double cmax = 0.0;
    for (int sb = 0; sb < 8; ++sb)
        for (int c = 0; c < 3; ++c)
            for (int d = 0; d < 3; ++d)
                for (int k = 0; k < 8; ++k)
                    for (int l = 0; l < 8; ++l) {
                        double sum = 0.0;
                        for (int a = 0; a < 8; ++a)
                            for (int b = 0; b < 8; ++b)
                                sum += Hm(k, a) * K[576 * sb + (3 * a + c) * 24 + 3 * b + d] * Hm(l, b);
                        const double v = sum / 64.0;
                        Cs[(size_t)sb * 576 + ((c * 3 + d) * 8 + k) * 8 + l] = v;
                        cmax = std::max(cmax, std::fabs(v));
                    }
    bool ok = true;
    for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d)
            for (int k = 0; k < 8; ++k)
                for (int l = 0; l < 8; ++l) {
                    const int m = k ^ l ^ bitk[c] ^ bitk[d];
                    double al = 0.0;
                    for (int sb = 0; sb < 8; ++sb) al += Cs[(size_t)sb * 576 + ((c * 3 + d) * 8 + k) * 8 + l] * Hm(m, sb);
                    al /= 8.0;
                    for (int sb = 0; sb < 8; ++sb)
                        ok = ok && std::fabs(Cs[(size_t)sb * 576 + ((c * 3 + d) * 8 + k) * 8 + l] - al * Hm(m, sb)) <=
                                       1e-13 * cmax;
                    if ((WHT_MR_NZ[8 * c + k] >> (8 * d + l)) & 1) {
                        t64.g[wht_mr_slot(c, k, d, l)] = al;
                        t32.g[wht_mr_slot(c, k, d, l)] = (float)al;
                    } else {
                        ok = ok && std::fabs(al) <= 1e-13 * cmax;
                    }
                }
    return ok;
}

// Galerkin integrands of the 8 fine elements of a coarse element: P_s' Ke P_s
// with P_s the trilinear interpolation (Implementation/Prolongation.m weights)
// from the coarse element's 8 nodes to fine element s's 8 nodes.
void galerkin_integrands(const double *Ke, double *K8) {
    for (int sb = 0; sb < 8; ++sb) {
        const int sy = sb & 1, sx = (sb >> 1) & 1, sz = sb >> 2;
        double P[24 * 24] = {0.0};
        for (int a = 0; a < 8; ++a) {
            const double pos[3] = {0.5 * (sx + ((a >> 1) & 1)), 0.5 * (sy + (a & 1)), 0.5 * (sz + (a >> 2))};
            for (int b = 0; b < 8; ++b) {
                const double bc[3] = {(double)((b >> 1) & 1), (double)(b & 1), (double)(b >> 2)};
                double w = 1.0;
                for (int ax = 0; ax < 3; ++ax) w *= 1.0 - std::fabs(pos[ax] - bc[ax]);
                for (int c = 0; c < 3; ++c) P[(3 * a + c) * 24 + 3 * b + c] = w;
            }
        }
        double KP[24 * 24];
        for (int i = 0; i < 24; ++i)
            for (int j = 0; j < 24; ++j) {
                double sum = 0.0;
                for (int q = 0; q < 24; ++q) sum += Ke[i * 24 + q] * P[q * 24 + j];
                KP[i * 24 + j] = sum;
            }
        for (int i = 0; i < 24; ++i)
            for (int j = 0; j < 24; ++j) {
                double sum = 0.0;
                for (int q = 0; q < 24; ++q) sum += P[q * 24 + i] * KP[q * 24 + j];
                K8[576 * sb + i * 24 + j] = sum;
            }
    }
}

// tile shape (TY in {8, 16, 32}, the fewest tiles per node plane) of the
// spectral apply on an nx x ny node plane
void wht_tiles(int nx, int ny, int &tyl, int &ntx, int &nty) {
    int64_t best = -1;
    for (int t = 3; t <= 5; ++t) {
        const int TY = 1 << t, TX = TPB / TY;
        const int ax = (nx + TX - 2) / (TX - 1), ay = (ny + TY - 2) / (TY - 1);
        const int64_t nt = (int64_t)ax * ay;
        if (best < 0 || nt < best || (nt == best && t == 4)) { best = nt; tyl = t; ntx = ax; nty = ay; }
    }
}

int64_t apply_blocks(tmg_solver *h, int l, bool dbl) {
    const DevLevel &lv = h->lev[l];
    if (l == 1 && h->l1_wht) {
        const int zc = dbl ? h->l1_zc64 : h->l1_zc32;
        return (int64_t)h->l1_ntx * h->l1_nty * ((lv.nz + zc - 1) / zc);
    }
    if (l == 0 && !h->fine_stencil) {
        int64_t pb = h->wht ? (int64_t)h->wht_ntx * h->wht_nty : nblk((int64_t)lv.nx * lv.ny);
        const int zc = dbl ? h->zchunk64 : h->zchunk;
        return pb * ((lv.nz + zc - 1) / zc);
    }
    return nblk(lv.nn);
}

template <typename T, typename TS>
EbeArgs<T, TS> ebe_args(tmg_solver *h) {
    const DevLevel &l0 = h->lev[0];
    EbeArgs<T, TS> a;
    memset(&a, 0, sizeof(a));
    a.nelx = l0.nelx; a.nely = l0.nely; a.nelz = l0.nelz;
    a.zchunk = std::is_same<T, double>::value ? h->zchunk64 : h->zchunk;
    a.fin = FIN_NONE;
    a.E = Eof<T>(h); a.K0 = K0of<T>(h); a.mask = l0.mask; a.invD = invof<T>(l0);
    a.cnx = h->L > 1 ? h->lev[1].nx : 1;
    a.cny = h->L > 1 ? h->lev[1].ny : 1;
    return a;
}

template <typename T>
const T *KWof(tmg_solver *h);
template <> const double *KWof<double>(tmg_solver *h) { return h->kw64_d; }
template <> const float *KWof<float>(tmg_solver *h) { return h->kw32_d; }

template <typename T, typename TS, int SRC, int EPI>
void launch_ebe(tmg_solver *h, const EbeArgs<T, TS> &a) {
    const DevLevel &lv = h->lev[0];
    if (h->wht && a.fin < 0 && (h->nsub == 1 || SRC == SRC_X)) {
        dim3 gw((unsigned)(h->wht_ntx * h->wht_nty), (unsigned)((lv.nz + a.zchunk - 1) / a.zchunk));
#define WHT_ARGS a.nelx, a.nely, a.nelz, a.zchunk, h->wht_tyl, h->wht_ntx, a.x, a.E, a.mask, a.b, a.invD, a.r, a.ec, \
                 a.cnx, a.cny, a.s, a.dold, a.dnew, a.sc, a.out, a.partial, a.done, KWof<T>(h), h->xcd_wht
        const bool pf = std::is_same<T, double>::value ? h->wht_pf64 : h->wht_pf32;
        if constexpr (SRC == SRC_X && EPI != EPI_DIRDOT) {
            if (h->ebe_lds && h->nsub == 1) {
                // FP64: registers for 3 waves per SIMD; FP32 Jacobi sweeps: 5
                constexpr int WPE = std::is_same<T, double>::value ? 3 : (EPI == EPI_JACOBI ? 5 : 1);
                hipLaunchKernelGGL((k_ebe_lds<T, EPI, 1, WPE>), gw, dim3(TPB), 0, h->st, a.nelx, a.nely, a.nelz, a.zchunk,
                                   h->wht_tyl, h->wht_ntx, a.x, a.E, a.mask, a.b, a.invD, a.r, a.out, a.partial, a.done,
                                   KWof<T>(h), h->xcd_wht);
                return;
            }
            if (std::is_same<T, double>::value && h->ebe_lds && h->mr_lds && h->nsub == 8) {
                hipLaunchKernelGGL((k_ebe_lds<T, EPI, 8, 1>), gw, dim3(TPB), 0, h->st, a.nelx, a.nely, a.nelz, a.zchunk,
                                   h->wht_tyl, h->wht_ntx, a.x, a.E, a.mask, a.b, a.invD, a.r, a.out, a.partial, a.done,
                                   KWof<T>(h), h->xcd_wht);
                return;
            }
        }
        if (h->nsub == 8) {
            // multiresolution: one form per precision (FP64 prefetching)
            if constexpr (SRC == SRC_X) {
                constexpr bool PF8 = std::is_same<T, double>::value;
                hipLaunchKernelGGL((k_ebe_wht<T, TS, SRC, EPI, PF8, 8>), gw, dim3(TPB), 0, h->st, WHT_ARGS);
            }
        } else if (pf) {
            hipLaunchKernelGGL((k_ebe_wht<T, TS, SRC, EPI, true>), gw, dim3(TPB), 0, h->st, WHT_ARGS);
        } else {
            hipLaunchKernelGGL((k_ebe_wht<T, TS, SRC, EPI, false>), gw, dim3(TPB), 0, h->st, WHT_ARGS);
        }
#undef WHT_ARGS
        return;
    }
    dim3 grid(nblk((int64_t)lv.nx * lv.ny), (lv.nz + a.zchunk - 1) / a.zchunk);
#define EBE_ARGS a.nelx, a.nely, a.nelz, a.zchunk, a.x, a.E, a.K0, a.mask, a.b, a.invD, a.r, a.ec, a.cnx, \
                 a.cny, a.s, a.dold, a.dnew, a.sc, a.out, a.partial, a.done, KCof<T>(h), h->xcd_ebe, a.fin
    if constexpr (SRC == SRC_X) {
        if (h->nsub == 8) {
            hipLaunchKernelGGL((k_ebe<T, TS, SRC, EPI, false, 8>), grid, dim3(TPB), 0, h->st, EBE_ARGS);
            return;
        }
    }
    if (h->cube && h->nsub == 1 && !h->ebe_tiled && h->ebe_prefetch)
        hipLaunchKernelGGL((k_ebe<T, TS, SRC, EPI, false, 1, 2>), grid, dim3(TPB), 0, h->st, EBE_ARGS);
    else if (h->cube && h->nsub == 1 && !h->ebe_tiled)
        hipLaunchKernelGGL((k_ebe<T, TS, SRC, EPI, false, 1, 1>), grid, dim3(TPB), 0, h->st, EBE_ARGS);
    else if (lv.ny <= EBE_TILE_MAXNY && h->ebe_tiled)
        hipLaunchKernelGGL((k_ebe<T, TS, SRC, EPI, true, 1>), grid, dim3(TPB), 0, h->st, EBE_ARGS);
    else
        hipLaunchKernelGGL((k_ebe<T, TS, SRC, EPI, false, 1>), grid, dim3(TPB), 0, h->st, EBE_ARGS);
#undef EBE_ARGS
}

#define ST_ARGS(a) (a).nx, (a).ny, (a).nz, (a).x, (a).S, (a).mask, (a).b, (a).invD, (a).ec, (a).cnx, (a).cny, (a).out, \
                   (a).partial, (a).done, h->xcd_remap

template <typename T>
StArgs<T> st_args(tmg_solver *h, int l) {
    const DevLevel &lv = h->lev[l];
    StArgs<T> a;
    memset(&a, 0, sizeof(a));
    a.nx = lv.nx; a.ny = lv.ny; a.nz = lv.nz;
    a.S = Sof<T>(lv); a.mask = lv.mask; a.invD = invof<T>(lv);
    a.cnx = l + 1 < h->L ? h->lev[l + 1].nx : 1;
    a.cny = l + 1 < h->L ? h->lev[l + 1].ny : 1;
    return a;
}

// Stencil sweep of level l: the three-wave split form on small levels.
template <typename T, int SRC, int EPI>
void launch_stencil(tmg_solver *h, int l, const StArgs<T> &a) {
    const DevLevel &lv = h->lev[l];
    if (EPI != EPI_DOT && lv.nn <= h->st3_maxnodes) {
        if constexpr (EPI != EPI_DOT)
            hipLaunchKernelGGL((k_stencil3<T, SRC, EPI>), dim3((unsigned)((lv.nn + 63) / 64)), dim3(ST3_THREADS), 0,
                               h->st, ST_ARGS(a));
        return;
    }
    hipLaunchKernelGGL((k_stencil<T, SRC, EPI>), dim3(nblk(lv.nn)), dim3(TPB), 0, h->st, ST_ARGS(a));
}

// x_l (+)= P e_{l+1} on level l's free DOFs (Implementation/Prolongation.m)
template <typename T>
void prolong_add(tmg_solver *h, int l, const T *ec, T *x, int add, const int *done) {
    const DevLevel &lf = h->lev[l], &lc = h->lev[l + 1];
    if (h->prolong4)
        hipLaunchKernelGGL((k_prolong_add4<T>), plane_grid(lc.nx, lc.ny, lf.nz), dim3(TPB), 0, h->st, lf.nx, lf.ny,
                           lf.nz, lc.nx, lc.ny, lc.nz, ec, lf.mask, x, add, done);
    else
        hipLaunchKernelGGL((k_prolong_add<T>), plane_grid(lf.nx, lf.ny, lf.nz), dim3(TPB), 0, h->st, lf.nx, lf.ny,
                           lf.nz, lc.nx, lc.ny, lc.nz, ec, lf.mask, x, add, done);
}

template <typename T>
const T *E1of(tmg_solver *h);
template <> const double *E1of<double>(tmg_solver *h) { return h->E1_64; }
template <> const float *E1of<float>(tmg_solver *h) { return h->E1_32; }
template <typename T>
const T *KWGof(tmg_solver *h);
template <> const double *KWGof<double>(tmg_solver *h) { return h->kwg64_d; }
template <> const float *KWGof<float>(tmg_solver *h) { return h->kwg32_d; }

template <typename T, int EPI>
void apply(tmg_solver *h, int l, const T *x, const T *b, T *out, double *partial, const int *done) {
    const DevLevel &lv = h->lev[l];
    if (l == 1 && h->l1_wht) {
        // level-1 Galerkin operator, matrix-free: the 8 children's moduli
        // times the Galerkin integrands' table (multiresolution form)
        const int zc = std::is_same<T, double>::value ? h->l1_zc64 : h->l1_zc32;
        constexpr bool PF8 = std::is_same<T, double>::value;
        dim3 gw((unsigned)(h->l1_ntx * h->l1_nty), (unsigned)((lv.nz + zc - 1) / zc));
        if constexpr (std::is_same<T, double>::value && EPI != EPI_DIRDOT && EPI != EPI_RESIDR) {
            if (h->ebe_lds) {   // measured faster in FP64; neutral in FP32 (k_ebe_wht kept there)
                hipLaunchKernelGGL((k_ebe_lds<T, EPI, 8, 1>), gw, dim3(TPB), 0, h->st, lv.nelx, lv.nely, lv.nelz, zc,
                                   h->l1_tyl, h->l1_ntx, x, E1of<T>(h), lv.mask, b, invof<T>(lv),
                                   (const double *)nullptr, out, partial, done, KWGof<T>(h), h->xcd_wht);
                return;
            }
        }
        hipLaunchKernelGGL((k_ebe_wht<T, T, SRC_X, EPI, PF8, 8>), gw, dim3(TPB), 0, h->st, lv.nelx, lv.nely, lv.nelz, zc,
                           h->l1_tyl, h->l1_ntx, x, E1of<T>(h), lv.mask, b, invof<T>(lv), (const double *)nullptr,
                           (const T *)nullptr, 1, 1, (const T *)nullptr, (const double *)nullptr, (double *)nullptr,
                           (PcgScalars *)nullptr, out, partial, done, KWGof<T>(h), h->xcd_wht);
        return;
    }
    if (l == 0 && !h->fine_stencil) {
        EbeArgs<T, T> a = ebe_args<T, T>(h);
        a.x = x; a.b = b; a.out = out; a.partial = partial; a.done = done;
        launch_ebe<T, T, SRC_X, EPI>(h, a);
    } else {
        StArgs<T> a = st_args<T>(h, l);
        a.x = x; a.b = b; a.out = out; a.partial = partial; a.done = done;
        launch_stencil<T, SRC_X, EPI>(h, l, a);
    }
}

// The PCG's q = A d with d'q. With the residual update deferred
// (k_r_update drops q at the fixed DOFs) the LDS kernel skips the free-DOF
// mask (EPI_DOTU); otherwise, or where that kernel does not serve the fine
// level, the masked apply gives the same free-DOF values.
void apply_pcg_q(tmg_solver *h, bool consumer_masks, const int *done) {
    const DevLevel &l0 = h->lev[0];
    if (consumer_masks && h->wht && h->ebe_lds && h->mr_lds && h->nsub == 8 && !h->fine_stencil) {
        dim3 gw((unsigned)(h->wht_ntx * h->wht_nty), (unsigned)((l0.nz + h->zchunk64 - 1) / h->zchunk64));
        hipLaunchKernelGGL((k_ebe_lds<double, EPI_DOTU, 8, 1>), gw, dim3(TPB), 0, h->st, l0.nelx, l0.nely, l0.nelz,
                           h->zchunk64, h->wht_tyl, h->wht_ntx, (const double *)h->pd, (const double *)h->E64, l0.mask,
                           (const double *)nullptr, (const double *)nullptr, (const double *)nullptr, h->pq, h->partial,
                           done, (const double *)h->kw64_d, h->xcd_wht);
        return;
    }
    if (consumer_masks && h->wht && h->ebe_lds && h->nsub == 1 && !h->fine_stencil) {
        dim3 gw((unsigned)(h->wht_ntx * h->wht_nty), (unsigned)((l0.nz + h->zchunk64 - 1) / h->zchunk64));
        hipLaunchKernelGGL((k_ebe_lds<double, EPI_DOTU, 1, 3>), gw, dim3(TPB), 0, h->st, l0.nelx, l0.nely, l0.nelz,
                           h->zchunk64, h->wht_tyl, h->wht_ntx, (const double *)h->pd, (const double *)h->E64, l0.mask,
                           (const double *)nullptr, (const double *)nullptr, (const double *)nullptr, h->pq, h->partial,
                           done, (const double *)h->kw64_d, h->xcd_wht);
        return;
    }
    apply<double, EPI_DOT>(h, 0, h->pd, nullptr, h->pq, h->partial, done);
}

template <typename T>
void coarse_solve(tmg_solver *h, const T *b, T *x, const int *done) {
    DevLevel &lc = h->lev[h->L - 1];
    if (h->coarse_nd) {
        nd_solve<T>(h, b, x, done);
        return;
    }
    const bool f32 = std::is_same<T, float>::value && h->coarse_f32;
    if (h->coarse_symv) {
        const int64_t nb = h->cs_nb;
        hipLaunchKernelGGL((k_gather<T>), dim3(nblk(h->nc)), dim3(TPB), 0, h->st, h->nc, lc.d_free, b, h->cb, done);
        if (f32)
            hipLaunchKernelGGL((k_coarse_symv_run<float, 4>), dim3((unsigned)h->cs_nruns), dim3(64), 0, h->st, h->nc,
                               h->cs_nb, (const float *)h->Ainv32, (const int4 *)h->cs_runs, (const double *)h->cb,
                               h->cs_part, h->cs_part + nb * (nb - 1) / 2 * CS_B, done);
        else
            hipLaunchKernelGGL((k_coarse_symv_run<double, 2>), dim3((unsigned)h->cs_nruns), dim3(64), 0, h->st, h->nc,
                               h->cs_nb, (const double *)h->Ainv, (const int4 *)h->cs_runs, (const double *)h->cb,
                               h->cs_part, h->cs_part + nb * (nb - 1) / 2 * CS_B, done);
        hipLaunchKernelGGL((k_coarse_symv_sum<T>), dim3((unsigned)h->cs_nb), dim3(TPB), 0, h->st, h->nc, h->cs_nb,
                           (const double *)h->cs_part, (const int *)h->cs_rs, lc.d_free, x, done);
        return;
    }
    if (f32)
        hipLaunchKernelGGL((k_coarse_solve<T, float>), dim3((unsigned)h->nc), dim3(TPB), 0, h->st, h->nc,
                           (const float *)h->Ainv32, lc.d_free, b, x, done);
    else
        hipLaunchKernelGGL((k_coarse_solve<T, double>), dim3((unsigned)h->nc), dim3(TPB), 0, h->st, h->nc,
                           (const double *)h->Ainv, lc.d_free, b, x, done);
}

// One V(nu1,nu2) cycle from a zero initial guess at level l
// (Implementation/VCycle.m:23-43). Returns the buffer holding the result.
template <typename T>
T *vcycle(tmg_solver *h, int l, const T *b, const int *done, T *x0 = nullptr, bool jz = false);

// Level m's cycle from zero takes the unfused path with a first sweep from
// zero (x = M \ b): the restriction into level m can form that sweep.
bool restrict_jz_ok(const tmg_solver *h, int m) {
    if (!h->restrict_jz || m <= 0 || m >= h->L - 1 || h->mg.smoothiter1 < 1) return false;
    const bool fused = h->mg.smoothiter1 == 1 && h->mg.smoothiter2 == 1 && h->fused_coarse &&
                       h->lev[m].nn <= h->fused_maxnodes;
    return !fused;
}

// rH = P' rh into level l + 1's b; with jz also x = M \ rH into its x.
template <typename T>
void restrict_to(tmg_solver *h, int l, const T *rf, bool jz, const int *done) {
    const DevLevel &lv = h->lev[l], &lc = h->lev[l + 1];
    if (jz)
        hipLaunchKernelGGL((k_restrict<T, true>), plane_grid(lc.nx, lc.ny, lc.nz), dim3(TPB), 0, h->st, lv.nx, lv.ny,
                           lv.nz, lc.nx, lc.ny, lc.nz, rf, lc.mask, (T *)lc.b, done, invof<T>(lc), (T *)lc.x);
    else
        hipLaunchKernelGGL((k_restrict<T>), plane_grid(lc.nx, lc.ny, lc.nz), dim3(TPB), 0, h->st, lv.nx, lv.ny, lv.nz,
                           lc.nx, lc.ny, lc.nz, rf, lc.mask, (T *)lc.b, done, nullptr, nullptr);
}

// Coarse correction of a cycle at level l: the direct solve below the
// coarsest-but-one level, else one cycle (V) or two, the second from the
// first's result (W, Implementation/WCycle.m:18-20).
template <typename T>
T *coarse_correction(tmg_solver *h, int l, const int *done, bool jz = false) {
    DevLevel &lc = h->lev[l + 1];
    if (l + 1 == h->L - 1) {
        coarse_solve<T>(h, (const T *)lc.b, (T *)lc.x, done);                      // eH = R \ (R' \ rH)
        return (T *)lc.x;
    }
    T *e = vcycle<T>(h, l + 1, (const T *)lc.b, done, nullptr, jz);
    if (h->mg.cycle == 1) e = vcycle<T>(h, l + 1, (const T *)lc.b, done, e);
    return e;
}

// One cycle at level l (Implementation/VCycle.m:23-43; with mg.cycle = 1 the
// recursion of Implementation/WCycle.m). x0: initial guess held in one of the
// level's two work vectors, or nullptr for zero. Returns the buffer holding
// the result (one of the level's two work vectors).
template <typename T>
T *vcycle(tmg_solver *h, int l, const T *b, const int *done, T *x0, bool jz) {
    DevLevel &lv = h->lev[l];
    DevLevel &lc = h->lev[l + 1];
    T *xa = (T *)lv.x, *xb = (T *)lv.t;
    const int nu1 = h->mg.smoothiter1, nu2 = h->mg.smoothiter2;
    // small stencil levels are launch-bound: fold the sweeps into the applies there
    if (!x0 && l > 0 && nu1 == 1 && nu2 == 1 && h->fused_coarse && lv.nn <= h->fused_maxnodes) {
        // stencil level, V(1,1): rh = b - A (M\b); then s = w + M\(b - A w), w = M\b + P eH
        StArgs<T> a = st_args<T>(h, l);
        a.b = b; a.out = xb; a.done = done;
        launch_stencil<T, SRC_JZ, EPI_RESID>(h, l, a);
        const bool jzc = restrict_jz_ok(h, l + 1);
        restrict_to<T>(h, l, xb, jzc, done);
        T *ec = coarse_correction<T>(h, l, done, jzc);
        StArgs<T> c = st_args<T>(h, l);
        c.b = b; c.ec = ec; c.out = xa; c.done = done;
        launch_stencil<T, SRC_POST, EPI_JACOBI>(h, l, c);
        return xa;
    }
    if (x0) {
        // sweeps from the given initial guess (Implementation/WCycle.m:3)
        xa = x0;
        xb = x0 == (T *)lv.x ? (T *)lv.t : (T *)lv.x;
        for (int s = 0; s < nu1; ++s) {
            apply<T, EPI_JACOBI>(h, l, xa, b, xb, nullptr, done);
            std::swap(xa, xb);
        }
    } else if (nu1 >= 1) {
        // first sweep from x0 = 0: x = M \ b (Implementation/Smooth.m:17 with x = 0);
        // with jz the restriction that produced b wrote it
        if (!jz)
            hipLaunchKernelGGL((k_jacobi_zero<T>), dim3(nblk(lv.ndof)), dim3(TPB), 0, h->st, lv.ndof, b, invof<T>(lv), xa,
                               done);
        for (int s = 1; s < nu1; ++s) {
            apply<T, EPI_JACOBI>(h, l, xa, b, xb, nullptr, done);
            std::swap(xa, xb);
        }
    } else {
        hipLaunchKernelGGL((k_fill<T>), dim3(nblk(lv.ndof)), dim3(TPB), 0, h->st, lv.ndof, T(0), xa);
    }
    apply<T, EPI_RESID>(h, l, xa, b, xb, nullptr, done);                          // rh = b - A x
    const bool jzc = restrict_jz_ok(h, l + 1);
    restrict_to<T>(h, l, xb, jzc, done);                                           // rH = P' rh
    T *ec = coarse_correction<T>(h, l, done, jzc);
    prolong_add<T>(h, l, (const T *)ec, xa, 1, done);                            // x += P eH
    for (int s = 0; s < nu2; ++s) {
        apply<T, EPI_JACOBI>(h, l, xa, b, xb, nullptr, done);
        std::swap(xa, xb);
    }
    return xa;
}

// Fine-level V(1,1) cycle with the first Jacobi sweep, the prolongation of
// the coarse correction and r's dot product folded into the two fine-level
// operator applies (Implementation/VCycle.m:23-43 with x0 = 0):
//   rh = r - A (M\r)                        (sweep from zero + residual)
//   eH = coarse cycle (P' rh)
//   s  = w + M\(r - A w), w = M\r + P eH    (prolongation + post-sweep)
// partial[] receives the per-block sums of r . s. Returns s.
template <typename T>
T *vcycle_fine_fused(tmg_solver *h, const double *r, const int *done) {
    DevLevel &l0 = h->lev[0];
    DevLevel &l1 = h->lev[1];
    EbeArgs<T, T> a = ebe_args<T, T>(h);
    a.r = r; a.done = done; a.out = (T *)l0.t;
    launch_ebe<T, T, SRC_JZ, EPI_RESIDR>(h, a);
    const bool jzc = restrict_jz_ok(h, 1);
    restrict_to<T>(h, 0, (const T *)l0.t, jzc, done);
    T *ec = coarse_correction<T>(h, 0, done, jzc);
    EbeArgs<T, T> b = ebe_args<T, T>(h);
    b.r = r; b.ec = ec; b.done = done; b.out = (T *)l0.x; b.partial = h->partial;
    launch_ebe<T, T, SRC_POST, EPI_POSTR>(h, b);
    return (T *)l0.x;
}

bool fused_ok(const tmg_solver *h) {
    return h->fused && h->nsub == 1 && h->mg.cycle != 2 && h->mg.smoothiter1 == 1 && h->mg.smoothiter2 == 1;
}

// One FMV cycle at level l (Implementation/FMVCycle.m:21-39): restrict b,
// solve the coarsest system directly or recurse, prolong, then one V-cycle
// at level l from that initial guess. Returns the buffer holding the result.
template <typename T>
T *fmv(tmg_solver *h, int l, const T *b, const int *done) {
    DevLevel &lv = h->lev[l];
    DevLevel &lc = h->lev[l + 1];
    hipLaunchKernelGGL((k_restrict<T>), plane_grid(lc.nx, lc.ny, lc.nz), dim3(TPB), 0, h->st, lv.nx, lv.ny, lv.nz,
                       lc.nx, lc.ny, lc.nz, b, lc.mask, (T *)lc.b, done);            // rH = P' b
    T *xc;
    if (l + 1 == h->L - 1) {
        coarse_solve<T>(h, (const T *)lc.b, (T *)lc.x, done);                          // x = R \ (R' \ rH)
        xc = (T *)lc.x;
    } else {
        xc = fmv<T>(h, l + 1, (const T *)lc.b, done);                                   // x = FMVCycle(..., nv + 1)
    }
    T *xa = (T *)lv.x;
    hipLaunchKernelGGL((k_fill<T>), dim3(nblk(lv.ndof)), dim3(TPB), 0, h->st, lv.ndof, T(0), xa);
    prolong_add<T>(h, l, (const T *)xc, xa, 1, done);                                  // x = P x
    return vcycle<T>(h, l, b, done, xa);                                                // VCycle(A, b, x, ...)
}

// The preconditioner cycle of mg.cycle at the finest level: V, W or FMV.
template <typename T>
T *fine_cycle(tmg_solver *h, const T *b, const int *done) {
    return h->mg.cycle == 2 ? fmv<T>(h, 0, b, done) : vcycle<T>(h, 0, b, done);
}

// Fine-level V(1,1) cycle on the FP64 PCG residual r with the conversion and
// the PCG's r's folded into kernels the cycle runs anyway
// (Implementation/VCycle.m:23-43 with x0 = 0):
//   x  = M \ r                     (first sweep from zero, reading r)
//   rh = r - A x                    (residual epilogue reading r)
//   eH = coarse correction (P' rh); x += P eH
//   s  = x + M \ (r - A x)          (post-sweep, partial[] += r . s)
// Returns s; the partial sums come from the post-sweep's blocks.
// x0_ready: the first sweep's x = M \ r is already in the level's x vector
// (written by the PCG's residual update).
template <typename T>
T *vcycle_fine_r(tmg_solver *h, const double *r, const int *done, bool x0_ready = false) {
    DevLevel &l0 = h->lev[0];
    DevLevel &l1 = h->lev[1];
    T *xa = (T *)l0.x, *xb = (T *)l0.t;
    if (!x0_ready)
        hipLaunchKernelGGL((k_jacobi_zero<T, double>), dim3(nblk(l0.ndof)), dim3(TPB), 0, h->st, l0.ndof, r,
                           invof<T>(l0), xa, done);
    EbeArgs<T, T> a = ebe_args<T, T>(h);
    a.x = xa; a.r = r; a.out = xb; a.done = done;
    launch_ebe<T, T, SRC_X, EPI_RESIDR>(h, a);
    const bool jzc = restrict_jz_ok(h, 1);
    restrict_to<T>(h, 0, (const T *)xb, jzc, done);
    T *ec = coarse_correction<T>(h, 0, done, jzc);
    prolong_add<T>(h, 0, (const T *)ec, xa, 1, done);
    EbeArgs<T, T> b = ebe_args<T, T>(h);
    b.x = xa; b.r = r; b.out = xb; b.partial = h->partial; b.done = done;
    if (h->fin_fused) { b.sc = h->d_sc; b.fin = FIN_RHO; }       // rho = r's, beta: in the last block
    kt_mark(h, KT_POST, false);
    launch_ebe<T, T, SRC_X, EPI_POSTR>(h, b);
    kt_mark(h, KT_POST, true);
    return xb;
}

// the fine level is matrix-free and the cycle is V(1,1) or W(1,1)
bool fine_r_ok(const tmg_solver *h) {
    return h->fine_r && !h->fine_stencil && h->mg.cycle != 2 && h->mg.smoothiter1 == 1 && h->mg.smoothiter2 == 1;
}

// Preconditioner application s = B r for the PCG; r is the FP64 residual.
// Returns s as a typed pointer through one of the two out-params; *nrho > 0
// when partial[0..*nrho) already hold the blocks' sums of r . s.
void precondition(tmg_solver *h, const double *r, const double **s64, const float **s32, const int *done,
                  int *nrho = nullptr, bool x0_ready = false) {
    DevLevel &l0 = h->lev[0];
    *s64 = nullptr;
    *s32 = nullptr;
    if (nrho) *nrho = 0;
    if (!fused_ok(h) && fine_r_ok(h)) {
        if (h->pprec == TMG_MIXED) *s32 = vcycle_fine_r<float>(h, r, done, x0_ready);
        else *s64 = vcycle_fine_r<double>(h, r, done, x0_ready);
        if (nrho) *nrho = (int)apply_blocks(h, 0, h->pprec != TMG_MIXED);
        return;
    }
    if (fused_ok(h)) {
        if (h->pprec == TMG_MIXED) *s32 = vcycle_fine_fused<float>(h, r, done);
        else *s64 = vcycle_fine_fused<double>(h, r, done);
    } else if (h->pprec == TMG_MIXED) {
        hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk(l0.ndof)), dim3(TPB), 0, h->st, l0.ndof, r, (float *)l0.b, done);
        *s32 = fine_cycle<float>(h, (const float *)l0.b, done);
    } else {
        *s64 = fine_cycle<double>(h, r, done);
    }
}

// the PCG runs with the deferred solution update (k_dir_xupdate, k_r_update)
bool xdefer_ok(const tmg_solver *h) { return h->xdefer && !fused_ok(h) && fine_r_ok(h) && !h->fin_fused; }

// x0 = M \ r at the fine level in the running preconditioner precision
void first_sweep_from_r(tmg_solver *h) {
    DevLevel &l0 = h->lev[0];
    if (h->pprec == TMG_MIXED)
        hipLaunchKernelGGL((k_jacobi_zero<float, double>), dim3(nblk(l0.ndof)), dim3(TPB), 0, h->st, l0.ndof, h->pr,
                           l0.inv32, (float *)l0.x, (const int *)nullptr);
    else
        hipLaunchKernelGGL((k_jacobi_zero<double, double>), dim3(nblk(l0.ndof)), dim3(TPB), 0, h->st, l0.ndof, h->pr,
                           l0.inv64, (double *)l0.x, (const int *)nullptr);
}

// q = A d with d = beta d_old + s formed on the fly; writes d_new, partial d.q.
template <typename TS>
void direction_apply(tmg_solver *h, const TS *s, const double *dold, double *dnew, const int *done) {
    EbeArgs<double, TS> a = ebe_args<double, TS>(h);
    a.s = s; a.dold = dold; a.dnew = dnew; a.sc = h->d_sc; a.out = h->pq; a.partial = h->partial; a.done = done;
    launch_ebe<double, TS, SRC_DIR, EPI_DIRDOT>(h, a);
}

// One PCG iteration of Implementation/SolveMG.m:30-52, all scalars on device.
// parity selects which of the two direction buffers holds d_old.
void pcg_iteration(tmg_solver *h, int parity) {
    const int *done = &h->d_sc->done;
    const int64_t n = h->lev[0].ndof;
    const double *s64;
    const float *s32;
    const int nb = h->npartial;
    const unsigned gdot = (unsigned)std::min<int64_t>(nblk(n), nb);
    const int napply = (int)apply_blocks(h, 0, true);
    const int finrr = h->fin_fused ? FIN_RR : FIN_NONE;
    int nrho = 0;
    const bool xd = xdefer_ok(h);
    precondition(h, h->pr, &s64, &s32, done, &nrho, xd);
    // rho/beta and alpha computed by every block of their consumers
    // (k_dir_xupdate, k_r_update) instead of one-block k_finalize launches
    const bool fc = xd && h->fin_consumer && !h->fin_fused && nrho > 0 && !fused_ok(h);
    const unsigned gfc = (unsigned)std::min<int64_t>(gdot, 2048);
    if (fused_ok(h)) {
        // rho = r' s (partials from the post-smoothing apply, preconditioner precision)
        const int npost = (int)apply_blocks(h, 0, s32 == nullptr);
        hipLaunchKernelGGL(k_finalize<FIN_RHO>, dim3(1), dim3(FIN_THREADS), 0, h->st, npost, h->partial, 0, nullptr, h->d_sc);
        double *dold = parity ? h->pd2 : h->pd, *dnew = parity ? h->pd : h->pd2;
        if (s32) direction_apply<float>(h, s32, dold, dnew, done);
        else direction_apply<double>(h, s64, dold, dnew, done);
        hipLaunchKernelGGL(k_finalize<FIN_ALPHA>, dim3(1), dim3(FIN_THREADS), 0, h->st, napply, h->partial, 0, nullptr, h->d_sc);
        hipLaunchKernelGGL(k_xr_update, dim3(gdot), dim3(TPB), 0, h->st, n, h->px, h->pr, dnew, h->pq, h->partial, h->d_sc,
                           finrr);
    } else {
        if (nrho > 0) {
            // rho = r' s from the post-sweep's partial sums (finished there when fin_fused)
            if (!h->fin_fused && !fc)
                hipLaunchKernelGGL(k_finalize<FIN_RHO>, dim3(1), dim3(FIN_THREADS), 0, h->st, nrho, h->partial, 0, nullptr,
                                   h->d_sc);
        } else {
            if (s32) hipLaunchKernelGGL((k_dot<double, float>), dim3(gdot), dim3(TPB), 0, h->st, n, h->pr, s32, h->partial, done);
            else hipLaunchKernelGGL((k_dot<double, double>), dim3(gdot), dim3(TPB), 0, h->st, n, h->pr, s64, h->partial, done);
            hipLaunchKernelGGL(k_finalize<FIN_RHO>, dim3(1), dim3(FIN_THREADS), 0, h->st, (int)gdot, h->partial, 0, nullptr, h->d_sc);
        }
        if (fc) {
            if (s32)
                hipLaunchKernelGGL((k_dir_xupdate<float, true>), dim3(gfc), dim3(TPB), 0, h->st, n, s32, h->pd, h->px,
                                   h->d_sc, nrho, (const double *)h->partial);
            else
                hipLaunchKernelGGL((k_dir_xupdate<double, true>), dim3(gfc), dim3(TPB), 0, h->st, n, s64, h->pd, h->px,
                                   h->d_sc, nrho, (const double *)h->partial);
        } else if (xd) {
            if (s32) hipLaunchKernelGGL((k_dir_xupdate<float>), dim3(nblk(n)), dim3(TPB), 0, h->st, n, s32, h->pd, h->px, h->d_sc, 0, (const double *)nullptr);
            else hipLaunchKernelGGL((k_dir_xupdate<double>), dim3(nblk(n)), dim3(TPB), 0, h->st, n, s64, h->pd, h->px, h->d_sc, 0, (const double *)nullptr);
        } else if (s32) {
            hipLaunchKernelGGL((k_dir_update<float>), dim3(nblk(n)), dim3(TPB), 0, h->st, n, s32, h->pd, h->d_sc);
        } else {
            hipLaunchKernelGGL((k_dir_update<double>), dim3(nblk(n)), dim3(TPB), 0, h->st, n, s64, h->pd, h->d_sc);
        }
        kt_mark(h, KT_APPLY, false);
        if (h->fin_fused && !h->fine_stencil) {
            // q = A d, d'q, alpha in the apply's last block
            EbeArgs<double, double> a = ebe_args<double, double>(h);
            a.x = h->pd; a.out = h->pq; a.partial = h->partial; a.done = done; a.sc = h->d_sc; a.fin = FIN_ALPHA;
            launch_ebe<double, double, SRC_X, EPI_DOT>(h, a);
            kt_mark(h, KT_APPLY, true);
        } else {
            apply_pcg_q(h, xd, done);      // q = A d, d'q
            kt_mark(h, KT_APPLY, true);
            if (!fc)
                hipLaunchKernelGGL(k_finalize<FIN_ALPHA>, dim3(1), dim3(FIN_THREADS), 0, h->st, napply, h->partial, 0,
                                   nullptr, h->d_sc);
        }
        if (fc) {
            // ||r||^2 partials into partial2: the blocks still read the apply's partials
            DevLevel &l0 = h->lev[0];
            if (h->pprec == TMG_MIXED)
                hipLaunchKernelGGL((k_r_update<float, true>), dim3(gfc), dim3(TPB), 0, h->st, n, h->pr, h->pq, l0.inv32,
                                   (float *)l0.x, h->partial2, h->d_sc, napply, (const double *)h->partial);
            else
                hipLaunchKernelGGL((k_r_update<double, true>), dim3(gfc), dim3(TPB), 0, h->st, n, h->pr, h->pq, l0.inv64,
                                   (double *)l0.x, h->partial2, h->d_sc, napply, (const double *)h->partial);
        } else if (xd) {
            DevLevel &l0 = h->lev[0];
            if (h->pprec == TMG_MIXED)
                hipLaunchKernelGGL((k_r_update<float>), dim3(gdot), dim3(TPB), 0, h->st, n, h->pr, h->pq, l0.inv32,
                                   (float *)l0.x, h->partial, h->d_sc, 0, (const double *)nullptr);
            else
                hipLaunchKernelGGL((k_r_update<double>), dim3(gdot), dim3(TPB), 0, h->st, n, h->pr, h->pq, l0.inv64,
                                   (double *)l0.x, h->partial, h->d_sc, 0, (const double *)nullptr);
        } else {
            hipLaunchKernelGGL(k_xr_update, dim3(gdot), dim3(TPB), 0, h->st, n, h->px, h->pr, h->pd, h->pq, h->partial,
                               h->d_sc, finrr);
        }
    }
    if (fc)
        hipLaunchKernelGGL(k_finalize<FIN_RR>, dim3(1), dim3(FIN_THREADS), 0, h->st, (int)gfc, h->partial2, 0, nullptr,
                           h->d_sc);
    else if (!h->fin_fused)
        hipLaunchKernelGGL(k_finalize<FIN_RR>, dim3(1), dim3(FIN_THREADS), 0, h->st, (int)gdot, h->partial, 0, nullptr, h->d_sc);
}

int check_launch(tmg_solver *h) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        h->err = std::string("kernel launch: ") + hipGetErrorString(e);
        return TMG_ERR_HIP;
    }
    return 0;
}

// Start a solve on the handle's px (x0) and pf: r = f - A x0, norms, and the
// PCG iteration graph (captured once per handle).
// The V-cycle work vectors are shared by the FP32 and FP64 cycles of a mixed
// handle. The coarse solve never writes fixed DOFs and relies on their being
// zero, which holds only in the view they were zeroed or written in: switching
// views re-zeroes every level's vectors (outside any graph capture).
void use_cycle_precision(tmg_solver *h, int p) {
    h->pprec = p;
    if (h->vec_prec == p) return;
    for (int l = 0; l < h->L; ++l) {
        DevLevel &lv = h->lev[l];
        for (void *v : {lv.b, lv.x, lv.t})
            hipLaunchKernelGGL((k_fill<double>), dim3(nblk(lv.ndof)), dim3(TPB), 0, h->st, lv.ndof, 0.0, (double *)v);
    }
    h->vec_prec = p;
}

int solve_begin(tmg_solver *h) {
    if (!h->have_op) {
        h->err = "no operator: call tmg_update_density first";
        return TMG_ERR_STATE;
    }
    const int64_t n = h->lev[0].ndof;
    HIPX(hipStreamSynchronize(h->st));  // h_sc is reused as the staging buffer
    memset(h->h_sc, 0, sizeof(PcgScalars));
    h->h_sc->tol = h->mg.tolMG;
    h->h_sc->maxiter = h->mg.maxiterMG;
    h->h_sc->restart = 1;
    use_cycle_precision(h, h->prec);
    h->stall_it = 0;
    h->stall_nr = 0.0;
    h->escalated_at = 0;
    HIPX(hipMemcpyAsync(h->d_sc, h->h_sc, sizeof(PcgScalars), hipMemcpyHostToDevice, h->st));
    // d = 0 so that the first direction update is exact
    hipLaunchKernelGGL((k_fill<double>), dim3(nblk(n)), dim3(TPB), 0, h->st, n, 0.0, h->pd);
    hipLaunchKernelGGL((k_fill<double>), dim3(nblk(n)), dim3(TPB), 0, h->st, n, 0.0, h->pd2);

    // r = f - A x0 (Implementation/SolveMG.m:26-27)
    apply<double, EPI_RESID>(h, 0, h->px, h->pf, h->pr, nullptr, nullptr);
    const unsigned gdot = (unsigned)std::min<int64_t>(nblk(n), h->npartial);
    hipLaunchKernelGGL((k_dot<double, double>), dim3(gdot), dim3(TPB), 0, h->st, n, h->pr, h->pr, h->partial, nullptr);
    hipLaunchKernelGGL((k_dot<double, double>), dim3(gdot), dim3(TPB), 0, h->st, n, h->pf, h->pf, h->partial2, nullptr);
    hipLaunchKernelGGL(k_finalize<FIN_INIT>, dim3(1), dim3(FIN_THREADS), 0, h->st, (int)gdot, h->partial, (int)gdot,
                       h->partial2, h->d_sc);
    if (xdefer_ok(h)) first_sweep_from_r(h);
    TRY(check_launch(h));
    if (h->use_graph && !h->graph_exec) {
        // capture GRAPH_ITERS consecutive iterations; iterations past
        // convergence are skipped on the device through the done flag
        HIPX(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < GRAPH_ITERS; ++k) pcg_iteration(h, k & 1);
        hipGraph_t g = nullptr;
        HIPX(hipStreamEndCapture(h->st, &g));
        h->graph = g;
        HIPX(hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
    }
    if (!h->ev) HIPX(hipEventCreateWithFlags(&h->ev, hipEventDisableTiming));
    return 0;
}

// Launch nchunk x GRAPH_ITERS iterations, then an asynchronous copy of the
// scalars to pinned memory and the completion event.
int solve_launch(tmg_solver *h, int nchunk) {
    hipGraphExec_t ge = !h->use_graph ? nullptr : h->pprec == h->prec ? h->graph_exec : h->graph_exec64;
    if (ge) {
        for (int k = 0; k < nchunk; ++k) HIPX(hipGraphLaunch(ge, h->st));
    } else {
        for (int k = 0; k < nchunk * GRAPH_ITERS; ++k) pcg_iteration(h, k & 1);
    }
    TRY(check_launch(h));
    HIPX(hipMemcpyAsync(h->h_sc, h->d_sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, h->st));
    HIPX(hipEventRecord(h->ev, h->st));
    return 0;
}

// Called between chunks of a mixed-precision solve (h_sc current, not done):
// when the recursive residual fell by less than 2x over the last 64 or more
// iterations, the FP32 preconditioner has stalled (on designs with large
// void regions its rounding hides the weak modes); the rest of the solve
// uses the FP64 hierarchy and restarts the conjugate directions.
int solve_check(tmg_solver *h) {
    if (h->prec != TMG_MIXED || h->pprec != TMG_MIXED || !h->escalate) return 0;
    const int it = h->h_sc->it;
    const double nr = h->h_sc->nr;
    if (h->stall_it == 0) { h->stall_it = it; h->stall_nr = nr; return 0; }
    if (it - h->stall_it < h->stall_window) return 0;
    if (!(nr > h->stall_factor * h->stall_nr)) { h->stall_it = it; h->stall_nr = nr; return 0; }
    use_cycle_precision(h, TMG_FP64);
    if (xdefer_ok(h)) first_sweep_from_r(h);   // the level vectors were cleared
    ++h->escalations;
    h->escalated_at = it;
    if (h->use_graph && !h->graph_exec64) {
        HIPX(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < GRAPH_ITERS; ++k) pcg_iteration(h, k & 1);
        hipGraph_t g = nullptr;
        HIPX(hipStreamEndCapture(h->st, &g));
        h->graph64 = g;
        HIPX(hipGraphInstantiate(&h->graph_exec64, h->graph64, nullptr, nullptr, 0));
    }
    hipLaunchKernelGGL(k_set_restart, dim3(1), dim3(64), 0, h->st, h->d_sc);
    return check_launch(h);
}

int solve_result(tmg_solver *h, int *it, double *relres) {
    *it = h->h_sc->it;
    *relres = h->h_sc->nr;
    if (!std::isfinite(h->h_sc->nr)) {
        h->err = "PCG breakdown (non-finite residual)";
        return TMG_ERR_NOT_SPD;
    }
    return (h->h_sc->nr > h->mg.tolMG) ? TMG_NOT_CONVERGED : TMG_OK;
}

// After the PCG has stopped: the last iteration's deferred solution update.
int solve_finish(tmg_solver *h) {
    if (!xdefer_ok(h)) return canary_check(h, "by the solve");
    const int64_t n = h->lev[0].ndof;
    hipLaunchKernelGGL(k_x_final, dim3(nblk(n)), dim3(TPB), 0, h->st, n, h->px, h->pd, h->d_sc);
    TRY(check_launch(h));
    return canary_check(h, "by the solve");
}

// Blocking device-resident solve on px (x0 in, x out) and pf.
int solve_resident(tmg_solver *h, int *it, double *relres) {
    TRY(solve_begin(h));
    int chunk = 1;
    for (;;) {
        TRY(solve_launch(h, chunk));
        HIPX(hipEventSynchronize(h->ev));
        if (h->h_sc->done) break;
        TRY(solve_check(h));
        chunk = std::min(chunk * 2, 8);
    }
    TRY(solve_finish(h));
    return solve_result(h, it, relres);
}

// ---------------------------------------------------------------------------
// setup / refresh
// ---------------------------------------------------------------------------
// fd1: the finest level's 1-based free-DOF list (l = 0), whose index maps
// are then built on the device; coarse levels build theirs on the host from
// the mask (1/8 of the size per level).
int alloc_level(tmg_solver *h, int l, const std::vector<uint8_t> &mask, const int64_t *fd1 = nullptr,
                int64_t nfree1 = 0) {
    DevLevel &lv = h->lev[l];
    lv.nx = lv.nelx + 1; lv.ny = lv.nely + 1; lv.nz = lv.nelz + 1;
    lv.nn = (int64_t)lv.nx * lv.ny * lv.nz;
    lv.ne = (int64_t)lv.nelx * lv.nely * lv.nelz;
    lv.ndof = 3 * lv.nn;
    lv.free_full.clear();
    if (fd1) {
        lv.nfree = nfree1;
        lv.free_full.resize((size_t)nfree1);
        for (int64_t i = 0; i < nfree1; ++i) lv.free_full[i] = fd1[i] - 1;
        int64_t *dfd1 = nullptr;
        TRY(dalloc(h, &lv.mask, lv.ndof));
        TRY(dalloc(h, &lv.d_free, lv.nfree));
        TRY(dalloc(h, &lv.d_cidx, lv.ndof));
        TRY(dalloc(h, &dfd1, std::max<int64_t>(nfree1, 1)));
        hipLaunchKernelGGL((k_fill<uint8_t>), dim3(nblk(lv.ndof)), dim3(TPB), 0, h->st, lv.ndof, (uint8_t)0, lv.mask);
        hipLaunchKernelGGL((k_fill<int64_t>), dim3(nblk(lv.ndof)), dim3(TPB), 0, h->st, lv.ndof, (int64_t)-1, lv.d_cidx);
        if (nfree1) {
            HIPX(hipMemcpyAsync(dfd1, fd1, sizeof(int64_t) * nfree1, hipMemcpyHostToDevice, h->st));
            hipLaunchKernelGGL(k_level_index, dim3(nblk(nfree1)), dim3(TPB), 0, h->st, nfree1, dfd1, lv.mask, lv.d_cidx,
                               lv.d_free);
        }
        HIPX(hipStreamSynchronize(h->st));
        dfree(h, dfd1);
    } else {
        std::vector<int64_t> cidx(lv.ndof, -1);
        for (int64_t d = 0; d < lv.ndof; ++d)
            if (mask[d]) { cidx[d] = (int64_t)lv.free_full.size(); lv.free_full.push_back(d); }
        lv.nfree = (int64_t)lv.free_full.size();
        TRY(dalloc(h, &lv.mask, lv.ndof));
        TRY(dalloc(h, &lv.d_free, lv.nfree));
        TRY(dalloc(h, &lv.d_cidx, lv.ndof));
        HIPX(hipMemcpyAsync(lv.mask, mask.data(), lv.ndof, hipMemcpyHostToDevice, h->st));
        if (lv.nfree)
            HIPX(hipMemcpyAsync(lv.d_free, lv.free_full.data(), sizeof(int64_t) * lv.nfree, hipMemcpyHostToDevice, h->st));
        HIPX(hipMemcpyAsync(lv.d_cidx, cidx.data(), sizeof(int64_t) * lv.ndof, hipMemcpyHostToDevice, h->st));
        HIPX(hipStreamSynchronize(h->st));  // host staging vectors go out of scope
    }
    // V-cycle work vectors are sized for FP64 in both modes: a mixed handle
    // runs its FP64 cycle in them when a solve escalates (solve_check)
    TRY(dalloc(h, (char **)&lv.b, (int64_t)(lv.ndof * sizeof(double))));
    TRY(dalloc(h, (char **)&lv.x, (int64_t)(lv.ndof * sizeof(double))));
    TRY(dalloc(h, (char **)&lv.t, (int64_t)(lv.ndof * sizeof(double))));
    // level vectors start at zero (all-zero bytes: zero in either view): fixed
    // DOFs of a level vector are never written afterwards except with zero
    // (the coarse solve skips them)
    for (void *v : {lv.b, lv.x, lv.t})
        hipLaunchKernelGGL((k_fill<double>), dim3(nblk(lv.ndof)), dim3(TPB), 0, h->st, lv.ndof, 0.0, (double *)v);
    TRY(dalloc(h, &lv.inv64, lv.ndof));
    if (h->prec == TMG_MIXED) TRY(dalloc(h, &lv.inv32, lv.ndof));
    if (l > 0 || h->fine_stencil) {
        const int64_t ns = ((lv.nn + 63) / 64) * 64 * NCOEF;  // wave-blocked symmetric stencil storage
        TRY(dalloc(h, &lv.S64, ns));
        if (h->prec == TMG_MIXED) TRY(dalloc(h, &lv.S32, ns));
    }
    return 0;
}

void destroy_all(tmg_solver *h) {
    if (h->st) (void)hipStreamSynchronize(h->st);
    for (int l = 0; l < MAXL; ++l) {
        DevLevel &lv = h->lev[l];
        dfree(h, lv.mask); dfree(h, lv.S64); dfree(h, lv.S32); dfree(h, lv.inv64); dfree(h, lv.inv32);
        dfree(h, lv.b); dfree(h, lv.x); dfree(h, lv.t); dfree(h, lv.d_free); dfree(h, lv.d_cidx);
    }
    dfree(h, h->K0_64); dfree(h, h->K0_32); dfree(h, h->kw64_d); dfree(h, h->kw32_d);
    dfree(h, h->kwg64_d); dfree(h, h->kwg32_d); dfree(h, h->cs_runs); dfree(h, h->cs_rs); dfree(h, h->cs_part); dfree(h, h->E1_64); dfree(h, h->E1_32); dfree(h, h->E64); dfree(h, h->E32); dfree(h, h->xdens);
    dfree(h, h->Ahat[0]); dfree(h, h->Ahat[1]);
    nd_free(h);
    dfree(h, h->Ainv32); dfree(h, h->Ainv); dfree(h, h->cb); dfree(h, h->cx); dfree(h, h->d_info);
    dfree(h, h->px); dfree(h, h->pr); dfree(h, h->pd); dfree(h, h->pd2); dfree(h, h->pq); dfree(h, h->pf);
    dfree(h, h->partial); dfree(h, h->partial2); dfree(h, h->d_sc);
    if (h->h_sc && h->own_hsc) (void)hipHostFree(h->h_sc);
    h->h_sc = nullptr;
    if (h->ev) (void)hipEventDestroy(h->ev);
    h->ev = nullptr;
    for (hipEvent_t e : h->kt_ev) (void)hipEventDestroy(e);
    h->kt_ev.clear();
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->graph_exec = nullptr;
    if (h->graph_exec64) (void)hipGraphExecDestroy(h->graph_exec64);
    if (h->graph64) (void)hipGraphDestroy(h->graph64);
    h->graph_exec64 = nullptr;
    h->graph64 = nullptr;
    h->graph = nullptr;
    if (h->rb && h->own_rb) rocblas_destroy_handle(h->rb);
    h->rb = nullptr;
    if (h->st) (void)hipStreamSynchronize(h->st);
    if (h->st && h->own_stream) (void)hipStreamDestroy(h->st);
    h->st = nullptr;
}

// Element stiffness by exact 2x2x2 Gauss integration of B'DB on the box
// (equal to the closed form of Implementation/ElemStiffMatrix.m, local DOF
// order 3*(dy + 2 dx + 4 dz) + c).
// Integral of B'DB over the sub-box [lo, lo + h]^3 of the reference cube
// [-1, 1]^3 (2-point Gauss per direction, exact for Q1 on boxes); the whole
// cube (lo = -1, h = 2) gives Ke, the 2x2x2 sub-boxes give the
// multiresolution integrands of Implementation/ElemStiffMatrixGQMR.m.
void element_stiffness_box(double el, double eh, double ew, double nu, const double lo[3], double hb, double *k) {
    const double a = 0.5 * el, b = 0.5 * eh, c = 0.5 * ew;
    const double g = 1.0 / std::sqrt(3.0);
    const double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = 1.0 / (2.0 * (1.0 + nu));
    std::fill(k, k + 576, 0.0);
    const double s = 0.5 * hb;  // half-width of the sub-box in reference coordinates
    for (int q = 0; q < 8; ++q) {
        const double xi = lo[0] + s + s * (q & 2 ? g : -g), et = lo[1] + s + s * (q & 1 ? g : -g),
                     ze = lo[2] + s + s * (q & 4 ? g : -g);
        double dN[8][3];
        for (int an = 0; an < 8; ++an) {
            const double sy = (an & 1) ? 1 : -1, sx = (an & 2) ? 1 : -1, sz = (an & 4) ? 1 : -1;
            dN[an][0] = sx * (1 + sy * et) * (1 + sz * ze) / (8 * a);
            dN[an][1] = sy * (1 + sx * xi) * (1 + sz * ze) / (8 * b);
            dN[an][2] = sz * (1 + sx * xi) * (1 + sy * et) / (8 * c);
        }
        const double w = a * b * c * s * s * s;
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 8; ++j) {
                const double *gi = dN[i], *gj = dN[j];
                const double dot = gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2];
                for (int ci = 0; ci < 3; ++ci)
                    for (int cj = 0; cj < 3; ++cj) {
                        // (B'DB)_{ci,cj} = lam gi_ci gj_cj + mu gi_cj gj_ci + mu delta (gi . gj)
                        double v = lam * gi[ci] * gj[cj] + mu * gi[cj] * gj[ci];
                        if (ci == cj) v += mu * dot;
                        k[(3 * i + ci) * 24 + 3 * j + cj] += w * v;
                    }
            }
    }
    for (int i = 0; i < 24; ++i)
        for (int j = i + 1; j < 24; ++j) k[j * 24 + i] = k[i * 24 + j];
}

void element_stiffness(double el, double eh, double ew, double nu, double *k) {
    const double lo[3] = {-1.0, -1.0, -1.0};
    element_stiffness_box(el, eh, ew, nu, lo, 2.0, k);
}

std::mutex &factor_mutex() {
    static std::mutex m;
    return m;
}

int refresh(tmg_solver *h) {
    DevLevel &l0 = h->lev[0];
    const double om = h->mg.omega;
    if (h->nsub == 1)
        hipLaunchKernelGGL(k_modulus, dim3(nblk(l0.ne)), dim3(TPB), 0, h->st, l0.ne, h->xdens, h->E0, h->emin, h->penal,
                           h->E64, h->prec == TMG_MIXED ? h->E32 : nullptr);
    else
        hipLaunchKernelGGL(k_modulus_mr, dim3(nblk(l0.ne * 8)), dim3(TPB), 0, h->st, l0.nelx, l0.nely, l0.nelz, h->xdens,
                           h->E0, h->emin, h->penal, h->E64, h->prec == TMG_MIXED ? h->E32 : nullptr);
    if (h->l1_wht) {
        const DevLevel &l1 = h->lev[1];
        hipLaunchKernelGGL((k_children_moduli<double>), dim3(nblk(l0.ne)), dim3(TPB), 0, h->st, l1.nelx, l1.nely, l1.nelz,
                           h->E64, h->E1_64);
        if (h->prec == TMG_MIXED)
            hipLaunchKernelGGL((k_children_moduli<float>), dim3(nblk(l0.ne)), dim3(TPB), 0, h->st, l1.nelx, l1.nely,
                               l1.nelz, h->E32, h->E1_32);
    }
    hipLaunchKernelGGL(k_diag_fine, dim3(nblk(l0.nn)), dim3(TPB), 0, h->st, l0.nelx, l0.nely, l0.nelz, h->nsub, h->E64,
                       h->K0_64, l0.mask, om, l0.inv64, h->prec == TMG_MIXED ? l0.inv32 : nullptr);
    if (h->fine_stencil)
        hipLaunchKernelGGL((k_stencil_assemble_mr<float>), dim3(nblk(l0.nn * 14)), dim3(TPB), 0, h->st, l0.nelx,
                           l0.nely, l0.nelz, h->E64, h->K0_64, l0.mask, l0.S64,
                           h->prec == TMG_MIXED ? l0.S32 : nullptr);
    for (int l = 1; l < h->L; ++l) {
        DevLevel &lv = h->lev[l];
        double *dst = h->Ahat[(l - 1) & 1];
        if (l == 1) {
            hipLaunchKernelGGL(k_galerkin_fine, dim3(nblk(lv.ne * 64)), dim3(TPB), 0, h->st, lv.nelx, lv.nely, lv.nelz,
                               h->nsub, h->E64, h->K0_64, l0.mask, dst);
        } else {
            hipLaunchKernelGGL(k_galerkin_coarse, dim3(nblk(lv.ne * 64)), dim3(TPB), 0, h->st, lv.nelx, lv.nely,
                               lv.nelz, h->Ahat[l & 1], h->lev[l - 1].mask, dst);
        }
        hipLaunchKernelGGL((k_stencil_assemble<float>), dim3(nblk(lv.nn * 14)), dim3(TPB), 0, h->st, lv.nelx, lv.nely,
                           lv.nelz, dst, lv.mask, lv.S64, h->prec == TMG_MIXED ? lv.S32 : nullptr);
        hipLaunchKernelGGL(k_diag_stencil, dim3(nblk(lv.nn)), dim3(TPB), 0, h->st, lv.nn, lv.S64, lv.mask, om, lv.inv64,
                           h->prec == TMG_MIXED ? lv.inv32 : nullptr);
    }
    // coarsest (Implementation/SetupGMG.m:51-60): multifrontal Cholesky over
    // a nested dissection, or the dense operator, Cholesky and inverse
    DevLevel &lc = h->lev[h->L - 1];
    const int64_t nc = h->nc;
    if (h->debug_fail_first) {
        // test hook for the callers' failure handling: the first build of
        // this handle reports a failed factorization (all kernels ran)
        h->debug_fail_first = false;
        HIPX(hipStreamSynchronize(h->st));
        h->err = "coarse factorization failure forced for testing (TMG_DEBUG_FAIL_FIRST_FACTOR)";
        return TMG_ERR_NOT_SPD;
    }
    // One coarse factorization at a time on the device, process-wide: the
    // rocSOLVER/rocBLAS calls of handles that set up concurrently (the
    // server's slots) never overlap. Held from the enqueue to the stream's
    // completion; the level operators above are finished first, so the wait
    // of another handle covers factorization work only.
    HIPX(hipStreamSynchronize(h->st));
    std::lock_guard<std::mutex> factor_lock(factor_mutex());
    if (h->coarse_nd) {
        TRY(nd_factor(h));
        HIPX(hipStreamSynchronize(h->st));
        TRY(canary_check(h, "by the density update"));
        h->have_op = true;
        return 0;
    }
    hipLaunchKernelGGL((k_fill<double>), dim3(nblk(nc * nc)), dim3(TPB), 0, h->st, nc * nc, 0.0, h->Ainv);
    hipLaunchKernelGGL(k_dense_from_stencil, dim3(nblk(lc.nn * 14)), dim3(TPB), 0, h->st, lc.nx, lc.ny, lc.nz, lc.S64,
                       lc.d_cidx, nc, h->Ainv);
    TRY(check_launch(h));
    if (nc > 0) {
        rocblas_set_stream(h->rb, h->st);  // the handle may be shared between solvers
        rocblas_status s1 = rocsolver_dpotrf(h->rb, rocblas_fill_upper, (rocblas_int)nc, h->Ainv, (rocblas_int)nc, h->d_info);
        int info = 0;
        HIPX(hipMemcpyAsync(&info, h->d_info, sizeof(int), hipMemcpyDeviceToHost, h->st));
        HIPX(hipStreamSynchronize(h->st));
        if (s1 != rocblas_status_success) {
            h->err = "rocsolver_dpotrf failed (status " + std::to_string((int)s1) + ")";
            return TMG_ERR_HIP;
        }
        if (info != 0) {
            h->err = "coarse Galerkin operator is not positive definite (potrf info " + std::to_string(info) + ")";
            return TMG_ERR_NOT_SPD;
        }
        rocblas_status s2 = rocsolver_dpotri(h->rb, rocblas_fill_upper, (rocblas_int)nc, h->Ainv, (rocblas_int)nc, h->d_info);
        if (s2 != rocblas_status_success) {
            h->err = "rocsolver_dpotri failed (status " + std::to_string((int)s2) + ")";
            return TMG_ERR_HIP;
        }
        hipLaunchKernelGGL(k_symmetrize_upper, dim3(nblk(nc * nc)), dim3(TPB), 0, h->st, nc, h->Ainv);
        if (h->coarse_f32)
            hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk(nc * nc)), dim3(TPB), 0, h->st, nc * nc,
                               (const double *)h->Ainv, h->Ainv32, (const int *)nullptr);
    }
    TRY(check_launch(h));
    HIPX(hipStreamSynchronize(h->st));
    TRY(canary_check(h, "by the density update"));
    h->have_op = true;
    return 0;
}

// Host source of an upload of n doubles: the lent pinned staging buffer
// (filled here) when there is one large enough, else hv itself. The staging
// buffer is reused only after the stream has been synchronised.
const double *host_src(tmg_solver *h, const double *hv, int64_t n) {
    if (!h->stage || (size_t)n > h->stage_cap) return hv;
    std::copy(hv, hv + n, h->stage);
    return h->stage;
}

// host free-vector helpers
int upload_free(tmg_solver *h, int l, const double *hv, double *d_full) {
    DevLevel &lv = h->lev[l];
    double *tmp = nullptr;
    TRY(dalloc(h, &tmp, lv.nfree));
    HIPX(hipMemcpyAsync(tmp, host_src(h, hv, lv.nfree), sizeof(double) * lv.nfree, hipMemcpyHostToDevice, h->st));
    hipLaunchKernelGGL((k_scatter_full<double>), dim3(nblk(lv.ndof)), dim3(TPB), 0, h->st, lv.ndof, lv.d_cidx, tmp,
                       d_full, nullptr);
    HIPX(hipStreamSynchronize(h->st));
    dfree(h, tmp);
    return 0;
}

int download_free(tmg_solver *h, int l, const double *d_full, double *hv) {
    DevLevel &lv = h->lev[l];
    double *tmp = nullptr;
    TRY(dalloc(h, &tmp, lv.nfree));
    hipLaunchKernelGGL((k_gather<double>), dim3(nblk(lv.nfree)), dim3(TPB), 0, h->st, lv.nfree, lv.d_free, d_full, tmp,
                       nullptr);
    const bool staged = h->stage && (size_t)lv.nfree <= h->stage_cap;
    HIPX(hipMemcpyAsync(staged ? h->stage : hv, tmp, sizeof(double) * lv.nfree, hipMemcpyDeviceToHost, h->st));
    HIPX(hipStreamSynchronize(h->st));
    if (staged) std::copy(h->stage, h->stage + lv.nfree, hv);
    dfree(h, tmp);
    return 0;
}

// Order-independent checksum (sum of the 64-bit words, modulo 2^64) of n
// doubles in device memory, computed on the handle's stream into its second
// partial-sum buffer; blocks until the value is on the host.
int device_checksum(tmg_solver *h, const double *d, int64_t n, uint64_t *out) {
    const unsigned nb = (unsigned)std::min<int64_t>(std::max<int64_t>(nblk(n), 1), std::min(h->npartial - 1, 1024));
    uint64_t *part = reinterpret_cast<uint64_t *>(h->partial2);
    hipLaunchKernelGGL(k_checksum_partial, dim3(nb), dim3(TPB), 0, h->st, n, reinterpret_cast<const uint64_t *>(d), part);
    hipLaunchKernelGGL(k_checksum_final, dim3(1), dim3(TPB), 0, h->st, (int)nb, part);
    uint64_t v = 0;
    HIPX(hipMemcpyAsync(&v, part + nb, sizeof(uint64_t), hipMemcpyDeviceToHost, h->st));
    HIPX(hipStreamSynchronize(h->st));
    *out = v;
    return 0;
}

uint64_t host_checksum(const double *v, int64_t n) {
    uint64_t s = 0;
    for (int64_t i = 0; i < n; ++i) {
        uint64_t w;
        std::memcpy(&w, v + i, sizeof(w));
        s += w;
    }
    return s;
}

int set_err(tmg_solver *h, int rc, const std::string &m) {
    h->err = m;
    return rc;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
namespace {

// Resources lent by a caller that runs many handles (the server): a stream,
// a rocBLAS handle, a pinned scalar block and a stream-ordered memory pool.
struct ExtRes {
    hipStream_t st = nullptr;
    rocblas_handle rb = nullptr;
    PcgScalars *h_sc = nullptr;
    hipMemPool_t pool = nullptr;
    char *arena = nullptr;            // device block the handle's buffers are carved from (optional)
    size_t arena_cap = 0;
    double *stage = nullptr;
    size_t stage_cap = 0;
};

int setup_impl(const tmg_problem *prob, const tmg_mg_params *mg, int precision, const ExtRes *ext, tmg_solver **out,
               int mr_n = 1) {
    g_setup_err.clear();
    if (!prob || !mg || !out) { g_setup_err = "null argument"; return TMG_ERR_ARG; }
    *out = nullptr;
    if (prob->nelx <= 0 || prob->nely <= 0 || prob->nelz <= 0 || !(prob->el > 0) || !(prob->eh > 0) || !(prob->ew > 0) ||
        !(prob->nu > -1.0 && prob->nu < 0.5) || !(prob->E0 > 0) || prob->emin < 0 || prob->nfree < 0 ||
        (prob->nfree > 0 && !prob->freedofs)) {
        g_setup_err = "invalid problem description";
        return TMG_ERR_ARG;
    }
    if (mg->cycle < 0 || mg->cycle > 2 || mg->smoother != 0 || mg->opchol != 0) {
        g_setup_err = "only cycle=0 (V), 1 (W) or 2 (FMV), smoother=0 (Jacobi), opchol=0 (coarse Cholesky) are supported";
        return TMG_ERR_UNSUPPORTED;
    }
    if (mg->ngrids < 2 || mg->ngrids > MAXL || mg->smoothiter1 < 0 || mg->smoothiter2 < 0 || !(mg->omega > 0) ||
        mg->maxiterMG < 0 || !(mg->tolMG >= 0)) {
        g_setup_err = "invalid multigrid parameters";
        return TMG_ERR_ARG;
    }
    const int d = 1 << (mg->ngrids - 1);
    if (prob->nelx % d || prob->nely % d || prob->nelz % d) {
        g_setup_err = "element counts must be divisible by 2^(ngrids-1) (Implementation/StartStrSLP.m:107-124)";
        return TMG_ERR_ARG;
    }
    if (precision != TMG_FP64 && precision != TMG_MIXED) { g_setup_err = "unknown precision"; return TMG_ERR_ARG; }
    if (mr_n != 1 && mr_n != 2) {
        g_setup_err = "multiresolution is supported with n = 2 (mr.n = mr.d = 2) only";
        return TMG_ERR_UNSUPPORTED;
    }

    tmg_solver *h = new tmg_solver();
    h->nelx = prob->nelx; h->nely = prob->nely; h->nelz = prob->nelz;
    h->el = prob->el; h->eh = prob->eh; h->ew = prob->ew;
    h->E0 = prob->E0; h->nu = prob->nu; h->penal = prob->penal; h->emin = prob->emin;
    h->mg = *mg;
    h->prec = precision;
    h->L = mg->ngrids;
    h->mr = mr_n;
    h->nsub = mr_n * mr_n * mr_n;
    h->pprec = precision;
    // multiresolution: matrix-free in spectral element form (a tenth of the
    // memory traffic of the assembled fine stencil); TMG_MR_STENCIL=1
    // assembles the fine level into a block stencil instead
    h->fine_stencil = h->nsub > 1 && getenv_flag("TMG_MR_STENCIL");
    if (h->nsub > 1) h->fused = false;
    if (const char *ev = getenv("TMG_XCD_REMAP")) h->xcd_remap = ev[0] == '1';
    h->xcd_ebe = getenv_flag("TMG_XCD_EBE") ? 1 : 0;
    if (const char *ev = getenv("TMG_NO_GRAPH")) h->use_graph = ev[0] == '0';
    h->escalate = !getenv_flag("TMG_NO_ESCALATE");
    if (const char *ev = getenv("TMG_FINE_R")) h->fine_r = ev[0] != '0';
    if (const char *ev = getenv("TMG_FIN_FUSED")) h->fin_fused = ev[0] != '0';
    if (const char *ev = getenv("TMG_FUSED_MAXNODES")) h->fused_maxnodes = atoll(ev);
    if (const char *ev = getenv("TMG_ST3_MAXNODES")) h->st3_maxnodes = atoll(ev);
    if (const char *ev = getenv("TMG_EBE_PREFETCH")) h->ebe_prefetch = ev[0] != '0';
    if (const char *ev = getenv("TMG_EBE_LDS")) h->ebe_lds = ev[0] != '0';
    if (const char *ev = getenv("TMG_MR_LDS")) h->mr_lds = ev[0] != '0';
    if (const char *ev = getenv("TMG_FIN_CONSUMER")) h->fin_consumer = ev[0] != '0';
    if (const char *ev = getenv("TMG_PROLONG4")) h->prolong4 = ev[0] != '0';
    if (const char *ev = getenv("TMG_RESTRICT_JZ")) h->restrict_jz = ev[0] != '0';
    h->debug_fail_first = getenv_flag("TMG_DEBUG_FAIL_FIRST_FACTOR");
    if (const char *ev = getenv("TMG_XDEFER")) h->xdefer = ev[0] != '0';
    if (const char *ev = getenv("TMG_STALL_WINDOW")) h->stall_window = std::max(1, atoi(ev));
    if (const char *ev = getenv("TMG_STALL_FACTOR")) h->stall_factor = atof(ev);
    if (const char *ev = getenv("TMG_EBE_UNTILED")) h->ebe_tiled = ev[0] == '0';
    if (const char *ev = getenv("TMG_FUSED_FINE")) h->fused = ev[0] == '1';
    if (const char *ev = getenv("TMG_UNFUSED_COARSE")) h->fused_coarse = ev[0] == '0';
    if (ext) {
        h->st = ext->st;
        h->rb = ext->rb;
        h->h_sc = ext->h_sc;
        h->pool = ext->pool;
        h->arena = ext->arena;
        h->arena_cap = ext->arena ? ext->arena_cap : 0;
        h->stage = ext->stage;
        h->stage_cap = ext->stage_cap;
        h->own_stream = h->own_rb = h->own_hsc = false;
    }
    auto fail = [&](int rc) {
        g_setup_err = h->err;
        destroy_all(h);
        delete h;
        return rc;
    };
    // level grids and free-DOF masks (Prolongation.m:38-46,54: coarse DOF free
    // iff the co-located fine DOF is free); validated before any device call
    std::vector<uint8_t> mask;
    {
        DevLevel &l0 = h->lev[0];
        l0.nelx = prob->nelx; l0.nely = prob->nely; l0.nelz = prob->nelz;
        const int64_t ndof = 3LL * (l0.nelx + 1) * (l0.nely + 1) * (l0.nelz + 1);
        mask.assign(ndof, 0);
        int64_t prev = 0;
        for (int64_t i = 0; i < prob->nfree; ++i) {
            const int64_t fd = prob->freedofs[i];
            if (fd < 1 || fd > ndof || fd <= prev) {
                h->err = "freedofs must be 1-based, in range and strictly increasing";
                return fail(TMG_ERR_ARG);
            }
            prev = fd;
            mask[fd - 1] = 1;
        }
    }
    (void)hipGetDevice(&h->dev);
    if (!h->st && hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) {
        h->err = "hipStreamCreate failed";
        return fail(TMG_ERR_HIP);
    }
    for (int l = 0; l < h->L; ++l) {
        DevLevel &lv = h->lev[l];
        if (l > 0) {
            const DevLevel &lf = h->lev[l - 1];
            lv.nelx = lf.nelx / 2; lv.nely = lf.nely / 2; lv.nelz = lf.nelz / 2;
            const int cnx = lv.nelx + 1, cny = lv.nely + 1, cnz = lv.nelz + 1;
            std::vector<uint8_t> cm(3LL * cnx * cny * cnz, 0);
            for (int iz = 0; iz < cnz; ++iz)
                for (int ix = 0; ix < cnx; ++ix)
                    for (int iy = 0; iy < cny; ++iy) {
                        const int64_t I = (int64_t)iz * cnx * cny + (int64_t)ix * cny + iy;
                        const int64_t fi = (int64_t)(2 * iz) * lf.nx * lf.ny + (int64_t)(2 * ix) * lf.ny + 2 * iy;
                        for (int c = 0; c < 3; ++c) cm[3 * I + c] = mask[3 * fi + c];
                    }
            if (l == 1) {
                // every fixed fine DOF may interpolate only from fixed coarse
                // DOFs: then M_f P M_c = P M_c and the level-1 operator is
                // M_c P' A P M_c, a sum over coarse elements
                bool ok = true;
                for (int fz = 0; ok && fz < lf.nz; ++fz)
                    for (int fx = 0; ok && fx < lf.nx; ++fx)
                        for (int fy = 0; ok && fy < lf.ny; ++fy) {
                            const int64_t fi = (int64_t)fz * lf.nx * lf.ny + (int64_t)fx * lf.ny + fy;
                            for (int c = 0; c < 3; ++c) {
                                if (mask[3 * fi + c]) continue;
                                for (int t = 0; t < 8; ++t) {
                                    const int X = (t & 2) ? (fx + 1) / 2 : fx / 2, Y = (t & 1) ? (fy + 1) / 2 : fy / 2;
                                    const int Z = (t & 4) ? (fz + 1) / 2 : fz / 2;
                                    ok = ok && !cm[3 * ((int64_t)Z * cnx * cny + (int64_t)X * cny + Y) + c];
                                }
                            }
                        }
                h->l1_maskok = ok;
            }
            mask.swap(cm);
        }
        int rc = l == 0 ? alloc_level(h, l, mask, prob->freedofs, prob->nfree) : alloc_level(h, l, mask);
        if (rc) return fail(rc);
    }
    h->nc = h->lev[h->L - 1].nfree;
    const DevLevel &l0 = h->lev[0];
    element_stiffness(h->el, h->eh, h->ew, h->nu, h->ke);
    // the operator's element matrices: Ke, or the 8 sub-element integrands
    // (order z, x, y with y fastest, Implementation/ElemStiffMatrixGQMR.m:54-56)
    std::vector<double> kop(576 * h->nsub);
    if (h->nsub == 1) {
        std::copy(h->ke, h->ke + 576, kop.begin());
    } else {
        for (int sidx = 0; sidx < 8; ++sidx) {
            const int sy = sidx & 1, sx = (sidx >> 1) & 1, sz = sidx >> 2;
            const double lo[3] = {-1.0 + sx, -1.0 + sy, -1.0 + sz};
            element_stiffness_box(h->el, h->eh, h->ew, h->nu, lo, 1.0, kop.data() + 576 * sidx);
        }
    }
    std::vector<float> ke32(kop.size());
    for (size_t i = 0; i < kop.size(); ++i) ke32[i] = (float)kop[i];
    // cube elements: Ke as ten scalars, taken from its own entries and
    // accepted only if every entry is reproduced
    h->cube = false;
    if (h->nsub == 1 && h->el == h->eh && h->el == h->ew && !getenv_flag("TMG_EBE_GENERIC")) {
        double v[10];
        bool seen[10] = {false};
        double kmax = 0.0;
        for (int i = 0; i < 576; ++i) kmax = std::max(kmax, std::fabs(kop[i]));
        for (int a = 0; a < 8; ++a)
            for (int c = 0; c < 3; ++c)
                for (int b = 0; b < 8; ++b)
                    for (int d = 0; d < 3; ++d) {
                        const int k = kc_class(a, c, b, d);
                        if (!seen[k]) { v[k] = kc_sign(a, c, b, d) * kop[(3 * a + c) * 24 + 3 * b + d]; seen[k] = true; }
                    }
        bool ok = true;
        for (int k = 0; k < 10; ++k) ok = ok && seen[k];
        for (int a = 0; ok && a < 8; ++a)
            for (int c = 0; c < 3; ++c)
                for (int b = 0; b < 8; ++b)
                    for (int d = 0; d < 3; ++d)
                        ok = ok && std::fabs(kc_sign(a, c, b, d) * v[kc_class(a, c, b, d)] -
                                             kop[(3 * a + c) * 24 + 3 * b + d]) <= 1e-13 * kmax;
        if (ok) {
            h->cube = true;
            for (int k = 0; k < 10; ++k) { h->kc64.v[k] = v[k]; h->kc32.v[k] = (float)v[k]; }
        }
    }
    // spectral element form: g_cd(o) = s_o(d) Ke[c][3o+d], G = shifted
    // transform of g / 8; accepted only if Ke has the reflection structure
    // (every entry reproduced) and G vanishes outside the WHT_NZ pattern
    h->wht = false;
    const bool wht_allowed = !getenv_flag("TMG_EBE_GENERIC") && !(getenv("TMG_EBE_WHT") && getenv("TMG_EBE_WHT")[0] == '0');
    if (h->nsub == 8 && !h->fine_stencil && wht_allowed) h->wht = wht_mr_table(kop.data(), h->kw64, h->kw32);
    if (h->nsub == 1 && wht_allowed) {
        const int bitc[3] = {1, 0, 2};
        auto sgn = [&](int a, int c) { return ((a >> bitc[c]) & 1) ? -1.0 : 1.0; };
        double kmax = 0.0;
        for (int i = 0; i < 576; ++i) kmax = std::max(kmax, std::fabs(kop[i]));
        double g[3][3][8];
        for (int c = 0; c < 3; ++c)
            for (int d = 0; d < 3; ++d)
                for (int o = 0; o < 8; ++o) g[c][d][o] = sgn(o, d) * kop[c * 24 + 3 * o + d];
        bool ok = true;
        for (int a = 0; ok && a < 8; ++a)
            for (int c = 0; c < 3; ++c)
                for (int b = 0; b < 8; ++b)
                    for (int d = 0; d < 3; ++d)
                        ok = ok && std::fabs(sgn(a, c) * sgn(b, d) * g[c][d][a ^ b] - kop[(3 * a + c) * 24 + 3 * b + d]) <=
                                       1e-13 * kmax;
        double G[3][3][8];
        double gmax = 0.0;
        for (int c = 0; c < 3; ++c)
            for (int d = 0; d < 3; ++d)
                for (int k = 0; k < 8; ++k) {
                    const int m = k ^ (1 << bitc[c]);
                    double sum = 0.0;
                    for (int o = 0; o < 8; ++o) sum += (__builtin_popcount(m & o) & 1 ? -1.0 : 1.0) * g[c][d][o];
                    G[c][d][k] = sum / 8.0;
                    gmax = std::max(gmax, std::fabs(G[c][d][k]));
                }
        for (int c = 0; c < 3; ++c)
            for (int d = 0; d < 3; ++d)
                for (int k = 0; k < 8; ++k) {
                    if ((WHT_NZ[c][d] >> k) & 1) {
                        h->kw64.g[wht_slot(c, d, k)] = G[c][d][k];
                        h->kw32.g[wht_slot(c, d, k)] = (float)G[c][d][k];
                    } else {
                        ok = ok && std::fabs(G[c][d][k]) <= 1e-13 * gmax;
                    }
                }
        h->wht = ok;
    }
    if (h->wht) {
        {
            if (const char *ev = getenv("TMG_WHT_PREFETCH")) h->wht_pf32 = h->wht_pf64 = ev[0] != '0';
            if (const char *ev = getenv("TMG_XCD_WHT")) h->xcd_wht = ev[0] != '0';
            // tile shape: TY in {8, 16, 32} with the fewest tiles per node plane
            int64_t best = -1;
            for (int tyl = 3; tyl <= 5; ++tyl) {
                const int TY = 1 << tyl, TX = TPB / TY;
                const int ntx = (l0.nx + TX - 2) / (TX - 1), nty = (l0.ny + TY - 2) / (TY - 1);
                const int64_t nt = (int64_t)ntx * nty;
                if (best < 0 || nt < best || (nt == best && tyl == 4)) {
                    best = nt; h->wht_tyl = tyl; h->wht_ntx = ntx; h->wht_nty = nty;
                }
            }
            if (const char *ev = getenv("TMG_WHT_TYL")) {
                const int tyl = std::max(3, std::min(atoi(ev), 5)), TY = 1 << tyl, TX = TPB / TY;
                h->wht_tyl = tyl;
                h->wht_ntx = (l0.nx + TX - 2) / (TX - 1);
                h->wht_nty = (l0.ny + TY - 2) / (TY - 1);
            }
        }
    }
    // level 1 matrix-free: single resolution, spectral fine apply, masks that
    // make P'AP element-decomposable (checked while the masks were built)
    h->l1_wht = false;
    if (h->wht && h->nsub == 1 && h->L >= 2 && h->l1_maskok && !getenv_flag("TMG_L1_STENCIL")) {
        std::vector<double> K8(8 * 576);
        galerkin_integrands(h->ke, K8.data());
        h->l1_wht = wht_mr_table(K8.data(), h->kwg64, h->kwg32);
        if (h->l1_wht) wht_tiles(h->lev[1].nx, h->lev[1].ny, h->l1_tyl, h->l1_ntx, h->l1_nty);
    }
    // the LDS-staged apply addresses its vectors through 32-bit buffer offsets
    if ((uint64_t)h->lev[0].ndof * 8 >= ((uint64_t)1 << 32) - 4096) h->ebe_lds = false;
    h->coarse_symv = h->nc >= 16384;
    h->coarse_f32 = precision == TMG_MIXED;
    if (const char *ev = getenv("TMG_COARSE_F32")) h->coarse_f32 = precision == TMG_MIXED && ev[0] != '0';
    if (const char *ev = getenv("TMG_COARSE_SYMV")) h->coarse_symv = ev[0] != '0';
    // large coarse systems: nested dissection and a multifrontal factorization
    // (falls back to the dense inverse when its fronts do not fit the solve)
    int64_t nd_min = 16384;
    if (const char *ev = getenv("TMG_ND_MIN_DOFS")) nd_min = std::max<int64_t>(1, atoll(ev));
    h->coarse_nd = h->nc >= nd_min;
    if (const char *ev = getenv("TMG_COARSE_ND")) h->coarse_nd = ev[0] != '0';
    NdHost ndh;
    if (h->coarse_nd) {
        int leaf = 256;
        if (const char *ev = getenv("TMG_ND_LEAF")) leaf = std::max(1, atoi(ev));
        if (!nd_symbolic(h->lev[h->L - 1], leaf, ndh)) h->coarse_nd = false;
    }
    if (h->coarse_nd) h->coarse_symv = h->coarse_f32 = false;
    const int64_t ncd = h->coarse_nd ? 1 : h->nc;   // dense inverse size (per side)
    h->cs_nb = (int)((h->nc + CS_B - 1) / CS_B);
    std::vector<int4> cs_runs;
    std::vector<int> cs_rs;
    if (h->coarse_symv) {
        // runs of up to CS_RUN consecutive tiles (I, J >= I), row by row
        // (k_coarse_symv_run keeps one run's column products in LDS)
        int run = CS_RUN;
        if (const char *ev = getenv("TMG_CS_RUN")) run = std::max(1, std::min(atoi(ev), CS_RUN));
        for (int I = 0; I < h->cs_nb; ++I) {
            cs_rs.push_back((int)cs_runs.size());
            for (int J = I; J < h->cs_nb; J += run) cs_runs.push_back(make_int4(I, J, std::min(J + run, h->cs_nb), 0));
        }
        cs_rs.push_back((int)cs_runs.size());
        h->cs_nruns = (int)cs_runs.size();
    }
    int rc = 0;
    const int64_t nem = l0.ne * h->nsub;
    if ((rc = dalloc(h, &h->K0_64, 576 * h->nsub)) || (rc = dalloc(h, &h->K0_32, 576 * h->nsub)) ||
        (rc = dalloc(h, &h->kw64_d, WHT_MR_NCOEF)) || (rc = dalloc(h, &h->kw32_d, WHT_MR_NCOEF)) ||
        (rc = dalloc(h, &h->kwg64_d, WHT_MR_NCOEF)) || (rc = dalloc(h, &h->kwg32_d, WHT_MR_NCOEF)) ||
        (rc = dalloc(h, &h->E1_64, h->l1_wht ? l0.ne : 1)) || (rc = dalloc(h, &h->E1_32, h->l1_wht ? l0.ne : 1)) ||
        (rc = dalloc(h, &h->E64, nem)) || (rc = dalloc(h, &h->E32, nem)) || (rc = dalloc(h, &h->xdens, nem)) ||
        (rc = dalloc(h, &h->Ahat[0], h->lev[1].ne * 576)) ||
        (rc = dalloc(h, &h->Ahat[1], h->L > 2 ? h->lev[2].ne * 576 : 1)) || (rc = dalloc(h, &h->Ainv, ncd * ncd)) ||
        (rc = dalloc(h, &h->Ainv32, h->coarse_f32 ? h->nc * h->nc : 1)) ||
        (rc = dalloc(h, &h->cb, h->nc)) || (rc = dalloc(h, &h->cx, h->nc)) || (rc = dalloc(h, &h->d_info, 1)) ||
        (rc = dalloc(h, &h->cs_runs, h->coarse_symv ? (int64_t)h->cs_nruns : 1)) ||
        (rc = dalloc(h, &h->cs_rs, h->coarse_symv ? (int64_t)h->cs_nb + 1 : 1)) ||
        (rc = dalloc(h, &h->cs_part, h->coarse_symv ? ((int64_t)h->cs_nb * (h->cs_nb - 1) / 2 + h->cs_nruns) * CS_B : 1)) ||
        (rc = dalloc(h, &h->px, l0.ndof)) || (rc = dalloc(h, &h->pr, l0.ndof)) || (rc = dalloc(h, &h->pd, l0.ndof)) ||
        (rc = dalloc(h, &h->pd2, l0.ndof)) ||
        (rc = dalloc(h, &h->pq, l0.ndof)) || (rc = dalloc(h, &h->pf, l0.ndof)))
        return fail(rc);
    if (h->coarse_nd && (rc = nd_upload(h, ndh)) != 0) return fail(rc);
    // z-chunk of the fine apply: enough blocks to fill the chip
    {
        const int64_t pb = nblk((int64_t)l0.nx * l0.ny);
        // FP32 sweeps: about 4096 blocks; FP64 apply (half the occupancy):
        // about 2048 blocks, i.e. longer z runs with fewer re-read halo planes
        int zc = (int)((l0.nz * pb + 4095) / 4096);
        h->zchunk = std::max(1, std::min(zc, 16));
        int zc64 = (int)((l0.nz * pb + 2047) / 2048);
        h->zchunk64 = std::max(1, std::min(zc64, 16));
        if (h->wht) {
            // each z chunk recomputes one element layer, and a grid just
            // above one resident round of blocks doubles the time: chunk
            // length from the occupancy of the FP32 post-sweep and the FP64
            // apply (the heaviest launches of each precision)
            const int64_t tp = (int64_t)h->wht_ntx * h->wht_nty;
            int dev = 0, ncu = 0, o32 = 0, o64 = 0;
            (void)hipGetDevice(&dev);
            (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
            const void *k32, *k64;
            if (h->nsub == 8) {
                k32 = reinterpret_cast<const void *>(&k_ebe_wht<float, float, SRC_X, EPI_POSTR, false, 8>);
                k64 = reinterpret_cast<const void *>(&k_ebe_wht<double, double, SRC_X, EPI_DOT, true, 8>);
            } else {
                k32 = h->wht_pf32 ? reinterpret_cast<const void *>(&k_ebe_wht<float, float, SRC_X, EPI_POSTR, true>)
                                  : reinterpret_cast<const void *>(&k_ebe_wht<float, float, SRC_X, EPI_POSTR, false>);
                k64 = h->wht_pf64 ? reinterpret_cast<const void *>(&k_ebe_wht<double, double, SRC_X, EPI_DOT, true>)
                                  : reinterpret_cast<const void *>(&k_ebe_wht<double, double, SRC_X, EPI_DOT, false>);
            }
            if (h->ebe_lds && h->nsub == 1) {
                k32 = reinterpret_cast<const void *>(&k_ebe_lds<float, EPI_POSTR, 1, 1>);
                k64 = reinterpret_cast<const void *>(&k_ebe_lds<double, EPI_DOT, 1, 3>);
            }
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&o32, k32, TPB, 0);
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&o64, k64, TPB, 0);
            ncu = std::max(ncu, 1);
            h->zchunk = wht_zchunk(tp, l0.nz, (int64_t)ncu * std::max(o32, 1));
            h->zchunk64 = wht_zchunk(tp, l0.nz, (int64_t)ncu * std::max(o64, 1));
            if (h->l1_wht) {
                int p32 = 0, p64 = 0;
                (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &p32, reinterpret_cast<const void *>(&k_ebe_wht<float, float, SRC_X, EPI_JACOBI, false, 8>), TPB, 0);
                (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &p64, h->ebe_lds ? reinterpret_cast<const void *>(&k_ebe_lds<double, EPI_JACOBI, 8, 1>)
                                     : reinterpret_cast<const void *>(&k_ebe_wht<double, double, SRC_X, EPI_JACOBI, true, 8>),
                    TPB, 0);
                const int64_t t1 = (int64_t)h->l1_ntx * h->l1_nty;
                h->l1_zc32 = wht_zchunk(t1, h->lev[1].nz, (int64_t)ncu * std::max(p32, 1));
                h->l1_zc64 = wht_zchunk(t1, h->lev[1].nz, (int64_t)ncu * std::max(p64, 1));
                if (const char *ev = getenv("TMG_L1_ZCHUNK")) h->l1_zc32 = h->l1_zc64 = std::max(1, std::min(atoi(ev), 64));
            }
        }
        if (const char *ev = getenv("TMG_ZCHUNK")) h->zchunk = std::max(1, std::min(atoi(ev), 64));
        if (const char *ev = getenv("TMG_ZCHUNK64")) h->zchunk64 = std::max(1, std::min(atoi(ev), 64));
    }
    h->npartial = (int)std::max<int64_t>(std::max(apply_blocks(h, 0, true), apply_blocks(h, 0, false)), 4096);
    if ((rc = dalloc(h, &h->partial, h->npartial)) || (rc = dalloc(h, &h->partial2, h->npartial)) ||
        (rc = dalloc(h, &h->d_sc, 1)))
        return fail(rc);
    if (!h->h_sc && hipHostMalloc((void **)&h->h_sc, sizeof(PcgScalars), hipHostMallocDefault) != hipSuccess) {
        h->err = "hipHostMalloc failed";
        return fail(TMG_ERR_OOM);
    }
    if (hipMemcpyAsync(h->K0_64, kop.data(), sizeof(double) * kop.size(), hipMemcpyHostToDevice, h->st) != hipSuccess ||
        hipMemcpyAsync(h->K0_32, ke32.data(), sizeof(float) * ke32.size(), hipMemcpyHostToDevice, h->st) != hipSuccess ||
        hipMemcpyAsync(h->kw64_d, h->kw64.g, sizeof(h->kw64.g), hipMemcpyHostToDevice, h->st) != hipSuccess ||
        hipMemcpyAsync(h->kw32_d, h->kw32.g, sizeof(h->kw32.g), hipMemcpyHostToDevice, h->st) != hipSuccess ||
        hipMemcpyAsync(h->kwg64_d, h->kwg64.g, sizeof(h->kwg64.g), hipMemcpyHostToDevice, h->st) != hipSuccess ||
        hipMemcpyAsync(h->kwg32_d, h->kwg32.g, sizeof(h->kwg32.g), hipMemcpyHostToDevice, h->st) != hipSuccess ||
        (!cs_runs.empty() && hipMemcpyAsync(h->cs_runs, cs_runs.data(), sizeof(int4) * cs_runs.size(),
                                            hipMemcpyHostToDevice, h->st) != hipSuccess) ||
        (!cs_rs.empty() && hipMemcpyAsync(h->cs_rs, cs_rs.data(), sizeof(int) * cs_rs.size(), hipMemcpyHostToDevice,
                                          h->st) != hipSuccess) ||
        hipStreamSynchronize(h->st) != hipSuccess) {
        h->err = "copy of Ke failed";
        return fail(TMG_ERR_HIP);
    }
    if (!h->rb) {
        if (rocblas_create_handle(&h->rb) != rocblas_status_success) {
            h->err = "rocblas_create_handle failed";
            return fail(TMG_ERR_HIP);
        }
        rocblas_set_stream(h->rb, h->st);
    }
    if ((rc = canary_check(h, "by the set-up")) != 0) return fail(rc);
    *out = h;
    return TMG_OK;
}

}  // namespace

extern "C" {

int tmg_setup(const tmg_problem *prob, const tmg_mg_params *mg, int precision, tmg_solver **out) {
    return setup_impl(prob, mg, precision, nullptr, out);
}

int tmg_setup_mr(const tmg_problem *prob, const tmg_mg_params *mg, int precision, int mr_n, tmg_solver **out) {
    return setup_impl(prob, mg, precision, nullptr, out, mr_n);
}

int tmg_update_density_device(tmg_solver *h, const double *d_x) {
    if (!h || !d_x) return TMG_ERR_ARG;
    HIPX(hipMemcpyAsync(h->xdens, d_x, sizeof(double) * h->lev[0].ne * h->nsub, hipMemcpyDeviceToDevice, h->st));
    return refresh(h);
}

int tmg_update_density(tmg_solver *h, const double *x) {
    if (!h || !x) return TMG_ERR_ARG;
    const int64_t n = h->lev[0].ne * h->nsub;
    // branch-free scan (vectorised): NaN fails both comparisons, +inf the second
    bool ok = true;
    for (int64_t e = 0; e < n; ++e) ok &= (x[e] >= 0.0) & (x[e] <= DBL_MAX);
    if (!ok) return set_err(h, TMG_ERR_ARG, "densities must be finite and >= 0");
    HIPX(hipMemcpyAsync(h->xdens, host_src(h, x, n), sizeof(double) * n, hipMemcpyHostToDevice, h->st));
    return refresh(h);
}

int tmg_solve_device(tmg_solver *h, double *d_x, const double *d_f, int *it, double *relres) {
    if (!h || !d_x || !d_f) return TMG_ERR_ARG;
    const int64_t n = h->lev[0].ndof;
    HIPX(hipMemcpyAsync(h->px, d_x, sizeof(double) * n, hipMemcpyDeviceToDevice, h->st));
    HIPX(hipMemcpyAsync(h->pf, d_f, sizeof(double) * n, hipMemcpyDeviceToDevice, h->st));
    h->have_rhs = true;
    int iters = 0;
    double rr = 0.0;
    int rc = solve_resident(h, &iters, &rr);
    if (rc < 0) return rc;
    HIPX(hipMemcpyAsync(d_x, h->px, sizeof(double) * n, hipMemcpyDeviceToDevice, h->st));
    HIPX(hipStreamSynchronize(h->st));
    if (it) *it = iters;
    if (relres) *relres = rr;
    return rc;
}

int tmg_solve(tmg_solver *h, const double *x0, const double *f, double *x_out, int *it, double *relres) {
    if (!h || !x0 || !f || !x_out) return TMG_ERR_ARG;
    TRY(upload_free(h, 0, x0, h->px));
    TRY(upload_free(h, 0, f, h->pf));
    h->have_rhs = true;
    int iters = 0;
    double rr = 0.0;
    int rc = solve_resident(h, &iters, &rr);
    if (rc < 0) return rc;
    TRY(download_free(h, 0, h->px, x_out));
    if (it) *it = iters;
    if (relres) *relres = rr;
    return rc;
}

int tmg_vcycle(tmg_solver *h, const double *b, double *x) {
    if (!h || !b || !x) return TMG_ERR_ARG;
    if (!h->have_op) return set_err(h, TMG_ERR_STATE, "no operator: call tmg_update_density first");
    DevLevel &l0 = h->lev[0];
    TRY(upload_free(h, 0, b, h->pr));
    const double *s64;
    const float *s32;
    use_cycle_precision(h, h->prec);
    precondition(h, h->pr, &s64, &s32, nullptr);
    if (s32) hipLaunchKernelGGL((k_convert<double, float>), dim3(nblk(l0.ndof)), dim3(TPB), 0, h->st, l0.ndof, s32, h->pq, nullptr);
    else HIPX(hipMemcpyAsync(h->pq, s64, sizeof(double) * l0.ndof, hipMemcpyDeviceToDevice, h->st));
    TRY(check_launch(h));
    return download_free(h, 0, h->pq, x);
}

// Level-wise helpers work in FP64 on scratch copies of the level vectors.
static int level_arg(tmg_solver *h, int level) {
    if (!h) return TMG_ERR_ARG;
    if (level < 1 || level > h->L) return set_err(h, TMG_ERR_ARG, "level out of range");
    if (!h->have_op) return set_err(h, TMG_ERR_STATE, "no operator: call tmg_update_density first");
    return 0;
}

int tmg_apply(tmg_solver *h, int level, const double *x, double *y) {
    TRY(level_arg(h, level));
    const int l = level - 1;
    DevLevel &lv = h->lev[l];
    double *dx = nullptr, *dy = nullptr;
    TRY(dalloc(h, &dx, lv.ndof));
    TRY(dalloc(h, &dy, lv.ndof));
    int rc = upload_free(h, l, x, dx);
    if (!rc) {
        apply<double, EPI_PLAIN>(h, l, dx, nullptr, dy, nullptr, nullptr);
        rc = check_launch(h);
    }
    if (!rc) rc = download_free(h, l, dy, y);
    dfree(h, dx); dfree(h, dy);
    return rc;
}

int tmg_smooth(tmg_solver *h, int level, const double *b, const double *x0, int niter, double *x) {
    TRY(level_arg(h, level));
    if (niter < 0) return set_err(h, TMG_ERR_ARG, "niter < 0");
    const int l = level - 1;
    DevLevel &lv = h->lev[l];
    double *db = nullptr, *xa = nullptr, *xb = nullptr;
    TRY(dalloc(h, &db, lv.ndof));
    TRY(dalloc(h, &xa, lv.ndof));
    TRY(dalloc(h, &xb, lv.ndof));
    int rc = upload_free(h, l, b, db);
    if (!rc) rc = upload_free(h, l, x0, xa);
    if (!rc) {
        for (int s = 0; s < niter; ++s) {  // x = x + M \ (b - A x) (Implementation/Smooth.m:16-18)
            apply<double, EPI_JACOBI>(h, l, xa, db, xb, nullptr, nullptr);
            std::swap(xa, xb);
        }
        rc = check_launch(h);
    }
    if (!rc) rc = download_free(h, l, xa, x);
    dfree(h, db); dfree(h, xa); dfree(h, xb);
    return rc;
}

int tmg_prolong(tmg_solver *h, int level, const double *xc, double *xf) {
    if (!h) return TMG_ERR_ARG;
    if (level < 1 || level >= h->L) return set_err(h, TMG_ERR_ARG, "level out of range");
    DevLevel &lf = h->lev[level - 1], &lc = h->lev[level];
    double *dc = nullptr, *df = nullptr;
    TRY(dalloc(h, &dc, lc.ndof));
    TRY(dalloc(h, &df, lf.ndof));
    int rc = upload_free(h, level, xc, dc);
    if (!rc) {
        prolong_add<double>(h, level - 1, dc, df, 0, nullptr);
        rc = check_launch(h);
    }
    if (!rc) rc = download_free(h, level - 1, df, xf);
    dfree(h, dc); dfree(h, df);
    return rc;
}

int tmg_restrict(tmg_solver *h, int level, const double *xf, double *xc) {
    if (!h) return TMG_ERR_ARG;
    if (level < 1 || level >= h->L) return set_err(h, TMG_ERR_ARG, "level out of range");
    DevLevel &lf = h->lev[level - 1], &lc = h->lev[level];
    double *dc = nullptr, *df = nullptr;
    TRY(dalloc(h, &dc, lc.ndof));
    TRY(dalloc(h, &df, lf.ndof));
    int rc = upload_free(h, level - 1, xf, df);
    if (!rc) {
        hipLaunchKernelGGL((k_restrict<double>), plane_grid(lc.nx, lc.ny, lc.nz), dim3(TPB), 0, h->st, lf.nx, lf.ny, lf.nz, lc.nx,
                           lc.ny, lc.nz, df, lc.mask, dc, nullptr);
        rc = check_launch(h);
    }
    if (!rc) rc = download_free(h, level, dc, xc);
    dfree(h, dc); dfree(h, df);
    return rc;
}

int tmg_level_info(tmg_solver *h, int level, int *nel3, int64_t *nfree, int64_t *freedofs1) {
    if (!h) return TMG_ERR_ARG;
    if (level < 1 || level > h->L) return set_err(h, TMG_ERR_ARG, "level out of range");
    const DevLevel &lv = h->lev[level - 1];
    if (nel3) { nel3[0] = lv.nelx; nel3[1] = lv.nely; nel3[2] = lv.nelz; }
    if (nfree) *nfree = lv.nfree;
    if (freedofs1)
        for (int64_t q = 0; q < lv.nfree; ++q) freedofs1[q] = lv.free_full[q] + 1;
    return 0;
}

int tmg_ngrids(const tmg_solver *h) { return h ? h->L : 0; }

int tmg_time_kernel(tmg_solver *h, int which, int nrep, double *avg_ms) {
    if (!h || !avg_ms || nrep <= 0 || which < 0 || which > 8) return TMG_ERR_ARG;
    if (!h->have_op) return set_err(h, TMG_ERR_STATE, "no operator: call tmg_update_density first");
    if (which == 2 && h->L < 3) return set_err(h, TMG_ERR_ARG, "level 2 is the coarsest level here");
    if (which >= 4 && h->nsub > 1)
        return set_err(h, TMG_ERR_UNSUPPORTED, "fused fine-level kernels have no multiresolution form");
    DevLevel &l0 = h->lev[0];
    const bool mixed = h->prec == TMG_MIXED;
    use_cycle_precision(h, h->prec);
    // deterministic non-trivial inputs
    hipLaunchKernelGGL((k_fill<double>), dim3(nblk(l0.ndof)), dim3(TPB), 0, h->st, l0.ndof, 1.0, h->pd);
    hipLaunchKernelGGL((k_fill<double>), dim3(nblk(l0.ndof)), dim3(TPB), 0, h->st, l0.ndof, 1.0, h->pr);
    HIPX(hipStreamSynchronize(h->st));
    memset(h->h_sc, 0, sizeof(PcgScalars));
    h->h_sc->it = 2;
    h->h_sc->beta = 0.5;
    HIPX(hipMemcpyAsync(h->d_sc, h->h_sc, sizeof(PcgScalars), hipMemcpyHostToDevice, h->st));
    auto launch = [&]() {
        if (which == 0) {
            apply_pcg_q(h, xdefer_ok(h), nullptr);
        } else if (which == 1 || which == 2) {
            const int l = which == 1 ? 0 : 1;
            DevLevel &lv = h->lev[l];
            if (mixed) apply<float, EPI_JACOBI>(h, l, (const float *)lv.x, (const float *)lv.b, (float *)lv.t, nullptr, nullptr);
            else apply<double, EPI_JACOBI>(h, l, (const double *)lv.x, (const double *)lv.b, (double *)lv.t, nullptr, nullptr);
        } else if (which == 7 || which == 8) {
            // the fine-level cycle's residual and post-sweep on the FP64 residual
            if (mixed) {
                EbeArgs<float, float> a = ebe_args<float, float>(h);
                a.x = (const float *)l0.x; a.r = h->pr; a.out = (float *)l0.t; a.partial = h->partial;
                if (which == 7) launch_ebe<float, float, SRC_X, EPI_RESIDR>(h, a);
                else launch_ebe<float, float, SRC_X, EPI_POSTR>(h, a);
            } else {
                EbeArgs<double, double> a = ebe_args<double, double>(h);
                a.x = (const double *)l0.x; a.r = h->pr; a.out = (double *)l0.t; a.partial = h->partial;
                if (which == 7) launch_ebe<double, double, SRC_X, EPI_RESIDR>(h, a);
                else launch_ebe<double, double, SRC_X, EPI_POSTR>(h, a);
            }
        } else if (which == 3) {
            const double *s64;
            const float *s32;
            precondition(h, h->pr, &s64, &s32, nullptr);
        } else if (which == 4 || which == 5) {
            if (mixed) {
                EbeArgs<float, float> a = ebe_args<float, float>(h);
                a.r = h->pr; a.ec = (const float *)h->lev[1].x; a.out = (float *)l0.t; a.partial = h->partial;
                if (which == 4) launch_ebe<float, float, SRC_JZ, EPI_RESIDR>(h, a);
                else launch_ebe<float, float, SRC_POST, EPI_POSTR>(h, a);
            } else {
                EbeArgs<double, double> a = ebe_args<double, double>(h);
                a.r = h->pr; a.ec = (const double *)h->lev[1].x; a.out = (double *)l0.t; a.partial = h->partial;
                if (which == 4) launch_ebe<double, double, SRC_JZ, EPI_RESIDR>(h, a);
                else launch_ebe<double, double, SRC_POST, EPI_POSTR>(h, a);
            }
        } else {
            if (mixed) direction_apply<float>(h, (const float *)l0.x, h->pd, h->pd2, nullptr);
            else direction_apply<double>(h, (const double *)l0.x, h->pd, h->pd2, nullptr);
        }
    };
    // level vectors need finite values for the sweeps (not the coarsest
    // level, whose fixed DOFs must stay zero)
    for (int l = 0; l < 2 && l < h->L - 1; ++l) {
        DevLevel &lv = h->lev[l];
        if (mixed) {
            hipLaunchKernelGGL((k_fill<float>), dim3(nblk(lv.ndof)), dim3(TPB), 0, h->st, lv.ndof, 1.0f, (float *)lv.x);
            hipLaunchKernelGGL((k_fill<float>), dim3(nblk(lv.ndof)), dim3(TPB), 0, h->st, lv.ndof, 1.0f, (float *)lv.b);
        } else {
            hipLaunchKernelGGL((k_fill<double>), dim3(nblk(lv.ndof)), dim3(TPB), 0, h->st, lv.ndof, 1.0, (double *)lv.x);
            hipLaunchKernelGGL((k_fill<double>), dim3(nblk(lv.ndof)), dim3(TPB), 0, h->st, lv.ndof, 1.0, (double *)lv.b);
        }
    }
    launch();  // warm-up
    hipEvent_t e0, e1;
    HIPX(hipEventCreate(&e0));
    HIPX(hipEventCreate(&e1));
    HIPX(hipEventRecord(e0, h->st));
    for (int r = 0; r < nrep; ++r) launch();
    HIPX(hipEventRecord(e1, h->st));
    HIPX(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPX(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    TRY(check_launch(h));
    *avg_ms = (double)ms / nrep;
    return TMG_OK;
}

int tmg_time_in_solve(tmg_solver *h, int which, double *avg_ms, int *nlaunch) {
    if (!h || !avg_ms || (which != KT_APPLY && which != KT_POST)) return TMG_ERR_ARG;
    if (!h->have_op) return set_err(h, TMG_ERR_STATE, "no operator: call tmg_update_density first");
    if (!h->have_rhs) return set_err(h, TMG_ERR_STATE, "no right-hand side: solve once first");
    if (!xdefer_ok(h) || fused_ok(h))
        return set_err(h, TMG_ERR_UNSUPPORTED, "in-solve timing covers the default fine-level schedule only");
    if (h->kt_ev.empty()) {
        h->kt_ev.resize(2 * KT_MAX_PAIRS, nullptr);
        for (auto &e : h->kt_ev) HIPX(hipEventCreate(&e));
    }
    const int64_t n = h->lev[0].ndof;
    hipLaunchKernelGGL((k_fill<double>), dim3(nblk(n)), dim3(TPB), 0, h->st, n, 0.0, h->px);
    const bool graph = h->use_graph;
    h->use_graph = false;
    h->kt_which = which;
    h->kt_n = 0;
    int it = 0;
    double rr = 0.0;
    const int rc = solve_resident(h, &it, &rr);
    h->kt_which = -1;
    h->use_graph = graph;
    if (rc < 0) return rc;
    HIPX(hipStreamSynchronize(h->st));
    // the PCG ran it - 1 iterations (the count starts at 1, as the
    // reference's); launches of the iterations past convergence exit at once
    const int m = std::min(h->kt_n, std::max(it - 1, 0));
    double sum = 0.0;
    for (int i = 0; i < m; ++i) {
        float ms = 0.f;
        HIPX(hipEventElapsedTime(&ms, h->kt_ev[2 * i], h->kt_ev[2 * i + 1]));
        sum += ms;
    }
    *avg_ms = m > 0 ? sum / m : 0.0;
    if (nlaunch) *nlaunch = m;
    return TMG_OK;
}

int64_t tmg_device_bytes(const tmg_solver *h) { return h ? (int64_t)h->peak_bytes : 0; }

int tmg_solve_stats(const tmg_solver *h, int *escalated_at, int *escalations) {
    if (!h) return TMG_ERR_ARG;
    if (escalated_at) *escalated_at = h->escalated_at;
    if (escalations) *escalations = h->escalations;
    return TMG_OK;
}

const char *tmg_last_error(const tmg_solver *h) { return h ? h->err.c_str() : g_setup_err.c_str(); }

void tmg_destroy(tmg_solver *h) {
    if (!h) return;
    destroy_all(h);
    delete h;
}

}  // extern "C"

// ===========================================================================
// Density filter and compliance gradient
// ===========================================================================
struct tmg_filter {
    int nelx, nely, nelz, noff;
    int *off = nullptr;
    double *wt = nullptr;
    uint8_t *fixed = nullptr;
    double *w = nullptr;
    hipStream_t st = nullptr;
    std::string err;
};

namespace {

template <typename T>
void hfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

void filter_free(tmg_filter *f) {
    if (f->st) (void)hipStreamSynchronize(f->st);
    hfree(f->off); hfree(f->wt); hfree(f->fixed); hfree(f->w);
    if (f->st) (void)hipStreamDestroy(f->st);
    f->st = nullptr;
}

// Neighbourhood offsets and weights (Implementation/ElemNeighborhood.m:10-39):
// |dz| <= floor(rmin/ew); |dx| <= floor(sqrt(rmin^2 - dz^2)/el);
// |dy| <= floor(sqrt(rmin^2 - dz^2 - dx^2)/eh); linear cone (rmin - d)/rmin
// (opfilter 2) or Gaussian exp(-d^2 alpha) beta (opfilter 1); self weight 1
// or beta (:60-64).
void filter_offsets(double el, double eh, double ew, double rmin, int opfilter, std::vector<int> &off,
                    std::vector<double> &wt) {
    const int frz = (int)std::floor(rmin / ew);
    const double rmin2 = rmin * rmin;
    const double alpha = 1.0 / (2.0 * (rmin / 3.0) * (rmin / 3.0));
    const double beta = 1.0 / (2.0 * M_PI * (rmin / 3.0));
    for (int adz = 0; adz <= frz; ++adz) {
        const double dz2 = (adz * ew) * (adz * ew);
        const int relx = (int)std::floor(std::sqrt(rmin2 - dz2) / el);
        for (int adx = 0; adx <= relx; ++adx) {
            const double dx2 = (adx * el) * (adx * el);
            const int rely = (int)std::floor(std::sqrt(rmin2 - dz2 - dx2) / eh);
            for (int ady = 0; ady <= rely; ++ady) {
                const double d2 = dz2 + dx2 + (ady * eh) * (ady * eh);
                double w;
                if (d2 == 0.0) w = opfilter == 1 ? beta : 1.0;
                else w = opfilter == 1 ? std::exp(-d2 * alpha) * beta : (rmin - std::sqrt(d2)) / rmin;
                for (int sz = (adz ? -1 : 1); sz <= 1; sz += 2)
                    for (int sx = (adx ? -1 : 1); sx <= 1; sx += 2)
                        for (int sy = (ady ? -1 : 1); sy <= 1; sy += 2) {
                            off.push_back(sx * adx);
                            off.push_back(sy * ady);
                            off.push_back(sz * adz);
                            wt.push_back(w);
                        }
            }
        }
    }
}

#define FHIPX(expr)                                                              \
    do {                                                                         \
        hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess) {                                                  \
            f->err = std::string(#expr) + ": " + hipGetErrorString(e_);          \
            return TMG_ERR_HIP;                                                  \
        }                                                                        \
    } while (0)

}  // namespace

extern "C" {

int tmg_filter_create(int nelx, int nely, int nelz, double el, double eh, double ew, double rmin, int opfilter,
                      const uint8_t *fixed, tmg_filter **out) {
    g_setup_err.clear();
    if (!out || nelx <= 0 || nely <= 0 || nelz <= 0 || !(el > 0) || !(eh > 0) || !(ew > 0) || !(rmin > 0) ||
        (opfilter != 1 && opfilter != 2)) {
        g_setup_err = "invalid filter arguments";
        return TMG_ERR_ARG;
    }
    *out = nullptr;
    tmg_filter *f = new tmg_filter();
    f->nelx = nelx; f->nely = nely; f->nelz = nelz;
    std::vector<int> off;
    std::vector<double> wt;
    filter_offsets(el, eh, ew, rmin, opfilter, off, wt);
    f->noff = (int)wt.size();
    const int64_t ne = (int64_t)nelx * nely * nelz;
    auto bad = [&](const char *m) {
        g_setup_err = m;
        filter_free(f);
        delete f;
        return TMG_ERR_HIP;
    };
    if (hipStreamCreateWithFlags(&f->st, hipStreamNonBlocking) != hipSuccess) return bad("hipStreamCreate failed");
    if (hipMalloc((void **)&f->off, sizeof(int) * off.size()) != hipSuccess ||
        hipMalloc((void **)&f->wt, sizeof(double) * wt.size()) != hipSuccess ||
        hipMalloc((void **)&f->w, sizeof(double) * ne) != hipSuccess)
        return bad("hipMalloc failed");
    if (fixed) {
        if (hipMalloc((void **)&f->fixed, ne) != hipSuccess) return bad("hipMalloc failed");
        if (hipMemcpy(f->fixed, fixed, ne, hipMemcpyHostToDevice) != hipSuccess) return bad("hipMemcpy failed");
    }
    if (hipMemcpy(f->off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(f->wt, wt.data(), sizeof(double) * wt.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bad("hipMemcpy failed");
    // w = row sums of W (Implementation/ElemNeighborhood.m:98)
    hipLaunchKernelGGL(k_filter, dim3(nblk(ne)), dim3(TPB), 0, f->st, nelx, nely, nelz, f->noff, f->off, f->wt,
                       f->fixed, nullptr, nullptr, nullptr, f->w);
    if (hipStreamSynchronize(f->st) != hipSuccess) return bad("filter setup failed");
    *out = f;
    return TMG_OK;
}

static int filter_apply_dev(tmg_filter *f, const double *dx, int trans, double *dy, hipStream_t st) {
    const int64_t ne = (int64_t)f->nelx * f->nely * f->nelz;
    // trans = 0: (W x)./w ; trans = 1: W'(x./w) = W (x./w) (W is symmetric)
    hipLaunchKernelGGL(k_filter, dim3(nblk(ne)), dim3(TPB), 0, st, f->nelx, f->nely, f->nelz, f->noff, f->off, f->wt,
                       f->fixed, dx, trans ? f->w : nullptr, trans ? nullptr : f->w, dy);
    return hipGetLastError() == hipSuccess ? 0 : TMG_ERR_HIP;
}

int tmg_filter_apply(tmg_filter *f, const double *x, int trans, double *y) {
    if (!f || !x || !y) return TMG_ERR_ARG;
    const int64_t ne = (int64_t)f->nelx * f->nely * f->nelz;
    double *dx = nullptr, *dy = nullptr;
    FHIPX(hipMalloc((void **)&dx, sizeof(double) * ne));
    FHIPX(hipMalloc((void **)&dy, sizeof(double) * ne));
    FHIPX(hipMemcpyAsync(dx, x, sizeof(double) * ne, hipMemcpyHostToDevice, f->st));
    int rc = filter_apply_dev(f, dx, trans, dy, f->st);
    if (!rc) FHIPX(hipMemcpyAsync(y, dy, sizeof(double) * ne, hipMemcpyDeviceToHost, f->st));
    FHIPX(hipStreamSynchronize(f->st));
    hfree(dx); hfree(dy);
    return rc;
}

int tmg_filter_rowsum(const tmg_filter *f, double *w) {
    if (!f || !w) return TMG_ERR_ARG;
    const int64_t ne = (int64_t)f->nelx * f->nely * f->nelz;
    if (hipMemcpy(w, f->w, sizeof(double) * ne, hipMemcpyDeviceToHost) != hipSuccess) return TMG_ERR_HIP;
    return TMG_OK;
}

void tmg_filter_destroy(tmg_filter *f) {
    if (!f) return;
    filter_free(f);
    delete f;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Compliance gradient. tmg_grad allocates its device buffers per call;
// tmg_grad_noalloc runs the same kernels on a file-scope device array instead,
// for callers that cannot allocate device memory.
// ---------------------------------------------------------------------------
// Largest gradient tmg_grad_noalloc takes: displacement (3 nnodes doubles),
// densities, g0 and the filtered gradient (n_el * nsub doubles each) and the
// free-element mask (n_el bytes), each section 256-byte aligned, in 96 MiB
// (enough for the 192x96x96 cantilever: 88 MB).
constexpr int64_t GRAD_SCRATCH_DOUBLES = (int64_t)12 << 20;
static __device__ __align__(256) double scratch[GRAD_SCRATCH_DOUBLES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[MAX_DEVICES] = {};

// Device address of device dev's copy of scratch, looked up on the first
// call for that device with the device current (the array's name is not a
// device address on the host). Returns nullptr with err set on failure.
double *scratch_on_device(int dev, std::string &err) {
    if (dev < 0 || dev >= MAX_DEVICES) {
        err = "device index out of range for the gradient scratch";
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess) {
            err = "hipGetDevice failed";
            return nullptr;
        }
        if (cur != dev && hipSetDevice(dev) != hipSuccess) {
            err = "hipSetDevice failed";
            return nullptr;
        }
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return nullptr;
        }
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    return g_scratch_ptr[dev];
}

inline int64_t round32(int64_t n) { return (n + 31) & ~(int64_t)31; }   // 256-byte sections of doubles

// g0 on the device (single resolution or multiresolution) and, with a filter,
// Fgrad = W'(g0 ./ w); the result is copied to the host Fgrad. Device buffers:
// du (ndof), dxe, dg, dF (nd each), dfd (n_el bytes, only with freedens).
int grad_impl(tmg_solver *h, const double *u, const double *x, const uint8_t *freedens, const tmg_filter *filt,
              double *Fgrad, double *du, double *dxe, double *dg, double *dF, uint8_t *dfd) {
    const DevLevel &l0 = h->lev[0];
    const int64_t nd = l0.ne * h->nsub;
    HIPX(hipMemcpyAsync(du, u, sizeof(double) * l0.ndof, hipMemcpyHostToDevice, h->st));
    HIPX(hipMemcpyAsync(dxe, x, sizeof(double) * nd, hipMemcpyHostToDevice, h->st));
    if (freedens) HIPX(hipMemcpyAsync(dfd, freedens, l0.ne, hipMemcpyHostToDevice, h->st));
    if (h->nsub == 1)
        hipLaunchKernelGGL(k_grad0, dim3(nblk(l0.ne)), dim3(TPB), 0, h->st, l0.nelx, l0.nely, l0.nelz, du, dxe,
                           freedens ? dfd : nullptr, h->K0_64, h->penal, h->E0, h->emin, dg);
    else
        hipLaunchKernelGGL(k_grad0_mr, dim3(nblk(nd)), dim3(TPB), 0, h->st, l0.nelx, l0.nely, l0.nelz, du, dxe,
                           freedens ? dfd : nullptr, h->K0_64, h->penal, h->E0, h->emin, dg);
    int rc = 0;
    if (filt) rc = filter_apply_dev(const_cast<tmg_filter *>(filt), dg, 1, dF, h->st);
    if (!rc) HIPX(hipMemcpyAsync(Fgrad, filt ? dF : dg, sizeof(double) * nd, hipMemcpyDeviceToHost, h->st));
    HIPX(hipStreamSynchronize(h->st));
    return rc ? set_err(h, rc, "grad kernel launch failed") : TMG_OK;
}

int grad_args(tmg_solver *h, const double *u, const double *x, const tmg_filter *filt, const double *Fgrad) {
    if (!h || !u || !x || !Fgrad) return TMG_ERR_ARG;
    const DevLevel &l0 = h->lev[0];
    const int m = h->mr;  // gradient lives on the density mesh (m = 2 with multiresolution)
    if (filt && (filt->nelx != m * l0.nelx || filt->nely != m * l0.nely || filt->nelz != m * l0.nelz))
        return set_err(h, TMG_ERR_ARG, "filter grid does not match the density grid");
    return 0;
}

}  // namespace

extern "C" {

int tmg_grad(tmg_solver *h, const double *u, const double *x, const uint8_t *freedens, const tmg_filter *filt,
             double *Fgrad) {
    TRY(grad_args(h, u, x, filt, Fgrad));
    const DevLevel &l0 = h->lev[0];
    const int64_t nd = l0.ne * h->nsub;
    double *du = nullptr, *dxe = nullptr, *dg = nullptr, *dF = nullptr;
    uint8_t *dfd = nullptr;
    int rc = dalloc(h, &du, l0.ndof);
    if (!rc) rc = dalloc(h, &dxe, nd);
    if (!rc) rc = dalloc(h, &dg, nd);
    if (!rc) rc = dalloc(h, &dF, nd);
    if (!rc && freedens) rc = dalloc(h, &dfd, l0.ne);
    if (!rc) rc = grad_impl(h, u, x, freedens, filt, Fgrad, du, dxe, dg, dF, dfd);
    (void)hipStreamSynchronize(h->st);
    dfree(h, du); dfree(h, dxe); dfree(h, dg); dfree(h, dF); dfree(h, dfd);
    return rc;
}

int tmg_grad_noalloc(tmg_solver *h, const double *u, const double *x, const uint8_t *freedens,
                     const tmg_filter *filt, double *Fgrad) {
    TRY(grad_args(h, u, x, filt, Fgrad));
    const DevLevel &l0 = h->lev[0];
    const int64_t nd = l0.ne * h->nsub;
    const int64_t need = round32(l0.ndof) + 3 * round32(nd) + round32((l0.ne + 7) / 8);
    if (need > GRAD_SCRATCH_DOUBLES)
        return set_err(h, TMG_ERR_ARG, "gradient of " + std::to_string(need * 8) + " bytes exceeds the " +
                                           std::to_string(GRAD_SCRATCH_DOUBLES * 8) +
                                           "-byte scratch of tmg_grad_noalloc (use tmg_grad)");
    double *base = scratch_on_device(h->dev, h->err);
    if (!base) return TMG_ERR_HIP;
    double *du = base;
    double *dxe = du + round32(l0.ndof);
    double *dg = dxe + round32(nd);
    double *dF = dg + round32(nd);
    uint8_t *dfd = reinterpret_cast<uint8_t *>(dF + round32(nd));
    return grad_impl(h, u, x, freedens, filt, Fgrad, du, dxe, dg, dF, dfd);
}

}  // extern "C"

// ===========================================================================
// Serving: continuous batching of many solves
// ===========================================================================
#include "tmg_server.inc.hpp"
