// tmg_kernels.hpp -- gfx950 kernels of the multigrid equilibrium solver.
//
// Structured grid of Q1 hexahedra, node-major vectors with the three
// displacement components interleaved (dof = 3*node + c), full length with
// fixed DOFs held at zero. node(ix,iy,iz) = iz*nx*ny + ix*ny + iy,
// elem(ex,ey,ez) = ez*nelx*nely + ex*nely + ey, element-local node
// a = dy + 2*dx + 4*dz (the reference's numbering, Implementation/EnterData.m:130,
// Implementation/DegreesOfFreedom.m:24-54, Implementation/ElemStiffMatrix.m:331-332).
//
// Every reduction is two-stage and in a fixed order, and no kernel scatters
// with atomics, so results are bitwise reproducible run to run and
// independent of what else runs on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tmg {

constexpr int TPB = 256;

// PCG state kept on the device so that an iteration needs no host round trip.
// restart: the iteration whose direction is d = s (beta = 0): 1, or the
// iteration at which a mixed-precision solve switched its preconditioner.
// ticket: blocks of a reduction kernel that finished (the last one completes
// the reduction and resets it to 0).
struct PcgScalars {
    double rho, gamma, alpha, beta, dq, rr, nf, nr, tol;
    int it, maxiter, done, status, restart;
    unsigned int ticket;
};

enum { FIN_NONE = -1, FIN_INIT = 0, FIN_RHO = 1, FIN_ALPHA = 2, FIN_RR = 3 };

// The scalar PCG recurrences (Implementation/SolveMG.m:38-50) once a
// reduction total is known; tot2 only for FIN_INIT (||f||^2).
__device__ inline void fin_update(int mode, double tot, double tot2, PcgScalars *sc) {
    if (mode == FIN_INIT) {               // r = f - A x0; nr = ||r||/||f||; it = 1
        sc->nf = sqrt(tot2);
        sc->rr = tot;
        sc->nr = sqrt(tot) / sc->nf;
        sc->it = 1;
        sc->done = !(sc->nr > sc->tol && sc->it <= sc->maxiter);
    } else if (mode == FIN_RHO) {         // rho = r' s; beta = rho / gamma2
        sc->rho = tot;
        sc->beta = (sc->it == sc->restart) ? 0.0 : tot / sc->gamma;
    } else if (mode == FIN_ALPHA) {       // alpha = rho / (d' q)
        sc->dq = tot;
        sc->alpha = sc->rho / tot;
    } else {                              // gamma2 = rho; nr = ||r||/||f||; it++
        sc->rr = tot;
        sc->gamma = sc->rho;
        sc->nr = sqrt(tot) / sc->nf;
        sc->it = sc->it + 1;
        sc->done = !(sc->nr > sc->tol && sc->it <= sc->maxiter);
    }
}

enum { EPI_PLAIN = 0, EPI_RESID = 1, EPI_JACOBI = 2, EPI_DOT = 3 };

// ---------------------------------------------------------------------------
// Block reduction of one double per thread into partial[block], fixed order.
// ---------------------------------------------------------------------------
__device__ inline void block_reduce_store(double v, double *partial, int64_t slot) {
    __shared__ double red[TPB / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < TPB / 64; ++w) s += red[w];
        partial[slot] = s;
    }
}

// block_reduce_store, then the last of the kernel's nb blocks to finish sums
// partial[0..nb) in a fixed order and applies the PCG recurrence `mode`
// (the work of k_finalize, without its launch). Ticket protocol: each block's
// thread 0 publishes its partial (release fence) before taking a ticket; the
// block holding the last ticket fences (acquire) and reads the partials
// through agent-scope loads, which bypass the non-coherent per-CU cache.
__device__ inline void finish_reduce(int mode, double v, double *partial, int64_t slot, int64_t nb, PcgScalars *sc) {
    block_reduce_store(v, partial, slot);
    __shared__ int last;
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned t = atomicAdd(&sc->ticket, 1u);
        last = t == (unsigned)(nb - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < nb; i += TPB) s += __hip_atomic_load(partial + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __shared__ double red[TPB / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
#pragma unroll
        for (int w = 0; w < TPB / 64; ++w) tot += red[w];
        fin_update(mode, tot, 0.0, sc);
        sc->ticket = 0u;
    }
}

// ---------------------------------------------------------------------------
// Fine-level operator: element-by-element y = sum_e E_e Ke x_e, gathered per
// node (no scatter). Each thread owns one (ix,iy) column of nodes and walks a
// chunk of z planes with a 3x3x3-node register window, so each x value is
// loaded once per column step and every Ke entry is a wave-uniform scalar.
// ---------------------------------------------------------------------------
// Window sources: what the operator is applied to, computed while a node
// plane is loaded.
enum {
    SRC_X = 0,     // x
    SRC_JZ = 1,    // invD * r            (first Jacobi sweep from x0 = 0)
    SRC_POST = 2,  // invD * r + P e_c    (that sweep plus the prolonged coarse correction)
    SRC_DIR = 3    // beta d + s          (PCG direction update; d = s on the first iteration)
};
// Epilogues beyond PLAIN / RESID / JACOBI / DOT (r is the FP64 PCG residual):
enum {
    EPI_RESIDR = 4,  // out = r - A w
    EPI_POSTR = 5,   // out = w + invD (r - A w);   partial r . out
    EPI_DIRDOT = 6,  // out = A w, dnew = w;         partial w . out
    // out = A w at every DOF, fixed ones included (no free-DOF mask);
    // partial w . out. For the PCG's q = A d when its consumer (k_r_update)
    // drops q at the fixed DOFs; w = d is 0 there, so w . out is unchanged.
    // Spectral LDS kernel (k_ebe_lds) only.
    EPI_DOTU = 7
};

template <typename T, typename TS>
struct EbeArgs {
    int nelx, nely, nelz, zchunk;
    const T *x;            // SRC_X input
    const T *E;            // element moduli
    const T *K0;           // 24x24 row-major element stiffness for E = 1
    const uint8_t *mask;   // free-DOF mask
    const T *b;            // right-hand side (RESID, JACOBI)
    const T *invD;         // omega / diag, 0 at fixed DOFs
    const double *r;       // PCG residual (SRC_JZ, SRC_POST, EPI_RESIDR, EPI_POSTR)
    const T *ec;           // coarse correction (SRC_POST)
    int cnx, cny;          // coarse node counts (SRC_POST)
    const TS *s;           // preconditioned residual (SRC_DIR)
    const double *dold;    // previous direction (SRC_DIR)
    double *dnew;          // new direction (EPI_DIRDOT)
    PcgScalars *sc;        // beta, it (SRC_DIR); reduction finish (fin >= 0)
    double dbeta;          // SRC_DIR: beta, 0 on the first iteration since a restart (d = s)
    int fin;               // FIN_* applied by the last block of a DOT-type epilogue
    T *out;
    double *partial;       // per-block partial sums (DOT, POSTR, DIRDOT)
    const int *done;       // skip once a PCG has converged
};

// Fine-level operator: element-by-element y = sum_e E_e Ke w_e, gathered per
// node (no scatter). Each thread owns one (ix, iy) column of nodes and walks a
// chunk of z planes with a 3x3x3-node register window; every Ke entry is a
// wave-uniform scalar. TILED = true stages each new node plane of the block's
// strip through LDS: the block's 256 consecutive plane indices p = ix*ny + iy
// need the plane indices [p0 - ny - 1, p0 + 256 + ny], computed once per node
// (source transform included) and loaded cooperatively, and the next plane's
// loads are issued before the current plane's products. Requires
// ny <= EBE_TILE_MAXNY; TILED = false builds the window directly.
constexpr int EBE_TILE_MAXNY = 127;
constexpr int EBE_TILE_W = TPB + 2 * EBE_TILE_MAXNY + 2;          // nodes per staged plane
constexpr int EBE_TILE_NODES_PER_THREAD = (EBE_TILE_W + TPB - 1) / TPB;

template <typename T, typename TS, int SRC>
__device__ inline void ebe_source(const EbeArgs<T, TS> &a, int jx, int jy, int jz, int64_t node, T (&v)[3]) {
    if (SRC == SRC_X) {
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = a.x[3 * node + c];
    } else if (SRC == SRC_JZ || SRC == SRC_POST) {
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = a.invD[3 * node + c] * (T)a.r[3 * node + c];
        if (SRC == SRC_POST) {
            // trilinear prolongation: 8 terms of weight 1/8; an even fine index
            // maps both of its terms to the co-located coarse node
            const int zl = jz >> 1, zh = (jz + 1) >> 1, xl = jx >> 1, xh = (jx + 1) >> 1;
            const int yl = jy >> 1, yh = (jy + 1) >> 1;
            const int64_t cpl = (int64_t)a.cnx * a.cny;
            T pe[3] = {T(0), T(0), T(0)};
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int Z = (t & 4) ? zh : zl, X = (t & 2) ? xh : xl, Y = (t & 1) ? yh : yl;
                const int64_t J = (int64_t)Z * cpl + (int64_t)X * a.cny + Y;
#pragma unroll
                for (int c = 0; c < 3; ++c) pe[c] += a.ec[3 * J + c];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] += a.mask[3 * node + c] ? T(0.125) * pe[c] : T(0);
        }
    } else {  // SRC_DIR (T = double); dbeta read once per thread (ebe_dir_scalars)
        // branch-free: a select between the two forms compiles to a branch
        // per value with its own load and wait. dbeta = 0 on the first
        // iteration since a restart, where d_old is finite (zeroed at the
        // start of a solve, else the previous direction).
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = (T)fma(a.dbeta, a.dold[3 * node + c], (double)a.s[3 * node + c]);
    }
}

// The direction update's scalars, read once before any store of the kernel
// (read per node, they would be re-loaded after every store that may alias).
template <typename T, typename TS, int SRC>
__device__ inline void ebe_dir_scalars(EbeArgs<T, TS> &a) {
    a.dbeta = 0.0;
    if (SRC == SRC_DIR) a.dbeta = a.sc->it == a.sc->restart ? 0.0 : a.sc->beta;
}

// Workgroups are dispatched to the 8 XCDs round-robin (block b runs on XCD
// b % 8). xcd_block maps the dispatch index to a logical block so that each
// XCD walks one contiguous range of blocks: neighbouring node planes are then
// read through the same L2. A bijection on [0, nb) for any nb.
__device__ inline int64_t xcd_block(int64_t b, int64_t nb) {
    const int64_t q = nb >> 3, r = nb & 7, k = b & 7, m = b >> 3;
    return k * q + (k < r ? k : r) + m;
}

// Cube elements (el = eh = ew): the 576 entries of Ke take only 10 distinct
// magnitudes. With s_k = 1 when local nodes a and b share their coordinate in
// direction k, sigma_k = 2 s_k - 1 and eps_a^k = 2 x_a^k - 1 (Q1 shape
// functions integrated exactly; lambda, mu the Lame constants for E = 1):
//   c == d:  h/36 [(l+2m) sig_c (1+s_k1)(1+s_k2) + m sig_k1 (1+s_c)(1+s_k2) + m sig_k2 (1+s_c)(1+s_k1)]
//   c != d:  h/24 (1+s_e) eps_a^c eps_b^d (l + m sig_c sig_d),  e the third direction.
// kc_class/kc_sign give each entry as sign * v[class]; after unrolling both
// are compile-time constants, so every FMA reads its coefficient from one of
// ten kernel-argument scalars that stay in SGPRs for the whole launch.
template <typename T>
struct KCube {
    T v[10];
};
__host__ __device__ constexpr int kc_pos(int a, int k) { return k == 0 ? (a >> 1) & 1 : (k == 1 ? a & 1 : (a >> 2) & 1); }
__host__ __device__ constexpr int kc_class(int a, int c, int b, int d) {
    const int sc = kc_pos(a, c) == kc_pos(b, c), sd = kc_pos(a, d) == kc_pos(b, d);
    if (c == d) {
        const int k1 = c == 0 ? 1 : 0, k2 = c == 2 ? 1 : 2;
        return sc * 3 + (kc_pos(a, k1) == kc_pos(b, k1)) + (kc_pos(a, k2) == kc_pos(b, k2));
    }
    const int e = 3 - c - d;
    return 6 + (sc == sd ? 0 : 2) + (kc_pos(a, e) == kc_pos(b, e));
}
__host__ __device__ constexpr int kc_sign(int a, int c, int b, int d) {
    return c == d ? 1 : (2 * kc_pos(a, c) - 1) * (2 * kc_pos(b, d) - 1);
}

// NSUB = 1: one modulus per element (E[e]); NSUB = 8: multiresolution with
// 2x2x2 density sub-elements, E[8e + s] and eight sub-element integrands
// K0[576 s + ...] (s = sy + 2 sx + 4 sz), element matrix sum_s E_s K_s.
// KC = 1: cube elements, Ke from the ten scalars of kc_ (NSUB = 1 only);
// KC = 2: the same with the next node plane loaded one step ahead.
template <typename T, typename TS, int SRC, int EPI, bool TILED, int NSUB, int KC = 0>
__global__ __launch_bounds__(TPB) void k_ebe(int nelx_, int nely_, int nelz_, int zchunk_, const T *__restrict__ x_,
                                             const T *__restrict__ E_, const T *__restrict__ K0_,
                                             const uint8_t *__restrict__ mask_, const T *__restrict__ b_,
                                             const T *__restrict__ invD_, const double *__restrict__ r_,
                                             const T *__restrict__ ec_, int cnx_, int cny_, const TS *__restrict__ s_,
                                             const double *__restrict__ dold_, double *__restrict__ dnew_,
                                             PcgScalars *__restrict__ sc_, T *__restrict__ out_,
                                             double *__restrict__ partial_, const int *__restrict__ done_,
                                             KCube<T> kc_, int xcd_, int fin_) {
    static_assert(!KC || NSUB == 1, "cube coefficients describe a single integrand");
    static_assert(KC != 2 || !TILED, "the prefetching form loads planes directly");
    constexpr bool PF = KC == 2;
    if (done_ && *done_) return;
    // logical block (bx = strip of the plane, bz = z chunk)
    int64_t bx = blockIdx.x, bz = blockIdx.y;
    if (xcd_) {
        const int64_t lb = xcd_block((int64_t)blockIdx.y * gridDim.x + blockIdx.x, (int64_t)gridDim.x * gridDim.y);
        bz = lb / gridDim.x;
        bx = lb - bz * gridDim.x;
    }
    // the arguments arrive as non-aliasing kernel parameters and are bundled
    // here only for the helpers
    EbeArgs<T, TS> a;
    a.nelx = nelx_; a.nely = nely_; a.nelz = nelz_; a.zchunk = zchunk_; a.x = x_; a.E = E_; a.K0 = K0_;
    a.mask = mask_; a.b = b_; a.invD = invD_; a.r = r_; a.ec = ec_; a.cnx = cnx_; a.cny = cny_; a.s = s_;
    a.dold = dold_; a.dnew = dnew_; a.sc = sc_; a.out = out_; a.partial = partial_; a.done = done_; a.fin = fin_;
    ebe_dir_scalars<T, TS, SRC>(a);
    const int nelx = a.nelx, nely = a.nely, nelz = a.nelz;
    const int nx = nelx + 1, ny = nely + 1, nz = nelz + 1;
    const int64_t plane = (int64_t)nx * ny;
    const int64_t p0 = bx * TPB + threadIdx.x;
    const bool act = p0 < plane;
    const int64_t p = act ? p0 : plane - 1;
    const int ix = (int)(p / ny);
    const int iy = (int)(p - (int64_t)ix * ny);
    const int z0 = (int)bz * a.zchunk;
    const int z1 = min(z0 + a.zchunk, nz);

    // Neighbour indices are clamped into the grid, so every load is in
    // bounds and needs no predicate; a clamped (nonexistent) neighbour only
    // belongs to nonexistent elements, whose modulus is forced to zero below.
    int cx[3], cy[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        cx[k] = min(max(ix - 1 + k, 0), nx - 1);
        cy[k] = min(max(iy - 1 + k, 0), ny - 1);
    }
    int ecx[2], ecy[2];
    T ewxy[2][2];  // 1 if the element column exists, else 0
#pragma unroll
    for (int o = 0; o < 2; ++o) {
        ecx[o] = min(max(ix - 1 + o, 0), nelx - 1);
        ecy[o] = min(max(iy - 1 + o, 0), nely - 1);
    }
#pragma unroll
    for (int ox = 0; ox < 2; ++ox)
#pragma unroll
        for (int oy = 0; oy < 2; ++oy)
            ewxy[ox][oy] = (ix - 1 + ox >= 0 && ix - 1 + ox < nelx && iy - 1 + oy >= 0 && iy - 1 + oy < nely) ? T(1) : T(0);

    // Element layers are streamed in z: layer L (between node planes L and
    // L+1) adds to its bottom node (plane L), which is then complete, and
    // starts its top node (plane L+1). Only two node planes are live.
    T xw[2][3][3][3];  // [bottom/top plane][x][y][c] around the column
    T ew[2][2][NSUB];  // moduli of the four elements of the current layer

    // ---- staging of node planes (TILED) ----
    __shared__ T tile[TILED ? 2 : 1][TILED ? 3 * EBE_TILE_W : 1];
    const int64_t tile_lo = bx * TPB - ny - 1;                   // first plane index of the strip
    const int wn = TPB + 2 * ny + 2;                             // nodes per staged plane
    T pre[TILED ? EBE_TILE_NODES_PER_THREAD : 1][3];
    auto fetch_plane = [&](int zz) {                             // sources -> registers
        const int jz = min(max(zz, 0), nz - 1);
        const int64_t zb = (int64_t)jz * plane;
#pragma unroll
        for (int r = 0; r < EBE_TILE_NODES_PER_THREAD; ++r) {
            const int w = r * TPB + (int)threadIdx.x;
            const int64_t pn = min(max(tile_lo + w, (int64_t)0), plane - 1);
            const int jx = (int)(pn / ny), jy = (int)(pn - (int64_t)jx * ny);
            if (w < wn) ebe_source<T, TS, SRC>(a, jx, jy, jz, zb + pn, pre[r]);
        }
    };
    auto stash_plane = [&](int buf) {                            // registers -> LDS
#pragma unroll
        for (int r = 0; r < EBE_TILE_NODES_PER_THREAD; ++r) {
            const int w = r * TPB + (int)threadIdx.x;
            if (w < wn) {
#pragma unroll
                for (int c = 0; c < 3; ++c) tile[buf][3 * w + c] = pre[r][c];
            }
        }
    };
    auto window_from_tile = [&](int buf) {                       // LDS -> top plane of the window
#pragma unroll
        for (int xx = 0; xx < 3; ++xx)
#pragma unroll
            for (int yy = 0; yy < 3; ++yy) {
                const int t = (int)((int64_t)cx[xx] * ny + cy[yy] - tile_lo);
#pragma unroll
                for (int c = 0; c < 3; ++c) xw[1][xx][yy][c] = tile[buf][3 * t + c];
            }
    };
    auto load_plane = [&](int zz) {                              // direct (untiled) top-plane load
        const int jz = min(max(zz, 0), nz - 1);
        const int64_t zb = (int64_t)jz * plane;
#pragma unroll
        for (int xx = 0; xx < 3; ++xx)
#pragma unroll
            for (int yy = 0; yy < 3; ++yy)
                ebe_source<T, TS, SRC>(a, cx[xx], cy[yy], jz, zb + (int64_t)cx[xx] * ny + cy[yy], xw[1][xx][yy]);
    };
    T xpre[PF ? 3 : 1][PF ? 3 : 1][3];                           // next plane, in flight during a step (PF)
    auto load_pre = [&](int zz) {
        const int jz = min(max(zz, 0), nz - 1);
        const int64_t zb = (int64_t)jz * plane;
#pragma unroll
        for (int xx = 0; xx < (PF ? 3 : 0); ++xx)
#pragma unroll
            for (int yy = 0; yy < 3; ++yy)
                ebe_source<T, TS, SRC>(a, cx[xx], cy[yy], jz, zb + (int64_t)cx[xx] * ny + cy[yy], xpre[xx][yy]);
    };
    // moduli of layer ez; single resolution prefetches one layer ahead,
    // multiresolution (8x the values) loads at use
    T epre[2][2][1];
    auto fetch_elems = [&](int ez, T (&dst)[2][2][1]) {
        const T vz = (ez >= 0 && ez < nelz) ? T(1) : T(0);
        const int64_t zb = (int64_t)min(max(ez, 0), nelz - 1) * nelx * nely;
#pragma unroll
        for (int ox = 0; ox < 2; ++ox)
#pragma unroll
            for (int oy = 0; oy < 2; ++oy)
                dst[ox][oy][0] = a.E[zb + (int64_t)ecx[ox] * nely + ecy[oy]] * (vz * ewxy[ox][oy]);
    };
    auto load_elems_mr = [&](int ez) {
        const T vz = (ez >= 0 && ez < nelz) ? T(1) : T(0);
        const int64_t zb = (int64_t)min(max(ez, 0), nelz - 1) * nelx * nely;
#pragma unroll
        for (int ox = 0; ox < 2; ++ox)
#pragma unroll
            for (int oy = 0; oy < 2; ++oy) {
                const int64_t e = zb + (int64_t)ecx[ox] * nely + ecy[oy];
#pragma unroll
                for (int s = 0; s < NSUB; ++s) ew[ox][oy][s] = a.E[NSUB * e + s] * (vz * ewxy[ox][oy]);
            }
    };

    // plane z0 - 1 into the top slot (it becomes the bottom of layer z0 - 1)
    if (TILED) {
        fetch_plane(z0 - 1);
        stash_plane(0);
        __syncthreads();
        window_from_tile(0);
        fetch_plane(z0);
    } else {
        load_plane(z0 - 1);
        if (PF) load_pre(z0);
    }
    if (NSUB == 1) fetch_elems(z0 - 1, epre);
    double dacc = 0.0;
    T ybot[3] = {T(0), T(0), T(0)};   // node L   (completed by layer L)
    // Ke is read through the constant address space: the loads are scalar
    // (s_load into SGPRs) and every VALU FMA takes its Ke operand from there.
    typedef const __attribute__((address_space(4))) T *cptr;
    cptr kp = (cptr)a.K0;
    for (int L = z0 - 1; L < z1; ++L) {
        // shift: the top plane becomes the bottom one; plane L+1 is the new top
#pragma unroll
        for (int xx = 0; xx < 3; ++xx)
#pragma unroll
            for (int yy = 0; yy < 3; ++yy)
#pragma unroll
                for (int c = 0; c < 3; ++c) xw[0][xx][yy][c] = xw[1][xx][yy][c];
        if (TILED) {
            // plane L+1 (fetched during the previous step) into LDS, then into
            // the window; buffers alternate so one barrier per step separates
            // every write from the reads of two steps before
            const int buf = (L - z0 + 2) & 1;
            stash_plane(buf);
            __syncthreads();
            window_from_tile(buf);
            fetch_plane(L + 2);       // in flight during this step's products
        } else if (PF) {
#pragma unroll
            for (int xx = 0; xx < 3; ++xx)
#pragma unroll
                for (int yy = 0; yy < 3; ++yy)
#pragma unroll
                    for (int c = 0; c < 3; ++c) xw[1][xx][yy][c] = xpre[PF ? xx : 0][PF ? yy : 0][c];
            load_pre(L + 2);          // in flight during this step's products
        } else {
            load_plane(L + 1);
        }
        if (NSUB == 1) {
#pragma unroll
            for (int ox = 0; ox < 2; ++ox)
#pragma unroll
                for (int oy = 0; oy < 2; ++oy) ew[ox][oy][0] = epre[ox][oy][0];
            fetch_elems(L + 1, epre);
        } else {
            load_elems_mr(L);
        }
        const bool do_bot = L >= z0;        // node L belongs to this chunk
        const bool do_top = L + 1 < z1;     // node L+1 belongs to this chunk
        T ytop[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int ox = 0; ox < 2; ++ox)
#pragma unroll
            for (int oy = 0; oy < 2; ++oy) {
                // the element at offset (ox, oy) of layer L: the column node is
                // its local node (1-oy) + 2(1-ox) at the bottom, +4 at the top
#pragma unroll
                for (int half = 0; half < 2; ++half) {
                    if (half == 0 ? !do_bot : !do_top) continue;
                    const int ai = (1 - oy) + 2 * (1 - ox) + 4 * half;
                    if constexpr (KC) {
                        const T Ee = ew[ox][oy][0];
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            T t = T(0);
#pragma unroll
                            for (int j = 0; j < 24; ++j) {
                                const int bn = j / 3, d = j % 3;
                                const int by = bn & 1, bx = (bn >> 1) & 1, bz = bn >> 2;
                                const T kv = kc_.v[kc_class(ai, c, bn, d)];
                                const T xv = xw[bz][ox + bx][oy + by][d];
                                t = fma(kc_sign(ai, c, bn, d) > 0 ? kv : -kv, xv, t);
                            }
                            if (half == 0) ybot[c] += Ee * t;
                            else ytop[c] += Ee * t;
                        }
                        continue;
                    }
#pragma unroll
                    for (int s = 0; s < NSUB; ++s) {
                        const T Ee = ew[ox][oy][s];
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            // Re-materialise the Ke base per row, ordered after
                            // the previous row's result, so one row (24 scalars)
                            // is live at a time: otherwise the loop-invariant
                            // scalar loads are hoisted en bloc and spill.
                            asm volatile("" : "+s"(kp) : "v"(half ? ytop[(c + 2) % 3] : ybot[(c + 2) % 3]));
                            T t = T(0);
#pragma unroll
                            for (int j = 0; j < 24; ++j) {
                                const int bn = j / 3, d = j % 3;
                                const int by = bn & 1, bx = (bn >> 1) & 1, bz = bn >> 2;
                                t += kp[576 * s + (3 * ai + c) * 24 + j] * xw[bz][ox + bx][oy + by][d];
                            }
                            if (half == 0) ybot[c] += Ee * t;
                            else ytop[c] += Ee * t;
                        }
                    }
                }
            }
        if (do_bot && act) {
            const int64_t node = (int64_t)L * plane + p;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int64_t dof = 3 * node + c;
                const bool m = a.mask[dof] != 0;
                const T w = xw[0][1][1][c];
                const T y = ybot[c];
                if (EPI == EPI_PLAIN) {
                    a.out[dof] = m ? y : T(0);
                } else if (EPI == EPI_RESID) {
                    a.out[dof] = m ? (a.b[dof] - y) : T(0);
                } else if (EPI == EPI_JACOBI) {
                    a.out[dof] = m ? (w + a.invD[dof] * (a.b[dof] - y)) : T(0);
                } else if (EPI == EPI_DOT) {
                    const T v = m ? y : T(0);
                    a.out[dof] = v;
                    dacc += (double)w * (double)v;
                } else if (EPI == EPI_RESIDR) {
                    a.out[dof] = m ? ((T)a.r[dof] - y) : T(0);
                } else if (EPI == EPI_POSTR) {
                    const double rv = a.r[dof];
                    const T v = m ? (w + a.invD[dof] * ((T)rv - y)) : T(0);
                    a.out[dof] = v;
                    dacc += rv * (double)v;
                } else {  // EPI_DIRDOT
                    const T v = m ? y : T(0);
                    a.out[dof] = v;
                    a.dnew[dof] = (double)w;
                    dacc += (double)w * (double)v;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) ybot[c] = ytop[c];
    }
    if (EPI == EPI_DOT || EPI == EPI_POSTR || EPI == EPI_DIRDOT) {
        if (a.fin >= 0)
            finish_reduce(a.fin, dacc, a.partial, bz * gridDim.x + bx, (int64_t)gridDim.x * gridDim.y, a.sc);
        else
            block_reduce_store(dacc, a.partial, bz * gridDim.x + bx);
    }
}

// ---------------------------------------------------------------------------
// Fine-level operator in spectral element form.
//
// The Q1 element matrix of a box is invariant under reflection of the box in
// each coordinate plane (a reflection in direction k maps local node a to
// a ^ e_k and flips displacement component k), so
//     Ke[3a+c][3b+d] = s_a(c) s_b(d) g_cd(a ^ b),   s_a(c) = (-1)^(bit e_c of a),
// with e_x = 2, e_y = 1, e_z = 4 the bits of a = dy + 2dx + 4dz. The 8-point
// Walsh-Hadamard transform H (H[k][b] = (-1)^popcount(k & b)) diagonalises
// every dyadic convolution, and the sign factors only shift frequencies:
//     (H y_c)(k) = sum_d G[c][d][k] (H x_d)(k ^ e_c ^ e_d),
//     G[c][d][k] = (H g_cd)(k ^ e_c) / 8,       y_c = H (H y_c).
// Only 45 of the 72 coefficients G are nonzero (WHT_NZ), so an element
// product costs two 24-value transforms and 45 multiply-adds instead of the
// 576 of a dense Ke. The transforms are separable: the in-plane 2x2 part of a
// node plane is formed once and serves the element layers below and above it,
// and the two layers' contributions to a node plane are summed before the
// plane's inverse transform.
//
// Work split: one thread per element column (ex, ey) of a TX x TY tile
// (TX * TY = TPB, y fastest), streaming the element layers of a z chunk.
// Node (ex, ey) sums corner (dx,dy) = (0,0) of its own element column with
// corners (0,1), (1,0), (1,1) of the columns at (ex, ey-1), (ex-1, ey) and
// (ex-1, ey-1), exchanged through LDS. Tiles overlap by one column in x and
// y; the first row and column of a tile only compute (halo). Elements
// outside the grid have modulus 0 and node loads are clamped into the grid,
// so their contributions are exact zeros. Every sum has a fixed order, so
// the result is bitwise reproducible.
// ---------------------------------------------------------------------------
constexpr unsigned char WHT_NZ[3][3] = {{0xFE, 0x66, 0x3C}, {0x66, 0xFE, 0x5A}, {0x3C, 0x5A, 0xFE}};
constexpr int WHT_BIT[3] = {2, 1, 4};
constexpr int WHT_NCOEF = 45;
__host__ __device__ constexpr int wht_slot(int c, int d, int k) {
    int s = 0;
    for (int i = 0; i < 3 * c + d; ++i)
        for (int kk = 0; kk < 8; ++kk) s += (WHT_NZ[i / 3][i % 3] >> kk) & 1;
    for (int kk = 0; kk < k; ++kk) s += (WHT_NZ[c][d] >> kk) & 1;
    return s;
}
// Multiresolution (2x2x2 density sub-elements, element matrix sum_s E_s K_s):
// each sub-element integrand is the reflection of sub-element 0's, so in the
// same basis the coupling becomes
//     (H y_c)(k) = sum_{d,l} A[c][d][k][l] (H E)(k ^ l ^ e_c ^ e_d) (H x_d)(l)
// with H E the transform of the 8 sub-element moduli (s = sy + 2sx + 4sz) and
// a fixed table A. Its nonzero pattern, identical for every box: row c*8+k,
// bit 8d+l (279 entries; the rows and columns of the translation mode k = 0
// are empty). The host derives A from the integrands and checks both the
// structure and this pattern.
constexpr uint32_t WHT_MR_NZ[24] = {0x000000, 0x00CCAA, 0xF0AACC, 0xF0EEEE, 0xCC00F0, 0xCCCCFA, 0xFCAAFC, 0xFCEEFE,
                                    0x000000, 0xF0AACC, 0x00CCAA, 0xF0EEEE, 0xAAF000, 0xFAFACC, 0xAAFCAA, 0xFAFEEE,
                                    0x000000, 0xAAF000, 0xCC00F0, 0xEEF0F0, 0xF0AACC, 0xFAFACC, 0xFCAAFC, 0xFEFAFC};
constexpr int WHT_MR_NCOEF = 279;
// slot of coefficient (row 8c+k, column 8d+l) in the table, -1 where zero;
// built at compile time so that unrolled kernels index it with constants
struct WhtMrSlots {
    short s[24][24];
};
constexpr WhtMrSlots wht_mr_slots() {
    WhtMrSlots t{};
    int n = 0;
    for (int r = 0; r < 24; ++r)
        for (int b = 0; b < 24; ++b) {
            const bool nz = (WHT_MR_NZ[r] >> b) & 1;
            t.s[r][b] = (short)(nz ? n : -1);
            n += nz;
        }
    return t;
}
constexpr WhtMrSlots WHT_MR_SLOT = wht_mr_slots();
static_assert(WHT_MR_SLOT.s[23][23] == WHT_MR_NCOEF - 1, "279 nonzero coefficients");
__host__ __device__ constexpr int wht_mr_slot(int c, int k, int d, int l) { return WHT_MR_SLOT.s[8 * c + k][8 * d + l]; }
// host-side coefficient tables (single resolution uses the first WHT_NCOEF)
template <typename T>
struct KWht {
    T g[WHT_MR_NCOEF];
};
// Smallest z chunk of the spectral apply whose grid fits in one resident
// round (cap blocks), else the chunk minimising rounds x (layers per block).
inline int wht_zchunk(int64_t tiles, int nz, int64_t cap) {
    int best = 2;
    int64_t bcost = -1;
    for (int zc = 2; zc <= 32; ++zc) {
        const int64_t blocks = tiles * ((nz + zc - 1) / zc);
        const int64_t rounds = (blocks + cap - 1) / (cap > 0 ? cap : 1);
        const int64_t cost = rounds * (zc + 1);
        if (bcost < 0 || cost < bcost) { bcost = cost; best = zc; }
    }
    return best;
}

// 2x2 transform of a node-plane patch, v[b] with b = dy + 2dx.
template <typename T>
__device__ inline void wht4(const T (&v)[4], T (&P)[4]) {
    const T s01 = v[0] + v[1], d01 = v[0] - v[1], s23 = v[2] + v[3], d23 = v[2] - v[3];
    P[0] = s01 + s23;
    P[1] = d01 + d23;
    P[2] = s01 - s23;
    P[3] = d01 - d23;
}

// tyl: log2 of the tile's y extent TY (TX = TPB / TY); ntx: tiles in x;
// kw_: the coefficients in slot order (G, or A for NSUB = 8). PF: the loads of
// the next node plane are issued before the current layer's arithmetic.
// NSUB = 8: multiresolution, moduli E[8e + s].
template <typename T, typename TS, int SRC, int EPI, bool PF, int NSUB = 1>
__global__ __launch_bounds__(TPB) void k_ebe_wht(int nelx_, int nely_, int nelz_, int zchunk_, int tyl_, int ntx_,
                                                 const T *__restrict__ x_, const T *__restrict__ E_,
                                                 const uint8_t *__restrict__ mask_, const T *__restrict__ b_,
                                                 const T *__restrict__ invD_, const double *__restrict__ r_,
                                                 const T *__restrict__ ec_, int cnx_, int cny_,
                                                 const TS *__restrict__ s_, const double *__restrict__ dold_,
                                                 double *__restrict__ dnew_, PcgScalars *__restrict__ sc_,
                                                 T *__restrict__ out_, double *__restrict__ partial_,
                                                 const int *__restrict__ done_, const T *__restrict__ kw_, int xcd_) {
    if (done_ && *done_) return;
    EbeArgs<T, TS> a;
    a.nelx = nelx_; a.nely = nely_; a.nelz = nelz_; a.zchunk = zchunk_; a.x = x_; a.E = E_; a.K0 = nullptr;
    a.mask = mask_; a.b = b_; a.invD = invD_; a.r = r_; a.ec = ec_; a.cnx = cnx_; a.cny = cny_; a.s = s_;
    a.dold = dold_; a.dnew = dnew_; a.sc = sc_; a.out = out_; a.partial = partial_; a.done = done_; a.fin = FIN_NONE;
    ebe_dir_scalars<T, TS, SRC>(a);
    const int nelx = nelx_, nely = nely_, nelz = nelz_;
    const int nx = nelx + 1, ny = nely + 1, nz = nelz + 1;
    const int64_t plane = (int64_t)nx * ny;
    const int TY = 1 << tyl_, TX = TPB >> tyl_;
    const int t = (int)threadIdx.x;
    const int tx = t >> tyl_, ty = t & (TY - 1);
    // logical block: tile (bt) and z chunk (bz); with xcd_ the blocks of one
    // XCD take a contiguous range, so neighbouring tiles share an L2
    int bt = (int)blockIdx.x, bz = (int)blockIdx.y;
    if (xcd_) {
        const int64_t lb = xcd_block((int64_t)blockIdx.y * gridDim.x + blockIdx.x, (int64_t)gridDim.x * gridDim.y);
        bz = (int)(lb / gridDim.x);
        bt = (int)(lb - (int64_t)bz * gridDim.x);
    }
    const int bxx = bt % ntx_, byy = bt / ntx_;
    const int ix = bxx * (TX - 1) + tx - 1, iy = byy * (TY - 1) + ty - 1;   // element column = its (0,0) node
    const bool out_node = tx >= 1 && ty >= 1 && ix < nx && iy < ny;
    const bool ecol = ix >= 0 && ix < nelx && iy >= 0 && iy < nely;
    const int z0 = bz * zchunk_;
    const int z1 = min(z0 + zchunk_, nz);
    // patch nodes b = dy + 2dx, clamped into the grid; in-plane offsets fit 32 bits
    int px[2], py[2];
#pragma unroll
    for (int o = 0; o < 2; ++o) {
        px[o] = min(max(ix + o, 0), nx - 1);
        py[o] = min(max(iy + o, 0), ny - 1);
    }
    int poff[4];
#pragma unroll
    for (int bb = 0; bb < 4; ++bb) poff[bb] = px[bb >> 1] * ny + py[bb & 1];
    const int ecol_idx = ecol ? ix * nely + iy : 0;
    const int own = poff[0];

    __shared__ T xch[2][3][3][TPB];   // [buffer][corner - 1][component][thread]

    T Pb[3][4], Pt[3][4];            // plane transforms below / above the current layer
    T wb[3], wt[3];                  // source values of node (ix, iy) on those planes
    T nv[4][3];                      // raw source values of the next plane
    auto load_raw = [&](int zz) {
        const int jz = min(max(zz, 0), nz - 1);
        const int64_t zb = (int64_t)jz * plane;
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) ebe_source<T, TS, SRC>(a, px[bb >> 1], py[bb & 1], jz, zb + poff[bb], nv[bb]);
    };
    auto transform = [&](T (&P)[3][4], T (&w)[3]) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T vc[4] = {nv[0][c], nv[1][c], nv[2][c], nv[3][c]};
            wht4(vc, P[c]);
            w[c] = nv[0][c];
        }
    };
    auto modulus = [&](int ez, T (&dst)[NSUB]) {
        const bool ok = ecol && ez >= 0 && ez < nelz;
        const int64_t e = (int64_t)min(max(ez, 0), nelz - 1) * nelx * nely + ecol_idx;
#pragma unroll
        for (int q = 0; q < NSUB; ++q) {
            const T v = E_[NSUB * e + q];
            dst[q] = ok ? v : T(0);
        }
    };
    // coefficients through the constant address space (scalar loads); the
    // base is re-materialised per component so that only a third of them
    // occupies SGPRs at a time
    typedef const __attribute__((address_space(4))) T *cptr;
    cptr kp = (cptr)kw_;

    load_raw(z0 - 1);
    transform(Pt, wt);
    if (PF) load_raw(z0);
    T Epre[NSUB], Ee[NSUB];
    modulus(z0 - 1, Epre);
    T Qtp[3][4];                     // top-plane spectra of the previous layer
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 4; ++k) Qtp[c][k] = T(0);
    double dacc = 0.0;
    for (int L = z0 - 1; L < z1; ++L) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            wb[c] = wt[c];
#pragma unroll
            for (int k = 0; k < 4; ++k) Pb[c][k] = Pt[c][k];
        }
        if (!PF) load_raw(L + 1);
        transform(Pt, wt);               // plane L + 1
        if (PF) load_raw(L + 2);         // in flight during this layer's arithmetic
#pragma unroll
        for (int q = 0; q < NSUB; ++q) Ee[q] = Epre[q];
        modulus(L + 1, Epre);
        // element spectrum x^(k), k = kxy + 4 kz (x^(0) is never used)
        T xh[3][8];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                xh[d][k] = Pb[d][k] + Pt[d][k];
                xh[d][k + 4] = Pb[d][k] - Pt[d][k];
            }
        // y^ = E G x^ (y^(0) = 0: translations are in the null space)
        T yh[3][8];
        if constexpr (NSUB == 1) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (c > 0) asm volatile("" : "+s"(kp) : "v"(yh[c - 1][7]));
                yh[c][0] = T(0);
#pragma unroll
                for (int k = 1; k < 8; ++k) {
                    T acc = T(0);
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                        if ((WHT_NZ[c][d] >> k) & 1)
                            acc = fma(kp[wht_slot(c, d, k)], xh[d][k ^ WHT_BIT[c] ^ WHT_BIT[d]], acc);
                    yh[c][k] = Ee[0] * acc;
                }
            }
        } else {
            // transform of the sub-element moduli, then each spectral row as
            // sum over modes m of (H E)(m) times its group of coefficients
            T Eh[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) Eh[q] = Ee[q];
#pragma unroll
            for (int bit = 1; bit < 8; bit <<= 1)
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (!(q & bit)) {
                        const T u = Eh[q], v = Eh[q | bit];
                        Eh[q] = u + v;
                        Eh[q | bit] = u - v;
                    }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                yh[c][0] = T(0);
#pragma unroll
                for (int k = 1; k < 8; ++k) {
                    // only one row's coefficients occupy SGPRs at a time
                    asm volatile("" : "+s"(kp) : "v"(k > 1 ? yh[c][k - 1] : (c > 0 ? yh[c - 1][7] : Eh[0])));
                    T grp[8];
                    bool has[8] = {false, false, false, false, false, false, false, false};
#pragma unroll
                    for (int d = 0; d < 3; ++d)
#pragma unroll
                        for (int l = 1; l < 8; ++l)
                            if ((WHT_MR_NZ[8 * c + k] >> (8 * d + l)) & 1) {
                                const int m = k ^ l ^ WHT_BIT[c] ^ WHT_BIT[d];
                                const T coef = kp[wht_mr_slot(c, k, d, l)];
                                grp[m] = has[m] ? fma(coef, xh[d][l], grp[m]) : coef * xh[d][l];
                                has[m] = true;
                            }
                    T acc = T(0);
#pragma unroll
                    for (int m = 0; m < 8; ++m)
                        if (has[m]) acc = fma(Eh[m], grp[m], acc);
                    yh[c][k] = acc;
                }
            }
        }
        asm volatile("" : "+s"(kp) : "v"(yh[2][7]));
        // inverse in z: bottom / top plane spectra; the bottom one completes
        // node plane L together with the previous layer's top one
        T cv[3][4];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T S[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                S[k] = Qtp[c][k] + (yh[c][k] + yh[c][k + 4]);
                Qtp[c][k] = yh[c][k] - yh[c][k + 4];
            }
            wht4(S, cv[c]);
        }
        if (L < z0) continue;   // uniform across the block: no barrier is skipped unevenly
        const int buf = (L - z0) & 1;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int c = 0; c < 3; ++c) xch[buf][j][c][t] = cv[c][j + 1];
        __syncthreads();
        if (out_node) {
            const int64_t node = (int64_t)L * plane + own;   // 64-bit: full-grid DOF index
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T y = cv[c][0] + xch[buf][0][c][t - 1] + xch[buf][1][c][t - TY] + xch[buf][2][c][t - TY - 1];
                const int64_t dof = 3 * node + c;
                const bool m = a.mask[dof] != 0;
                const T w = wb[c];
                if (EPI == EPI_PLAIN) {
                    a.out[dof] = m ? y : T(0);
                } else if (EPI == EPI_RESID) {
                    a.out[dof] = m ? (a.b[dof] - y) : T(0);
                } else if (EPI == EPI_JACOBI) {
                    a.out[dof] = m ? (w + a.invD[dof] * (a.b[dof] - y)) : T(0);
                } else if (EPI == EPI_DOT) {
                    const T v = m ? y : T(0);
                    a.out[dof] = v;
                    dacc += (double)w * (double)v;
                } else if (EPI == EPI_RESIDR) {
                    a.out[dof] = m ? ((T)a.r[dof] - y) : T(0);
                } else if (EPI == EPI_POSTR) {
                    const double rv = a.r[dof];
                    const T v = m ? (w + a.invD[dof] * ((T)rv - y)) : T(0);
                    a.out[dof] = v;
                    dacc += rv * (double)v;
                } else {  // EPI_DIRDOT
                    const T v = m ? y : T(0);
                    a.out[dof] = v;
                    a.dnew[dof] = (double)w;
                    dacc += (double)w * (double)v;
                }
            }
        }
    }
    // partial slots in logical (spatial) order: the reduction order does
    // not depend on the block mapping
    if (EPI == EPI_DOT || EPI == EPI_POSTR || EPI == EPI_DIRDOT)
        block_reduce_store(dacc, a.partial, (int64_t)bz * gridDim.x + bt);
}

// ---------------------------------------------------------------------------
// The spectral element apply with the node plane staged through LDS
// (k_ebe_lds; same arithmetic as k_ebe_wht, SRC_X sources). Each thread loads
// only its own node of a plane (threads 0..NQ-TPB-1 also one halo node of the
// tile's (TX+1) x (TY+1)-node region) and writes it to LDS; the 2x2 patch of
// an element column is then read from LDS instead of being loaded four times
// from memory. All memory accesses are buffer loads and stores: per-thread
// 32-bit offsets fixed for the launch plus one scalar plane offset per step
// (no per-step 64-bit address arithmetic; out-of-range buffer accesses read 0
// and are dropped, so every access stays inside its buffer). The next plane,
// the next layer's moduli and the next output node's epilogue operands are
// loaded one step ahead. One barrier per step: the exchange of plane L
// between neighbouring element columns and the staging of plane L + 2 share it
// (double-buffered). The exchange is separable: the y neighbour's part moves
// one lane (DPP), the x neighbour's part (one value per component) goes
// through LDS. Requires every
// vector to be smaller than 4 GiB (checked by the host). Jacobi-type
// epilogues (JACOBI, POSTR) do not read the free-DOF mask: at fixed DOFs
// invD = 0 and the source value is 0, so the sweep writes 0 there; DOTU
// leaves the masking to its consumer.
// WPE: waves per SIMD the register allocation must allow (1 = no limit).
// ---------------------------------------------------------------------------
typedef unsigned int bu2 __attribute__((__vector_size__(8)));    // operand types of the buffer builtins
typedef unsigned int bu4 __attribute__((__vector_size__(16)));
typedef float f2v __attribute__((ext_vector_type(2)));            // element access goes through these
typedef double d2v __attribute__((ext_vector_type(2)));
typedef __amdgpu_buffer_rsrc_t brsrc_t;

__device__ inline brsrc_t buf_rsrc(const void *p, uint32_t nbytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)nbytes, 0x00020000);
}
// three consecutive values at byte offset v + s
__device__ inline void bld3(brsrc_t r, uint32_t v, uint32_t s, double (&o)[3]) {
    const d2v a = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(r, v, s, 0));
    o[0] = a.x;
    o[1] = a.y;
    o[2] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, v + 16, s, 0));
}
__device__ inline void bld3(brsrc_t r, uint32_t v, uint32_t s, float (&o)[3]) {
    const f2v a = __builtin_bit_cast(f2v, __builtin_amdgcn_raw_buffer_load_b64(r, v, s, 0));
    o[0] = a.x;
    o[1] = a.y;
    o[2] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, v + 8, s, 0));
}
__device__ inline void bst3(brsrc_t r, uint32_t v, uint32_t s, const double (&o)[3]) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(bu4, d2v{o[0], o[1]}), r, v, s, 0);
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(bu2, o[2]), r, v + 16, s, 0);
}
__device__ inline void bst3(brsrc_t r, uint32_t v, uint32_t s, const float (&o)[3]) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(bu2, f2v{o[0], o[1]}), r, v, s, 0);
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, o[2]), r, v + 8, s, 0);
}
__device__ inline double bld1(brsrc_t r, uint32_t v, uint32_t s, double) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, v, s, 0));
}
__device__ inline float bld1(brsrc_t r, uint32_t v, uint32_t s, float) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, v, s, 0));
}

// Value of lane l - 1 (0 in lane 0), DPP wave_shr:1
__device__ inline float wave_shr1(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, true));
}
__device__ inline double wave_shr1(double v) {
    int2 iv = __builtin_bit_cast(int2, v);
    iv.x = __builtin_amdgcn_update_dpp(0, iv.x, 0x138, 0xF, 0xF, true);
    iv.y = __builtin_amdgcn_update_dpp(0, iv.y, 0x138, 0xF, 0xF, true);
    return __builtin_bit_cast(double, iv);
}

constexpr int LDS_REGION_MAX = 300;   // (TX + 1)(TY + 1) for TY in {8, 16, 32}: at most 297

template <typename T>
struct EpiOps {
    unsigned m[3];
    T b[3], inv[3];
    double r[3];
};

template <typename T, int EPI, int NSUB = 1, int WPE = 1>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(WPE))) void k_ebe_lds(int nelx, int nely, int nelz, int zchunk, int tyl, int ntx,
                                                 const T *__restrict__ x_, const T *__restrict__ E_,
                                                 const uint8_t *__restrict__ mask_, const T *__restrict__ b_,
                                                 const T *__restrict__ invD_, const double *__restrict__ r_,
                                                 T *__restrict__ out_, double *__restrict__ partial_,
                                                 const int *__restrict__ done_, const T *__restrict__ kw_, int xcd_) {
    if (done_ && *done_) return;
    const int nx = nelx + 1, ny = nely + 1, nz = nelz + 1;
    const uint32_t plane = (uint32_t)nx * ny;
    const int TY = 1 << tyl, TX = TPB >> tyl, RY = TY + 1, NQ = (TX + 1) * RY;
    const int t = (int)threadIdx.x;
    const int tx = t >> tyl, ty = t & (TY - 1);
    int bt = (int)blockIdx.x, bz = (int)blockIdx.y;
    if (xcd_) {
        const int64_t lb = xcd_block((int64_t)blockIdx.y * gridDim.x + blockIdx.x, (int64_t)gridDim.x * gridDim.y);
        bz = (int)(lb / gridDim.x);
        bt = (int)(lb - (int64_t)bz * gridDim.x);
    }
    const int bxx = bt % ntx, byy = bt / ntx;
    const int ix0 = bxx * (TX - 1) - 1, iy0 = byy * (TY - 1) - 1;   // node origin of the tile's region
    const int ix = ix0 + tx, iy = iy0 + ty;                         // element column = its (0,0) node
    const bool out_node = tx >= 1 && ty >= 1 && ix < nx && iy < ny;
    const bool ecol = ix >= 0 && ix < nelx && iy >= 0 && iy < nely;
    const int z0 = bz * zchunk, z1 = min(z0 + zchunk, nz);
    // region nodes this thread stages: q = t and q = t + TPB, clamped into the grid
    auto qoff = [&](int q) {
        const int a = q / RY, bq = q - a * RY;
        const int gx = min(max(ix0 + a, 0), nx - 1), gy = min(max(iy0 + bq, 0), ny - 1);
        return (uint32_t)(gx * ny + gy) * 3u * (uint32_t)sizeof(T);
    };
    const uint32_t vq0 = qoff(t);
    const bool has_q1 = t + TPB < NQ;
    const uint32_t vq1 = has_q1 ? qoff(t + TPB) : 0u;
    const uint32_t own = (uint32_t)(min(max(ix, 0), nx - 1) * ny + min(max(iy, 0), ny - 1));
    const uint32_t pT = plane * 3u * (uint32_t)sizeof(T);   // bytes of one node plane of a T vector
    const uint32_t nbT = pT * (uint32_t)nz;
    const brsrc_t rx = buf_rsrc(x_, nbT);
    const brsrc_t rE = buf_rsrc(E_, (uint32_t)nelx * nely * nelz * NSUB * (uint32_t)sizeof(T));
    const brsrc_t rm = buf_rsrc(mask_, plane * 3u * (uint32_t)nz);
    const brsrc_t rb = buf_rsrc(b_, b_ ? nbT : 0u);
    const brsrc_t ri = buf_rsrc(invD_, invD_ ? nbT : 0u);
    const brsrc_t rr = buf_rsrc(r_, r_ ? plane * 24u * (uint32_t)nz : 0u);
    const brsrc_t ro = buf_rsrc(out_, nbT);
    const uint32_t vE = (ecol ? (uint32_t)(ix * nely + iy) : 0u) * NSUB * (uint32_t)sizeof(T);
    const uint32_t sEz = (uint32_t)nelx * nely * NSUB * (uint32_t)sizeof(T);

    __shared__ T lin[2][3][LDS_REGION_MAX];
    __shared__ T xch[2][3][TPB];   // [buffer][component][thread]: x-stage values for column tx + 1
    const int xq = max(t - TY, 0);

    T R0[3], R1[3];
    auto load_raw = [&](int zz) {
        const uint32_t s = (uint32_t)min(max(zz, 0), nz - 1) * pT;
        bld3(rx, vq0, s, R0);
        if (has_q1) bld3(rx, vq1, s, R1);
    };
    auto store_lin = [&](int p) {
#pragma unroll
        for (int c = 0; c < 3; ++c) lin[p][c][t] = R0[c];
        if (has_q1) {
#pragma unroll
            for (int c = 0; c < 3; ++c) lin[p][c][t + TPB] = R1[c];
        }
    };
    auto patch_transform = [&](int p, T (&P)[3][4], T (&w)[3]) {
        const int q = tx * RY + ty;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T v[4] = {lin[p][c][q], lin[p][c][q + 1], lin[p][c][q + RY], lin[p][c][q + RY + 1]};
            wht4(v, P[c]);
            w[c] = v[0];
        }
    };
    auto load_E = [&](int ez, T (&dst)[NSUB]) {
        const bool ok = ecol && ez >= 0 && ez < nelz;
        const uint32_t s = (uint32_t)min(max(ez, 0), nelz - 1) * sEz;
#pragma unroll
        for (int q = 0; q < NSUB; ++q) {
            const T v = bld1(rE, vE + q * (uint32_t)sizeof(T), s, T(0));
            dst[q] = ok ? v : T(0);
        }
    };
    auto load_epi = [&](int L, EpiOps<T> &e) {
        const uint32_t zl = (uint32_t)min(L, nz - 1);
#pragma unroll
        for (int c = 0; c < 3; ++c)
            e.m[c] = (EPI == EPI_JACOBI || EPI == EPI_POSTR || EPI == EPI_DOTU) ? 1u
                     : __builtin_amdgcn_raw_buffer_load_b8(rm, own * 3u + c, zl * plane * 3u, 0);
        if (EPI == EPI_RESID || EPI == EPI_JACOBI) bld3(rb, own * 3u * (uint32_t)sizeof(T), zl * pT, e.b);
        if (EPI == EPI_JACOBI || EPI == EPI_POSTR) bld3(ri, own * 3u * (uint32_t)sizeof(T), zl * pT, e.inv);
        if (EPI == EPI_RESIDR || EPI == EPI_POSTR) bld3(rr, own * 24u, zl * plane * 24u, e.r);
    };

    typedef const __attribute__((address_space(4))) T *cptr;
    cptr kp = (cptr)kw_;
    T Qtp[3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 4; ++k) Qtp[c][k] = T(0);
    double dacc = 0.0;

    auto step = [&](const int L, const int p, T (&Pb)[3][4], T (&Pt)[3][4], T (&wb)[3], T (&wt)[3], T (&Ee)[NSUB],
                    T (&En)[NSUB], EpiOps<T> &ec, EpiOps<T> &en) {
        patch_transform(p, Pt, wt);   // plane L + 1
        load_E(L + 1, En);
        T xh[3][8];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                xh[d][k] = Pb[d][k] + Pt[d][k];
                xh[d][k + 4] = Pb[d][k] - Pt[d][k];
            }
        T yh[3][8];
        if constexpr (NSUB == 1) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (c > 0) asm volatile("" : "+s"(kp) : "v"(yh[c - 1][7]));
                yh[c][0] = T(0);
#pragma unroll
                for (int k = 1; k < 8; ++k) {
                    T acc = T(0);
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                        if ((WHT_NZ[c][d] >> k) & 1)
                            acc = fma(kp[wht_slot(c, d, k)], xh[d][k ^ WHT_BIT[c] ^ WHT_BIT[d]], acc);
                    yh[c][k] = Ee[0] * acc;
                }
            }
        } else {
            T Eh[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) Eh[q] = Ee[q];
#pragma unroll
            for (int bit = 1; bit < 8; bit <<= 1)
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (!(q & bit)) {
                        const T u = Eh[q], v = Eh[q | bit];
                        Eh[q] = u + v;
                        Eh[q | bit] = u - v;
                    }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                yh[c][0] = T(0);
#pragma unroll
                for (int k = 1; k < 8; ++k) {
                    asm volatile("" : "+s"(kp) : "v"(k > 1 ? yh[c][k - 1] : (c > 0 ? yh[c - 1][7] : Eh[0])));
                    T grp[8];
                    bool has[8] = {false, false, false, false, false, false, false, false};
#pragma unroll
                    for (int d = 0; d < 3; ++d)
#pragma unroll
                        for (int l = 1; l < 8; ++l)
                            if ((WHT_MR_NZ[8 * c + k] >> (8 * d + l)) & 1) {
                                const int m = k ^ l ^ WHT_BIT[c] ^ WHT_BIT[d];
                                const T coef = kp[wht_mr_slot(c, k, d, l)];
                                grp[m] = has[m] ? fma(coef, xh[d][l], grp[m]) : coef * xh[d][l];
                                has[m] = true;
                            }
                    T acc = T(0);
#pragma unroll
                    for (int m = 0; m < 8; ++m)
                        if (has[m]) acc = fma(Eh[m], grp[m], acc);
                    yh[c][k] = acc;
                }
            }
        }
        asm volatile("" : "+s"(kp) : "v"(yh[2][7]));
        // node plane L: S = this layer's bottom spectra plus the previous
        // layer's top ones; the inverse 2x2 transform and the sum over the
        // four element columns around a node are done separably: in y (S
        // index k = ky + 2 kx) with a one-lane wave shift, since t - 1 is
        // column (tx, ty - 1) whenever ty >= 1, then in x through LDS
        T C[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T S[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                S[k] = Qtp[c][k] + (yh[c][k] + yh[c][k + 4]);
                Qtp[c][k] = yh[c][k] - yh[c][k + 4];
            }
            const T A0 = S[0] + S[1], B0 = S[0] - S[1], A1 = S[2] + S[3], B1 = S[2] - S[3];
            const T U0 = A0 + wave_shr1(B0), U1 = A1 + wave_shr1(B1);
            C[c] = U0 + U1;
            xch[p][c][t] = U0 - U1;
        }
        store_lin(p ^ 1);          // plane L + 2
        load_raw(L + 3);
        load_epi(L + 1, en);
        __syncthreads();
        if (L >= z0 && out_node) {
            T o[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T y = C[c] + xch[p][c][xq];
                const bool m = ec.m[c] != 0;
                const T w = wb[c];
                if (EPI == EPI_PLAIN) {
                    o[c] = m ? y : T(0);
                } else if (EPI == EPI_RESID) {
                    o[c] = m ? (ec.b[c] - y) : T(0);
                } else if (EPI == EPI_JACOBI) {
                    o[c] = m ? (w + ec.inv[c] * (ec.b[c] - y)) : T(0);
                } else if (EPI == EPI_DOT || EPI == EPI_DOTU) {
                    o[c] = m ? y : T(0);
                    dacc += (double)w * (double)o[c];
                } else if (EPI == EPI_RESIDR) {
                    o[c] = m ? ((T)ec.r[c] - y) : T(0);
                } else {  // EPI_POSTR
                    o[c] = m ? (w + ec.inv[c] * ((T)ec.r[c] - y)) : T(0);
                    dacc += ec.r[c] * (double)o[c];
                }
            }
            bst3(ro, own * 3u * (uint32_t)sizeof(T), (uint32_t)L * pT, o);
        }
    };

    T PA[3][4], PB[3][4], wA[3], wB[3], EA[NSUB], EB[NSUB];
    EpiOps<T> epA, epB;
    load_raw(z0 - 1);
    store_lin(0);
    load_raw(z0);
    load_E(z0 - 1, EA);
    __syncthreads();
    patch_transform(0, PA, wA);
    store_lin(1);
    load_raw(z0 + 1);
    __syncthreads();
    int L = z0 - 1;
    for (; L + 1 < z1; L += 2) {
        step(L, 1, PA, PB, wA, wB, EA, EB, epA, epB);
        step(L + 1, 0, PB, PA, wB, wA, EB, EA, epB, epA);
    }
    if (L < z1) step(L, 1, PA, PB, wA, wB, EA, EB, epA, epB);
    if (EPI == EPI_DOT || EPI == EPI_DOTU || EPI == EPI_POSTR) block_reduce_store(dacc, partial_, (int64_t)bz * gridDim.x + bt);
}


// ---------------------------------------------------------------------------
// Coarse-level operator: assembled 27-point block stencil of the Galerkin
// operator P'AP, stored symmetrically: only the centre block and the 13
// "forward" neighbour blocks (offset index k = (dz+1)*9 + (dx+1)*3 + (dy+1)
// > 13) are kept, 14 blocks x 9 = 126 coefficients per node; the block of a
// backward neighbour j is the transpose of j's forward block towards i.
// Layout is structure-of-arrays blocked by wavefront: coefficient
// f*9 + c*3 + d (f = k - 13) of node i sits at
// S[(i/64)*126*64 + (f*9 + c*3 + d)*64 + i%64], so a wave's loads of one
// coefficient are one contiguous 64-element run.
// ---------------------------------------------------------------------------
constexpr int NCOEF = 126;
__host__ __device__ inline int64_t sidx(int64_t i, int coef) { return (i >> 6) * (NCOEF * 64) + coef * 64 + (i & 63); }


template <typename T>
struct StArgs {
    int nx, ny, nz;
    const T *x;            // SRC_X input
    const T *S;            // symmetric block stencil
    const uint8_t *mask;
    const T *b;            // right-hand side
    const T *invD;
    const T *ec;           // next-coarser level's correction (SRC_POST)
    int cnx, cny;
    T *out;
    double *partial;
    const int *done;
};

// Source value of node j (SRC_X: x; SRC_JZ: M\b, the first Jacobi sweep
// from zero; SRC_POST: M\b + P e_c, that sweep plus the prolonged coarse
// correction).
template <typename T, int SRC>
__device__ inline void stencil_source(const StArgs<T> &a, int jx, int jy, int jz, int64_t j, T (&v)[3]) {
    if (SRC == SRC_X) {
#pragma unroll
        for (int d = 0; d < 3; ++d) v[d] = a.x[3 * j + d];
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) v[d] = a.invD[3 * j + d] * a.b[3 * j + d];
        if (SRC == SRC_POST) {
            const int zl = jz >> 1, zh = (jz + 1) >> 1, xl = jx >> 1, xh = (jx + 1) >> 1;
            const int yl = jy >> 1, yh = (jy + 1) >> 1;
            const int64_t cpl = (int64_t)a.cnx * a.cny;
            T pe[3] = {T(0), T(0), T(0)};
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int Z = (t & 4) ? zh : zl, X = (t & 2) ? xh : xl, Y = (t & 1) ? yh : yl;
                const int64_t J = (int64_t)Z * cpl + (int64_t)X * a.cny + Y;
#pragma unroll
                for (int d = 0; d < 3; ++d) pe[d] += a.ec[3 * J + d];
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) v[d] += a.mask[3 * j + d] ? T(0.125) * pe[d] : T(0);
        }
    }
}

// y += contributions of the 9 neighbours in plane dz = DZ.
template <typename T, int SRC, int DZ>
__device__ inline void stencil_plane(const StArgs<T> &a, int64_t i, int ix, int iy, int iz, int64_t plane, T (&y)[3]) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int jz = iz + DZ;
    const bool vz = jz >= 0 && jz < nz;
    const int jzc = min(max(jz, 0), nz - 1);
    const int64_t jzb = (int64_t)jzc * plane;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int dx = q / 3 - 1, dy = q % 3 - 1;
        const int k = (DZ + 1) * 9 + q;
        const int jx = ix + dx, jy = iy + dy;
        const bool v = vz && jx >= 0 && jx < nx && jy >= 0 && jy < ny;
        const int jxc = min(max(jx, 0), nx - 1), jyc = min(max(jy, 0), ny - 1);
        const int64_t j = jzb + (int64_t)jxc * ny + jyc;
        const T vf = v ? T(1) : T(0);   // a clamped neighbour contributes nothing
        T xv[3];
        stencil_source<T, SRC>(a, jxc, jyc, jzc, j, xv);
#pragma unroll
        for (int d = 0; d < 3; ++d) xv[d] *= vf;
        if (k >= 13) {
            const T *Sw = a.S + sidx(i, (k - 13) * 9);
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int d = 0; d < 3; ++d) y[c] += Sw[(c * 3 + d) * 64] * xv[d];
        } else {
            const T *Sw = a.S + sidx(j, (13 - k) * 9);   // (A_ji)^T
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int d = 0; d < 3; ++d) y[c] += Sw[(d * 3 + c) * 64] * xv[d];
        }
    }
}

template <typename T, int SRC, int EPI>
__global__ __launch_bounds__(TPB) void k_stencil(int nx_, int ny_, int nz_, const T *__restrict__ x_, const T *__restrict__ S_,
                                                 const uint8_t *__restrict__ mask_, const T *__restrict__ b_,
                                                 const T *__restrict__ invD_, const T *__restrict__ ec_, int cnx_, int cny_,
                                                 T *__restrict__ out_, double *__restrict__ partial_,
                                                 const int *__restrict__ done_, int xcd_) {
    if (done_ && *done_) return;
    // the arguments arrive as non-aliasing kernel parameters and are bundled
    // here only for the helpers
    StArgs<T> a;
    a.nx = nx_; a.ny = ny_; a.nz = nz_; a.x = x_; a.S = S_; a.mask = mask_; a.b = b_; a.invD = invD_;
    a.ec = ec_; a.cnx = cnx_; a.cny = cny_; a.out = out_; a.partial = partial_; a.done = done_;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t plane = (int64_t)nx * ny;
    const int64_t nn = plane * nz;
    const int64_t blk = xcd_ ? xcd_block(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
    const int64_t i0 = blk * TPB + threadIdx.x;
    const bool act = i0 < nn;
    double dacc = 0.0;
    if (act) {
        int64_t i = i0;
        int iz = (int)(i / plane);
        int ix = (int)((i / ny) % nx);
        int iy = (int)(i % ny);
        T y[3] = {T(0), T(0), T(0)};
        // one plane of neighbours at a time: opaque copies of the coordinates
        // keep the next plane's loads from being hoisted (and held in
        // registers) before this plane's products are done
        stencil_plane<T, SRC, -1>(a, i, ix, iy, iz, plane, y);
        asm volatile("" : "+v"(i), "+v"(ix), "+v"(iy), "+v"(iz) : "v"(y[0]), "v"(y[1]), "v"(y[2]));
        stencil_plane<T, SRC, 0>(a, i, ix, iy, iz, plane, y);
        asm volatile("" : "+v"(i), "+v"(ix), "+v"(iy), "+v"(iz) : "v"(y[0]), "v"(y[1]), "v"(y[2]));
        stencil_plane<T, SRC, 1>(a, i, ix, iy, iz, plane, y);
        T w[3];
        stencil_source<T, SRC>(a, (int)((i0 / ny) % nx), (int)(i0 % ny), (int)(i0 / plane), i0, w);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int64_t dof = 3 * i0 + c;
            const bool m = a.mask[dof] != 0;
            if (EPI == EPI_PLAIN) {
                a.out[dof] = m ? y[c] : T(0);
            } else if (EPI == EPI_RESID) {
                a.out[dof] = m ? (a.b[dof] - y[c]) : T(0);
            } else if (EPI == EPI_JACOBI) {
                a.out[dof] = m ? (w[c] + a.invD[dof] * (a.b[dof] - y[c])) : T(0);
            } else {
                const T v = m ? y[c] : T(0);
                a.out[dof] = v;
                dacc += (double)w[c] * (double)v;
            }
        }
    }
    if (EPI == EPI_DOT) block_reduce_store(dacc, a.partial, blk);
}

// Small stencil levels (too few nodes to fill the chip, so latency-bound):
// one workgroup per 64 nodes and one wave per neighbour plane (dz = -1, 0, 1),
// the three partial sums added through LDS in that order. Three times the
// waves in flight, each with a third of the dependent loads.
constexpr int ST3_THREADS = 192;
template <typename T, int SRC, int EPI>
__global__ __launch_bounds__(ST3_THREADS) void k_stencil3(int nx_, int ny_, int nz_, const T *__restrict__ x_,
                                                          const T *__restrict__ S_, const uint8_t *__restrict__ mask_,
                                                          const T *__restrict__ b_, const T *__restrict__ invD_,
                                                          const T *__restrict__ ec_, int cnx_, int cny_,
                                                          T *__restrict__ out_, double *__restrict__ partial_,
                                                          const int *__restrict__ done_, int xcd_) {
    static_assert(EPI != EPI_DOT, "the split form has no dot-product epilogue");
    if (done_ && *done_) return;
    StArgs<T> a;
    a.nx = nx_; a.ny = ny_; a.nz = nz_; a.x = x_; a.S = S_; a.mask = mask_; a.b = b_; a.invD = invD_;
    a.ec = ec_; a.cnx = cnx_; a.cny = cny_; a.out = out_; a.partial = partial_; a.done = done_;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t plane = (int64_t)nx * ny;
    const int64_t nn = plane * nz;
    const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;   // wave wv: neighbour plane dz = wv - 1
    const int64_t blk = xcd_ ? xcd_block(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
    const int64_t i0 = blk * 64 + lane;
    const bool act = i0 < nn;
    __shared__ T part[2][3][64];
    T y[3] = {T(0), T(0), T(0)};
    const int iz = (int)(i0 / plane), ix = (int)((i0 / ny) % nx), iy = (int)(i0 % ny);
    if (act) {
        if (wv == 0) stencil_plane<T, SRC, -1>(a, i0, ix, iy, iz, plane, y);
        else if (wv == 1) stencil_plane<T, SRC, 0>(a, i0, ix, iy, iz, plane, y);
        else stencil_plane<T, SRC, 1>(a, i0, ix, iy, iz, plane, y);
    }
    if (wv > 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) part[wv - 1][c][lane] = y[c];
    }
    __syncthreads();
    if (wv != 0 || !act) return;
    T w[3];
    stencil_source<T, SRC>(a, ix, iy, iz, i0, w);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const T yc = (y[c] + part[0][c][lane]) + part[1][c][lane];
        const int64_t dof = 3 * i0 + c;
        const bool m = a.mask[dof] != 0;
        if (EPI == EPI_PLAIN) {
            a.out[dof] = m ? yc : T(0);
        } else if (EPI == EPI_RESID) {
            a.out[dof] = m ? (a.b[dof] - yc) : T(0);
        } else {  // EPI_JACOBI
            a.out[dof] = m ? (w[c] + a.invD[dof] * (a.b[dof] - yc)) : T(0);
        }
    }
}

// ---------------------------------------------------------------------------
// Transfers: trilinear prolongation P (vertex-centred, weights 1, 1/2, 1/4,
// 1/8, Implementation/Prolongation.m:15-37) restricted to free DOFs, and
// restriction P'.
// ---------------------------------------------------------------------------
// Launch shape of both transfers: blockIdx.y = z plane of the output level,
// blockIdx.x * TPB + threadIdx.x = in-plane index x * ny + y (32-bit).
// Index arithmetic is per thread, once; boundary neighbours are clamped and
// given weight zero, so there are no branches.
// JZ: also the coarse level's first Jacobi sweep from zero, xc = invc * bc
// (the product k_jacobi_zero forms, on the value just stored).
template <typename T, bool JZ = false>
__global__ __launch_bounds__(TPB) void k_restrict(int fnx, int fny, int fnz, int cnx, int cny, int cnz,
                                                  const T *__restrict__ rf, const uint8_t *__restrict__ cmask,
                                                  T *__restrict__ bc, const int *done, const T *__restrict__ invc = nullptr,
                                                  T *__restrict__ xc = nullptr) {
    if (done && *done) return;
    const int p = (int)blockIdx.x * TPB + (int)threadIdx.x;
    if (p >= cnx * cny) return;
    const int Iz = (int)blockIdx.y, Ix = p / cny, Iy = p - Ix * cny;
    const int64_t fplane = (int64_t)fnx * fny;
    // fine neighbours 2I-1, 2I, 2I+1 per direction, weights 1/2, 1, 1/2
    // (Implementation/Prolongation.m:15,34: w = 2^-(squared offset))
    const T wlz = Iz > 0 ? T(0.5) : T(0), whz = Iz < cnz - 1 ? T(0.5) : T(0);
    const T wlx = Ix > 0 ? T(0.5) : T(0), whx = Ix < cnx - 1 ? T(0.5) : T(0);
    const T wly = Iy > 0 ? T(0.5) : T(0), why = Iy < cny - 1 ? T(0.5) : T(0);
    const int fz = 2 * Iz, fx = 2 * Ix, fy = 2 * Iy;
    const int64_t oz[3] = {(int64_t)max(fz - 1, 0) * fplane, (int64_t)fz * fplane, (int64_t)min(fz + 1, fnz - 1) * fplane};
    const int64_t ox[3] = {(int64_t)max(fx - 1, 0) * fny, (int64_t)fx * fny, (int64_t)min(fx + 1, fnx - 1) * fny};
    const int oy[3] = {max(fy - 1, 0), fy, min(fy + 1, fny - 1)};
    const T wz[3] = {wlz, T(1), whz}, wx[3] = {wlx, T(1), whx}, wy[3] = {wly, T(1), why};
    T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        T sx[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            T sy[3] = {T(0), T(0), T(0)};
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int64_t j = oz[a] + ox[b] + oy[d];
#pragma unroll
                for (int c = 0; c < 3; ++c) sy[c] += wy[d] * rf[3 * j + c];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) sx[c] += wx[b] * sy[c];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] += wz[a] * sx[c];
    }
    const int64_t I = (int64_t)Iz * cnx * cny + p;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const T v = cmask[3 * I + c] ? acc[c] : T(0);
        bc[3 * I + c] = v;
        if (JZ) xc[3 * I + c] = invc[3 * I + c] * v;
    }
}

// x_f (+)= P e_c on free fine DOFs: each fine node takes its co-located or
// two/four/eight neighbouring coarse nodes with weights 1, 1/2, 1/4, 1/8.
template <typename T>
__global__ __launch_bounds__(TPB) void k_prolong_add(int fnx, int fny, int fnz, int cnx, int cny, int cnz,
                                                     const T *__restrict__ ec, const uint8_t *__restrict__ fmask,
                                                     T *__restrict__ xf, int add, const int *done) {
    if (done && *done) return;
    const int p = (int)blockIdx.x * TPB + (int)threadIdx.x;
    if (p >= fnx * fny) return;
    const int iz = (int)blockIdx.y, ix = p / fny, iy = p - ix * fny;
    const int64_t cplane = (int64_t)cnx * cny;
    // even index: the co-located coarse node (weight 1); odd: the two
    // neighbours (1/2 each)
    const int zl = iz >> 1, zh = (iz + 1) >> 1, xl = ix >> 1, xh = (ix + 1) >> 1, yl = iy >> 1, yh = (iy + 1) >> 1;
    const T wz0 = (iz & 1) ? T(0.5) : T(1), wz1 = (iz & 1) ? T(0.5) : T(0);
    const T wx0 = (ix & 1) ? T(0.5) : T(1), wx1 = (ix & 1) ? T(0.5) : T(0);
    const T wy0 = (iy & 1) ? T(0.5) : T(1), wy1 = (iy & 1) ? T(0.5) : T(0);
    const int64_t Zo[2] = {(int64_t)zl * cplane, (int64_t)zh * cplane};
    const int64_t Xo[2] = {(int64_t)xl * cny, (int64_t)xh * cny};
    const int Yo[2] = {yl, yh};
    const T wz[2] = {wz0, wz1}, wx[2] = {wx0, wx1}, wy[2] = {wy0, wy1};
    T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const int64_t J = Zo[a] + Xo[b] + Yo[d];
                const T w = wz[a] * wx[b] * wy[d];
#pragma unroll
                for (int c = 0; c < 3; ++c) acc[c] += w * ec[3 * J + c];
            }
    // the node's three values are read and written as one 3-vector (one
    // vector access per node instead of three strided ones)
    const int64_t i = (int64_t)iz * fnx * fny + p;
    T *xn = xf + 3 * i;
    T xo[3] = {T(0), T(0), T(0)};
    if (add) {
        xo[0] = xn[0]; xo[1] = xn[1]; xo[2] = xn[2];
    }
    T xv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) xv[c] = xo[c] + (fmask[3 * i + c] ? acc[c] : T(0));
    xn[0] = xv[0]; xn[1] = xv[1]; xn[2] = xv[2];
}

// The same x_f (+)= P e_c with one thread per coarse in-plane node (X, Y) and
// fine plane iz (blockIdx.y): the thread loads the 2 x 2 x 2 coarse nodes
// around its fine nodes (2X + ox, 2Y + oy), ox, oy in {0, 1}, once and writes
// those up to four fine nodes. Each fine value is the same sum, in the same
// order and with the same weights, as in the one-thread-per-fine-node form
// above (bitwise equal results), from a quarter of the coarse loads.
// Requires fnx = 2 cnx - 1, fny = 2 cny - 1, fnz = 2 cnz - 1.
template <typename T>
__global__ __launch_bounds__(TPB) void k_prolong_add4(int fnx, int fny, int fnz, int cnx, int cny, int cnz,
                                                      const T *__restrict__ ec, const uint8_t *__restrict__ fmask,
                                                      T *__restrict__ xf, int add, const int *done) {
    if (done && *done) return;
    const int p = (int)blockIdx.x * TPB + (int)threadIdx.x;
    if (p >= cnx * cny) return;
    const int iz = (int)blockIdx.y, X = p / cny, Y = p - X * cny;
    const int64_t cplane = (int64_t)cnx * cny;
    const int zl = iz >> 1, zh = (iz + 1) >> 1;
    const int X1 = min(X + 1, cnx - 1), Y1 = min(Y + 1, cny - 1);
    const T wz[2] = {(iz & 1) ? T(0.5) : T(1), (iz & 1) ? T(0.5) : T(0)};
    // e[a][b][d]: coarse node (z_a, X + b, Y + d), clamped at the last node
    T e[2][2][2][3];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const int64_t J = (int64_t)(a ? zh : zl) * cplane + (int64_t)(b ? X1 : X) * cny + (d ? Y1 : Y);
#pragma unroll
                for (int c = 0; c < 3; ++c) e[a][b][d][c] = ec[3 * J + c];
            }
    // the fine operands of all four nodes are loaded before any is used (at
    // clamped, in-range nodes: the last coarse node of a row has one fine
    // node in that direction), so their loads are in flight together
    const int64_t fplane = (int64_t)fnx * fny;
    int64_t fi[2][2];
    T xo[2][2][3];
    uint8_t mk[2][2][3];
#pragma unroll
    for (int ox = 0; ox < 2; ++ox)
#pragma unroll
        for (int oy = 0; oy < 2; ++oy) {
            const int ix = min(2 * X + ox, fnx - 1), iy = min(2 * Y + oy, fny - 1);
            fi[ox][oy] = (int64_t)iz * fplane + (int64_t)ix * fny + iy;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                xo[ox][oy][c] = xf[3 * fi[ox][oy] + c];
                mk[ox][oy][c] = fmask[3 * fi[ox][oy] + c];
            }
        }
#pragma unroll
    for (int ox = 0; ox < 2; ++ox) {
        // an even fine index has its co-located coarse node twice (weights
        // 1 and 0), an odd one its two neighbours (1/2 each)
        const T wx[2] = {ox ? T(0.5) : T(1), ox ? T(0.5) : T(0)};
#pragma unroll
        for (int oy = 0; oy < 2; ++oy) {
            const T wy[2] = {oy ? T(0.5) : T(1), oy ? T(0.5) : T(0)};
            T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int d = 0; d < 2; ++d) {
                        const T w = wz[a] * wx[b] * wy[d];
#pragma unroll
                        for (int c = 0; c < 3; ++c) acc[c] += w * e[a][ox ? b : 0][oy ? d : 0][c];
                    }
            if (2 * X + ox < fnx && 2 * Y + oy < fny) {
                T *xn = xf + 3 * fi[ox][oy];
#pragma unroll
                for (int c = 0; c < 3; ++c) xn[c] = (add ? xo[ox][oy][c] : T(0)) + (mk[ox][oy][c] ? acc[c] : T(0));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Vector kernels.
// ---------------------------------------------------------------------------
template <typename T, typename TB = T>
__global__ __launch_bounds__(TPB) void k_jacobi_zero(int64_t n, const TB *__restrict__ b, const T *__restrict__ invD,
                                                     T *__restrict__ x, const int *done) {
    if (done && *done) return;
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) x[i] = invD[i] * (T)b[i];
}

template <typename TD, typename TS>
__global__ __launch_bounds__(TPB) void k_convert(int64_t n, const TS *__restrict__ s, TD *__restrict__ d, const int *done) {
    if (done && *done) return;
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) d[i] = (TD)s[i];
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_fill(int64_t n, T v, T *__restrict__ d) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) d[i] = v;
}

// partial[block] = sum a_i * b_i over a grid-stride range
template <typename TA, typename TB>
__global__ __launch_bounds__(TPB) void k_dot(int64_t n, const TA *__restrict__ a, const TB *__restrict__ b,
                                             double *partial, const int *done) {
    if (done && *done) return;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
        s += (double)a[i] * (double)b[i];
    block_reduce_store(s, partial, blockIdx.x);
}

// d = beta d + s  (d = s on the first iteration, Implementation/SolveMG.m:39-44)
template <typename TS>
__global__ __launch_bounds__(TPB) void k_dir_update(int64_t n, const TS *__restrict__ s, double *__restrict__ d,
                                                    const PcgScalars *sc) {
    if (sc->done) return;
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    if (sc->it == sc->restart) d[i] = (double)s[i];
    else d[i] = sc->beta * d[i] + (double)s[i];
}

// The next iteration starts a new conjugate direction sequence.
__global__ void k_set_restart(PcgScalars *sc) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && !sc->done) sc->restart = sc->it;
}

// x += alpha d; r -= alpha q; partial ||r||^2  (Implementation/SolveMG.m:47-50)
__global__ __launch_bounds__(TPB) void k_xr_update(int64_t n, double *__restrict__ x, double *__restrict__ r,
                                                   const double *__restrict__ d, const double *__restrict__ q,
                                                   double *partial, PcgScalars *sc, int fin) {
    if (sc->done) return;
    const double al = sc->alpha;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
        x[i] = x[i] + al * d[i];
        const double ri = r[i] - al * q[i];
        r[i] = ri;
        s += ri * ri;
    }
    if (fin == FIN_RR) finish_reduce(FIN_RR, s, partial, blockIdx.x, gridDim.x, sc);
    else block_reduce_store(s, partial, blockIdx.x);
}

// Fixed-order total of np partial sums by one block of TPB threads, the
// same order as k_finalize's (strided per-thread sums, wave shuffles, the
// wave sums in order); every thread gets the total. A consumer kernel that
// needs a PCG scalar computes it this way in every block, instead of waiting
// for a one-block k_finalize launch.
__device__ inline double fin_block_total(int np, const double *__restrict__ partial) {
    __shared__ double red[TPB / 64];
    __shared__ double tot;
    double s = 0.0;
    for (int i = (int)threadIdx.x; i < np; i += TPB) s += partial[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < TPB / 64; ++w) t += red[w];
        tot = t;
    }
    __syncthreads();
    return tot;
}

// The PCG's vector passes with the solution update deferred by one iteration
// (x_{k+1} = x_k + alpha_k d_k is the same sum, done when d_k is read anyway):
// k_r_update: r -= alpha q; partial ||r||^2; and the next preconditioner
// cycle's first Jacobi sweep from zero, x0 = invD r (in its precision TX).
// invD is 0 exactly at the fixed DOFs (and positive elsewhere), so r stays 0
// there whatever q holds: q may come from the unmasked apply (EPI_DOTU).
// FUSE: alpha = rho / (d'q) from the apply's npa partial sums pa in every
// block (block 0 stores dq and alpha); the block sums go to partial, which
// must not be pa.
template <typename TX, bool FUSE = false>
__global__ __launch_bounds__(TPB) void k_r_update(int64_t n, double *__restrict__ r, const double *__restrict__ q,
                                                  const TX *__restrict__ invD, TX *__restrict__ x0, double *partial,
                                                  PcgScalars *sc, int npa = 0, const double *pa = nullptr) {
    if (sc->done) return;
    double al;
    if constexpr (FUSE) {
        const double dq = fin_block_total(npa, pa);
        al = sc->rho / dq;                      // the FIN_ALPHA recurrence
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            sc->dq = dq;
            sc->alpha = al;
        }
    } else {
        al = sc->alpha;
    }
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
        const TX iv = invD[i];
        const double ri = iv != TX(0) ? r[i] - al * q[i] : 0.0;
        r[i] = ri;
        s += ri * ri;
        x0[i] = iv * (TX)ri;
    }
    block_reduce_store(s, partial, blockIdx.x);
}

// k_dir_xupdate: x += alpha d with the previous iteration's alpha and d (0
// before the first iteration: alpha = 0 and d = 0 at the start of a solve),
// then d = s + beta d (d = s on a restart iteration). FUSE: rho = r's from
// the post-sweep's npr partial sums pr and beta = rho / gamma in every block
// (block 0 stores rho and beta; gamma, it and restart are not written here).
template <typename TS, bool FUSE = false>
__global__ __launch_bounds__(TPB) void k_dir_xupdate(int64_t n, const TS *__restrict__ s, double *__restrict__ d,
                                                     double *__restrict__ x, PcgScalars *sc, int npr = 0,
                                                     const double *pr = nullptr) {
    if (sc->done) return;
    const bool restart = sc->it == sc->restart;
    double beta;
    if constexpr (FUSE) {
        const double rho = fin_block_total(npr, pr);
        beta = restart ? 0.0 : rho / sc->gamma;  // the FIN_RHO recurrence
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            sc->rho = rho;
            sc->beta = beta;
        }
    } else {
        beta = sc->beta;
    }
    const double al = sc->alpha;
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
        const double di = d[i];
        x[i] = x[i] + al * di;
        d[i] = restart ? (double)s[i] : beta * di + (double)s[i];
    }
}

// The deferred update of the last iteration, once the PCG has stopped.
__global__ __launch_bounds__(TPB) void k_x_final(int64_t n, double *__restrict__ x, const double *__restrict__ d,
                                                 const PcgScalars *sc) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) x[i] = x[i] + sc->alpha * d[i];
}

// Single-block, fixed-order final reduction and the scalar PCG recurrences:
// strided per-thread sums, then wave shuffles and the 4 wave sums in order
// (one barrier).
constexpr int FIN_THREADS = 256;
template <int MODE>
__global__ __launch_bounds__(FIN_THREADS) void k_finalize(int np, const double *partial, int np2, const double *partial2,
                                                          PcgScalars *sc) {
    if (MODE != FIN_INIT && sc->done) return;
    __shared__ double red[FIN_THREADS / 64], red2[FIN_THREADS / 64];
    double s = 0.0, s2 = 0.0;
    for (int i = threadIdx.x; i < np; i += FIN_THREADS) s += partial[i];
    if (MODE == FIN_INIT)
        for (int i = threadIdx.x; i < np2; i += FIN_THREADS) s2 += partial2[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_down(s, off, 64);
        s2 += __shfl_down(s2, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = s;
        red2[threadIdx.x >> 6] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0, t2 = 0.0;
#pragma unroll
        for (int w = 0; w < FIN_THREADS / 64; ++w) {
            t += red[w];
            t2 += red2[w];
        }
        fin_update(MODE, t, t2, sc);
    }
}

// ---------------------------------------------------------------------------
// Coarsest grid: gather free DOFs, y = Ainv b (FP64, one wave per row),
// scatter back. Ainv is the inverse of the coarsest Galerkin operator from its
// Cholesky factorization.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(TPB) void k_gather(int64_t n, const int64_t *__restrict__ idx, const T *__restrict__ v,
                                                double *__restrict__ out, const int *done) {
    if (done && *done) return;
    const int64_t q = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (q < n) out[q] = (double)v[idx[q]];
}

// Finest-level index maps from the 1-based free-DOF list (strictly
// increasing, validated on the host): mask, full -> compact index and
// compact -> full. mask and cidx are pre-filled with 0 and -1.
__global__ __launch_bounds__(TPB) void k_level_index(int64_t nfree, const int64_t *__restrict__ fd1,
                                                     uint8_t *__restrict__ mask, int64_t *__restrict__ cidx,
                                                     int64_t *__restrict__ dfree) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= nfree) return;
    const int64_t d = fd1[i] - 1;
    mask[d] = 1;
    cidx[d] = i;
    dfree[i] = d;
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_scatter_full(int64_t ndof, const int64_t *__restrict__ cidx,
                                                      const double *__restrict__ v, T *__restrict__ out, const int *done) {
    if (done && *done) return;
    const int64_t d = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (d >= ndof) return;
    const int64_t q = cidx[d];
    out[d] = q >= 0 ? (T)v[q] : T(0);
}

// Coarsest-grid solve x = A^-1 b with the dense inverse (FP64, or its FP32
// copy for the FP32 cycle; products and sums in FP64): one block per row,
// every thread's loads issued together, fixed-order block reduction.
// Fixed DOFs of x are never written (they stay zero).
template <typename T, typename TA = double>
__global__ __launch_bounds__(TPB) void k_coarse_solve(int64_t n, const TA *__restrict__ A,
                                                      const int64_t *__restrict__ freeidx, const T *__restrict__ b,
                                                      T *__restrict__ x, const int *done) {
    if (done && *done) return;
    const int64_t row = blockIdx.x;
    const TA *Ar = A + row * n;
    // four independent partial sums keep four index -> value -> product
    // chains in flight per thread
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int64_t j = threadIdx.x;
    for (; j + 3 * TPB < n; j += 4 * TPB) {
        const int64_t i0 = freeidx[j], i1 = freeidx[j + TPB], i2 = freeidx[j + 2 * TPB], i3 = freeidx[j + 3 * TPB];
        s0 += Ar[j] * (double)b[i0];
        s1 += Ar[j + TPB] * (double)b[i1];
        s2 += Ar[j + 2 * TPB] * (double)b[i2];
        s3 += Ar[j + 3 * TPB] * (double)b[i3];
    }
    for (; j < n; j += TPB) s0 += Ar[j] * (double)b[freeidx[j]];
    double s = (s0 + s1) + (s2 + s3);
    __shared__ double red[TPB / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < TPB / 64; ++w) t += red[w];
        x[freeidx[row]] = (T)t;
    }
}

// Coarse solve reading half of the symmetric inverse (large coarse systems).
// A (the full n x n inverse, row-major) is cut into 64x64 tiles (I, J); only
// the tiles with J >= I are read, once each. xg holds b compacted to the free
// DOFs (k_gather). Rows and columns past n read as zero.
//
// Pass 1, k_coarse_symv_run: one wave per run of up to RUN consecutive tiles
// (I, Ja..Jb-1) of one block row. Lane c owns column c of every tile; the
// wave reads the run in four 16-row groups, row by row (one 512-byte row
// segment of a tile per load). Each lane accumulates the column products
// v_J = A_IJ' x_I of its columns over the four groups (stored per tile pair,
// J > I), and per group 16 partial row products A_I,J[r][c] x_J[c] summed
// over the run's tiles; a transposing butterfly over lane bits 3..0 and two
// shuffle sums over bits 4 and 5 leave the group's row products in lanes
// 0..15.
//
// Pass 2, k_coarse_symv_sum: block row K sums the column products of the
// tiles (I, K), I < K, then the row products of its own runs, each part split
// over the block's 4 waves in contiguous ranges and the 4 wave sums added in
// order; every sum has a fixed order.
constexpr int CS_B = 64;
constexpr int CS_RUN = 8;

// index of the tile pair (I, J), I < J, among the strictly upper tiles
__host__ __device__ inline int64_t cs_pair(int64_t I, int64_t J, int64_t nb) {
    return I * (nb - 1) - I * (I - 1) / 2 + (J - I - 1);
}

// the 16 values of one tile column for rows k = 0..15 of a row group
template <typename TA>
__device__ __forceinline__ void cs_tile_load(brsrc_t r, uint32_t rowbytes, uint32_t bo, TA (&a)[16]) {
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = bld1(r, (uint32_t)k * rowbytes + bo, 0u, TA(0));
}
template <typename TA>
__device__ __forceinline__ double cs_tile_fma(const TA (&a)[16], const double (&xi)[16], double xj, double (&u)[16]) {
    double vp = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const double ak = (double)a[k];
        vp += ak * xi[k];
        u[k] += ak * xj;
    }
    return vp;
}

// TA: the stored type of A (double, or the float copy the FP32 cycle reads;
// products and sums are FP64 either way); TS tiles are loaded per step
// (2 of double, 4 of float: 16 KB of row segments in flight per wave).
template <typename TA, int TS>
__global__ __launch_bounds__(64) void k_coarse_symv_run(int64_t n, int nb, const TA *__restrict__ A,
                                                       const int4 *__restrict__ runs, const double *__restrict__ xg,
                                                       double *__restrict__ V, double *__restrict__ U, const int *done) {
    if (done && *done) return;
    const int4 rn = runs[blockIdx.x];
    const int I = rn.x, Ja = rn.y, nt = rn.z - rn.y;     // nt <= CS_RUN
    const int c = (int)threadIdx.x;
    const int64_t I0 = (int64_t)I * CS_B;
    const int nr = (int)min<int64_t>(CS_B, n - I0);     // rows of this block row inside the matrix
    constexpr uint32_t SA = sizeof(TA);
    const uint32_t rowbytes = (uint32_t)n * SA;          // n * 8 * 16 < 2^32
    __shared__ double vl[CS_RUN][CS_B];                 // column products of the run's tiles (lane c: column c)
    for (int j = 0; j < nt; ++j) vl[j][c] = 0.0;
    // A is read through one buffer per 16-row group that ends at the last row
    // inside the matrix: rows past n read 0. Columns past n read finite
    // entries of the next row (or 0); their products only reach rows past n,
    // which pass 2 drops, or meet an x entry of 0.
#pragma unroll 1
    for (int r0 = 0; r0 < CS_B; r0 += 16) {
        const int nrow = max(0, min(16, nr - r0));
        // the group's base address and extent are wave-uniform (scalar registers)
        const uint64_t ab = reinterpret_cast<uint64_t>(A + (I0 + r0) * n);
        const uint64_t abu = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(ab >> 32)) << 32) |
                             (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)ab);
        const brsrc_t rA = buf_rsrc(reinterpret_cast<const void *>(abu),
                                    (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)nrow * rowbytes));
        double xi[16], u[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const double xr = xg[min<int64_t>(I0 + r0 + k, n - 1)];
            xi[k] = k < nrow ? xr : 0.0;
            u[k] = 0.0;
        }
        int j = 0;
#pragma unroll 1
        for (; j + TS <= nt; j += TS) {
            TA a[TS][16];
            double xjv[TS];
#pragma unroll
            for (int t = 0; t < TS; ++t) {
                const int64_t ct = (int64_t)(Ja + j + t) * CS_B + c;
                cs_tile_load<TA>(rA, rowbytes, (uint32_t)ct * SA, a[t]);
                xjv[t] = xg[min<int64_t>(ct, n - 1)] * (ct < n ? 1.0 : 0.0);
            }
#pragma unroll
            for (int t = 0; t < TS; ++t) vl[j + t][c] += cs_tile_fma<TA>(a[t], xi, xjv[t], u);
        }
#pragma unroll 1
        for (; j < nt; ++j) {
            const int64_t ct = (int64_t)(Ja + j) * CS_B + c;
            TA a0[16];
            cs_tile_load<TA>(rA, rowbytes, (uint32_t)ct * SA, a0);
            const double x0 = xg[min<int64_t>(ct, n - 1)] * (ct < n ? 1.0 : 0.0);
            vl[j][c] += cs_tile_fma<TA>(a0, xi, x0, u);
        }
        // transposing butterfly: at each step a lane keeps the half of its
        // rows selected by its lane bit and adds its partner's copy of it
#pragma unroll
        for (int h = 8; h >= 1; h >>= 1) {
            const bool hi = (c & h) != 0;
#pragma unroll
            for (int i = 0; i < h; ++i) {
                const double keep = hi ? u[i + h] : u[i];
                const double send = hi ? u[i] : u[i + h];
                u[i] = keep + __shfl_xor(send, h, 64);
            }
        }
        double s = u[0];
        s += __shfl_xor(s, 16, 64);
        s += __shfl_xor(s, 32, 64);
        if (c < 16) U[(int64_t)blockIdx.x * CS_B + r0 + c] = s;
    }
    for (int j = 0; j < nt; ++j)
        if (Ja + j > I) V[cs_pair(I, Ja + j, nb) * CS_B + c] = vl[j][c];
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_coarse_symv_sum(int64_t n, int nb, const double *__restrict__ P,
                                                         const int *__restrict__ rs,
                                                         const int64_t *__restrict__ freeidx, T *__restrict__ x,
                                                         const int *done) {
    if (done && *done) return;
    // P: the column products of the nb(nb-1)/2 strictly upper tiles, then the
    // row products of the runs (row K's runs are rs[K] .. rs[K+1]-1)
    const int K = (int)blockIdx.x, r = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
    const int nv = K, nu = rs[K + 1] - rs[K];
    const int nt = nv + nu, t0 = nt * w / 4, t1 = nt * (w + 1) / 4;
    const int64_t ubase = ((int64_t)nb * (nb - 1) / 2 + rs[K] - nv) * CS_B;
    double y = 0.0;
#pragma unroll 8
    for (int t = t0; t < t1; ++t) y += P[(t < nv ? cs_pair(t, K, nb) * CS_B : ubase + (int64_t)t * CS_B) + r];
    __shared__ double red[4][CS_B];
    red[w][r] = y;
    __syncthreads();
    if (w == 0) {
        const int64_t i = (int64_t)K * CS_B + r;
        if (i < n) x[freeidx[i]] = (T)(((red[0][r] + red[1][r]) + red[2][r]) + red[3][r]);
    }
}

// ---------------------------------------------------------------------------
// Coarsest grid by nested dissection (large coarse systems): the free coarse
// DOFs are ordered by a geometric nested dissection of the node grid (node
// planes as separators); each tree node is a front with p pivots (its own
// DOFs) and b boundary DOFs (the free DOFs of the ancestor separators next to
// its region). The multifrontal Cholesky factorization leaves in the first p
// columns of each front's column-major (p+b) x (p+b) block
//   M = [L_PP^-1 ; L_BP L_PP^-1],
// so that both triangular solves become matrix-vector products:
//   forward  (leaves first):  [y_P ; t] = M (b_P - children's updates on P)
//   backward (root first):    x_P = M' [y_P ; -x_B]
// where t (plus the children's updates on B) is the front's update vector
// that its parent gathers. The forward product reads a row-major copy of M
// (Mt), the backward one M itself, so that both read contiguous rows or
// columns one wave at a time. Every sum has a fixed order.
// ---------------------------------------------------------------------------
struct NdFront {
    int p, b, ld;            // pivots, boundary DOFs, leading dimension p + b
    int child[2];            // children fronts or -1
    int64_t P_off, B_off;    // compact coarse DOF indices of the pivots / boundary (nd_P, nd_B)
    int64_t F_off;           // column-major ld x ld block in nd_F
    int64_t T_off;           // update vector (b entries) in nd_T
    int64_t Mt_off;          // row-major copy of M (ld rows of p entries) in nd_Mt
    int64_t cmap_off[2];     // per child: ld entries, the position in the child's boundary or -1
};

// original operator entries of every front (lower triangle; F zeroed first)
__global__ __launch_bounds__(TPB) void k_nd_assemble(int64_t nent, const int64_t *__restrict__ src,
                                                     const int64_t *__restrict__ dst, const double *__restrict__ S,
                                                     double *__restrict__ F) {
    const int64_t e = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (e < nent) F[dst[e]] = S[src[e]];
}

// extend-add of the children's update matrices (lower triangles) into the
// fronts of one height: one block per (front, column l), rows k >= l
__global__ __launch_bounds__(TPB) void k_nd_extend_add(const int2 *__restrict__ work, const NdFront *__restrict__ fr,
                                                       const int *__restrict__ cmap, double *__restrict__ F) {
    const int2 w = work[blockIdx.x];
    const NdFront f = fr[w.x];
    const int l = w.y;
    double *Fs = F + f.F_off + (int64_t)l * f.ld;
    for (int c = 0; c < 2; ++c) {
        if (f.child[c] < 0) continue;
        const NdFront fc = fr[f.child[c]];
        const int *cm = cmap + f.cmap_off[c];
        const int jl = cm[l];
        if (jl < 0) continue;
        const double *Uc = F + fc.F_off + fc.p + (int64_t)(fc.p + jl) * fc.ld;   // column jl of the child's update
        for (int k = l + (int)threadIdx.x; k < f.ld; k += TPB) {
            const int ik = cm[k];
            if (ik >= 0) Fs[k] += Uc[ik];
        }
    }
}

// the children's update vectors at position k of front f, in child order
// (tsrc: per child and front position, the entry of the child's update vector
// in Tv, or -1)
__device__ __forceinline__ double nd_child_sum(const NdFront &f, const int64_t *__restrict__ tsrc,
                                               const double *__restrict__ Tv, int k) {
    double a = 0.0;
    for (int c = 0; c < 2; ++c) {
        if (f.child[c] < 0) continue;
        const int64_t i = tsrc[f.cmap_off[c] + k];
        if (i >= 0) a += Tv[i];
    }
    return a;
}

// row-major copy of M: one block per (front, pivot column j), Mt[j + k p] = M[k][j]
__global__ __launch_bounds__(TPB) void k_nd_transpose(const int2 *__restrict__ work, const NdFront *__restrict__ fr,
                                                      const double *__restrict__ F, double *__restrict__ Mt) {
    const int2 w = work[blockIdx.x];
    const NdFront f = fr[w.x];
    const int j = w.y;
    const double *Mj = F + f.F_off + (int64_t)j * f.ld;
    double *out = Mt + f.Mt_off + j;
    for (int k = (int)threadIdx.x; k < f.ld; k += TPB) out[(int64_t)k * f.p] = Mj[k];
}

// Each wave first loads its share of M into registers (up to ND_NREG values
// per lane; longer rows or columns finish from memory), so that the matrix
// loads overlap the dependent loads that gather the right-hand side.
constexpr int ND_NREG = 16;

// forward solve of the fronts of one height: one row k of M per wave (4 per
// block), the lanes over the row's contiguous pivot columns (j <= k on the
// lower-triangular pivot rows). z (the front's pivot right-hand side) is
// formed in LDS by every block of the front; bsrc holds the level-vector
// index of each pivot.
template <typename T>
__global__ __launch_bounds__(TPB) void k_nd_fwd(const int2 *__restrict__ work, const NdFront *__restrict__ fr,
                                                const int64_t *__restrict__ Pidx, const int64_t *__restrict__ bsrc,
                                                const int64_t *__restrict__ tsrc, const double *__restrict__ Mt,
                                                const T *__restrict__ bl, double *__restrict__ y,
                                                double *__restrict__ Tv, const int *done) {
    if (done && *done) return;
    extern __shared__ double nd_sh[];
    const int2 w = work[blockIdx.x];
    const NdFront f = fr[w.x];
    const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
    const int k = w.y + wv;
    const int jn = k >= f.ld ? 0 : (k < f.p ? k + 1 : f.p);
    const double *Mk = Mt + f.Mt_off + (int64_t)k * f.p;
    double m[ND_NREG];
#pragma unroll
    for (int c = 0; c < ND_NREG; ++c) m[c] = lane + 64 * c < jn ? Mk[lane + 64 * c] : 0.0;
    for (int j = (int)threadIdx.x; j < f.p; j += TPB)
        nd_sh[j] = (double)bl[bsrc[f.P_off + j]] - nd_child_sum(f, tsrc, Tv, j);
    __syncthreads();
    if (k >= f.ld) return;
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < ND_NREG; ++c)
        if (lane + 64 * c < jn) s += m[c] * nd_sh[lane + 64 * c];
    for (int j = lane + 64 * ND_NREG; j < jn; j += 64) s += Mk[j] * nd_sh[j];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) {
        if (k < f.p) y[Pidx[f.P_off + k]] = s;
        else Tv[f.T_off + (k - f.p)] = s + nd_child_sum(f, tsrc, Tv, k);
    }
}

// backward solve of the fronts of one height: one wave per pivot column j
// (4 per block): x_P[j] = sum_{k >= j} M[k][j] q[k], q = [y_P ; -x_B]
template <typename T>
__global__ __launch_bounds__(TPB) void k_nd_bwd(const int2 *__restrict__ work, const NdFront *__restrict__ fr,
                                                const int64_t *__restrict__ Pidx, const int64_t *__restrict__ Bidx,
                                                const double *__restrict__ F, const int64_t *__restrict__ bsrc,
                                                const double *__restrict__ y, double *__restrict__ xc,
                                                T *__restrict__ xl, const int *done) {
    if (done && *done) return;
    extern __shared__ double nd_sh[];
    const int2 w = work[blockIdx.x];
    const NdFront f = fr[w.x];
    const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
    const int j = w.y + wv;
    const int kn = j < f.p ? f.ld - j : 0;              // rows j .. ld-1 of column j
    const double *Mj = F + f.F_off + (int64_t)j * f.ld + j;
    double m[ND_NREG];
#pragma unroll
    for (int c = 0; c < ND_NREG; ++c) m[c] = lane + 64 * c < kn ? Mj[lane + 64 * c] : 0.0;
    const int64_t *P = Pidx + f.P_off;
    const int64_t *B = Bidx + f.B_off;
    for (int k = (int)threadIdx.x; k < f.ld; k += TPB) nd_sh[k] = k < f.p ? y[P[k]] : -xc[B[k - f.p]];
    __syncthreads();
    if (j >= f.p) return;
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < ND_NREG; ++c)
        if (lane + 64 * c < kn) s += m[c] * nd_sh[j + lane + 64 * c];
    for (int k = lane + 64 * ND_NREG; k < kn; k += 64) s += Mj[k] * nd_sh[j + k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) {
        xc[P[j]] = s;
        xl[bsrc[f.P_off + j]] = (T)s;
    }
}

__global__ __launch_bounds__(TPB) void k_gemv(int64_t n, const double *__restrict__ A, const double *__restrict__ x,
                                              double *__restrict__ y, const int *done) {
    if (done && *done) return;
    const int64_t row = (int64_t)blockIdx.x * (TPB / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const double *Ar = A + row * n;
    double s = 0.0;
    for (int64_t j = lane; j < n; j += 64) s += Ar[j] * x[j];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) y[row] = s;
}

// Checksum of n 64-bit words (their sum modulo 2^64, so the order of the
// additions does not matter): per-block sums of a block-strided range into
// part[block], then one block adds the nb partials into part[nb].
__device__ inline uint64_t block_sum_u64(uint64_t v) {
    __shared__ uint64_t red[TPB / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    uint64_t s = 0;
#pragma unroll
    for (int w = 0; w < TPB / 64; ++w) s += red[w];
    return s;
}

__global__ __launch_bounds__(TPB) void k_checksum_partial(int64_t n, const uint64_t *__restrict__ v,
                                                          uint64_t *__restrict__ part) {
    uint64_t s = 0;
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) s += v[i];
    s = block_sum_u64(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(TPB) void k_checksum_final(int nb, uint64_t *__restrict__ part) {
    uint64_t s = 0;
    for (int i = threadIdx.x; i < nb; i += TPB) s += part[i];
    s = block_sum_u64(s);
    if (threadIdx.x == 0) part[nb] = s;
}

// ---------------------------------------------------------------------------
// Setup: moduli, diagonals, Galerkin element matrices, stencils, dense coarse.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_modulus(int64_t ne, const double *__restrict__ x, double E0, double emin,
                                                 double p, double *__restrict__ E64, float *__restrict__ E32) {
    const int64_t e = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (e >= ne) return;
    const double v = emin + (E0 - emin) * pow(x[e], p);   // Implementation/GlobalStiffMatrix.m:17
    E64[e] = v;
    if (E32) E32[e] = (float)v;
}

// Multiresolution moduli: E[8e + s] of sub-element s = sy + 2 sx + 4 sz of
// element e, from densities x on the density mesh (2nelx x 2nely x 2nelz,
// y fastest) (Implementation/GlobalStiffMatrix.m:19-62).
// Moduli of the 8 fine elements of each coarse element, in child order
// s = sy + 2sx + 4sz: E1[8C + s] = E0[child s of C] (the level-1 Galerkin
// operator applied in the multiresolution form).
template <typename T>
__global__ __launch_bounds__(TPB) void k_children_moduli(int cnelx, int cnely, int cnelz, const T *__restrict__ E0,
                                                         T *__restrict__ E1) {
    const int64_t q = (int64_t)blockIdx.x * TPB + threadIdx.x;
    const int64_t nc = (int64_t)cnelx * cnely * cnelz;
    if (q >= 8 * nc) return;
    const int64_t C = q >> 3;
    const int sidx = (int)(q & 7);
    const int cz = (int)(C / ((int64_t)cnelx * cnely));
    const int64_t r = C - (int64_t)cz * cnelx * cnely;
    const int cx = (int)(r / cnely), cy = (int)(r - (int64_t)cx * cnely);
    const int fx = 2 * cx + ((sidx >> 1) & 1), fy = 2 * cy + (sidx & 1), fz = 2 * cz + (sidx >> 2);
    const int fnelx = 2 * cnelx, fnely = 2 * cnely;
    E1[q] = E0[(int64_t)fz * fnelx * fnely + (int64_t)fx * fnely + fy];
}

__global__ __launch_bounds__(TPB) void k_modulus_mr(int nelx, int nely, int nelz, const double *__restrict__ x,
                                                    double E0, double emin, double p, double *__restrict__ E64,
                                                    float *__restrict__ E32) {
    const int64_t ne = (int64_t)nelx * nely * nelz;
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= ne * 8) return;
    const int64_t e = t >> 3;
    const int s = (int)(t & 7);
    const int ez = (int)(e / ((int64_t)nelx * nely)), ex = (int)((e / nely) % nelx), ey = (int)(e % nely);
    const int sy = s & 1, sx = (s >> 1) & 1, sz = s >> 2;
    const int64_t NX = 2 * (int64_t)nelx, NY = 2 * (int64_t)nely;
    const int64_t v = (2 * (int64_t)ez + sz) * NX * NY + (2 * (int64_t)ex + sx) * NY + (2 * (int64_t)ey + sy);
    const double val = emin + (E0 - emin) * pow(x[v], p);
    E64[t] = val;
    if (E32) E32[t] = (float)val;
}

// Jacobi at the fine level: invD = 1 / ((1/omega) diag(K)), diag(K) gathered
// element by element (the recipe of Implementation/WOK/DiagPrec.m).
__global__ __launch_bounds__(TPB) void k_diag_fine(int nelx, int nely, int nelz, int nsub, const double *__restrict__ E,
                                                   const double *__restrict__ K0, const uint8_t *__restrict__ mask,
                                                   double omega, double *__restrict__ inv64, float *__restrict__ inv32) {
    const int nx = nelx + 1, ny = nely + 1, nz = nelz + 1;
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= (int64_t)nx * ny * nz) return;
    const int iz = (int)(i / ((int64_t)nx * ny)), ix = (int)((i / ny) % nx), iy = (int)(i % ny);
    double dg[3] = {0, 0, 0};
    for (int oz = 0; oz < 2; ++oz)
        for (int ox = 0; ox < 2; ++ox)
            for (int oy = 0; oy < 2; ++oy) {
                const int ex = ix - 1 + ox, ey = iy - 1 + oy, ez = iz - 1 + oz;
                if (ex < 0 || ex >= nelx || ey < 0 || ey >= nely || ez < 0 || ez >= nelz) continue;
                const int64_t e = (int64_t)ez * nelx * nely + (int64_t)ex * nely + ey;
                const int ai = (1 - oy) + 2 * (1 - ox) + 4 * (1 - oz);
                for (int s = 0; s < nsub; ++s)
                    for (int c = 0; c < 3; ++c) dg[c] += E[nsub * e + s] * K0[576 * s + (3 * ai + c) * 24 + 3 * ai + c];
            }
    for (int c = 0; c < 3; ++c) {
        const int64_t dof = 3 * i + c;
        const double v = mask[dof] ? 1.0 / ((1.0 / omega) * dg[c]) : 0.0;   // M = (1/omega) diag(A)
        inv64[dof] = v;
        if (inv32) inv32[dof] = (float)v;
    }
}

__global__ __launch_bounds__(TPB) void k_diag_stencil(int64_t nn, const double *__restrict__ S,
                                                      const uint8_t *__restrict__ mask, double omega,
                                                      double *__restrict__ inv64, float *__restrict__ inv32) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= nn) return;
    for (int c = 0; c < 3; ++c) {
        const int64_t dof = 3 * i + c;
        const double dg = S[sidx(i, c * 3 + c)];
        const double v = mask[dof] ? 1.0 / ((1.0 / omega) * dg) : 0.0;
        inv64[dof] = v;
        if (inv32) inv32[dof] = (float)v;
    }
}

__device__ inline double tri_w(int off) {  // 1-D trilinear weight of a fine node at offset off from a coarse vertex
    return off == 0 ? 1.0 : ((off == 1 || off == -1) ? 0.5 : 0.0);
}

// Galerkin element matrices of the first coarse level from the fine moduli:
// Ahat_E = sum_k T_k' D_k (E_k Ke) D_k T_k over the 8 children k, D_k the
// fine free-DOF mask, T_k the trilinear interpolation of the coarse element
// onto child k. One thread per (coarse element, 3x3 output block); output
// structure-of-arrays Ahat[(row*24 + col)*ne + E].
__global__ __launch_bounds__(TPB) void k_galerkin_fine(int cnelx, int cnely, int cnelz, int nsub,
                                                       const double *__restrict__ E, const double *__restrict__ K0,
                                                       const uint8_t *__restrict__ fmask, double *__restrict__ Ahat) {
    const int64_t cne = (int64_t)cnelx * cnely * cnelz;
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= cne * 64) return;
    const int blk = (int)(t / cne);
    const int64_t Ec = t - (int64_t)blk * cne;
    const int A = blk >> 3, B = blk & 7;
    const int Ez = (int)(Ec / ((int64_t)cnelx * cnely)), Ex = (int)((Ec / cnely) % cnelx), Ey = (int)(Ec % cnely);
    const int Ay = A & 1, Ax = (A >> 1) & 1, Az = A >> 2;
    const int By = B & 1, Bx = (B >> 1) & 1, Bz = B >> 2;
    const int fnelx = 2 * cnelx, fnely = 2 * cnely;
    const int fnx = fnelx + 1, fny = fnely + 1;
    double R[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int k = 0; k < 8; ++k) {
        const int ky = k & 1, kx = (k >> 1) & 1, kz = k >> 2;
        const int64_t ek = (int64_t)(2 * Ez + kz) * fnelx * fnely + (int64_t)(2 * Ex + kx) * fnely + (2 * Ey + ky);
        for (int s = 0; s < nsub; ++s) {
        const double Ek = E[nsub * ek + s];
        const double *Ks = K0 + 576 * s;
#pragma unroll
        for (int an = 0; an < 8; ++an) {
            const int ay = an & 1, ax = (an >> 1) & 1, az = an >> 2;
            const double tA = tri_w(kx + ax - 2 * Ax) * tri_w(ky + ay - 2 * Ay) * tri_w(kz + az - 2 * Az);
            if (tA == 0.0) continue;
            const int64_t na = (int64_t)(2 * Ez + kz + az) * fnx * fny + (int64_t)(2 * Ex + kx + ax) * fny + (2 * Ey + ky + ay);
            double ma[3];
            for (int c = 0; c < 3; ++c) ma[c] = fmask[3 * na + c] ? tA * Ek : 0.0;
#pragma unroll
            for (int bn = 0; bn < 8; ++bn) {
                const int by = bn & 1, bx = (bn >> 1) & 1, bz = bn >> 2;
                const double tB = tri_w(kx + bx - 2 * Bx) * tri_w(ky + by - 2 * By) * tri_w(kz + bz - 2 * Bz);
                if (tB == 0.0) continue;
                const int64_t nb = (int64_t)(2 * Ez + kz + bz) * fnx * fny + (int64_t)(2 * Ex + kx + bx) * fny + (2 * Ey + ky + by);
                double mb[3];
                for (int d = 0; d < 3; ++d) mb[d] = fmask[3 * nb + d] ? tB : 0.0;
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int d = 0; d < 3; ++d) R[c][d] += ma[c] * mb[d] * Ks[(3 * an + c) * 24 + 3 * bn + d];
            }
        }
        }
    }
    for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) Ahat[(int64_t)((3 * A + c) * 24 + 3 * B + d) * cne + Ec] = R[c][d];
}

// Galerkin element matrices of level l+1 from those of level l (same
// recursion with the child matrices read from Ahat_f and the level-l mask).
__global__ __launch_bounds__(TPB) void k_galerkin_coarse(int cnelx, int cnely, int cnelz, const double *__restrict__ Ahat_f,
                                                         const uint8_t *__restrict__ fmask, double *__restrict__ Ahat) {
    const int64_t cne = (int64_t)cnelx * cnely * cnelz;
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= cne * 64) return;
    const int blk = (int)(t / cne);
    const int64_t Ec = t - (int64_t)blk * cne;
    const int A = blk >> 3, B = blk & 7;
    const int Ez = (int)(Ec / ((int64_t)cnelx * cnely)), Ex = (int)((Ec / cnely) % cnelx), Ey = (int)(Ec % cnely);
    const int Ay = A & 1, Ax = (A >> 1) & 1, Az = A >> 2;
    const int By = B & 1, Bx = (B >> 1) & 1, Bz = B >> 2;
    const int fnelx = 2 * cnelx, fnely = 2 * cnely, fnelz = 2 * cnelz;
    const int64_t fne = (int64_t)fnelx * fnely * fnelz;
    const int fnx = fnelx + 1, fny = fnely + 1;
    double R[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int k = 0; k < 8; ++k) {
        const int ky = k & 1, kx = (k >> 1) & 1, kz = k >> 2;
        const int64_t ek = (int64_t)(2 * Ez + kz) * fnelx * fnely + (int64_t)(2 * Ex + kx) * fnely + (2 * Ey + ky);
        for (int an = 0; an < 8; ++an) {
            const int ay = an & 1, ax = (an >> 1) & 1, az = an >> 2;
            const double tA = tri_w(kx + ax - 2 * Ax) * tri_w(ky + ay - 2 * Ay) * tri_w(kz + az - 2 * Az);
            if (tA == 0.0) continue;
            const int64_t na = (int64_t)(2 * Ez + kz + az) * fnx * fny + (int64_t)(2 * Ex + kx + ax) * fny + (2 * Ey + ky + ay);
            double ma[3];
            for (int c = 0; c < 3; ++c) ma[c] = fmask[3 * na + c] ? tA : 0.0;
            for (int bn = 0; bn < 8; ++bn) {
                const int by = bn & 1, bx = (bn >> 1) & 1, bz = bn >> 2;
                const double tB = tri_w(kx + bx - 2 * Bx) * tri_w(ky + by - 2 * By) * tri_w(kz + bz - 2 * Bz);
                if (tB == 0.0) continue;
                const int64_t nb = (int64_t)(2 * Ez + kz + bz) * fnx * fny + (int64_t)(2 * Ex + kx + bx) * fny + (2 * Ey + ky + by);
                double mb[3];
                for (int d = 0; d < 3; ++d) mb[d] = fmask[3 * nb + d] ? tB : 0.0;
                for (int c = 0; c < 3; ++c)
                    for (int d = 0; d < 3; ++d)
                        R[c][d] += ma[c] * mb[d] * Ahat_f[(int64_t)((3 * an + c) * 24 + 3 * bn + d) * fne + ek];
            }
        }
    }
    for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) Ahat[(int64_t)((3 * A + c) * 24 + 3 * B + d) * cne + Ec] = R[c][d];
}

// Assembled block stencil of a level from its element matrices, masked by the
// level's free-DOF mask on both sides: the centre and forward blocks only
// (k = 13..26). One thread per (node, forward neighbour).
template <typename T32>
__global__ __launch_bounds__(TPB) void k_stencil_assemble(int nelx, int nely, int nelz, const double *__restrict__ Ahat,
                                                          const uint8_t *__restrict__ mask, double *__restrict__ S64,
                                                          T32 *__restrict__ S32) {
    const int nx = nelx + 1, ny = nely + 1, nz = nelz + 1;
    const int64_t nn = (int64_t)nx * ny * nz, ne = (int64_t)nelx * nely * nelz;
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= nn * 14) return;
    const int f = (int)(t / nn);
    const int k = 13 + f;
    const int64_t i = t - (int64_t)f * nn;
    const int iz = (int)(i / ((int64_t)nx * ny)), ix = (int)((i / ny) % nx), iy = (int)(i % ny);
    const int jz = iz + k / 9 - 1, jx = ix + (k / 3) % 3 - 1, jy = iy + k % 3 - 1;
    double acc[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const bool v = jz >= 0 && jz < nz && jx >= 0 && jx < nx && jy >= 0 && jy < ny;
    if (v) {
        const int64_t j = (int64_t)jz * nx * ny + (int64_t)jx * ny + jy;
        for (int ez = max(iz, jz) - 1; ez <= min(iz, jz); ++ez)
            for (int ex = max(ix, jx) - 1; ex <= min(ix, jx); ++ex)
                for (int ey = max(iy, jy) - 1; ey <= min(iy, jy); ++ey) {
                    if (ez < 0 || ez >= nelz || ex < 0 || ex >= nelx || ey < 0 || ey >= nely) continue;
                    const int64_t e = (int64_t)ez * nelx * nely + (int64_t)ex * nely + ey;
                    const int ai = (iy - ey) + 2 * (ix - ex) + 4 * (iz - ez);
                    const int aj = (jy - ey) + 2 * (jx - ex) + 4 * (jz - ez);
                    for (int c = 0; c < 3; ++c)
                        for (int d = 0; d < 3; ++d) acc[c][d] += Ahat[(int64_t)((3 * ai + c) * 24 + 3 * aj + d) * ne + e];
                }
        for (int c = 0; c < 3; ++c)
            for (int d = 0; d < 3; ++d)
                if (!(mask[3 * i + c] && mask[3 * j + d])) acc[c][d] = 0.0;
    }
    for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) {
            const int64_t o = sidx(i, f * 9 + c * 3 + d);
            S64[o] = acc[c][d];
            if (S32) S32[o] = (T32)acc[c][d];
        }
}

// Multiresolution fine level assembled into the symmetric block stencil: the
// block between nodes i and j sums, over the elements containing both, the
// element matrix sum_s E[8e+s] K_s (Implementation/GlobalStiffMatrix.m:19-62;
// the same element matrices Implementation/WOK/KelemProd.m:58-83 applies).
// Couplings to fixed DOFs are zeroed as on the Galerkin levels. One thread per
// (node, forward block).
template <typename T32>
__global__ __launch_bounds__(TPB) void k_stencil_assemble_mr(int nelx, int nely, int nelz, const double *__restrict__ E,
                                                             const double *__restrict__ K0,
                                                             const uint8_t *__restrict__ mask, double *__restrict__ S64,
                                                             T32 *__restrict__ S32) {
    const int nx = nelx + 1, ny = nely + 1, nz = nelz + 1;
    const int64_t nn = (int64_t)nx * ny * nz;
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= nn * 14) return;
    const int f = (int)(t / nn);
    const int k = 13 + f;
    const int64_t i = t - (int64_t)f * nn;
    const int iz = (int)(i / ((int64_t)nx * ny)), ix = (int)((i / ny) % nx), iy = (int)(i % ny);
    const int jz = iz + k / 9 - 1, jx = ix + (k / 3) % 3 - 1, jy = iy + k % 3 - 1;
    double acc[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const bool v = jz >= 0 && jz < nz && jx >= 0 && jx < nx && jy >= 0 && jy < ny;
    if (v) {
        const int64_t j = (int64_t)jz * nx * ny + (int64_t)jx * ny + jy;
        for (int ez = max(iz, jz) - 1; ez <= min(iz, jz); ++ez)
            for (int ex = max(ix, jx) - 1; ex <= min(ix, jx); ++ex)
                for (int ey = max(iy, jy) - 1; ey <= min(iy, jy); ++ey) {
                    if (ez < 0 || ez >= nelz || ex < 0 || ex >= nelx || ey < 0 || ey >= nely) continue;
                    const int64_t e = (int64_t)ez * nelx * nely + (int64_t)ex * nely + ey;
                    const int ai = (iy - ey) + 2 * (ix - ex) + 4 * (iz - ez);
                    const int aj = (jy - ey) + 2 * (jx - ex) + 4 * (jz - ez);
                    double es[8];
#pragma unroll
                    for (int sb = 0; sb < 8; ++sb) es[sb] = E[8 * e + sb];
                    for (int c = 0; c < 3; ++c)
                        for (int d = 0; d < 3; ++d) {
                            const int o = (3 * ai + c) * 24 + 3 * aj + d;
                            double a = 0.0;
#pragma unroll
                            for (int sb = 0; sb < 8; ++sb) a += es[sb] * K0[576 * sb + o];
                            acc[c][d] += a;
                        }
                }
        for (int c = 0; c < 3; ++c)
            for (int d = 0; d < 3; ++d)
                if (!(mask[3 * i + c] && mask[3 * j + d])) acc[c][d] = 0.0;
    }
    for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) {
            const int64_t o = sidx(i, f * 9 + c * 3 + d);
            S64[o] = acc[c][d];
            if (S32) S32[o] = (T32)acc[c][d];
        }
}

// Dense coarsest operator (free DOFs, row-major) from the coarsest stencil:
// each forward block fills its entries and, mirrored, the transposed ones.
__global__ __launch_bounds__(TPB) void k_dense_from_stencil(int nx, int ny, int nz, const double *__restrict__ S,
                                                            const int64_t *__restrict__ cidx, int64_t n,
                                                            double *__restrict__ Ad) {
    const int64_t nn = (int64_t)nx * ny * nz;
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= nn * 14) return;
    const int f = (int)(t / nn);
    const int k = 13 + f;
    const int64_t i = t - (int64_t)f * nn;
    const int iz = (int)(i / ((int64_t)nx * ny)), ix = (int)((i / ny) % nx), iy = (int)(i % ny);
    const int jz = iz + k / 9 - 1, jx = ix + (k / 3) % 3 - 1, jy = iy + k % 3 - 1;
    if (jz < 0 || jz >= nz || jx < 0 || jx >= nx || jy < 0 || jy >= ny) return;
    const int64_t j = (int64_t)jz * nx * ny + (int64_t)jx * ny + jy;
    for (int c = 0; c < 3; ++c) {
        const int64_t row = cidx[3 * i + c];
        if (row < 0) continue;
        for (int d = 0; d < 3; ++d) {
            const int64_t col = cidx[3 * j + d];
            if (col < 0) continue;
            const double v = S[sidx(i, f * 9 + c * 3 + d)];
            Ad[row * n + col] = v;
            if (f > 0) Ad[col * n + row] = v;
        }
    }
}

// Fill the lower triangle of a column-major symmetric matrix from the upper
// one (rocSOLVER potri writes one triangle).
__global__ __launch_bounds__(TPB) void k_symmetrize_upper(int64_t n, double *__restrict__ A) {
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= n * n) return;
    const int64_t r = t % n, c = t / n;  // column-major (r, c)
    if (r > c) A[c * n + r] = A[r * n + c];
}

// ---------------------------------------------------------------------------
// Compliance gradient and density filter.
// ---------------------------------------------------------------------------
// g0_e = -p x_e^(p-1) (E0 - emin) u_e' Ke u_e (Implementation/Grad.m:18-23)
__global__ __launch_bounds__(TPB) void k_grad0(int nelx, int nely, int nelz, const double *__restrict__ u,
                                               const double *__restrict__ x, const uint8_t *__restrict__ freedens,
                                               const double *__restrict__ K0, double p, double E0, double emin,
                                               double *__restrict__ g0) {
    const int64_t ne = (int64_t)nelx * nely * nelz;
    const int64_t e = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (e >= ne) return;
    if (freedens && !freedens[e]) { g0[e] = 0.0; return; }
    const int nx = nelx + 1, ny = nely + 1;
    const int ez = (int)(e / ((int64_t)nelx * nely)), ex = (int)((e / nely) % nelx), ey = (int)(e % nely);
    double ue[24];
#pragma unroll
    for (int an = 0; an < 8; ++an) {
        const int dy = an & 1, dx = (an >> 1) & 1, dz = an >> 2;
        const int64_t nd = (int64_t)(ez + dz) * nx * ny + (int64_t)(ex + dx) * ny + (ey + dy);
#pragma unroll
        for (int c = 0; c < 3; ++c) ue[3 * an + c] = u[3 * nd + c];
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 24; ++i) {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < 24; ++j) t += K0[i * 24 + j] * ue[j];
        s += ue[i] * t;
    }
    g0[e] = -p * pow(x[e], p - 1.0) * (E0 - emin) * s;
}

// Multiresolution g0 on the density mesh: for each sub-element v = (e, s),
// g0_v = -p x_v^(p-1) (E0 - emin) u_e' K_s u_e (Implementation/Grad.m:24-49,
// mr.interp = false); freedens masks displacement elements (str.freeDensG).
__global__ __launch_bounds__(TPB) void k_grad0_mr(int nelx, int nely, int nelz, const double *__restrict__ u,
                                                  const double *__restrict__ xd, const uint8_t *__restrict__ freedens,
                                                  const double *__restrict__ Ks, double p, double E0, double emin,
                                                  double *__restrict__ g0) {
    const int64_t ne = (int64_t)nelx * nely * nelz;
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= ne * 8) return;
    const int64_t e = t >> 3;
    const int s = (int)(t & 7);
    const int nx = nelx + 1, ny = nely + 1;
    const int ez = (int)(e / ((int64_t)nelx * nely)), ex = (int)((e / nely) % nelx), ey = (int)(e % nely);
    const int sy = s & 1, sx = (s >> 1) & 1, sz = s >> 2;
    const int64_t NX = 2 * (int64_t)nelx, NY = 2 * (int64_t)nely;
    const int64_t v = (2 * (int64_t)ez + sz) * NX * NY + (2 * (int64_t)ex + sx) * NY + (2 * (int64_t)ey + sy);
    if (freedens && !freedens[e]) { g0[v] = 0.0; return; }
    double ue[24];
#pragma unroll
    for (int an = 0; an < 8; ++an) {
        const int dy = an & 1, dx = (an >> 1) & 1, dz = an >> 2;
        const int64_t nd = (int64_t)(ez + dz) * nx * ny + (int64_t)(ex + dx) * ny + (ey + dy);
#pragma unroll
        for (int c = 0; c < 3; ++c) ue[3 * an + c] = u[3 * nd + c];
    }
    const double *K = Ks + 576 * s;
    double q = 0.0;
    for (int i = 0; i < 24; ++i) {
        double tt = 0.0;
#pragma unroll
        for (int j = 0; j < 24; ++j) tt += K[i * 24 + j] * ue[j];
        q += ue[i] * tt;
    }
    g0[v] = -p * pow(xd[v], p - 1.0) * (E0 - emin) * q;
}

// Filter row sums w and weighted neighbourhood products on the element grid.
// Offsets/weights depend only on the offset (Implementation/ElemNeighborhood.m:25-39);
// fixed elements are decoupled with unit self weight (:90-96).
__global__ __launch_bounds__(TPB) void k_filter(int nelx, int nely, int nelz, int noff, const int *__restrict__ off,
                                                const double *__restrict__ wt, const uint8_t *__restrict__ fixed,
                                                const double *__restrict__ x, const double *__restrict__ scale_in,
                                                const double *__restrict__ scale_out, double *__restrict__ y) {
    const int64_t ne = (int64_t)nelx * nely * nelz;
    const int64_t e = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (e >= ne) return;
    const int ez = (int)(e / ((int64_t)nelx * nely)), ex = (int)((e / nely) % nelx), ey = (int)(e % nely);
    const bool ef = fixed && fixed[e];
    double s = 0.0;
    for (int k = 0; k < noff; ++k) {
        const int dx = off[3 * k], dy = off[3 * k + 1], dz = off[3 * k + 2];
        const int jx = ex + dx, jy = ey + dy, jz = ez + dz;
        if (jx < 0 || jx >= nelx || jy < 0 || jy >= nely || jz < 0 || jz >= nelz) continue;
        const int64_t j = (int64_t)jz * nelx * nely + (int64_t)jx * nely + jy;
        double w;
        if (j == e) w = ef ? 1.0 : wt[k];
        else {
            if (ef || (fixed && fixed[j])) continue;
            w = wt[k];
        }
        const double xv = x ? x[j] : 1.0;
        s += w * (scale_in ? xv / scale_in[j] : xv);
    }
    y[e] = scale_out ? s / scale_out[e] : s;
}

}  // namespace tmg
