/*
 * tmg.h -- C-ABI of the MI355X multigrid equilibrium solver (libtmg.so).
 *
 * The solver replaces the equilibrium solve of TopOpt-SLP-3D: the geometric
 * multigrid preconditioned CG for K(rho) u = f on a structured grid of Q1
 * hexahedra, with the multigrid hierarchy resident in device memory.
 *
 * Each entry point below replaces one MATLAB function of the reference (cited
 * per declaration). Plain C types only: host pointers are caller-owned,
 * vectors are double, free-DOF vectors are compact in str.freedofs order (the
 * reference's u(str.freedofs)), indices are 1-based where the reference's are.
 * Device memory is owned by the handle. Calls on one handle are serialised on
 * the handle's own HIP stream; different handles are independent.
 *
 * Return convention: 0 = ok, > 0 = solve did not reach the tolerance
 * (TMG_NOT_CONVERGED), < 0 = error; tmg_last_error() gives the message.
 */
#ifndef TMG_H
#define TMG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tmg_solver tmg_solver;   /* device-resident GMG hierarchy        */
typedef struct tmg_filter tmg_filter;   /* device-resident density filter       */

/* Precision of the preconditioner. The PCG recurrences, dot products, the
 * outer operator apply q = K d, the coarsest-grid factorization and the
 * gradient are FP64 in every mode. */
enum {
    TMG_FP64 = 0,   /* V-cycle in FP64                                          */
    TMG_MIXED = 1   /* V-cycle (smoother, transfers, level operators) in FP32   */
};

enum {
    TMG_OK = 0,
    TMG_NOT_CONVERGED = 1,
    TMG_ERR_ARG = -1,
    TMG_ERR_UNSUPPORTED = -2,
    TMG_ERR_OOM = -3,
    TMG_ERR_NOT_SPD = -4,
    TMG_ERR_HIP = -5,
    TMG_ERR_STATE = -6,
    TMG_ERR_BUDGET = -7
};

/* The fields of the reference's `str` struct that the solve reads
 * (Implementation/TopOpt3D.m:4-43, Implementation/EnterData.m:79-86,147-156). */
typedef struct tmg_problem {
    int nelx, nely, nelz;       /* str.nelx, str.nely, str.nelz                  */
    double el, eh, ew;          /* str.el, str.eh, str.ew                         */
    double E0;                  /* str.E                                          */
    double nu;                  /* str.nu                                         */
    double penal;               /* p (SIMP penalty)                               */
    double emin;                /* emin (Young's modulus of void)                 */
    const int64_t *freedofs;    /* str.freedofs, 1-based, strictly increasing     */
    int64_t nfree;
} tmg_problem;

/* The `mg` struct (Implementation/TopOpt3D.m:171-188). cycle = 0 (V), 1 (W)
 * or 2 (FMV), smoother = 0 (Jacobi) and opchol = 0 (coarse Cholesky) are
 * accepted; the SOR/SSOR smoothers and the coarse PCG return
 * TMG_ERR_UNSUPPORTED. */
typedef struct tmg_mg_params {
    int ngrids;          /* mg.ngrids                                             */
    int cycle;           /* mg.cycle        (0 = V, 1 = W, 2 = FMV)               */
    int smoother;        /* mg.smoother     (0 = Jacobi)                          */
    double omega;        /* mg.omega                                              */
    int smoothiter1;     /* mg.smoothiter1                                        */
    int smoothiter2;     /* mg.smoothiter2                                        */
    double tolMG;        /* mg.tolMG                                              */
    int maxiterMG;       /* mg.maxiterMG                                          */
    int opchol;          /* mg.opchol       (0 = Cholesky on the coarsest grid)   */
    int nmaxchol;        /* mg.nmaxchol                                           */
} tmg_mg_params;

/* Level structure, prolongations, free-DOF sets per level and device memory
 * for the hierarchy. Replaces the structural part of
 * [A,P,M,L,U,R,perm] = SetupGMG(K,str,mg,elem) (Implementation/SetupGMG.m:16-44).
 * The operators are built by tmg_update_density. mg->ngrids must already obey
 * the level rules of Implementation/StartStrSLP.m:107-133. */
int tmg_setup(const tmg_problem *prob, const tmg_mg_params *mg, int precision, tmg_solver **out);

/* Multiresolution variant of tmg_setup (mr.op = true, mr.n = mr.d = 2,
 * mr.interp = false): each displacement element carries 2x2x2 density
 * sub-elements and K_e = sum_s (emin + (E0-emin) x_s^p) K_s with the
 * sub-element integrands of Implementation/ElemStiffMatrixGQMR.m
 * (Implementation/GlobalStiffMatrix.m:19-62, Implementation/WOK/KelemProd.m:58-83).
 * Densities (tmg_update_density) and gradients (tmg_grad) then live on the
 * density mesh (2nelx x 2nely x 2nelz, y fastest). mr_n = 1 is tmg_setup. */
int tmg_setup_mr(const tmg_problem *prob, const tmg_mg_params *mg, int precision, int mr_n, tmg_solver **out);

/* Operator refresh for new filtered densities x (n_el, element order):
 * K = sum_e (emin + (E0-emin) x_e^p) Ke, Galerkin A{i} = P'A{i-1}P, Jacobi
 * M{i} = (1/omega) diag(A{i}), coarse Cholesky. Replaces
 * GlobalStiffMatrix (Implementation/GlobalStiffMatrix.m:16-18,64-65) plus the
 * numeric part of SetupGMG (Implementation/SetupGMG.m:45-75) and the refresh
 * loop of Implementation/StrSLP.m:319,345-365. Returns TMG_ERR_NOT_SPD when
 * the coarse factorization fails. */
int tmg_update_density(tmg_solver *h, const double *x);
/* Same with x already in device memory. */
int tmg_update_density_device(tmg_solver *h, const double *d_x);

/* [x,it] = SolveMG(x0,f,Ac,P,M,L,U,R,perm,ngrids,cycle,smoother,
 *                  smoothiter1,smoothiter2,tolMG,maxiterMG,opchol,opcond,pcgp)
 * (Implementation/SolveMG.m:1-55). x0, f, x_out: nfree host vectors.
 * *it gets the reference's count (PCG iterations + 1); *relres the final
 * recursive ||r||/||f||. Returns TMG_NOT_CONVERGED when maxiterMG was hit. */
int tmg_solve(tmg_solver *h, const double *x0, const double *f, double *x_out, int *it, double *relres);
/* Same on device-resident full-length vectors (3*nnodes, fixed DOFs zero):
 * d_x holds x0 on entry and the solution on exit. Blocks until done. */
int tmg_solve_device(tmg_solver *h, double *d_x, const double *d_f, int *it, double *relres);

/* x = VCycle(A,b,zeros,niter1,niter2,ngrids,1,...) (Implementation/VCycle.m:1-45),
 * WCycle(...) (Implementation/WCycle.m:1-27) for a handle set up with
 * mg->cycle = 1, or FMVCycle(A,b,niter1,niter2,ngrids,1,...)
 * (Implementation/FMVCycle.m:1-41) with mg->cycle = 2: one preconditioner
 * cycle at the finest level on nfree host vectors. */
int tmg_vcycle(tmg_solver *h, const double *b, double *x);

/* x = Smooth(A{level},M{level},L,U,b,x0,niter,0) (Implementation/Smooth.m:1-24),
 * level 1-based as the reference's cell index; vectors are that level's free vectors. */
int tmg_smooth(tmg_solver *h, int level, const double *b, const double *x0, int niter, double *x);

/* y = A{level} * x (the matrix products of Implementation/SolveMG.m:26,45 and
 * Implementation/VCycle.m:24). Level 1-based; free vectors of that level. */
int tmg_apply(tmg_solver *h, int level, const double *x, double *y);

/* xf = P{level} * xc and xc = P{level}' * xf with P from
 * [P,freedofsc] = Prolongation(nxf,nyf,nzf,nxc,nyc,nzc,nc,freedofsf)
 * (Implementation/Prolongation.m:1-57, used at Implementation/VCycle.m:25,41).
 * level is the fine level (1-based); vectors are free vectors of the levels. */
int tmg_prolong(tmg_solver *h, int level, const double *xc, double *xf);
int tmg_restrict(tmg_solver *h, int level, const double *xf, double *xc);

/* Level dimensions and free DOFs: nel3 = {nelx,nely,nelz} of the level,
 * freedofs1 (optional, nfree entries) the 1-based free DOF list of the level
 * (the freedofsc output of Prolongation.m). Level 1-based. */
int tmg_level_info(tmg_solver *h, int level, int *nel3, int64_t *nfree, int64_t *freedofs1);
int tmg_ngrids(const tmg_solver *h);
/* Mixed precision (no counterpart in the reference): iteration at which the
 * last solve switched its preconditioner from the FP32 to the FP64 hierarchy
 * because the residual stalled (0: it did not), and the number of such
 * switches over the handle's lifetime. A switch restarts the conjugate
 * directions (beta = 0) and leaves the stop rule unchanged. */
int tmg_solve_stats(const tmg_solver *h, int *escalated_at, int *escalations);

/* Fgrad = Grad(str,u,p,emin,k,x,Dofs,W,w,opfilter,mr,strDens)
 * (Implementation/Grad.m:18-62; with a multiresolution handle x, Fgrad and
 * the filter are on the density mesh, mr.interp = false). u: full-length displacement
 * (3*nnodes, fixed DOFs 0); x: filtered densities (n_el); freedens: optional
 * 0/1 element mask (str.freeDensG; NULL = all); filt: the filter (NULL =
 * opfilter 0). Fgrad: n_el. All host pointers. */
int tmg_grad(tmg_solver *h, const double *u, const double *x, const uint8_t *freedens,
             const tmg_filter *filt, double *Fgrad);

/* tmg_grad without device allocation. No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same arguments,
 * kernels and results as tmg_grad; its device buffers (displacement,
 * densities, g0, filtered gradient, mask) live in one 96 MiB array per GPU
 * that the library holds until the process ends. Returns TMG_ERR_ARG when
 * they need more (3*nnodes + 3*n_el*nsub doubles + n_el bytes, each section
 * rounded to 256 bytes; the 192x96x96 cantilever needs 88 MB). Every call on
 * a GPU uses that GPU's array: calls that overlap on one GPU (from several
 * threads or handles) race; nothing in it needs to start at zero. */
int tmg_grad_noalloc(tmg_solver *h, const double *u, const double *x, const uint8_t *freedens,
                     const tmg_filter *filt, double *Fgrad);

/* Density filter [W,w] = ElemNeighborhood(str,rmin,opfilter)
 * (Implementation/ElemNeighborhood.m:1-100) for opfilter 1 or 2 on a
 * structured grid, fixed: optional 0/1 element mask (str.fixedDens).
 * tmg_filter_apply is xfil = ApplyFilter(x,W,w,opfilter,trans)
 * (Implementation/ApplyFilter.m:1-21). Host vectors of n_el. */
int tmg_filter_create(int nelx, int nely, int nelz, double el, double eh, double ew,
                      double rmin, int opfilter, const uint8_t *fixed, tmg_filter **out);
int tmg_filter_apply(tmg_filter *f, const double *x, int trans, double *y);
/* w = full(sum(W,2)), the row sums of W (Implementation/ElemNeighborhood.m:98). */
int tmg_filter_rowsum(const tmg_filter *f, double *w);
void tmg_filter_destroy(tmg_filter *f);

/* The LP of one SLP step, [s,~,~,~,lambda] = linprog(c,A,-b,[],[],sL,sU)
 * (Implementation/StrSLP.m:265-269; equality form with Aeq at :271): minimise
 * c's subject to the single row a's <= r (equality != 0: a's = r) and
 * sL <= s <= sU. Solved exactly as a continuous knapsack on the host.
 * lambda receives the row's multiplier (lambda.ineqlin / lambda.eqlin).
 * Returns TMG_OK, or TMG_NOT_CONVERGED when the row cannot be met inside the
 * box (s is then the box point closest to feasibility), or TMG_ERR_ARG.
 * Host vectors of n. */
int tmg_lp_volume(int64_t n, const double *c, const double *a, const double *sL, const double *sU, double r,
                  int equality, double *s, double *lambda);

/* Instrumentation (no counterpart in the reference): average device time in
 * milliseconds of one launch of a solver kernel, measured with HIP events on
 * the handle's stream over nrep back-to-back launches on the handle's own
 * buffers. which: 0 = fine-level operator apply in FP64 (the PCG's q = K d),
 * 1 = fine-level Jacobi sweep in the preconditioner precision, 2 = level-2
 * stencil Jacobi sweep in the preconditioner precision, 3 = one V-cycle
 * (preconditioner application), 4 = fine-level residual with the first
 * Jacobi sweep folded in, 5 = fine-level post-sweep with the prolongation
 * folded in (both in the preconditioner precision), 6 = FP64 fine-level
 * apply with the PCG direction update folded in, 7 and 8 = the fine-level
 * residual sweep and post-sweep on the FP64 PCG residual (the V-cycle's own
 * forms, preconditioner precision). Requires a prior tmg_update_density. */
int tmg_time_kernel(tmg_solver *h, int which, int nrep, double *avg_ms);

/* Instrumentation (no counterpart in the reference): average device time in
 * milliseconds of kernel which (0 = the FP64 fine-level apply q = K d with
 * d'q, 8 = the fine-level post-sweep on the FP64 PCG residual) as it runs
 * inside a solve: one solve from x0 = 0 on the right-hand side of the last
 * solve, launched without the graph, with HIP events on the handle's stream
 * around each launch of that kernel in the iterations the PCG ran (*nlaunch
 * of them). Unlike tmg_time_kernel's back-to-back launches, the kernel's
 * inputs here have passed through the caches the way the solve leaves them.
 * Overwrites the handle's solution vector. */
int tmg_time_in_solve(tmg_solver *h, int which, double *avg_ms, int *nlaunch);

/* Instrumentation (no counterpart in the reference): the most device bytes
 * the handle has held at once, each allocation rounded up to 256 bytes. */
int64_t tmg_device_bytes(const tmg_solver *h);

/* ---------------------------------------------------------------------------
 * Serving (no counterpart in the reference, which solves one system at a time
 * inside its SLP loop): a continuous-batching server for streams of
 * equilibrium solves of uneven grid sizes. Requests are admitted FIFO as
 * slots and device memory (estimated per request, against hbm_budget_bytes)
 * free up; each of the max_active slots has its own worker thread, stream,
 * memory pool and rocBLAS handle, each admitted request gets its own
 * device-resident hierarchy (carved from one block of its slot's pool), and
 * the PCG iterations of all running requests overlap on the device. Each
 * result is bitwise identical to the same solve run alone. With the
 * environment variable TMG_SERVER_TRACE=1 each request's set-up phase times
 * are printed to stderr.
 * ------------------------------------------------------------------------- */
typedef struct tmg_server tmg_server;

int tmg_server_create(double hbm_budget_bytes, int max_active, tmg_server **out);
/* Queue one solve: prob/mg/precision as tmg_setup, x the filtered densities
 * (n_el), f the load (nfree), x0 the initial guess (nfree, NULL = 0),
 * t_arrival the arrival time on the server clock (< 0 = now). Inputs are
 * copied. Returns TMG_ERR_BUDGET if the request can never fit the budget. */
int tmg_server_submit(tmg_server *s, const tmg_problem *prob, const tmg_mg_params *mg, int precision,
                      const double *x, const double *f, const double *x0, double t_arrival, int64_t *id);
/* Wait for a request (timeout_s < 0: no limit). Returns 0 when it has
 * finished (state 3) or failed (state < 0), 1 on timeout. */
int tmg_server_wait(tmg_server *s, int64_t id, double timeout_s, int *state);
/* Solution (nfree), reference iteration count, final relative residual and
 * times[4] = {arrival, admission, start of PCG, completion} on the server clock. */
int tmg_server_result(tmg_server *s, int64_t id, double *u, int *it, double *relres, double *times);
int tmg_server_release(tmg_server *s, int64_t id);
/* stats[8] = {bytes reserved, peak bytes reserved, queued, running, done, clock,
 * peak number of requests solving at the same time, number of requests whose
 * hierarchy was rebuilt after a failed first coarse factorization with their
 * densities verified intact on the device}. */
int tmg_server_stats(tmg_server *s, double *stats);
double tmg_server_clock(const tmg_server *s);
/* Device bytes the server reserves for a request of this shape. */
int64_t tmg_server_estimate(const tmg_problem *prob, const tmg_mg_params *mg, int precision);
const char *tmg_server_last_error(const tmg_server *s);
void tmg_server_destroy(tmg_server *s);

/* Message of the last failed call on h (or of the last failed tmg_setup when h is NULL). */
const char *tmg_last_error(const tmg_solver *h);

/* Frees the hierarchy and all its device memory. */
void tmg_destroy(tmg_solver *h);

#ifdef __cplusplus
}
#endif
#endif
