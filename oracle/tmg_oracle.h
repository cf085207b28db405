/*
 * tmg_oracle.h -- FP64 CPU restatement of TopOpt-SLP-3D's assembled
 * geometric-multigrid PCG path (test infrastructure only).
 *
 * This library is the parity oracle and the CPU baseline for the HIP solver.
 * It is only ever loaded by tests, the smoke check and the CPU-baseline leg of
 * the benchmark; the product path never links or calls it.
 *
 * Conventions (all 0-based here; the reference is 1-based MATLAB):
 *   node(ix,iy,iz) = iz*nx*ny + ix*ny + iy        (Implementation/EnterData.m:130)
 *   dof            = 3*node + c, c = 0:x 1:y 2:z  (Implementation/EnterData.m:146-150)
 *   elem(ex,ey,ez) = ez*nelx*nely + ex*nely + ey  (Implementation/DegreesOfFreedom.m:40-54)
 *   element-local node a = dy + 2*dx + 4*dz, local dof 3*a + c
 *                                                  (Implementation/ElemStiffMatrix.m:331-332)
 * Vectors passed in and out of the solver are compact free-DOF vectors, as
 * MATLAB's u(str.freedofs).
 */
#ifndef TMG_ORACLE_H
#define TMG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

/* 24x24 Q1 element stiffness, row-major, E = 1 (Implementation/ElemStiffMatrix.m). */
void or_elem_stiff(double el, double eh, double ew, double nu, double *k576);
/* n^3 multiresolution sub-element integrands, each 24x24 row-major, sub-element
 * order z, x, y (y fastest) (Implementation/ElemStiffMatrixGQMR.m:54-236). */
void or_elem_stiff_mr(double el, double eh, double ew, double nu, int n, double *kout);

/* Problem: grid, material and 1-based free DOF list (str.freedofs). */
or_ctx *or_create(int nelx, int nely, int nelz, double el, double eh, double ew,
                  double E0, double nu, double p, double Emin,
                  const int64_t *freedofs1, int64_t nfree);
void or_destroy(or_ctx *c);
int64_t or_nfree(const or_ctx *c, int level);
int64_t or_ndof(const or_ctx *c, int level);

/* K = sum_e (Emin + (E0-Emin) x_e^p) Ke, restricted to free DOFs
 * (Implementation/GlobalStiffMatrix.m:16-18,64-65). */
int or_assemble(or_ctx *c, const double *x);
/* Multiresolution assembly, n^3 density sub-elements per element, xd indexed
 * on the density mesh (Implementation/GlobalStiffMatrix.m:19-62). */
int or_assemble_mr(or_ctx *c, int n, const double *xd);
/* SetupGMG: P, Galerkin A, Jacobi M, coarse Cholesky (Implementation/SetupGMG.m:16-75). */
int or_setup_gmg(or_ctx *c, int ngrids, double omega);
/* Galerkin refresh after a new assembly (Implementation/StrSLP.m:345-365). */
int or_refresh_gmg(or_ctx *c);

/* y = A_level * x (compact free vectors). */
void or_spmv(const or_ctx *c, int level, const double *x, double *y);
/* Matrix-free y = K x following Implementation/WOK/KelemProd.m (fine level). */
void or_kelemprod(const or_ctx *c, const double *x, double *y);
/* x = Smooth(A{l}, M{l}, b, x0, niter) (Implementation/Smooth.m:14-18). */
void or_smooth(const or_ctx *c, int level, const double *b, const double *x0, int niter, double *x);
/* x = VCycle(A, b, 0, nu1, nu2, ..., nv=level) (Implementation/VCycle.m:23-43). */
void or_vcycle(const or_ctx *c, int level, const double *b, int nu1, int nu2, double *x);
/* [x,it] = SolveMG(x, f, ...) with cycle=0, smoother=0 (Implementation/SolveMG.m:26-53).
 * Returns it with the reference's meaning (iterations + 1); relres gets the
 * final recursive ||r||/||f||. */
int or_solve_mg(const or_ctx *c, double *x, const double *f, int nu1, int nu2,
                double tol, int maxiter, double *relres);
/* x = WCycle(A, b, x0, nu1, nu2, ..., nv=level) (Implementation/WCycle.m:1-27);
 * x0 may be NULL for a zero initial guess. */
void or_wcycle(const or_ctx *c, int level, const double *b, const double *x0, int nu1, int nu2, double *x);
/* x = VCycle(A, b, x0, nu1, nu2, ..., nv) from an initial guess (x0 NULL = zero). */
void or_vcycle_x0(const or_ctx *c, int level, const double *b, const double *x0, int nu1, int nu2, double *x);
/* x = FMVCycle(A, b, nu1, nu2, ..., nv) (Implementation/FMVCycle.m:1-41). */
void or_fmvcycle(const or_ctx *c, int level, const double *b, int nu1, int nu2, double *x);
/* SolveMG with the cycle switch of Implementation/SolveMG.m:31-37 (0 V, 1 W, 2 FMV). */
int or_solve_mg_cycle(const or_ctx *c, double *x, const double *f, int cycle, int nu1, int nu2,
                      double tol, int maxiter, double *relres);

/* Export a level's operator / prolongation as CSR (compact indices). */
int64_t or_csr_nnz(const or_ctx *c, int which, int level); /* which: 0=A, 1=P */
void or_csr_export(const or_ctx *c, int which, int level, int64_t *rp, int32_t *ci, double *v);
/* Free DOF map of a level (compact -> full 0-based dof). */
void or_level_free(const or_ctx *c, int level, int64_t *full_of_free);
void or_level_dims(const or_ctx *c, int level, int *nel3);
int or_ngrids(const or_ctx *c);
/* OpenMP threads of the following calls; returns the previous count. */
int or_set_threads(int n);
/* Jacobi diagonal M{l} = (1/omega) diag(A{l}). */
void or_jacobi_diag(const or_ctx *c, int level, double *m);

/* Density filter W (opfilter 1 or 2), w = row sums (Implementation/ElemNeighborhood.m:10-98).
 * fixed: optional per-element 0/1 mask (NULL = none). Returns nnz; call with
 * rp == NULL to size, then again to fill. */
int64_t or_filter_build(int nelx, int nely, int nelz, double el, double eh, double ew,
                        double rmin, int opfilter, const uint8_t *fixed,
                        int64_t *rp, int32_t *ci, double *v, double *w);
/* ApplyFilter (Implementation/ApplyFilter.m:11-19): trans=0 -> (W x)./w, trans=1 -> W'(x./w). */
void or_filter_apply(int64_t n, const int64_t *rp, const int32_t *ci, const double *v,
                     const double *w, const double *x, int trans, double *y);

/* Unfiltered compliance sensitivity g0 (Implementation/Grad.m:18-23): u is the
 * full-length displacement vector (fixed DOFs 0), x the filtered densities,
 * freedens an optional 0/1 element mask (NULL = all). */
void or_grad0(const or_ctx *c, const double *u_full, const double *x, const uint8_t *freedens,
              double *g0);
/* Multiresolution g0 on the density mesh (Implementation/Grad.m:24-49, mr.interp=false). */
void or_grad0_mr(const or_ctx *c, int n, const double *u_full, const double *xd, double *g0d);

#ifdef __cplusplus
}
#endif
#endif
