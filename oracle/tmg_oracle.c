/*
 * tmg_oracle.c -- FP64 CPU restatement of TopOpt-SLP-3D's equilibrium solve.
 *
 * TEST INFRASTRUCTURE ONLY. This is the parity oracle and the CPU baseline:
 * tests, the smoke check and the benchmark's cpu_baseline leg load it as the
 * checker. The HIP product path never links, calls or falls back to it.
 *
 * It restates the reference's default assembled path (genK = true,
 * opsolver = 4, smoother = 0, cycle = 0):
 *   element stiffness        Implementation/ElemStiffMatrix.m (via exact 2x2x2 Gauss,
 *                            the recipe of Implementation/ElemStiffMatrixGQMR.m with n = 1)
 *   global stiffness         Implementation/GlobalStiffMatrix.m:16-18,64-65
 *   prolongation             Implementation/Prolongation.m:15-55
 *   Galerkin hierarchy       Implementation/SetupGMG.m:21-75
 *   Jacobi smoother          Implementation/Smooth.m:14-18
 *   V-cycle                  Implementation/VCycle.m:23-43
 *   MG-preconditioned CG     Implementation/SolveMG.m:26-53
 *   compliance gradient      Implementation/Grad.m:18-62
 *   density filter           Implementation/ElemNeighborhood.m:10-98, Implementation/ApplyFilter.m:11-19
 * The coarse Cholesky uses the natural DOF order where the reference uses
 * amd(); ordering changes rounding only.
 */
#include "tmg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAXL 12

typedef struct {
    int64_t n, m, nnz;
    int64_t *rp;
    int32_t *ci;
    double *v;
} csr_t;

typedef struct {
    int nelx, nely, nelz;
    int64_t nx, ny, nz, nn, ndof, nfree;
    int64_t *c_of_f; /* full dof -> compact index, -1 if fixed */
    int64_t *f_of_c; /* compact -> full dof */
} level_t;

struct or_ctx {
    double el, eh, ew, E0, nu, p, Emin, omega;
    double ke[576];
    int L;
    level_t lev[OR_MAXL];
    csr_t A[OR_MAXL];
    csr_t P[OR_MAXL];
    csr_t PT[OR_MAXL];
    double *Mdiag[OR_MAXL];
    double *R; /* dense upper Cholesky factor of A{L}, row-major */
    int64_t nR;
    double *Elast; /* moduli of the last single-resolution assembly (for the EbE product) */
};

/* ------------------------------------------------------------------------- */
/* small helpers                                                              */
/* ------------------------------------------------------------------------- */

static void csr_free(csr_t *a) {
    free(a->rp); free(a->ci); free(a->v);
    memset(a, 0, sizeof(*a));
}

/* Deterministic dot product: fixed 4096-blocks, partials summed in order. */
static double ddot(int64_t n, const double *a, const double *b) {
    const int64_t B = 4096;
    int64_t nb = (n + B - 1) / B;
    double *part = (double *)malloc(sizeof(double) * (size_t)(nb > 0 ? nb : 1));
    #pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nb; ++k) {
        int64_t lo = k * B, hi = lo + B < n ? lo + B : n;
        double s = 0.0;
        for (int64_t i = lo; i < hi; ++i) s += a[i] * b[i];
        part[k] = s;
    }
    double s = 0.0;
    for (int64_t k = 0; k < nb; ++k) s += part[k];
    free(part);
    return s;
}

static void spmv(const csr_t *a, const double *x, double *y) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < a->n; ++i) {
        double s = 0.0;
        for (int64_t k = a->rp[i]; k < a->rp[i + 1]; ++k) s += a->v[k] * x[a->ci[k]];
        y[i] = s;
    }
}

static int cmp_i32(const void *x, const void *y) {
    int32_t a = *(const int32_t *)x, b = *(const int32_t *)y;
    return (a > b) - (a < b);
}

/* C = A*B (Gustavson, per-row dense accumulators, sorted columns). */
static int spgemm(const csr_t *A, const csr_t *B, csr_t *C) {
    int64_t n = A->n, m = B->m;
    C->n = n; C->m = m;
    C->rp = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    if (!C->rp) return -1;
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    int32_t *markall = (int32_t *)malloc(sizeof(int32_t) * (size_t)m * nthr);
    if (!markall) return -1;
    /* pass 1: count */
    #pragma omp parallel
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        int32_t *mark = markall + (size_t)t * m;
        for (int64_t j = 0; j < m; ++j) mark[j] = -1;
        #pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < n; ++i) {
            int64_t cnt = 0;
            for (int64_t ka = A->rp[i]; ka < A->rp[i + 1]; ++ka) {
                int32_t k = A->ci[ka];
                for (int64_t kb = B->rp[k]; kb < B->rp[k + 1]; ++kb) {
                    int32_t j = B->ci[kb];
                    if (mark[j] != (int32_t)(i & 0x7fffffff)) { mark[j] = (int32_t)(i & 0x7fffffff); ++cnt; }
                }
            }
            C->rp[i + 1] = cnt;
        }
    }
    for (int64_t i = 0; i < n; ++i) C->rp[i + 1] += C->rp[i];
    C->nnz = C->rp[n];
    C->ci = (int32_t *)malloc(sizeof(int32_t) * (size_t)(C->nnz > 0 ? C->nnz : 1));
    C->v = (double *)malloc(sizeof(double) * (size_t)(C->nnz > 0 ? C->nnz : 1));
    double *accall = (double *)malloc(sizeof(double) * (size_t)m * nthr);
    if (!C->ci || !C->v || !accall) return -1;
    /* pass 2: fill */
    #pragma omp parallel
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        int32_t *mark = markall + (size_t)t * m;
        double *acc = accall + (size_t)t * m;
        for (int64_t j = 0; j < m; ++j) { mark[j] = -1; acc[j] = 0.0; }
        #pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < n; ++i) {
            int64_t pos = C->rp[i];
            for (int64_t ka = A->rp[i]; ka < A->rp[i + 1]; ++ka) {
                int32_t k = A->ci[ka];
                double av = A->v[ka];
                for (int64_t kb = B->rp[k]; kb < B->rp[k + 1]; ++kb) {
                    int32_t j = B->ci[kb];
                    if (mark[j] != (int32_t)(i & 0x7fffffff)) {
                        mark[j] = (int32_t)(i & 0x7fffffff);
                        C->ci[pos++] = j;
                        acc[j] = 0.0;
                    }
                    acc[j] += av * B->v[kb];
                }
            }
            qsort(C->ci + C->rp[i], (size_t)(C->rp[i + 1] - C->rp[i]), sizeof(int32_t), cmp_i32);
            for (int64_t q = C->rp[i]; q < C->rp[i + 1]; ++q) C->v[q] = acc[C->ci[q]];
        }
    }
    free(markall); free(accall);
    return 0;
}

static int csr_transpose(const csr_t *A, csr_t *T) {
    T->n = A->m; T->m = A->n; T->nnz = A->nnz;
    T->rp = (int64_t *)calloc((size_t)T->n + 1, sizeof(int64_t));
    T->ci = (int32_t *)malloc(sizeof(int32_t) * (size_t)(A->nnz > 0 ? A->nnz : 1));
    T->v = (double *)malloc(sizeof(double) * (size_t)(A->nnz > 0 ? A->nnz : 1));
    if (!T->rp || !T->ci || !T->v) return -1;
    for (int64_t k = 0; k < A->nnz; ++k) T->rp[A->ci[k] + 1]++;
    for (int64_t i = 0; i < T->n; ++i) T->rp[i + 1] += T->rp[i];
    int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)(T->n + 1));
    memcpy(pos, T->rp, sizeof(int64_t) * (size_t)(T->n + 1));
    for (int64_t i = 0; i < A->n; ++i)
        for (int64_t k = A->rp[i]; k < A->rp[i + 1]; ++k) {
            int64_t q = pos[A->ci[k]]++;
            T->ci[q] = (int32_t)i;
            T->v[q] = A->v[k];
        }
    free(pos);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* element stiffness                                                          */
/* ------------------------------------------------------------------------- */

/* Integrate B^T D B over the sub-box [x0,x1]x[y0,y1]x[z0,z1] of the reference
 * cube [-1,1]^3 with 2-point Gauss per direction (exact for Q1 on boxes), for
 * a physical element with half-sizes a, b, c. Accumulates into k (24x24). */
static void integrate_box(double a, double b, double c, double nu,
                          double x0, double x1, double y0, double y1, double z0, double z1,
                          double *k) {
    const double g = 1.0 / sqrt(3.0);
    const double gp[2] = {-g, g};
    double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double mu = 1.0 / (2.0 * (1.0 + nu));
    /* D for E = 1: normal block lam + 2mu on diagonal, lam off; shear mu */
    double hx = 0.5 * (x1 - x0), hy = 0.5 * (y1 - y0), hz = 0.5 * (z1 - z0);
    double wdet = hx * hy * hz * a * b * c; /* sub-box map times element jacobian */
    for (int gz = 0; gz < 2; ++gz)
        for (int gx = 0; gx < 2; ++gx)
            for (int gy = 0; gy < 2; ++gy) {
                double xi = hx * gp[gx] + 0.5 * (x1 + x0);
                double et = hy * gp[gy] + 0.5 * (y1 + y0);
                double ze = hz * gp[gz] + 0.5 * (z1 + z0);
                double dN[8][3];
                for (int an = 0; an < 8; ++an) {
                    int dy = an & 1, dx = (an >> 1) & 1, dz = (an >> 2) & 1;
                    double sx = dx ? 1.0 : -1.0, sy = dy ? 1.0 : -1.0, sz = dz ? 1.0 : -1.0;
                    dN[an][0] = sx * (1.0 + sy * et) * (1.0 + sz * ze) / (8.0 * a);
                    dN[an][1] = sy * (1.0 + sx * xi) * (1.0 + sz * ze) / (8.0 * b);
                    dN[an][2] = sz * (1.0 + sx * xi) * (1.0 + sy * et) / (8.0 * c);
                }
                /* B is 6x24: rows exx, eyy, ezz, gxy, gyz, gxz */
                double B[6][24];
                memset(B, 0, sizeof(B));
                for (int an = 0; an < 8; ++an) {
                    B[0][3 * an + 0] = dN[an][0];
                    B[1][3 * an + 1] = dN[an][1];
                    B[2][3 * an + 2] = dN[an][2];
                    B[3][3 * an + 0] = dN[an][1];
                    B[3][3 * an + 1] = dN[an][0];
                    B[4][3 * an + 1] = dN[an][2];
                    B[4][3 * an + 2] = dN[an][1];
                    B[5][3 * an + 0] = dN[an][2];
                    B[5][3 * an + 2] = dN[an][0];
                }
                double DB[6][24];
                for (int j = 0; j < 24; ++j) {
                    double e0 = B[0][j], e1 = B[1][j], e2 = B[2][j];
                    DB[0][j] = (lam + 2 * mu) * e0 + lam * e1 + lam * e2;
                    DB[1][j] = lam * e0 + (lam + 2 * mu) * e1 + lam * e2;
                    DB[2][j] = lam * e0 + lam * e1 + (lam + 2 * mu) * e2;
                    DB[3][j] = mu * B[3][j];
                    DB[4][j] = mu * B[4][j];
                    DB[5][j] = mu * B[5][j];
                }
                for (int i = 0; i < 24; ++i)
                    for (int j = 0; j < 24; ++j) {
                        double s = 0.0;
                        for (int r = 0; r < 6; ++r) s += B[r][i] * DB[r][j];
                        k[i * 24 + j] += wdet * s;
                    }
            }
}

void or_elem_stiff(double el, double eh, double ew, double nu, double *k) {
    memset(k, 0, sizeof(double) * 576);
    integrate_box(0.5 * el, 0.5 * eh, 0.5 * ew, nu, -1, 1, -1, 1, -1, 1, k);
    for (int i = 0; i < 24; ++i) /* exact symmetry, as ElemStiffMatrix.m:325-329 */
        for (int j = i + 1; j < 24; ++j) k[j * 24 + i] = k[i * 24 + j];
}

void or_elem_stiff_mr(double el, double eh, double ew, double nu, int n, double *kout) {
    int idx = 0;
    for (int sz = 0; sz < n; ++sz)
        for (int sx = 0; sx < n; ++sx)
            for (int sy = 0; sy < n; ++sy) {
                double *k = kout + (size_t)576 * idx++;
                memset(k, 0, sizeof(double) * 576);
                double cx = -1.0 + (2.0 * sx + 1.0) / n, cy = -1.0 + (2.0 * sy + 1.0) / n,
                       cz = -1.0 + (2.0 * sz + 1.0) / n;
                integrate_box(0.5 * el, 0.5 * eh, 0.5 * ew, nu, cx - 1.0 / n, cx + 1.0 / n,
                              cy - 1.0 / n, cy + 1.0 / n, cz - 1.0 / n, cz + 1.0 / n, k);
                for (int i = 0; i < 24; ++i) /* ElemStiffMatrixGQMR.m:234 */
                    for (int j = i + 1; j < 24; ++j) {
                        double s = 0.5 * (k[i * 24 + j] + k[j * 24 + i]);
                        k[i * 24 + j] = s; k[j * 24 + i] = s;
                    }
            }
}

/* ------------------------------------------------------------------------- */
/* context / levels                                                           */
/* ------------------------------------------------------------------------- */

static void level_dims(level_t *lv, int nelx, int nely, int nelz) {
    lv->nelx = nelx; lv->nely = nely; lv->nelz = nelz;
    lv->nx = nelx + 1; lv->ny = nely + 1; lv->nz = nelz + 1;
    lv->nn = lv->nx * lv->ny * lv->nz;
    lv->ndof = 3 * lv->nn;
}

static int level_set_free(level_t *lv, const uint8_t *freemask) {
    lv->c_of_f = (int64_t *)malloc(sizeof(int64_t) * (size_t)lv->ndof);
    int64_t nf = 0;
    for (int64_t d = 0; d < lv->ndof; ++d) nf += freemask[d] ? 1 : 0;
    lv->nfree = nf;
    lv->f_of_c = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nf > 0 ? nf : 1));
    if (!lv->c_of_f || !lv->f_of_c) return -1;
    int64_t q = 0;
    for (int64_t d = 0; d < lv->ndof; ++d) {
        if (freemask[d]) { lv->c_of_f[d] = q; lv->f_of_c[q++] = d; }
        else lv->c_of_f[d] = -1;
    }
    return 0;
}

or_ctx *or_create(int nelx, int nely, int nelz, double el, double eh, double ew,
                  double E0, double nu, double p, double Emin,
                  const int64_t *freedofs1, int64_t nfree) {
    or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
    if (!c) return NULL;
    c->el = el; c->eh = eh; c->ew = ew; c->E0 = E0; c->nu = nu; c->p = p; c->Emin = Emin;
    or_elem_stiff(el, eh, ew, nu, c->ke);
    level_dims(&c->lev[0], nelx, nely, nelz);
    uint8_t *mask = (uint8_t *)calloc((size_t)c->lev[0].ndof, 1);
    for (int64_t i = 0; i < nfree; ++i) {
        int64_t d = freedofs1[i] - 1;
        if (d >= 0 && d < c->lev[0].ndof) mask[d] = 1;
    }
    level_set_free(&c->lev[0], mask);
    free(mask);
    c->L = 1;
    return c;
}

void or_destroy(or_ctx *c) {
    if (!c) return;
    for (int l = 0; l < OR_MAXL; ++l) {
        free(c->lev[l].c_of_f); free(c->lev[l].f_of_c);
        csr_free(&c->A[l]); csr_free(&c->P[l]); csr_free(&c->PT[l]);
        free(c->Mdiag[l]);
    }
    free(c->R);
    free(c->Elast);
    free(c);
}

int64_t or_nfree(const or_ctx *c, int level) { return c->lev[level].nfree; }
int64_t or_ndof(const or_ctx *c, int level) { return c->lev[level].ndof; }
int or_ngrids(const or_ctx *c) { return c->L; }

/* OpenMP threads of the calls that follow (the CPU baseline's all-core and
 * one-core legs); returns the previous count. */
int or_set_threads(int n) {
    const int prev = omp_get_max_threads();
    if (n > 0) omp_set_num_threads(n);
    return prev;
}
void or_level_dims(const or_ctx *c, int level, int *nel3) {
    nel3[0] = c->lev[level].nelx; nel3[1] = c->lev[level].nely; nel3[2] = c->lev[level].nelz;
}
void or_level_free(const or_ctx *c, int level, int64_t *out) {
    memcpy(out, c->lev[level].f_of_c, sizeof(int64_t) * (size_t)c->lev[level].nfree);
}
void or_jacobi_diag(const or_ctx *c, int level, double *m) {
    memcpy(m, c->Mdiag[level], sizeof(double) * (size_t)c->lev[level].nfree);
}

/* ------------------------------------------------------------------------- */
/* global stiffness (node-row assembly of sum_e E_e Ke, free rows/cols)       */
/* ------------------------------------------------------------------------- */

/* Element matrix accessor: either E_e * ke (single resolution) or the
 * per-element sum of weighted sub-element integrands (multiresolution). */
typedef struct {
    const double *E;     /* n_el moduli (single resolution) */
    const double *kemr;  /* n^3 x 576 integrands (multiresolution) */
    const double *Ed;    /* density-mesh moduli (multiresolution) */
    int n;               /* multiresolution factor (0 = off) */
} elemsrc_t;

static int assemble_generic(or_ctx *c, const elemsrc_t *src) {
    level_t *lv = &c->lev[0];
    csr_free(&c->A[0]);
    csr_t *A = &c->A[0];
    int64_t nx = lv->nx, ny = lv->ny, nz = lv->nz;
    int nelx = lv->nelx, nely = lv->nely;
    A->n = A->m = lv->nfree;
    A->rp = (int64_t *)calloc((size_t)A->n + 1, sizeof(int64_t));
    if (!A->rp) return -1;
    /* count */
    #pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < A->n; ++r) {
        int64_t d = lv->f_of_c[r], node = d / 3;
        int64_t iz = node / (nx * ny), ix = (node / ny) % nx, iy = node % ny;
        int64_t cnt = 0;
        for (int64_t jz = iz - 1; jz <= iz + 1; ++jz)
            for (int64_t jx = ix - 1; jx <= ix + 1; ++jx)
                for (int64_t jy = iy - 1; jy <= iy + 1; ++jy) {
                    if (jz < 0 || jz >= nz || jx < 0 || jx >= nx || jy < 0 || jy >= ny) continue;
                    int64_t jn = jz * nx * ny + jx * ny + jy;
                    for (int dd = 0; dd < 3; ++dd) cnt += lv->c_of_f[3 * jn + dd] >= 0;
                }
        A->rp[r + 1] = cnt;
    }
    for (int64_t r = 0; r < A->n; ++r) A->rp[r + 1] += A->rp[r];
    A->nnz = A->rp[A->n];
    A->ci = (int32_t *)malloc(sizeof(int32_t) * (size_t)A->nnz);
    A->v = (double *)malloc(sizeof(double) * (size_t)A->nnz);
    if (!A->ci || !A->v) return -1;
    int nmr = src->n, nmr3 = nmr * nmr * nmr;
    int64_t NXd = (int64_t)nelx * nmr, NYd = (int64_t)nely * nmr;
    #pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < A->n; ++r) {
        int64_t d = lv->f_of_c[r], node = d / 3;
        int ci_ = (int)(d % 3);
        int64_t iz = node / (nx * ny), ix = (node / ny) % nx, iy = node % ny;
        int64_t pos = A->rp[r];
        for (int64_t jz = iz - 1; jz <= iz + 1; ++jz)
            for (int64_t jx = ix - 1; jx <= ix + 1; ++jx)
                for (int64_t jy = iy - 1; jy <= iy + 1; ++jy) {
                    if (jz < 0 || jz >= nz || jx < 0 || jx >= nx || jy < 0 || jy >= ny) continue;
                    int64_t jn = jz * nx * ny + jx * ny + jy;
                    double blk[3] = {0, 0, 0};
                    /* elements containing both nodes */
                    int64_t ez0 = (iz > jz ? iz : jz) - 1, ez1 = iz < jz ? iz : jz;
                    int64_t ex0 = (ix > jx ? ix : jx) - 1, ex1 = ix < jx ? ix : jx;
                    int64_t ey0 = (iy > jy ? iy : jy) - 1, ey1 = iy < jy ? iy : jy;
                    for (int64_t ez = ez0; ez <= ez1; ++ez)
                        for (int64_t ex = ex0; ex <= ex1; ++ex)
                            for (int64_t ey = ey0; ey <= ey1; ++ey) {
                                if (ez < 0 || ez >= nz - 1 || ex < 0 || ex >= nx - 1 || ey < 0 || ey >= ny - 1)
                                    continue;
                                int ai = (int)((iy - ey) + 2 * (ix - ex) + 4 * (iz - ez));
                                int aj = (int)((jy - ey) + 2 * (jx - ex) + 4 * (jz - ez));
                                if (!nmr) {
                                    double Ee = src->E[ez * nelx * nely + ex * nely + ey];
                                    for (int dd = 0; dd < 3; ++dd)
                                        blk[dd] += Ee * c->ke[(3 * ai + ci_) * 24 + 3 * aj + dd];
                                } else {
                                    for (int s = 0; s < nmr3; ++s) {
                                        int sy = s % nmr, sx = (s / nmr) % nmr, sz = s / (nmr * nmr);
                                        int64_t v = (ez * nmr + sz) * NXd * NYd + (ex * nmr + sx) * NYd + (ey * nmr + sy);
                                        double Ev = src->Ed[v];
                                        const double *k = src->kemr + (size_t)576 * s;
                                        for (int dd = 0; dd < 3; ++dd)
                                            blk[dd] += Ev * k[(3 * ai + ci_) * 24 + 3 * aj + dd];
                                    }
                                }
                            }
                    for (int dd = 0; dd < 3; ++dd) {
                        int64_t cc = lv->c_of_f[3 * jn + dd];
                        if (cc >= 0) { A->ci[pos] = (int32_t)cc; A->v[pos] = blk[dd]; ++pos; }
                    }
                }
    }
    return 0;
}

int or_assemble(or_ctx *c, const double *x) {
    int64_t ne = (int64_t)c->lev[0].nelx * c->lev[0].nely * c->lev[0].nelz;
    double *E = (double *)malloc(sizeof(double) * (size_t)ne);
    for (int64_t e = 0; e < ne; ++e) E[e] = c->Emin + (c->E0 - c->Emin) * pow(x[e], c->p);
    elemsrc_t s = {E, NULL, NULL, 0};
    int rc = assemble_generic(c, &s);
    free(c->Elast);
    c->Elast = E;
    return rc;
}

int or_assemble_mr(or_ctx *c, int n, const double *xd) {
    int64_t ne = (int64_t)c->lev[0].nelx * c->lev[0].nely * c->lev[0].nelz * n * n * n;
    double *Ed = (double *)malloc(sizeof(double) * (size_t)ne);
    double *kmr = (double *)malloc(sizeof(double) * 576 * (size_t)(n * n * n));
    for (int64_t e = 0; e < ne; ++e) Ed[e] = c->Emin + (c->E0 - c->Emin) * pow(xd[e], c->p);
    or_elem_stiff_mr(c->el, c->eh, c->ew, c->nu, n, kmr);
    elemsrc_t s = {NULL, kmr, Ed, n};
    int rc = assemble_generic(c, &s);
    free(Ed); free(kmr);
    return rc;
}

/* Matrix-free K x on free vectors, element by element, scattering in element
 * order as the reference loop does (Implementation/WOK/KelemProd.m:14-16,57-100,185).
 * Uses the moduli of the last or_assemble call. Serial. */
void or_kelemprod(const or_ctx *c, const double *x, double *y) {
    const level_t *lv = &c->lev[0];
    double *xc = (double *)calloc((size_t)lv->ndof, sizeof(double));
    double *yc = (double *)calloc((size_t)lv->ndof, sizeof(double));
    for (int64_t q = 0; q < lv->nfree; ++q) xc[lv->f_of_c[q]] = x[q];   /* xc(freedofs) = x */
    int64_t ne = (int64_t)lv->nelx * lv->nely * lv->nelz;
    for (int64_t e = 0; e < ne; ++e) {
        int64_t ez = e / ((int64_t)lv->nelx * lv->nely), ex = (e / lv->nely) % lv->nelx, ey = e % lv->nely;
        int64_t dofs[24];
        double xe[24];
        for (int a = 0; a < 8; ++a) {
            int dy = a & 1, dx = (a >> 1) & 1, dz = (a >> 2) & 1;
            int64_t node = (ez + dz) * lv->nx * lv->ny + (ex + dx) * lv->ny + (ey + dy);
            for (int cc = 0; cc < 3; ++cc) { dofs[3 * a + cc] = 3 * node + cc; xe[3 * a + cc] = c->Elast[e] * xc[3 * node + cc]; }
        }
        for (int i = 0; i < 24; ++i) {
            double t = 0.0;
            for (int j = 0; j < 24; ++j) t += c->ke[i * 24 + j] * xe[j];
            yc[dofs[i]] += t;
        }
    }
    for (int64_t q = 0; q < lv->nfree; ++q) y[q] = yc[lv->f_of_c[q]];   /* y = y(freedofs) */
    free(xc); free(yc);
}

/* ------------------------------------------------------------------------- */
/* prolongation (Implementation/Prolongation.m:15-55)                          */
/* ------------------------------------------------------------------------- */

static int build_prolongation(const level_t *f, level_t *cl, csr_t *P) {
    /* coarse dof free iff the co-located fine dof (2I, c) is free
     * (Prolongation.m:38-46,54) */
    uint8_t *cm = (uint8_t *)malloc((size_t)cl->ndof);
    for (int64_t iz = 0; iz < cl->nz; ++iz)
        for (int64_t ix = 0; ix < cl->nx; ++ix)
            for (int64_t iy = 0; iy < cl->ny; ++iy) {
                int64_t I = iz * cl->nx * cl->ny + ix * cl->ny + iy;
                int64_t fi = (2 * iz) * f->nx * f->ny + (2 * ix) * f->ny + 2 * iy;
                for (int cc = 0; cc < 3; ++cc) cm[3 * I + cc] = f->c_of_f[3 * fi + cc] >= 0;
            }
    level_set_free(cl, cm);
    free(cm);
    P->n = f->nfree; P->m = cl->nfree;
    P->rp = (int64_t *)calloc((size_t)P->n + 1, sizeof(int64_t));
    /* at most 8 coarse nodes per fine dof */
    int32_t *ci = (int32_t *)malloc(sizeof(int32_t) * 8 * (size_t)(P->n > 0 ? P->n : 1));
    double *v = (double *)malloc(sizeof(double) * 8 * (size_t)(P->n > 0 ? P->n : 1));
    int64_t pos = 0;
    for (int64_t r = 0; r < P->n; ++r) {
        int64_t d = f->f_of_c[r], node = d / 3;
        int cc = (int)(d % 3);
        int64_t iz = node / (f->nx * f->ny), ix = (node / f->ny) % f->nx, iy = node % f->ny;
        int64_t z0 = iz / 2, z1 = (iz + 1) / 2, x0 = ix / 2, x1 = (ix + 1) / 2, y0 = iy / 2, y1 = (iy + 1) / 2;
        for (int64_t Z = z0; Z <= z1; ++Z)
            for (int64_t X = x0; X <= x1; ++X)
                for (int64_t Y = y0; Y <= y1; ++Y) {
                    int64_t I = Z * cl->nx * cl->ny + X * cl->ny + Y;
                    int64_t col = cl->c_of_f[3 * I + cc];
                    if (col < 0) continue;
                    /* w = [1, .5, .25, .125] indexed by squared offset + 1 (Prolongation.m:15,34) */
                    int64_t dz = iz - 2 * Z, dx = ix - 2 * X, dy = iy - 2 * Y;
                    int val = (int)(dx * dx + dy * dy + dz * dz);
                    static const double w[4] = {1.0, 0.5, 0.25, 0.125};
                    ci[pos] = (int32_t)col; v[pos] = w[val]; ++pos;
                }
        P->rp[r + 1] = pos;
    }
    P->nnz = pos;
    P->ci = ci; P->v = v;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* SetupGMG                                                                   */
/* ------------------------------------------------------------------------- */

static int dense_cholesky_upper(int64_t n, double *a) {
    /* a (row-major, full symmetric) -> upper R with R^T R = A, lower part zeroed */
    for (int64_t k = 0; k < n; ++k) {
        double d = a[k * n + k];
        if (!(d > 0.0)) return -2;
        d = sqrt(d);
        a[k * n + k] = d;
        for (int64_t j = k + 1; j < n; ++j) a[k * n + j] /= d;
        #pragma omp parallel for schedule(static)
        for (int64_t i = k + 1; i < n; ++i) {
            double rki = a[k * n + i];
            if (rki == 0.0) continue;
            double *ri = a + i * n;
            const double *rk = a + k * n;
            for (int64_t j = i; j < n; ++j) ri[j] -= rki * rk[j];
        }
    }
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < i; ++j) a[i * n + j] = 0.0;
    return 0;
}

static int galerkin_and_factor(or_ctx *c) {
    for (int l = 1; l < c->L; ++l) {
        csr_t AP;
        memset(&AP, 0, sizeof(AP));
        csr_free(&c->A[l]);
        if (spgemm(&c->A[l - 1], &c->P[l - 1], &AP)) return -1;
        if (spgemm(&c->PT[l - 1], &AP, &c->A[l])) return -1;
        csr_free(&AP);
    }
    for (int l = 0; l < c->L; ++l) {
        free(c->Mdiag[l]);
        int64_t n = c->lev[l].nfree;
        c->Mdiag[l] = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        for (int64_t i = 0; i < n; ++i) {
            double dg = 0.0;
            for (int64_t k = c->A[l].rp[i]; k < c->A[l].rp[i + 1]; ++k)
                if (c->A[l].ci[k] == i) dg = c->A[l].v[k];
            c->Mdiag[l][i] = (1.0 / c->omega) * dg; /* SetupGMG.m:67 */
        }
    }
    int Lc = c->L - 1;
    int64_t n = c->lev[Lc].nfree;
    free(c->R);
    c->nR = n;
    c->R = (double *)calloc((size_t)(n * n > 0 ? n * n : 1), sizeof(double));
    if (!c->R) return -1;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = c->A[Lc].rp[i]; k < c->A[Lc].rp[i + 1]; ++k) c->R[i * n + c->A[Lc].ci[k]] = c->A[Lc].v[k];
    return dense_cholesky_upper(n, c->R);
}

int or_setup_gmg(or_ctx *c, int ngrids, double omega) {
    if (ngrids < 2 || ngrids > OR_MAXL) return -1;
    c->omega = omega;
    for (int l = 1; l < c->L; ++l) {
        free(c->lev[l].c_of_f); free(c->lev[l].f_of_c);
        csr_free(&c->P[l - 1]); csr_free(&c->PT[l - 1]);
    }
    c->L = ngrids;
    for (int l = 1; l < ngrids; ++l) {
        const level_t *f = &c->lev[l - 1];
        if (f->nelx % 2 || f->nely % 2 || f->nelz % 2) return -1;
        level_dims(&c->lev[l], f->nelx / 2, f->nely / 2, f->nelz / 2);
        build_prolongation(f, &c->lev[l], &c->P[l - 1]);
        csr_transpose(&c->P[l - 1], &c->PT[l - 1]);
    }
    return galerkin_and_factor(c);
}

int or_refresh_gmg(or_ctx *c) { return galerkin_and_factor(c); }

/* ------------------------------------------------------------------------- */
/* smoother, V-cycle, PCG                                                     */
/* ------------------------------------------------------------------------- */

void or_spmv(const or_ctx *c, int level, const double *x, double *y) { spmv(&c->A[level], x, y); }

void or_smooth(const or_ctx *c, int level, const double *b, const double *x0, int niter, double *x) {
    int64_t n = c->lev[level].nfree;
    double *t = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (x != x0) memcpy(x, x0, sizeof(double) * (size_t)n);
    for (int it = 0; it < niter; ++it) {
        spmv(&c->A[level], x, t);
        const double *M = c->Mdiag[level];
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) x[i] = x[i] + (b[i] - t[i]) / M[i]; /* x + M\(b - A x) */
    }
    free(t);
}

static void coarse_solve(const or_ctx *c, const double *b, double *x) {
    /* eH = R \ (R' \ rH) (VCycle.m:35-36) */
    int64_t n = c->nR;
    const double *R = c->R;
    double *y = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) { /* R' y = b (R' lower) */
        double s = b[i];
        for (int64_t k = 0; k < i; ++k) s -= R[k * n + i] * y[k];
        y[i] = s / R[i * n + i];
    }
    for (int64_t i = n - 1; i >= 0; --i) { /* R x = y */
        double s = y[i];
        for (int64_t k = i + 1; k < n; ++k) s -= R[i * n + k] * x[k];
        x[i] = s / R[i * n + i];
    }
    free(y);
}

/* x = VCycle(A, b, x0, nu1, nu2, ..., nv) (Implementation/VCycle.m:23-43);
 * x0 may be NULL (zero initial guess, as SolveMG calls it). */
void or_vcycle_x0(const or_ctx *c, int nv, const double *b, const double *x0, int nu1, int nu2, double *x) {
    int64_t n = c->lev[nv].nfree, nc = c->lev[nv + 1].nfree;
    double *zero = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double *rh = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double *rH = (double *)malloc(sizeof(double) * (size_t)(nc > 0 ? nc : 1));
    double *eH = (double *)calloc((size_t)(nc > 0 ? nc : 1), sizeof(double));
    double *eh = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    or_smooth(c, nv, b, x0 ? x0 : zero, nu1, x);       /* VCycle.m:23 */
    spmv(&c->A[nv], x, rh);                             /* VCycle.m:24 */
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) rh[i] = b[i] - rh[i];
    spmv(&c->PT[nv], rh, rH);                           /* VCycle.m:25 */
    if (nv == c->L - 2) coarse_solve(c, rH, eH);        /* VCycle.m:35-36 */
    else or_vcycle_x0(c, nv + 1, rH, NULL, nu1, nu2, eH);   /* VCycle.m:39 */
    spmv(&c->P[nv], eH, eh);                            /* VCycle.m:41 */
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) x[i] = x[i] + eh[i];
    or_smooth(c, nv, b, x, nu2, x);                     /* VCycle.m:43 */
    free(zero); free(rh); free(rH); free(eH); free(eh);
}

void or_vcycle(const or_ctx *c, int nv, const double *b, int nu1, int nu2, double *x) {
    or_vcycle_x0(c, nv, b, NULL, nu1, nu2, x);
}

/* x = FMVCycle(A, b, nu1, nu2, ..., nv) (Implementation/FMVCycle.m:21-39):
 * restrict b, solve the coarsest system directly or recurse, prolong, then
 * one V-cycle from that initial guess. */
void or_fmvcycle(const or_ctx *c, int nv, const double *b, int nu1, int nu2, double *x) {
    int64_t n = c->lev[nv].nfree, nc = c->lev[nv + 1].nfree;
    double *rH = (double *)malloc(sizeof(double) * (size_t)(nc > 0 ? nc : 1));
    double *xc = (double *)calloc((size_t)(nc > 0 ? nc : 1), sizeof(double));
    double *xp = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    spmv(&c->PT[nv], b, rH);                            /* FMVCycle.m:21 */
    if (nv == c->L - 2) coarse_solve(c, rH, xc);        /* FMVCycle.m:22-33 */
    else or_fmvcycle(c, nv + 1, rH, nu1, nu2, xc);      /* FMVCycle.m:35 */
    spmv(&c->P[nv], xc, xp);                            /* FMVCycle.m:37 */
    or_vcycle_x0(c, nv, b, xp, nu1, nu2, x);            /* FMVCycle.m:38 */
    free(rH); free(xc); free(xp);
}

/* x = WCycle(A, b, x0, nu1, nu2, ..., nv) (Implementation/WCycle.m:1-27):
 * the coarse correction recurses twice, the second time from the first
 * result; the level above the coarsest solves directly once. x0 may be NULL
 * (zero initial guess). */
void or_wcycle(const or_ctx *c, int nv, const double *b, const double *x0, int nu1, int nu2, double *x) {
    int64_t n = c->lev[nv].nfree, nc = c->lev[nv + 1].nfree;
    double *zero = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double *rh = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double *rH = (double *)malloc(sizeof(double) * (size_t)(nc > 0 ? nc : 1));
    double *eH = (double *)calloc((size_t)(nc > 0 ? nc : 1), sizeof(double));
    double *eh = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    or_smooth(c, nv, b, x0 ? x0 : zero, nu1, x);       /* WCycle.m:3 */
    spmv(&c->A[nv], x, rh);                             /* WCycle.m:4 */
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) rh[i] = b[i] - rh[i];
    spmv(&c->PT[nv], rh, rH);                           /* WCycle.m:5 */
    if (nv == c->L - 2) {
        coarse_solve(c, rH, eH);                        /* WCycle.m:15-16 */
    } else {
        double *e1 = (double *)calloc((size_t)(nc > 0 ? nc : 1), sizeof(double));
        or_wcycle(c, nv + 1, rH, NULL, nu1, nu2, e1);   /* WCycle.m:19 */
        or_wcycle(c, nv + 1, rH, e1, nu1, nu2, eH);     /* WCycle.m:20 */
        free(e1);
    }
    spmv(&c->P[nv], eH, eh);                            /* WCycle.m:22 */
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) x[i] = x[i] + eh[i];
    or_smooth(c, nv, b, x, nu2, x);                     /* WCycle.m:24 */
    free(zero); free(rh); free(rH); free(eH); free(eh);
}

int or_solve_mg(const or_ctx *c, double *x, const double *f, int nu1, int nu2,
                double tol, int maxiter, double *relres) {
    return or_solve_mg_cycle(c, x, f, 0, nu1, nu2, tol, maxiter, relres);
}

int or_solve_mg_cycle(const or_ctx *c, double *x, const double *f, int cycle, int nu1, int nu2,
                      double tol, int maxiter, double *relres) {
    int64_t n = c->lev[0].nfree;
    size_t bytes = sizeof(double) * (size_t)(n > 0 ? n : 1);
    double *r = (double *)malloc(bytes), *s = (double *)malloc(bytes);
    double *d = (double *)malloc(bytes), *q = (double *)malloc(bytes);
    spmv(&c->A[0], x, r);                               /* SolveMG.m:26 */
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) r[i] = f[i] - r[i];
    double nf = sqrt(ddot(n, f, f));
    double nr = sqrt(ddot(n, r, r)) / nf;               /* SolveMG.m:27 */
    int it = 1;
    double gamma2 = 0.0;
    while (nr > tol && it <= maxiter) {                 /* SolveMG.m:30 */
        if (c->L >= 2) {                                /* SolveMG.m:31-37 */
            if (cycle == 1) or_wcycle(c, 0, r, NULL, nu1, nu2, s);
            else if (cycle == 2) or_fmvcycle(c, 0, r, nu1, nu2, s);
            else or_vcycle(c, 0, r, nu1, nu2, s);
        } else {
            memcpy(s, r, bytes);
        }
        double rho = ddot(n, r, s);                     /* SolveMG.m:38 */
        if (it == 1) memcpy(d, s, bytes);
        else {
            double beta = rho / gamma2;                 /* SolveMG.m:42-43 */
            #pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) d[i] = beta * d[i] + s[i];
        }
        spmv(&c->A[0], d, q);                           /* SolveMG.m:45 */
        double alpha = rho / ddot(n, d, q);             /* SolveMG.m:46 */
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) { x[i] = x[i] + alpha * d[i]; r[i] = r[i] - alpha * q[i]; }
        gamma2 = rho;
        nr = sqrt(ddot(n, r, r)) / nf;                  /* SolveMG.m:50 */
        ++it;
    }
    if (relres) *relres = nr;
    free(r); free(s); free(d); free(q);
    return it;
}

/* ------------------------------------------------------------------------- */
/* CSR export                                                                 */
/* ------------------------------------------------------------------------- */

int64_t or_csr_nnz(const or_ctx *c, int which, int level) {
    return which == 0 ? c->A[level].nnz : c->P[level].nnz;
}

void or_csr_export(const or_ctx *c, int which, int level, int64_t *rp, int32_t *ci, double *v) {
    const csr_t *a = which == 0 ? &c->A[level] : &c->P[level];
    memcpy(rp, a->rp, sizeof(int64_t) * (size_t)(a->n + 1));
    memcpy(ci, a->ci, sizeof(int32_t) * (size_t)a->nnz);
    memcpy(v, a->v, sizeof(double) * (size_t)a->nnz);
}

/* ------------------------------------------------------------------------- */
/* density filter (Implementation/ElemNeighborhood.m:10-98)                    */
/* ------------------------------------------------------------------------- */

int64_t or_filter_build(int nelx, int nely, int nelz, double el, double eh, double ew,
                        double rmin, int opfilter, const uint8_t *fixed,
                        int64_t *rp, int32_t *ci, double *v, double *w) {
    int frz = (int)floor(rmin / ew);
    double rmin2 = rmin * rmin;
    double alpha = 0, beta = 0;
    if (opfilter == 1) {
        alpha = 1.0 / (2.0 * ((rmin / 3.0) * (rmin / 3.0)));
        beta = 1.0 / (2.0 * M_PI * (rmin / 3.0));
    }
    /* offset table per |dz|, |dx| (ElemNeighborhood.m:25-39) */
    int relx[64];
    int rely[64][64];
    for (int k1 = 0; k1 <= frz && k1 < 64; ++k1) {
        double dz2 = (k1 * ew) * (k1 * ew);
        relx[k1] = (int)floor(sqrt(rmin2 - dz2) / el);
        for (int i1 = 0; i1 <= relx[k1] && i1 < 64; ++i1) {
            double dx2 = (i1 * el) * (i1 * el);
            rely[k1][i1] = (int)floor(sqrt(rmin2 - dz2 - dx2) / eh);
        }
    }
    int64_t ne = (int64_t)nelx * nely * nelz;
    int64_t nnz = 0;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1 && !rp) return nnz;
        nnz = 0;
        for (int64_t e = 0; e < ne; ++e) {
            int64_t ez = e / ((int64_t)nelx * nely), ex = (e / nely) % nelx, ey = e % nely;
            if (pass == 1) rp[e] = nnz;
            double rs = 0.0;
            int efix = fixed && fixed[e];
            for (int64_t kz = ez - frz; kz <= ez + frz; ++kz) {
                if (kz < 0 || kz >= nelz) continue;
                int adz = (int)(kz > ez ? kz - ez : ez - kz);
                for (int64_t kx = ex - relx[adz]; kx <= ex + relx[adz]; ++kx) {
                    if (kx < 0 || kx >= nelx) continue;
                    int adx = (int)(kx > ex ? kx - ex : ex - kx);
                    for (int64_t ky = ey - rely[adz][adx]; ky <= ey + rely[adz][adx]; ++ky) {
                        if (ky < 0 || ky >= nely) continue;
                        int64_t e2 = kz * (int64_t)nelx * nely + kx * nely + ky;
                        int ady = (int)(ky > ey ? ky - ey : ey - ky);
                        double wt;
                        if (e2 == e) {
                            wt = efix ? 1.0 : (opfilter == 1 ? beta : 1.0); /* :60-64, :90-96 */
                        } else {
                            if (efix || (fixed && fixed[e2])) continue;     /* :90-92 */
                            double dd = (adz * ew) * (adz * ew) + (adx * el) * (adx * el) + (ady * eh) * (ady * eh);
                            if (opfilter == 1) wt = exp(-dd * alpha) * beta; /* :33-34 */
                            else wt = (rmin - sqrt(dd)) / rmin;             /* :35-36 */
                        }
                        if (pass == 1) { ci[nnz] = (int32_t)e2; v[nnz] = wt; }
                        rs += wt;
                        ++nnz;
                    }
                }
            }
            if (pass == 1 && w) w[e] = rs; /* w = full(sum(W,2)) (:98) */
        }
        if (pass == 1) rp[ne] = nnz;
    }
    return nnz;
}

void or_filter_apply(int64_t n, const int64_t *rp, const int32_t *ci, const double *v,
                     const double *w, const double *x, int trans, double *y) {
    if (!trans) {
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            double s = 0.0;
            for (int64_t k = rp[i]; k < rp[i + 1]; ++k) s += v[k] * x[ci[k]];
            y[i] = s / w[i];
        }
    } else {
        /* W'(x./w); W is symmetric here, but the transpose is taken literally */
        double *t = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        for (int64_t i = 0; i < n; ++i) { t[i] = x[i] / w[i]; y[i] = 0.0; }
        for (int64_t i = 0; i < n; ++i)
            for (int64_t k = rp[i]; k < rp[i + 1]; ++k) y[ci[k]] += v[k] * t[i];
        free(t);
    }
}

/* ------------------------------------------------------------------------- */
/* compliance gradient (Implementation/Grad.m)                                 */
/* ------------------------------------------------------------------------- */

void or_grad0(const or_ctx *c, const double *u, const double *x, const uint8_t *freedens, double *g0) {
    const level_t *lv = &c->lev[0];
    int64_t ne = (int64_t)lv->nelx * lv->nely * lv->nelz;
    #pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; ++e) {
        if (freedens && !freedens[e]) { g0[e] = 0.0; continue; }
        int64_t ez = e / ((int64_t)lv->nelx * lv->nely), ex = (e / lv->nely) % lv->nelx, ey = e % lv->nely;
        double ue[24];
        for (int a = 0; a < 8; ++a) {
            int dy = a & 1, dx = (a >> 1) & 1, dz = (a >> 2) & 1;
            int64_t node = (ez + dz) * lv->nx * lv->ny + (ex + dx) * lv->ny + (ey + dy);
            for (int cc = 0; cc < 3; ++cc) ue[3 * a + cc] = u[3 * node + cc];
        }
        double s = 0.0;
        for (int i = 0; i < 24; ++i) {
            double t = 0.0;
            for (int j = 0; j < 24; ++j) t += c->ke[i * 24 + j] * ue[j];
            s += ue[i] * t;
        }
        g0[e] = -c->p * pow(x[e], c->p - 1.0) * (c->E0 - c->Emin) * s; /* Grad.m:21 */
    }
}

void or_grad0_mr(const or_ctx *c, int n, const double *u, const double *xd, double *g0d) {
    const level_t *lv = &c->lev[0];
    int64_t ne = (int64_t)lv->nelx * lv->nely * lv->nelz;
    int n3 = n * n * n;
    int64_t NXd = (int64_t)lv->nelx * n, NYd = (int64_t)lv->nely * n;
    double *kmr = (double *)malloc(sizeof(double) * 576 * (size_t)n3);
    or_elem_stiff_mr(c->el, c->eh, c->ew, c->nu, n, kmr);
    #pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; ++e) {
        int64_t ez = e / ((int64_t)lv->nelx * lv->nely), ex = (e / lv->nely) % lv->nelx, ey = e % lv->nely;
        double ue[24];
        for (int a = 0; a < 8; ++a) {
            int dy = a & 1, dx = (a >> 1) & 1, dz = (a >> 2) & 1;
            int64_t node = (ez + dz) * lv->nx * lv->ny + (ex + dx) * lv->ny + (ey + dy);
            for (int cc = 0; cc < 3; ++cc) ue[3 * a + cc] = u[3 * node + cc];
        }
        for (int sidx = 0; sidx < n3; ++sidx) {
            int sy = sidx % n, sx = (sidx / n) % n, sz = sidx / (n * n);
            int64_t v = (ez * n + sz) * NXd * NYd + (ex * n + sx) * NYd + (ey * n + sy);
            const double *k = kmr + (size_t)576 * sidx;
            double s = 0.0;
            for (int i = 0; i < 24; ++i) {
                double t = 0.0;
                for (int j = 0; j < 24; ++j) t += k[i * 24 + j] * ue[j];
                s += ue[i] * t;
            }
            g0d[v] = -c->p * pow(xd[v], c->p - 1.0) * (c->E0 - c->Emin) * s; /* Grad.m:46 */
        }
    }
    free(kmr);
}
