/*
 * Oracle: FP64 CPU restatement of one StaRMAP time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is part of the oracle: it is built into
 * oracle/liboracle_stepper.so and loaded only by tests/, by
 * __graft_entry__.smoke() and by the CPU-baseline leg of bench.py, as the
 * checker.  The product path never links or calls it.
 *
 * It restates the body of the reference time loop, starmap_solver.m:307-421,
 * statement by statement:
 *   - derivatives  dU = diff(U(extend{g},:,:),[],d)/h(d)          :313-348, :361-396
 *   - flux         W  = -sumcell([dxU(Ix),dyU(Iy),dzU(Iz)], w)      :349-351, :397-399, :516-519
 *   - half step    U  = U + dt/2*(W+Q - s.*U).*E                     :352-354, :402, :404-406
 *   - filters      U  = U.*F{..,k} (substep)  /  U.*F (full)         :355-358, :407-411, :417-421
 * with the Strang order [1 2 1] of :310 and the filter-slot index k of
 * :309 and :400 (k=1 for substep 1, k=2 for substep 3 and for the second
 * application in substep 2).
 *
 * Arrays use MATLAB's column-major layout with per-component extents
 * (first solver axis fastest), exactly as U{j} in the reference.  It must be
 * compiled with -ffp-contract=off so that no multiply-add is fused; the
 * association of every expression follows MATLAB's left-to-right evaluation.
 *
 * Pinned by tests/test_oracle_stepper.py (Fourier amplification matrix,
 * dense Kronecker operator, mass conservation and exact decay, P_N == SP_N).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int32_t n_sys;
    int32_t n[3];          /* solver-axis cell counts (par.n after lifting)     */
    int32_t bc[3];         /* 0 periodic, 1 extrapolation                       */
    double h[3];           /* grid spacing per solver axis                      */
    const int32_t *gt;     /* [3*n_sys] grid type (1 or 2) per axis             */
    const int32_t *rp[3];  /* CSR row pointers (n_sys+1) of Mx, My, Mz          */
    const int32_t *ci[3];  /* CSR column indices, ascending within a row        */
    const double *cv[3];   /* CSR values                                        */
    const int32_t *even;   /* [c111 c221 c212 c122]                             */
    int32_t n_even;
    const int32_t *odd;    /* [c211 c121 c112 c222]                             */
    int32_t n_odd;
    int32_t f_position;    /* 0 none, 1 'substep', 2 'full'                    */
} oracle_desc;

/* per-component coefficient: a field of the component's extent, or a scalar */
typedef struct {
    const double *p;
    int32_t is_field;
} oracle_coef;

static int64_t ext_of(const oracle_desc *d, int j, int ax)
{
    /* size of grid x{g} along an axis: n for cell centres, n+bc for faces
     * (starmap_solver.m:113-118, n2 = size(X{2,2,2}) at :126) */
    return d->gt[3 * j + ax] == 1 ? d->n[ax] : d->n[ax] + d->bc[ax];
}

static int64_t numel(const oracle_desc *d, int j)
{
    return ext_of(d, j, 0) * ext_of(d, j, 1) * ext_of(d, j, 2);
}

/* extension index vector (0-based), starmap_solver.m:127-132:
 * g=1: [1:bc, 1:n, bc*(n-1)+1]   g=2: [n2*(1:1-bc), 1:n2] */
static int build_extend(int n, int bc, int g, int32_t *out)
{
    int len = 0;
    if (g == 1) {
        if (bc) out[len++] = 0;
        for (int i = 0; i < n; ++i) out[len++] = i;
        out[len++] = bc * (n - 1);
    } else {
        int n2 = n + bc;
        if (!bc) out[len++] = n2 - 1;
        for (int i = 0; i < n2; ++i) out[len++] = i;
    }
    return len;
}

/* D = diff(U(extend),[],ax)/h along one axis */
static void derivative(const oracle_desc *d, int k, int ax, const double *U, double *D)
{
    int64_t e[3] = {ext_of(d, k, 0), ext_of(d, k, 1), ext_of(d, k, 2)};
    int g = d->gt[3 * k + ax];
    int32_t *ext = (int32_t *)malloc(sizeof(int32_t) * (size_t)(d->n[ax] + 3));
    int len = build_extend(d->n[ax], d->bc[ax], g, ext);
    int64_t o[3] = {e[0], e[1], e[2]};
    o[ax] = len - 1;
    double hh = d->h[ax];
    for (int64_t i3 = 0; i3 < o[2]; ++i3)
        for (int64_t i2 = 0; i2 < o[1]; ++i2)
            for (int64_t i1 = 0; i1 < o[0]; ++i1) {
                int64_t a[3] = {i1, i2, i3}, b[3] = {i1, i2, i3};
                int64_t id = a[ax];
                a[ax] = ext[id + 1];
                b[ax] = ext[id];
                double ua = U[a[0] + e[0] * (a[1] + e[1] * a[2])];
                double ub = U[b[0] + e[0] * (b[1] + e[1] * b[2])];
                D[i1 + o[0] * (i2 + o[1] * i3)] = (ua - ub) / hh;
            }
    free(ext);
}

static inline double cval(const oracle_coef *c, int64_t i)
{
    return c->is_field ? c->p[i] : c->p[0];
}

/* One half step of target j with a precomputed W, starmap_solver.m:352-354 */
static void half_step(double *Uj, const double *W, const double *Q, const oracle_coef *s,
                      const oracle_coef *E, double dt, int64_t m)
{
    double c = dt / 2;
    for (int64_t i = 0; i < m; ++i) {
        double a = Q ? W[i] + Q[i] : W[i] + 0.0;
        double b = cval(s, i) * Uj[i];
        double t = c * (a - b);
        Uj[i] = Uj[i] + t * cval(E, i);
    }
}

static void apply_filter(double *Uj, const oracle_coef *F, int64_t m)
{
    for (int64_t i = 0; i < m; ++i) Uj[i] = Uj[i] * cval(F, i);
}

/* W = -sumcell(...) for target j; dU[ax][k] holds the derivative fields */
static void flux(const oracle_desc *d, int j, double **dU[3], double *W, int64_t m)
{
    int first = 1;
    for (int ax = 0; ax < 3; ++ax) {
        for (int32_t p = d->rp[ax][j]; p < d->rp[ax][j + 1]; ++p) {
            int k = d->ci[ax][p];
            double w = d->cv[ax][p];
            const double *Dk = dU[ax][k];
            if (first) {
                for (int64_t i = 0; i < m; ++i) W[i] = Dk[i] * w;
                first = 0;
            } else {
                for (int64_t i = 0; i < m; ++i) W[i] = W[i] + Dk[i] * w;
            }
        }
    }
    if (first) memset(W, 0, sizeof(double) * (size_t)m);
    for (int64_t i = 0; i < m; ++i) W[i] = -W[i];
}

/* does any target row of direction ax reference column k? */
static int used_as_source(const oracle_desc *d, int ax, int k, const int32_t *tg, int nt)
{
    for (int t = 0; t < nt; ++t) {
        int j = tg[t];
        for (int32_t p = d->rp[ax][j]; p < d->rp[ax][j + 1]; ++p)
            if (d->ci[ax][p] == k) return 1;
    }
    return 0;
}

static int half_substep(const oracle_desc *d, double dt, double **U, double **Q,
                        const oracle_coef *S, const oracle_coef *E,
                        const oracle_coef *F, int two_apps, const oracle_coef *F2,
                        const int32_t *src, int ns, const int32_t *tg, int nt)
{
    int n_sys = d->n_sys;
    double **dU[3];
    for (int ax = 0; ax < 3; ++ax) {
        dU[ax] = (double **)calloc((size_t)n_sys, sizeof(double *));
        if (!dU[ax]) return -1;
    }
    int rc = 0;
    /* derivatives of every source component that some target couples to */
    for (int ax = 0; ax < 3 && !rc; ++ax) {
        #pragma omp parallel for schedule(dynamic)
        for (int s = 0; s < ns; ++s) {
            int k = src[s];
            if (!used_as_source(d, ax, k, tg, nt)) continue;
            int64_t o[3] = {ext_of(d, k, 0), ext_of(d, k, 1), ext_of(d, k, 2)};
            int g = d->gt[3 * k + ax];
            o[ax] = (g == 1) ? d->n[ax] + d->bc[ax] : d->n[ax];
            double *D = (double *)malloc(sizeof(double) * (size_t)(o[0] * o[1] * o[2]));
            if (!D) { rc = -1; continue; }
            derivative(d, k, ax, U[k], D);
            dU[ax][k] = D;
        }
    }
    if (!rc) {
        #pragma omp parallel for schedule(dynamic)
        for (int t = 0; t < nt; ++t) {
            int j = tg[t];
            int64_t m = numel(d, j);
            double *W = (double *)malloc(sizeof(double) * (size_t)m);
            if (!W) { rc = -1; continue; }
            flux(d, j, dU, W, m);
            if (!two_apps) {
                half_step(U[j], W, Q[j], &S[j], &E[j], dt, m);
                if (d->f_position == 1) apply_filter(U[j], &F[j], m);
            } else {
                for (int kk = 1; kk <= 2; ++kk) {
                    half_step(U[j], W, Q[j], &S[j], &E[j], dt, m);
                    if (j != 0 && d->f_position == 1)
                        apply_filter(U[j], kk == 1 ? &F[j] : &F2[j], m);
                }
            }
            free(W);
        }
    }
    for (int ax = 0; ax < 3; ++ax) {
        for (int k = 0; k < n_sys; ++k) free(dU[ax][k]);
        free(dU[ax]);
    }
    return rc;
}

/*
 * One full time step (starmap_solver.m:308-421).
 *   U[j]  state of component j (modified in place)
 *   Q[j]  source field of component j, or NULL where Q{j} is zero
 *   S[j]  decay rate: sA1 for j==0, sT1{gt(j),mom_order(j)} otherwise
 *   E[j]  expm1div(-S*dt/2), same shapes
 *   F1[j], F2[j]  substep filter slots k=1,2 ('substep'); F1 is the full
 *                 filter term when f_position is 'full'
 * Returns 0 on success, -1 on allocation failure.
 */
int oracle_step(const oracle_desc *d, double dt, double **U, double **Q,
                const oracle_coef *S, const oracle_coef *E,
                const oracle_coef *F1, const oracle_coef *F2)
{
    int rc;
    /* substep 1: sources even, targets odd, filter slot k=1 */
    rc = half_substep(d, dt, U, Q, S, E, F1, 0, F2, d->even, d->n_even, d->odd, d->n_odd);
    if (rc) return rc;
    /* substep 2: sources odd, targets even, two applications with k=1,2 */
    rc = half_substep(d, dt, U, Q, S, E, F1, 1, F2, d->odd, d->n_odd, d->even, d->n_even);
    if (rc) return rc;
    /* substep 3: as substep 1 but with filter slot k=2 (k left at 2 by :400) */
    rc = half_substep(d, dt, U, Q, S, E, F2, 0, F2, d->even, d->n_even, d->odd, d->n_odd);
    if (rc) return rc;
    if (d->f_position == 2) {
        #pragma omp parallel for schedule(dynamic)
        for (int j = 1; j < d->n_sys; ++j) apply_filter(U[j], &F1[j], numel(d, j));
    }
    return 0;
}

/* (e^x - 1)/x with the reference's series branch, starmap_solver.m:531-535 */
void oracle_expm1div(const double *x, double *y, int64_t m);
#include <math.h>
void oracle_expm1div(const double *x, double *y, int64_t m)
{
    for (int64_t i = 0; i < m; ++i) {
        double v = x[i];
        if (fabs(v) > 2e-4)
            y[i] = (exp(v) - 1) / v;
        else
            y[i] = 1 + v * .5 + (v * v) / 6;
    }
}
