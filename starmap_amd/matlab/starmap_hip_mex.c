/*
 * starmap_hip_mex.c -- MATLAB gateway of the MI355X StaRMAP time-stepping
 * core (include/starmap_hip.h).  starmap_solver_hip.m calls it; it only
 * marshals: MATLAB arrays in and out, sparse closure matrices to CSR, status
 * codes to MATLAB errors with identifiers starmap:hip:<code>.
 *
 * Build where MATLAB is installed (this repository's image has no MATLAB, so
 * the file is not compiled here):
 *     mex -R2018a -I<repo>/include starmap_hip_mex.c \
 *         -L<repo>/starmap_amd -lstarmap_hip
 *
 * Calls (component indices j are 1-based here, as in MATLAB; the C ABI is
 * 0-based):
 *   h = starmap_hip_mex('create', Mx, My, Mz, n, bc, h, mom_order, f_order,
 *                       f_position, precision, device)
 *   starmap_hip_mex('destroy', h)
 *   gt = starmap_hip_mex('grid_types', h)           n_sys x 3 (1 or 2)
 *   sz = starmap_hip_mex('shape', h, j)             size of U{j}, 1 x 3
 *   starmap_hip_mex('set_state', h, j, U)
 *   U = starmap_hip_mex('get_state', h, j)
 *   starmap_hip_mex('set_coeff', h, kind, gt, order, slot, value)
 *        kind 'sigma_a1' | 'sigma_t' | 'source' | 'filter'; value a scalar
 *        or an array of the grid's size; for 'source' order is j
 *   starmap_hip_mex('step', h, dt, n_steps)
 *   bad = starmap_hip_mex('check_finite', h)
 *   starmap_hip_mex('snapshot', h, comps)
 *   Us = starmap_hip_mex('interp', h, lambda, comps)   cell array, 1 per comp
 *   [sums, sq, nonfinite] = starmap_hip_mex('observables', h)
 *   [sumsq, diffsq] = starmap_hip_mex('moment_norms', h)
 *   v = starmap_hip_mex('probe', h, lo, point)         n_sys x 1
 *   t = starmap_hip_mex('timing', h)    [total_ms substep_ms launches bytes]
 *   starmap_hip_mex('set_schedule', h, tile, blocks_per_cu, cell_waves)
 *   starmap_hip_mex('adjoint_reset', h)
 *   starmap_hip_mex('adjoint_seed', h, j, dJdU)
 *   starmap_hip_mex('adjoint_memory', h, bytes)      device-memory budget (0: none)
 *   starmap_hip_mex('adjoint', h, dt, n_steps, checkpoint_every)
 *   s = starmap_hip_mex('adjoint_stats', h)          1 x 8: segments, steps per segment,
 *                                                    recorded in the forward pass, checkpoint
 *                                                    slots, recomputed, further advances,
 *                                                    buffers and bytes at the peak
 *   G = starmap_hip_mex('adjoint_get', h, what, j)   what 0 lambda, 1 dJ/dsigma_a,
 *                                                    2 dJ/dsigma_s0
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"
#include "starmap_hip.h"

static void check(int rc) {
    static const char *names[] = {"ok", "invalid", "hip", "oom", "state", "unstable"};
    char id[64];
    if (rc == SMP_OK) return;
    snprintf(id, sizeof id, "starmap:hip:%s", (rc > 0 && rc < 6) ? names[rc] : "unknown");
    mexErrMsgIdAndTxt(id, "%s", smp_last_error());
}

static void usage(const char *msg) { mexErrMsgIdAndTxt("starmap:hip:usage", "%s", msg); }

static double scalar_arg(const mxArray *a, const char *what) {
    if (!mxIsNumeric(a) || mxGetNumberOfElements(a) != 1) usage(what);
    return mxGetScalar(a);
}

/* a real double array's data (copied to a new buffer for other classes) */
static double *doubles_of(const mxArray *a, int64_t *count, int *owned) {
    const size_t n = mxGetNumberOfElements(a);
    *count = (int64_t)n;
    if (mxIsDouble(a) && !mxIsComplex(a) && !mxIsSparse(a)) {
        *owned = 0;
        return mxGetDoubles(a);
    }
    if (!mxIsNumeric(a) || mxIsComplex(a) || mxIsSparse(a)) usage("expected a real numeric array");
    double *out = (double *)mxMalloc((n ? n : 1) * sizeof(double));
    {
        /* other numeric classes: converted by MATLAB's double() */
        mxArray *d = NULL;
        mxArray *in[1];
        in[0] = (mxArray *)a;
        mexCallMATLAB(1, &d, 1, in, "double");
        memcpy(out, mxGetDoubles(d), n * sizeof(double));
        mxDestroyArray(d);
    }
    *owned = 1;
    return out;
}

static void int3_arg(const mxArray *a, int32_t out[3], const char *what) {
    int64_t n;
    int owned;
    double *v = doubles_of(a, &n, &owned);
    if (n != 3) usage(what);
    for (int k = 0; k < 3; ++k) out[k] = (int32_t)v[k];
    if (owned) mxFree(v);
}

/* Rows of a closure matrix (n_sys x n_sys, sparse or full) in CSR with
 * ascending columns.  MATLAB stores sparse matrices by columns (CSC), so
 * the rows are gathered by counting entries per row first. */
static void csr_of(const mxArray *M, int32_t n_sys, int32_t **rp, int32_t **ci, double **cv) {
    if (mxGetM(M) != (size_t)n_sys || mxGetN(M) != (size_t)n_sys)
        usage("coupling matrices must be n_sys x n_sys");
    *rp = (int32_t *)mxCalloc((size_t)n_sys + 1, sizeof(int32_t));
    if (mxIsSparse(M)) {
        const mwIndex *jc = mxGetJc(M), *ir = mxGetIr(M);
        const double *pr = mxGetDoubles(M);
        const mwIndex nnz = jc[n_sys];
        int32_t *fill = (int32_t *)mxCalloc((size_t)n_sys, sizeof(int32_t));
        for (mwIndex k = 0; k < nnz; ++k)
            if (pr[k] != 0) (*rp)[ir[k] + 1]++;
        for (int32_t r = 0; r < n_sys; ++r) (*rp)[r + 1] += (*rp)[r];
        *ci = (int32_t *)mxMalloc(((size_t)(*rp)[n_sys] + 1) * sizeof(int32_t));
        *cv = (double *)mxMalloc(((size_t)(*rp)[n_sys] + 1) * sizeof(double));
        for (int32_t c = 0; c < n_sys; ++c)          /* columns in order: rows */
            for (mwIndex k = jc[c]; k < jc[c + 1]; ++k) { /* get ascending columns */
                if (pr[k] == 0) continue;
                const int32_t r = (int32_t)ir[k];
                const int32_t at = (*rp)[r] + fill[r]++;
                (*ci)[at] = c;
                (*cv)[at] = pr[k];
            }
        mxFree(fill);
    } else {
        int64_t cnt;
        int owned;
        const double *a = doubles_of(M, &cnt, &owned);
        for (int32_t r = 0; r < n_sys; ++r)
            for (int32_t c = 0; c < n_sys; ++c)
                if (a[r + (size_t)c * n_sys] != 0) (*rp)[r + 1]++;
        for (int32_t r = 0; r < n_sys; ++r) (*rp)[r + 1] += (*rp)[r];
        *ci = (int32_t *)mxMalloc(((size_t)(*rp)[n_sys] + 1) * sizeof(int32_t));
        *cv = (double *)mxMalloc(((size_t)(*rp)[n_sys] + 1) * sizeof(double));
        for (int32_t r = 0; r < n_sys; ++r) {
            int32_t at = (*rp)[r];
            for (int32_t c = 0; c < n_sys; ++c) {
                const double v = a[r + (size_t)c * n_sys];
                if (v != 0) {
                    (*ci)[at] = c;
                    (*cv)[at++] = v;
                }
            }
        }
        if (owned) mxFree((void *)a);
    }
}

static smp_problem *handle_arg(const mxArray *a) {
    if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1) usage("expected a handle from 'create'");
    return (smp_problem *)(uintptr_t)(*(uint64_t *)mxGetData(a));
}

static int32_t comp_arg(const mxArray *a) { return (int32_t)scalar_arg(a, "component index") - 1; }

/* a new double array of component j's size (2D problems: [ny nz] as the
 * reference reshapes U{j} for its output, :489-494) */
static mxArray *field_array(smp_problem *h, int32_t j, int dim2) {
    int64_t s[3];
    check(smp_component_shape(h, j, s));
    if (dim2) {
        mwSize d[2] = {(mwSize)s[1], (mwSize)s[2]};
        return mxCreateNumericArray(2, d, mxDOUBLE_CLASS, mxREAL);
    }
    mwSize d[3] = {(mwSize)s[0], (mwSize)s[1], (mwSize)s[2]};
    return mxCreateNumericArray(3, d, mxDOUBLE_CLASS, mxREAL);
}

static void cmd_create(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    char fpos[16], prec[8];
    smp_desc d;
    int32_t *rp[3], *ci[3];
    double *cv[3];
    int64_t cnt;
    int owned;
    (void)nlhs;
    if (nrhs != 12) usage("create: Mx, My, Mz, n, bc, h, mom_order, f_order, f_position, precision, device");
    memset(&d, 0, sizeof d);
    d.n_sys = (int32_t)mxGetM(prhs[3]);
    for (int a = 0; a < 3; ++a) {
        csr_of(prhs[1 + a], d.n_sys, &rp[a], &ci[a], &cv[a]);
        d.row_ptr[a] = rp[a];
        d.col[a] = ci[a];
        d.val[a] = cv[a];
    }
    int3_arg(prhs[4], d.n, "n must have 3 entries (the lifted par.n)");
    int3_arg(prhs[5], d.bc, "bc must have 3 entries (the lifted par.bc)");
    {
        double *hv = doubles_of(prhs[6], &cnt, &owned);
        if (cnt != 3) usage("h must have 3 entries");
        for (int a = 0; a < 3; ++a) d.h[a] = hv[a];
        if (owned) mxFree(hv);
    }
    {
        double *mo = doubles_of(prhs[7], &cnt, &owned);
        if (cnt != d.n_sys) usage("mom_order must have n_sys entries");
        int32_t *m = (int32_t *)mxMalloc((size_t)d.n_sys * sizeof(int32_t));
        for (int32_t j = 0; j < d.n_sys; ++j) m[j] = (int32_t)mo[j];
        d.mom_order = m;
        if (owned) mxFree(mo);
    }
    if (!mxIsEmpty(prhs[8])) {
        double *fo = doubles_of(prhs[8], &cnt, &owned);
        if (cnt != d.n_sys) usage("f_order must have n_sys entries or be empty");
        int32_t *f = (int32_t *)mxMalloc((size_t)d.n_sys * sizeof(int32_t));
        for (int32_t j = 0; j < d.n_sys; ++j) f[j] = (int32_t)fo[j];
        d.f_order = f;
        if (owned) mxFree(fo);
    }
    if (mxGetString(prhs[9], fpos, sizeof fpos)) fpos[0] = 0;
    d.f_position = !strcmp(fpos, "substep") ? SMP_FILTER_SUBSTEP
                   : !strcmp(fpos, "full")  ? SMP_FILTER_FULL
                                            : SMP_FILTER_NONE;
    if (mxGetString(prhs[10], prec, sizeof prec)) usage("precision must be 'fp64' or 'fp32'");
    d.precision = !strcmp(prec, "fp32") ? SMP_FP32 : SMP_FP64;
    d.device = (int32_t)scalar_arg(prhs[11], "device");
    smp_problem *h = NULL;
    const int rc = smp_create(&d, &h);
    for (int a = 0; a < 3; ++a) {
        mxFree(rp[a]);
        mxFree(ci[a]);
        mxFree(cv[a]);
    }
    mxFree((void *)d.mom_order);
    if (d.f_order) mxFree((void *)d.f_order);
    check(rc);
    plhs[0] = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
    *(uint64_t *)mxGetData(plhs[0]) = (uint64_t)(uintptr_t)h;
}

static int coef_kind(const mxArray *a) {
    char k[16];
    if (mxGetString(a, k, sizeof k)) usage("coefficient kind must be a string");
    if (!strcmp(k, "sigma_a1")) return SMP_COEF_SIGMA_A1;
    if (!strcmp(k, "sigma_t")) return SMP_COEF_SIGMA_T;
    if (!strcmp(k, "source")) return SMP_COEF_SOURCE;
    if (!strcmp(k, "filter")) return SMP_COEF_FILTER;
    usage("coefficient kind: 'sigma_a1', 'sigma_t', 'source' or 'filter'");
    return -1;
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    char cmd[32];
    if (nrhs < 1 || mxGetString(prhs[0], cmd, sizeof cmd)) usage("first argument: a command string");
    if (!strcmp(cmd, "create")) {
        cmd_create(nlhs, plhs, nrhs, prhs);
        return;
    }
    if (nrhs < 2) usage("missing handle");
    smp_problem *h = handle_arg(prhs[1]);
    int64_t cnt;
    int owned;

    if (!strcmp(cmd, "destroy")) {
        check(smp_destroy(h));
    } else if (!strcmp(cmd, "grid_types")) {
        int32_t n_sys = 0;
        check(smp_num_components(h, &n_sys));
        int32_t *gt = (int32_t *)mxMalloc((size_t)3 * (n_sys ? n_sys : 1) * sizeof(int32_t));
        check(smp_grid_types(h, gt));
        plhs[0] = mxCreateDoubleMatrix((mwSize)n_sys, 3, mxREAL);
        double *o = mxGetDoubles(plhs[0]);
        for (int32_t j = 0; j < n_sys; ++j)
            for (int a = 0; a < 3; ++a) o[j + (size_t)a * n_sys] = gt[3 * j + a];
        mxFree(gt);
    } else if (!strcmp(cmd, "shape")) {
        int64_t s[3];
        check(smp_component_shape(h, comp_arg(prhs[2]), s));
        plhs[0] = mxCreateDoubleMatrix(1, 3, mxREAL);
        for (int a = 0; a < 3; ++a) mxGetDoubles(plhs[0])[a] = (double)s[a];
    } else if (!strcmp(cmd, "set_state")) {
        double *v = doubles_of(prhs[3], &cnt, &owned);
        const int rc = smp_set_state(h, comp_arg(prhs[2]), v, cnt);
        if (owned) mxFree(v);
        check(rc);
    } else if (!strcmp(cmd, "get_state")) {
        const int32_t j = comp_arg(prhs[2]);
        plhs[0] = field_array(h, j, 0);
        check(smp_get_state(h, j, mxGetDoubles(plhs[0]), (int64_t)mxGetNumberOfElements(plhs[0])));
    } else if (!strcmp(cmd, "set_coeff")) {
        if (nrhs != 7) usage("set_coeff: h, kind, gt, order, slot, value");
        const int kind = coef_kind(prhs[2]);
        int32_t gt[3];
        int3_arg(prhs[3], gt, "gt must have 3 entries");
        int32_t order = (int32_t)scalar_arg(prhs[4], "order");
        if (kind == SMP_COEF_SOURCE) order -= 1; /* component j, 1-based in MATLAB */
        const int32_t slot = (int32_t)scalar_arg(prhs[5], "slot");
        double *v = doubles_of(prhs[6], &cnt, &owned);
        const int rc = cnt == 1 ? smp_set_coeff(h, kind, gt, order, slot, NULL, v[0], 0)
                                : smp_set_coeff(h, kind, gt, order, slot, v, 0.0, cnt);
        if (owned) mxFree(v);
        check(rc);
    } else if (!strcmp(cmd, "step")) {
        check(smp_step(h, scalar_arg(prhs[2], "dt"), (int64_t)scalar_arg(prhs[3], "n_steps")));
    } else if (!strcmp(cmd, "check_finite")) {
        int32_t bad = 0;
        const int rc = smp_check_finite(h, &bad);
        if (rc != SMP_OK && rc != SMP_E_UNSTABLE) check(rc);
        plhs[0] = mxCreateLogicalScalar(bad != 0);
    } else if (!strcmp(cmd, "snapshot") || !strcmp(cmd, "interp")) {
        const int snap = !strcmp(cmd, "snapshot");
        const mxArray *ca = prhs[snap ? 2 : 3];
        double *c = doubles_of(ca, &cnt, &owned);
        int32_t *comps = (int32_t *)mxMalloc((size_t)(cnt ? cnt : 1) * sizeof(int32_t));
        for (int64_t i = 0; i < cnt; ++i) comps[i] = (int32_t)c[i] - 1;
        if (owned) mxFree(c);
        if (snap) {
            const int rc = smp_snapshot(h, comps, (int32_t)cnt);
            mxFree(comps);
            check(rc);
            return;
        }
        int64_t total = 0, s[3];
        for (int64_t i = 0; i < cnt; ++i) {
            check(smp_component_shape(h, comps[i], s));
            total += s[0] * s[1] * s[2];
        }
        double *buf = (double *)mxMalloc((size_t)(total ? total : 1) * sizeof(double));
        const int rc = smp_interp(h, scalar_arg(prhs[2], "lambda"), buf, total);
        if (rc != SMP_OK) {
            mxFree(buf);
            mxFree(comps);
            check(rc);
        }
        plhs[0] = mxCreateCellMatrix(1, (mwSize)cnt);
        int64_t pos = 0;
        for (int64_t i = 0; i < cnt; ++i) {
            mxArray *u = field_array(h, comps[i], 0);
            const size_t m = mxGetNumberOfElements(u);
            memcpy(mxGetDoubles(u), buf + pos, m * sizeof(double));
            pos += (int64_t)m;
            mxSetCell(plhs[0], (mwIndex)i, u);
        }
        mxFree(buf);
        mxFree(comps);
    } else if (!strcmp(cmd, "observables")) {
        int32_t n_sys = 0, bad = 0;
        check(smp_num_components(h, &n_sys));
        plhs[0] = mxCreateDoubleMatrix((mwSize)n_sys, 1, mxREAL);
        mxArray *sq = mxCreateDoubleMatrix((mwSize)n_sys, 1, mxREAL);
        check(smp_observables(h, mxGetDoubles(plhs[0]), mxGetDoubles(sq), &bad));
        if (nlhs > 1) plhs[1] = sq; else mxDestroyArray(sq);
        if (nlhs > 2) plhs[2] = mxCreateLogicalScalar(bad != 0);
    } else if (!strcmp(cmd, "moment_norms")) {
        int32_t n_sys = 0;
        check(smp_num_components(h, &n_sys));
        plhs[0] = mxCreateDoubleMatrix((mwSize)n_sys, 1, mxREAL);
        mxArray *ds = mxCreateDoubleMatrix(3, (mwSize)n_sys, mxREAL);
        check(smp_moment_norms(h, mxGetDoubles(plhs[0]), mxGetDoubles(ds)));
        if (nlhs > 1) plhs[1] = ds; else mxDestroyArray(ds);
    } else if (!strcmp(cmd, "probe")) {
        int32_t n_sys = 0;
        check(smp_num_components(h, &n_sys));
        double lo[3], pt[3];
        double *a = doubles_of(prhs[2], &cnt, &owned);
        if (cnt != 3) usage("lo must have 3 entries");
        memcpy(lo, a, sizeof lo);
        if (owned) mxFree(a);
        a = doubles_of(prhs[3], &cnt, &owned);
        if (cnt != 3) usage("point must have 3 entries (solver axes)");
        memcpy(pt, a, sizeof pt);
        if (owned) mxFree(a);
        plhs[0] = mxCreateDoubleMatrix((mwSize)n_sys, 1, mxREAL);
        check(smp_probe(h, lo, pt, mxGetDoubles(plhs[0])));
    } else if (!strcmp(cmd, "timing")) {
        double tot, sub, by;
        int64_t nl;
        check(smp_last_step_timing(h, &tot, &sub, &nl, &by));
        plhs[0] = mxCreateDoubleMatrix(1, 4, mxREAL);
        double *o = mxGetDoubles(plhs[0]);
        o[0] = tot;
        o[1] = sub;
        o[2] = (double)nl;
        o[3] = by;
    } else if (!strcmp(cmd, "set_schedule")) {
        int32_t tile[3];
        int3_arg(prhs[2], tile, "tile must have 3 entries");
        check(smp_set_schedule(h, tile, (int32_t)scalar_arg(prhs[3], "blocks_per_cu"),
                               (int32_t)scalar_arg(prhs[4], "cell_waves")));
    } else if (!strcmp(cmd, "adjoint_reset")) {
        check(smp_adjoint_reset(h));
    } else if (!strcmp(cmd, "adjoint_seed")) {
        double *v = doubles_of(prhs[3], &cnt, &owned);
        const int rc = smp_adjoint_seed(h, comp_arg(prhs[2]), v, cnt);
        if (owned) mxFree(v);
        check(rc);
    } else if (!strcmp(cmd, "adjoint")) {
        check(smp_adjoint(h, scalar_arg(prhs[2], "dt"), (int64_t)scalar_arg(prhs[3], "n_steps"),
                          (int32_t)scalar_arg(prhs[4], "checkpoint_every")));
    } else if (!strcmp(cmd, "adjoint_memory")) {
        if (nrhs != 3) usage("adjoint_memory: h, bytes");
        check(smp_adjoint_set_memory(h, (int64_t)scalar_arg(prhs[2], "bytes")));
    } else if (!strcmp(cmd, "adjoint_stats")) {
        int64_t st[8];
        check(smp_adjoint_stats(h, st));
        plhs[0] = mxCreateDoubleMatrix(1, 8, mxREAL);
        for (int k = 0; k < 8; ++k) mxGetDoubles(plhs[0])[k] = (double)st[k];
    } else if (!strcmp(cmd, "adjoint_get")) {
        const int32_t what = (int32_t)scalar_arg(prhs[2], "what");
        const int32_t j = comp_arg(prhs[3]);
        plhs[0] = field_array(h, j, 0);
        check(smp_adjoint_get(h, what, j, mxGetDoubles(plhs[0]),
                              (int64_t)mxGetNumberOfElements(plhs[0])));
    } else {
        usage("unknown command");
    }
}
