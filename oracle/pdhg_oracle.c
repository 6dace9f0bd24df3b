/*
 * CPU oracle for the PDHG LP path — TEST INFRASTRUCTURE ONLY (see
 * pdhg_oracle.h).  Plain C99, built with -ffp-contract=off so every
 * floating-point expression rounds exactly as the reference's g++ build on
 * x86-64 does: the op order of each function follows the cited reference
 * lines one for one, which is what lets tests compare trajectories bitwise
 * against the compiled reference.
 *
 * Pinning (what checks this file):
 *   - ops: the reference's hand-computed values (tests/test_backend.cpp) and
 *     dense brute force on small random CSR;
 *   - scaling: the literal O(n*nnz) loop vs the O(nnz) pass, and a dense
 *     restatement;
 *   - the solve: bitwise-equal iterates and status against the reference's
 *     own sources compiled under oracle/_ref, analytic optima, and the scipy
 *     HiGHS optimum of the 1000x3000 LP committed under tests/golden.
 */
#include "pdhg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* std::max / std::min as the reference calls them: max(a,b) = (a<b)?b:a. */
static double smax(double a, double b) { return (a < b) ? b : a; }
static double smin(double a, double b) { return (b < a) ? b : a; }

/* ------------------------------------------------------------------ ops -- */

/* cpu_ops.cpp:331-343: y = alpha*x + beta*y, literally (no beta==0 case). */
void orc_axpby(int64_t n, double alpha, const double *x, double beta, double *y) {
    for (int64_t i = 0; i < n; ++i) y[i] = alpha * x[i] + beta * y[i];
}

/* cpu_ops.cpp:346-355: std::inner_product with init 0.0. */
double orc_dot(int64_t n, const double *x, const double *y) {
    double acc = 0.0;
    for (int64_t i = 0; i < n; ++i) acc = acc + x[i] * y[i];
    return acc;
}

/* cpu_ops.cpp:358-364. */
double orc_norm_inf(int64_t n, const double *x) {
    double r = 0.0;
    for (int64_t i = 0; i < n; ++i) r = smax(r, fabs(x[i]));
    return r;
}

/* cpu_ops.cpp:367-383: x = min(max(x, lo), hi). */
void orc_project_box(int64_t n, double *x, const double *lo, const double *hi) {
    for (int64_t i = 0; i < n; ++i) x[i] = smin(smax(x[i], lo[i]), hi[i]);
}

/* cpu_ops.cpp:393-460.  N: sequential row sum then alpha*acc + beta*y.
 * T: y *= beta first, then scatter y[col] += v * (alpha*x[row]) in row order. */
void orc_spmv(int64_t rows, int64_t cols, const int64_t *rp, const int64_t *ci,
              const double *v, double alpha, const double *x, double beta,
              double *y, int transpose) {
    if (!transpose) {
        for (int64_t r = 0; r < rows; ++r) {
            double acc = 0.0;
            for (int64_t k = rp[r]; k < rp[r + 1]; ++k) acc += v[k] * x[ci[k]];
            y[r] = alpha * acc + beta * y[r];
        }
    } else {
        for (int64_t j = 0; j < cols; ++j) y[j] *= beta;
        for (int64_t r = 0; r < rows; ++r) {
            const double sx = alpha * x[r];
            for (int64_t k = rp[r]; k < rp[r + 1]; ++k) y[ci[k]] += v[k] * sx;
        }
    }
}

/* cpu_ops.cpp:38-63 (error codes instead of exceptions). */
int orc_validate_pattern(int64_t rows, int64_t cols, int64_t offsets_size,
                         const int64_t *rp, int64_t indices_size, const int64_t *ci) {
    if (rows < 0 || cols < 0) return -1;
    if (rows == INT64_MAX || offsets_size != rows + 1) return -2;
    if (rp[0] != 0 || rp[rows] != indices_size) return -3;
    for (int64_t r = 0; r < rows; ++r) {
        const int64_t b = rp[r], e = rp[r + 1];
        if (b < 0 || b > e || e > indices_size) return -4;
        int64_t prev = -1;
        for (int64_t k = b; k < e; ++k) {
            if (ci[k] < 0 || ci[k] >= cols || ci[k] <= prev) return -5;
            prev = ci[k];
        }
    }
    return 0;
}

/* cpu_ops.cpp:142-158 */
static void csr_matvec(int64_t n, const int64_t *rp, const int64_t *ci,
                       const double *v, const double *x, double *y) {
    for (int64_t r = 0; r < n; ++r) {
        double acc = 0.0;
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) acc += v[k] * x[ci[k]];
        y[r] = acc;
    }
}

/* cpu_ops.cpp:160-267: BiCGSTAB from x = 0 with fixed shadow residual. */
int orc_bicgstab(int64_t n, const int64_t *rp, const int64_t *ci, const double *v,
                 const double *b, double *x, double tol, int64_t max_iter,
                 int64_t *iterations, double *residual) {
    double *r = calloc((size_t)n + 1, sizeof(double));
    double *rh = calloc((size_t)n + 1, sizeof(double));
    double *p = calloc((size_t)n + 1, sizeof(double));
    double *vv = calloc((size_t)n + 1, sizeof(double));
    double *s = calloc((size_t)n + 1, sizeof(double));
    double *t = calloc((size_t)n + 1, sizeof(double));
    double *tmp = calloc((size_t)n + 1, sizeof(double));
    int converged = 0;
    int64_t its = 0;
    double res_norm = INFINITY;
    for (int64_t i = 0; i < n; ++i) { x[i] = 0.0; r[i] = b[i]; rh[i] = b[i]; }
    double rho_prev = 1.0, alpha = 1.0, omega = 1.0;
    double rhr = orc_dot(n, rh, r);
    double init = 0.0;
    for (int64_t i = 0; i < n; ++i) init = smax(init, fabs(r[i]));
    if (init <= tol) {
        converged = 1; its = 0; res_norm = init;
        goto done;
    }
    for (int64_t iter = 0; iter < max_iter; ++iter) {
        its = iter + 1;
        const double rho = rhr;
        if (rho == 0.0) break;
        if (iter == 0) {
            memcpy(p, r, (size_t)n * sizeof(double));
        } else {
            const double beta = (rho / rho_prev) * (alpha / omega);
            for (int64_t i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * vv[i]);
        }
        csr_matvec(n, rp, ci, v, p, vv);
        const double da = orc_dot(n, rh, vv);
        if (da == 0.0) break;
        alpha = rho / da;
        for (int64_t i = 0; i < n; ++i) s[i] = r[i] - alpha * vv[i];
        double ns = 0.0;
        for (int64_t i = 0; i < n; ++i) ns = smax(ns, fabs(s[i]));
        if (ns <= tol) {
            for (int64_t i = 0; i < n; ++i) x[i] += alpha * p[i];
            converged = 1; its = iter + 1; res_norm = ns;
            goto done;
        }
        csr_matvec(n, rp, ci, v, s, t);
        const double tt = orc_dot(n, t, t);
        if (tt == 0.0) break;
        omega = orc_dot(n, t, s) / tt;
        if (omega == 0.0) {
            for (int64_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] = s[i]; }
            break;
        }
        for (int64_t i = 0; i < n; ++i) x[i] += alpha * p[i] + omega * s[i];
        for (int64_t i = 0; i < n; ++i) r[i] = s[i] - omega * t[i];
        double nr = 0.0;
        for (int64_t i = 0; i < n; ++i) nr = smax(nr, fabs(r[i]));
        if (nr <= tol) {
            converged = 1; its = iter + 1; res_norm = nr;
            goto done;
        }
        rhr = orc_dot(n, rh, r);
        rho_prev = rho;
    }
    csr_matvec(n, rp, ci, v, x, tmp);
    res_norm = 0.0;
    for (int64_t i = 0; i < n; ++i) res_norm = smax(res_norm, fabs(b[i] - tmp[i]));
    converged = 0;
done:
    if (iterations) *iterations = its;
    if (residual) *residual = res_norm;
    free(r); free(rh); free(p); free(vv); free(s); free(t); free(tmp);
    return converged;
}

/* -------------------------------------------------------------- scaling -- */

/* continuous_common.hpp:128-131 */
static double scale_bound(double value, double scale, int multiply) {
    if (!isfinite(value)) return value;
    return multiply ? value * scale : value / scale;
}

/* continuous_common.hpp:133-205.  Row factor r = 1/max|a_i.| when that max
 * exceeds 1; column factor s = 1/max|a_.j| over the ORIGINAL values when
 * that max exceeds 1; values a*(r*s); c*s; l/s, u/s; rl*r, ru*r. */
void orc_scale_lp(const orc_lp *lp, const double *lower, const double *upper,
                  int literal, double *values_out, double *objective_out,
                  double *row_lower_out, double *row_upper_out,
                  double *var_lower_out, double *var_upper_out,
                  double *row_scale, double *col_scale) {
    const int64_t m = lp->m, n = lp->n;
    const int64_t *rp = lp->row_offsets, *ci = lp->column_indices;
    const double *a = lp->values;
    for (int64_t r = 0; r < m; ++r) row_scale[r] = 1.0;
    for (int64_t j = 0; j < n; ++j) col_scale[j] = 1.0;
    for (int64_t r = 0; r < m; ++r) {
        double mx = 0.0;
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) mx = smax(mx, fabs(a[k]));
        if (mx > 1.0) row_scale[r] = 1.0 / mx;
    }
    if (literal) {
        /* :162-176 verbatim, O(n*nnz). */
        for (int64_t j = 0; j < n; ++j) {
            double mx = 0.0;
            for (int64_t r = 0; r < m; ++r)
                for (int64_t k = rp[r]; k < rp[r + 1]; ++k)
                    if (ci[k] == j) mx = smax(mx, fabs(a[k]));
            if (mx > 1.0) col_scale[j] = 1.0 / mx;
        }
    } else {
        /* Same factors in one pass: max over a set is order-independent. */
        double *cm = calloc((size_t)n + 1, sizeof(double));
        for (int64_t k = 0; k < rp[m]; ++k) cm[ci[k]] = smax(cm[ci[k]], fabs(a[k]));
        for (int64_t j = 0; j < n; ++j)
            if (cm[j] > 1.0) col_scale[j] = 1.0 / cm[j];
        free(cm);
    }
    /* :178-187 */
    for (int64_t r = 0; r < m; ++r)
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k)
            values_out[k] = a[k] * (row_scale[r] * col_scale[ci[k]]);
    /* :188-196 */
    for (int64_t j = 0; j < n; ++j) {
        objective_out[j] = lp->objective[j] * col_scale[j];
        var_lower_out[j] = scale_bound(lower[j], col_scale[j], 0);
        var_upper_out[j] = scale_bound(upper[j], col_scale[j], 0);
    }
    /* :197-203 */
    for (int64_t r = 0; r < m; ++r) {
        row_lower_out[r] = scale_bound(lp->row_lower[r], row_scale[r], 1);
        row_upper_out[r] = scale_bound(lp->row_upper[r], row_scale[r], 1);
    }
}

/* continuous_common.hpp:207-225: sqrt(max row abs-sum * max column abs-sum). */
double orc_operator_bound(int64_t m, int64_t n, const int64_t *rp,
                          const int64_t *ci, const double *v) {
    double max_row = 0.0;
    double *cols = calloc((size_t)n + 1, sizeof(double));
    for (int64_t r = 0; r < m; ++r) {
        double rs = 0.0;
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
            rs += fabs(v[k]);
            cols[ci[k]] += fabs(v[k]);
        }
        max_row = smax(max_row, rs);
    }
    double max_col = 0.0;
    for (int64_t j = 0; j < n; ++j) max_col = smax(max_col, cols[j]);
    free(cols);
    return sqrt(max_row * max_col);
}

/* ----------------------------------------------------- relative KKT (new) -- */

/* The reference has no relative KKT; this is the definition the HIP path
 * also evaluates, in original coordinates with the reference's dual sign
 * (stationarity g = c + A'y, pdhg.cpp:189-198):
 *   rel_p = ||Ax - P_[rl,ru](Ax)||_2 / (1 + ||bhat||_2)
 *   rel_d = ||x - P_X(x - g)||_2     / (1 + ||c||_2)
 *   gap   = |c'x - D(y)| / (1 + |c'x| + |D(y)|)
 *   D(y)  = sum_j min(g_j l_j, g_j u_j) - sum_i (max(y_i,0) ru_i - max(-y_i,0) rl_i)
 * with 0*inf = 0; bhat_i is the finite row bound (the larger magnitude when
 * both are finite); D = -inf gives gap = 1.  Objective offset excluded. */
void orc_relative_kkt(const orc_lp *lp, const double *x, const double *y,
                      double *rel_primal, double *rel_dual, double *rel_gap,
                      double *primal_obj, double *dual_obj) {
    const int64_t m = lp->m, n = lp->n;
    const int64_t *rp = lp->row_offsets, *ci = lp->column_indices;
    const double *a = lp->values;
    double *g = calloc((size_t)n + 1, sizeof(double));
    double sp = 0.0, sb = 0.0, dual = 0.0;
    int dual_neg_inf = 0;
    for (int64_t r = 0; r < m; ++r) {
        double ax = 0.0;
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
            ax += a[k] * x[ci[k]];
            g[ci[k]] += a[k] * y[r];
        }
        const double rl = lp->row_lower[r], ru = lp->row_upper[r];
        const double d = ax - smin(smax(ax, rl), ru);
        sp += d * d;
        double bh = 0.0;
        if (isfinite(rl) && isfinite(ru)) bh = smax(fabs(rl), fabs(ru));
        else if (isfinite(rl)) bh = rl;
        else if (isfinite(ru)) bh = ru;
        sb += bh * bh;
        if (y[r] > 0.0) {
            if (isinf(ru)) dual_neg_inf = 1; else dual -= y[r] * ru;
        } else if (y[r] < 0.0) {
            if (isinf(rl)) dual_neg_inf = 1; else dual -= y[r] * rl;
        }
    }
    double sd = 0.0, sc = 0.0, pobj = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        const double c = lp->objective[j];
        const double gj = c + g[j];
        const double l = lp->var_lower[j], u = lp->var_upper[j];
        const double d = x[j] - smin(smax(x[j] - gj, l), u);
        sd += d * d;
        sc += c * c;
        pobj += c * x[j];
        if (gj > 0.0) {
            if (isinf(l)) dual_neg_inf = 1; else dual += gj * l;
        } else if (gj < 0.0) {
            if (isinf(u)) dual_neg_inf = 1; else dual += gj * u;
        }
    }
    free(g);
    *rel_primal = sqrt(sp) / (1.0 + sqrt(sb));
    *rel_dual = sqrt(sd) / (1.0 + sqrt(sc));
    if (dual_neg_inf) {
        *rel_gap = 1.0;
        dual = -INFINITY;
    } else {
        *rel_gap = fabs(pobj - dual) / (1.0 + fabs(pobj) + fabs(dual));
    }
    if (primal_obj) *primal_obj = pobj;
    if (dual_obj) *dual_obj = dual;
}

/* --------------------------------------------------------------- solve -- */

typedef struct {
    int64_t m, n, nnz;
    const int64_t *rp, *ci;
    double *v, *c, *l, *u, *rl, *ru, *rs, *cs;
} scaled_lp;

static void scaled_alloc(scaled_lp *s, const orc_lp *lp) {
    s->m = lp->m; s->n = lp->n; s->nnz = lp->row_offsets[lp->m];
    s->rp = lp->row_offsets; s->ci = lp->column_indices;
    s->v = malloc(((size_t)s->nnz + 1) * sizeof(double));
    s->c = malloc(((size_t)s->n + 1) * sizeof(double));
    s->l = malloc(((size_t)s->n + 1) * sizeof(double));
    s->u = malloc(((size_t)s->n + 1) * sizeof(double));
    s->cs = malloc(((size_t)s->n + 1) * sizeof(double));
    s->rl = malloc(((size_t)s->m + 1) * sizeof(double));
    s->ru = malloc(((size_t)s->m + 1) * sizeof(double));
    s->rs = malloc(((size_t)s->m + 1) * sizeof(double));
}

static void scaled_free(scaled_lp *s) {
    free(s->v); free(s->c); free(s->l); free(s->u); free(s->cs);
    free(s->rl); free(s->ru); free(s->rs);
}

/* Iterates of pdhg.cpp:37-49, zero-initialised like CpuVector. */
typedef struct {
    double *x, *x_next, *x_bar, *aty, *grad, *px, *res;   /* n */
    double *y, *ax_bar, *y_arg, *y_proj, *ax, *pax;      /* m */
} work;

static void work_alloc(work *w, int64_t m, int64_t n) {
    double **nv[] = {&w->x, &w->x_next, &w->x_bar, &w->aty, &w->grad, &w->px, &w->res};
    double **mv[] = {&w->y, &w->ax_bar, &w->y_arg, &w->y_proj, &w->ax, &w->pax};
    for (size_t i = 0; i < sizeof(nv) / sizeof(nv[0]); ++i)
        *nv[i] = calloc((size_t)n + 1, sizeof(double));
    for (size_t i = 0; i < sizeof(mv) / sizeof(mv[0]); ++i)
        *mv[i] = calloc((size_t)m + 1, sizeof(double));
}

static void work_free(work *w) {
    free(w->x); free(w->x_next); free(w->x_bar); free(w->aty); free(w->grad);
    free(w->px); free(w->res); free(w->y); free(w->ax_bar); free(w->y_arg);
    free(w->y_proj); free(w->ax); free(w->pax);
}

typedef struct {
    double tau, sigma, max_step, prev_metric, last_metric;
    int restart;
} loop_state;

/* One iteration of pdhg.cpp:92-130; returns the metric of :131. */
static double pdhg_iteration(const scaled_lp *s, work *w, loop_state *st) {
    const int64_t m = s->m, n = s->n;
    const double tau = st->tau, sigma = st->sigma;
    /* :92 */
    orc_spmv(m, n, s->rp, s->ci, s->v, 1.0, w->y, 0.0, w->aty, 1);
    /* :93-96 */
    memcpy(w->x_next, w->x, (size_t)n * sizeof(double));
    orc_axpby(n, -tau, s->c, 1.0, w->x_next);
    orc_axpby(n, -tau, w->aty, 1.0, w->x_next);
    orc_project_box(n, w->x_next, s->l, s->u);
    /* :98-104 */
    if (st->restart) {
        memcpy(w->x_bar, w->x_next, (size_t)n * sizeof(double));
        st->restart = 0;
    } else {
        memcpy(w->x_bar, w->x, (size_t)n * sizeof(double));
        orc_axpby(n, 2.0, w->x_next, -1.0, w->x_bar);
    }
    /* :105-113 */
    orc_spmv(m, n, s->rp, s->ci, s->v, 1.0, w->x_bar, 0.0, w->ax_bar, 0);
    memcpy(w->y_arg, w->y, (size_t)m * sizeof(double));
    orc_axpby(m, sigma, w->ax_bar, 1.0, w->y_arg);
    memcpy(w->y_proj, w->y_arg, (size_t)m * sizeof(double));
    orc_axpby(m, 1.0 / sigma, w->y_proj, 0.0, w->y_proj);
    orc_project_box(m, w->y_proj, s->rl, s->ru);
    memcpy(w->y, w->y_arg, (size_t)m * sizeof(double));
    orc_axpby(m, -sigma, w->y_proj, 1.0, w->y);
    memcpy(w->x, w->x_next, (size_t)n * sizeof(double));
    /* :116-120 */
    orc_spmv(m, n, s->rp, s->ci, s->v, 1.0, w->x, 0.0, w->ax, 0);
    memcpy(w->pax, w->ax, (size_t)m * sizeof(double));
    orc_project_box(m, w->pax, s->rl, s->ru);
    orc_axpby(m, 1.0, w->pax, -1.0, w->ax);
    const double primal = orc_norm_inf(m, w->ax);
    /* :122-130 */
    orc_spmv(m, n, s->rp, s->ci, s->v, 1.0, w->y, 0.0, w->aty, 1);
    memcpy(w->grad, s->c, (size_t)n * sizeof(double));
    orc_axpby(n, 1.0, w->aty, 1.0, w->grad);
    memcpy(w->px, w->x, (size_t)n * sizeof(double));
    orc_axpby(n, -1.0, w->grad, 1.0, w->px);
    orc_project_box(n, w->px, s->l, s->u);
    memcpy(w->res, w->x, (size_t)n * sizeof(double));
    orc_axpby(n, -1.0, w->px, 1.0, w->res);
    const double stationarity = orc_norm_inf(n, w->res);
    return smax(primal, stationarity);
}

/* pdhg.cpp:133-145 */
static void step_rule(loop_state *st, int64_t iter, double metric) {
    if (iter > 0 && iter % 50 == 0) {
        if (metric > st->prev_metric * 1.25) {
            st->tau *= 0.5;
            st->sigma *= 0.5;
            st->restart = 1;
        } else if (metric < st->prev_metric * 0.1) {
            st->tau = smin(st->tau * 1.05, st->max_step);
            st->sigma = st->tau;
        }
        st->prev_metric = metric;
    } else if (iter == 0) {
        st->prev_metric = metric;
    }
}

static void unscale(const scaled_lp *s, const work *w, double *x, double *y) {
    for (int64_t j = 0; j < s->n; ++j) x[j] = w->x[j] * s->cs[j];
    for (int64_t r = 0; r < s->m; ++r) y[r] = w->y[r] * s->rs[r];
}

static int valid_options(const orc_options *o) {
    /* continuous_common.hpp:50-56 */
    return o->iteration_limit >= 0 && isfinite(o->primal_tolerance) &&
           isfinite(o->dual_tolerance) && o->primal_tolerance >= 0.0 &&
           o->dual_tolerance >= 0.0 &&
           (!isfinite(o->time_limit_seconds) || o->time_limit_seconds >= 0.0);
}

static void result_init(orc_result *r) {
    memset(r, 0, sizeof(*r));
    r->objective_value = NAN;
    r->primal_residual = INFINITY;
    r->stationarity_residual = INFINITY;
    r->complementarity_residual = INFINITY;
    r->absolute_gap = INFINITY;
    r->relative_gap = INFINITY;
    r->rel_primal = r->rel_dual = r->rel_gap = r->rel_kkt = INFINITY;
}

int orc_solve_lp(const orc_lp *lp, const double *lower_override,
                 const double *upper_override, const double *warm_start,
                 const orc_options *options, double *x_out, double *y_out,
                 orc_result *result) {
    result_init(result);
    const int64_t m = lp->m, n = lp->n;
    const double *lower = lower_override ? lower_override : lp->var_lower;
    const double *upper = upper_override ? upper_override : lp->var_upper;
    /* pdhg.cpp:269-293 with copy_bounds/validate (continuous_common.hpp:80-126) */
    for (int64_t j = 0; j < n; ++j)
        if (isnan(lower[j]) || isnan(upper[j]) || lower[j] > upper[j]) {
            result->status = ORC_INVALID_MODEL;
            return result->status;
        }
    for (int64_t r = 0; r < m; ++r)
        if (isnan(lp->row_lower[r]) || isnan(lp->row_upper[r]) ||
            lp->row_lower[r] > lp->row_upper[r]) {
            result->status = ORC_INVALID_MODEL;
            return result->status;
        }
    if (orc_validate_pattern(m, n, m + 1, lp->row_offsets, lp->row_offsets[m],
                             lp->column_indices) != 0 || n == 0 ||
        !valid_options(options)) {
        result->status = ORC_INVALID_MODEL;
        return result->status;
    }
    const double started = now_seconds();
    scaled_lp s;
    scaled_alloc(&s, lp);
    orc_scale_lp(lp, lower, upper, 0, s.v, s.c, s.rl, s.ru, s.l, s.u, s.rs, s.cs);
    if (warm_start) {
        for (int64_t j = 0; j < n; ++j)
            if (!isfinite(warm_start[j])) {
                scaled_free(&s);
                result->status = ORC_INVALID_MODEL;
                return result->status;
            }
    }
    work w;
    work_alloc(&w, m, n);
    /* pdhg.cpp:61-72 */
    if (warm_start)
        for (int64_t j = 0; j < n; ++j) w.x[j] = warm_start[j] / s.cs[j];
    orc_project_box(n, w.x, s.l, s.u);

    /* pdhg.cpp:75-80 */
    const double norm = orc_operator_bound(m, n, s.rp, s.ci, s.v);
    loop_state st;
    st.max_step = 0.99 / smax(1.0, norm);
    st.tau = st.max_step;
    st.sigma = st.max_step;
    st.prev_metric = INFINITY;
    st.last_metric = INFINITY;
    st.restart = 0;

    orc_lp orig = *lp;
    orig.var_lower = lower;
    orig.var_upper = upper;
    double *xo = malloc(((size_t)n + 1) * sizeof(double));
    double *yo = malloc(((size_t)m + 1) * sizeof(double));
    const double ktol = options->relative_kkt_tolerance;
    const int64_t kper = options->kkt_check_period > 0 ? options->kkt_check_period : 1;

    int status = ORC_NOT_SOLVED;
    int64_t iterations = 0;
    for (int64_t iter = 0; iter < options->iteration_limit; ++iter) {
        if (isfinite(options->time_limit_seconds) &&
            now_seconds() - started >= options->time_limit_seconds) {
            status = ORC_TIME_LIMIT;
            break;
        }
        const double metric = pdhg_iteration(&s, &w, &st);
        iterations = iter + 1;
        st.last_metric = metric;
        step_rule(&st, iter, metric);
        if (ktol > 0.0) {
            if ((iter + 1) % kper == 0) {
                double rp, rd, gp;
                unscale(&s, &w, xo, yo);
                orc_relative_kkt(&orig, xo, yo, &rp, &rd, &gp, NULL, NULL);
                if (smax(rp, smax(rd, gp)) <= ktol) {
                    status = ORC_OPTIMAL;
                    break;
                }
            }
        } else if (metric <= smax(options->primal_tolerance, options->dual_tolerance)) {
            status = ORC_OPTIMAL;
            break;
        }
    }
    result->iterations = iterations;
    result->tau = st.tau;
    result->sigma = st.sigma;
    result->last_metric = st.last_metric;

    /* postsolve: pdhg.cpp:153-264 */
    unscale(&s, &w, xo, yo);
    result->objective_value = lp->objective_offset + orc_dot(n, lp->objective, xo);
    if (!isfinite(result->objective_value)) status = ORC_NUMERICAL_FAILURE;
    double *row_values = calloc((size_t)m + 1, sizeof(double));
    for (int64_t r = 0; r < m; ++r)
        for (int64_t k = lp->row_offsets[r]; k < lp->row_offsets[r + 1]; ++k)
            row_values[r] += lp->values[k] * xo[lp->column_indices[k]];
    /* box_violation (continuous_common.hpp:317-326) */
    double pv = 0.0;
    for (int64_t r = 0; r < m; ++r) {
        if (isfinite(lp->row_lower[r])) pv = smax(pv, lp->row_lower[r] - row_values[r]);
        if (isfinite(lp->row_upper[r])) pv = smax(pv, row_values[r] - lp->row_upper[r]);
    }
    pv = smax(0.0, pv);
    double bv = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        if (isfinite(lower[j])) bv = smax(bv, lower[j] - xo[j]);
        if (isfinite(upper[j])) bv = smax(bv, xo[j] - upper[j]);
    }
    bv = smax(0.0, bv);
    result->primal_residual = smax(pv, bv);
    /* :189-198 */
    double *stat = malloc(((size_t)n + 1) * sizeof(double));
    memcpy(stat, lp->objective, (size_t)n * sizeof(double));
    for (int64_t r = 0; r < m; ++r)
        for (int64_t k = lp->row_offsets[r]; k < lp->row_offsets[r + 1]; ++k)
            stat[lp->column_indices[k]] += yo[r] * lp->values[k];
    /* :200-243 */
    double sr = 0.0, cr = 0.0;
    const double active = smax(10.0 * options->primal_tolerance, 1e-8);
    for (int64_t j = 0; j < n; ++j) {
        double lo = lower[j], hi = upper[j];
        double t = xo[j] - stat[j];
        /* std::clamp(v, lo, hi) = v < lo ? lo : (hi < v ? hi : v) */
        const double proj = (t < lo) ? lo : ((hi < t) ? hi : t);
        sr = smax(sr, fabs(xo[j] - proj));
        if (isfinite(lo) && xo[j] - lo <= active) {
            const double ld = smax(0.0, stat[j]);
            cr = smax(cr, fabs((xo[j] - lo) * ld));
        }
        if (isfinite(hi) && hi - xo[j] <= active) {
            const double ud = smax(0.0, -stat[j]);
            cr = smax(cr, fabs((hi - xo[j]) * ud));
        }
    }
    for (int64_t r = 0; r < m; ++r) {
        const double v = row_values[r], lo = lp->row_lower[r], hi = lp->row_upper[r];
        const double ld = smax(0.0, -yo[r]), ud = smax(0.0, yo[r]);
        if (isfinite(lo)) cr = smax(cr, fabs((v - lo) * ld));
        if (isfinite(hi)) cr = smax(cr, fabs((hi - v) * ud));
    }
    result->stationarity_residual = sr;
    result->complementarity_residual = cr;
    if (status == ORC_NOT_SOLVED) status = ORC_ITERATION_LIMIT;
    int finite = 1;
    for (int64_t j = 0; j < n; ++j) if (!isfinite(xo[j])) { finite = 0; break; }
    if (!finite) {
        status = ORC_NUMERICAL_FAILURE;
    } else if (status == ORC_OPTIMAL && ktol <= 0.0 &&
               (result->primal_residual > options->primal_tolerance ||
                result->stationarity_residual > options->dual_tolerance)) {
        status = ORC_NUMERICAL_FAILURE;
    }
    if (status == ORC_OPTIMAL) {
        result->absolute_gap = 0.0;
        result->relative_gap = 0.0;
    }
    orc_relative_kkt(&orig, xo, yo, &result->rel_primal, &result->rel_dual,
                     &result->rel_gap, &result->primal_objective,
                     &result->dual_objective);
    result->rel_kkt = smax(result->rel_primal, smax(result->rel_dual, result->rel_gap));
    result->status = status;
    result->solve_seconds = now_seconds() - started;
    if (x_out) memcpy(x_out, xo, (size_t)n * sizeof(double));
    if (y_out) memcpy(y_out, yo, (size_t)m * sizeof(double));
    free(stat); free(row_values); free(xo); free(yo);
    work_free(&w);
    scaled_free(&s);
    return status;
}

double orc_time_iterations(const orc_lp *lp, int64_t iterations) {
    scaled_lp s;
    scaled_alloc(&s, lp);
    orc_scale_lp(lp, lp->var_lower, lp->var_upper, 0, s.v, s.c, s.rl, s.ru, s.l,
                 s.u, s.rs, s.cs);
    work w;
    work_alloc(&w, s.m, s.n);
    orc_project_box(s.n, w.x, s.l, s.u);
    loop_state st;
    st.max_step = 0.99 / smax(1.0, orc_operator_bound(s.m, s.n, s.rp, s.ci, s.v));
    st.tau = st.sigma = st.max_step;
    st.prev_metric = INFINITY;
    st.restart = 0;
    const double t0 = now_seconds();
    for (int64_t it = 0; it < iterations; ++it) {
        const double metric = pdhg_iteration(&s, &w, &st);
        step_rule(&st, it, metric);
    }
    const double dt = now_seconds() - t0;
    work_free(&w);
    scaled_free(&s);
    return dt;
}

/* ------------------------------------------------- unrolled PDHG + adjoint --
 * The fixed-step iteration a differentiable layer unrolls: the op sequence of
 * pdhg.cpp:92-113 with tau = sigma = 0.99 / max(1, bound) (pdhg.cpp:75-78)
 * held fixed and no restart branch (pdhg.cpp:98-100 never taken), from the
 * projected zero start (pdhg.cpp:71) and y = 0.  The adjoint is derived here
 * from that sequence; nothing in the reference differentiates.  Per step, with
 * z = x - tau c - tau A'y, x+ = P_X(z), xbar = 2 x+ - x, w = y + sigma A xbar,
 * t = w / sigma, p = P_[rl,ru](t), y+ = w - sigma p:
 *   gw   = [row clamped] gy+            (y+ = w - sigma p: d/dw is 0 inside)
 *   grl -= sigma [t < rl] gy+,  gru -= sigma [t > ru] gy+
 *   gxb  = sigma A' gw
 *   gz   = [l <= z <= u] (gx+ + 2 gxb)
 *   gx   = gz - gxb,  gc -= tau gz,  gy = gw - tau A gz
 * Masks follow project_box = min(max(v, lo), hi): v == lo or v == hi count
 * as inside. */

static void unrolled_step(const scaled_lp *s, double tau, double sigma, double *x, double *y,
                          double *x_next, double *x_bar, double *aty, double *ax_bar, double *y_arg,
                          double *y_proj, unsigned char *mx, unsigned char *my) {
    const int64_t m = s->m, n = s->n;
    orc_spmv(m, n, s->rp, s->ci, s->v, 1.0, y, 0.0, aty, 1);           /* :92 */
    memcpy(x_next, x, (size_t)n * sizeof(double));                     /* :93-96 */
    orc_axpby(n, -tau, s->c, 1.0, x_next);
    orc_axpby(n, -tau, aty, 1.0, x_next);
    if (mx)
        for (int64_t j = 0; j < n; ++j) mx[j] = !(x_next[j] < s->l[j]) && !(s->u[j] < x_next[j]);
    orc_project_box(n, x_next, s->l, s->u);
    memcpy(x_bar, x, (size_t)n * sizeof(double));                      /* :102-103 */
    orc_axpby(n, 2.0, x_next, -1.0, x_bar);
    orc_spmv(m, n, s->rp, s->ci, s->v, 1.0, x_bar, 0.0, ax_bar, 0);   /* :105-113 */
    memcpy(y_arg, y, (size_t)m * sizeof(double));
    orc_axpby(m, sigma, ax_bar, 1.0, y_arg);
    memcpy(y_proj, y_arg, (size_t)m * sizeof(double));
    orc_axpby(m, 1.0 / sigma, y_proj, 0.0, y_proj);
    if (my)
        for (int64_t i = 0; i < m; ++i)
            my[i] = (y_proj[i] < s->rl[i]) ? 1 : ((s->ru[i] < y_proj[i]) ? 2 : 0);
    orc_project_box(m, y_proj, s->rl, s->ru);
    memcpy(y, y_arg, (size_t)m * sizeof(double));
    orc_axpby(m, -sigma, y_proj, 1.0, y);
    memcpy(x, x_next, (size_t)n * sizeof(double));
}

int orc_unrolled(const orc_lp *lp, int64_t iterations, const double *grad_x, const double *grad_y,
                 double *x_out, double *y_out, double *grad_c, double *grad_row_lower,
                 double *grad_row_upper, double *step_out) {
    if (iterations < 0 || lp->n <= 0) return -1;
    scaled_lp s;
    scaled_alloc(&s, lp);
    orc_scale_lp(lp, lp->var_lower, lp->var_upper, 0, s.v, s.c, s.rl, s.ru, s.l, s.u, s.rs, s.cs);
    const int64_t m = s.m, n = s.n;
    const double step = 0.99 / smax(1.0, orc_operator_bound(m, n, s.rp, s.ci, s.v));
    const double tau = step, sigma = step;
    if (step_out) *step_out = step;
    double *x = calloc((size_t)n + 1, sizeof(double)), *y = calloc((size_t)m + 1, sizeof(double));
    double *xn = calloc((size_t)n + 1, sizeof(double)), *xb = calloc((size_t)n + 1, sizeof(double));
    double *aty = calloc((size_t)n + 1, sizeof(double)), *axb = calloc((size_t)m + 1, sizeof(double));
    double *ya = calloc((size_t)m + 1, sizeof(double)), *yp = calloc((size_t)m + 1, sizeof(double));
    const int want_grad = grad_x || grad_y;
    unsigned char *mx = want_grad ? malloc((size_t)iterations * n + 1) : NULL;
    unsigned char *my = want_grad ? malloc((size_t)iterations * m + 1) : NULL;
    orc_project_box(n, x, s.l, s.u);                                     /* :71 */
    for (int64_t k = 0; k < iterations; ++k)
        unrolled_step(&s, tau, sigma, x, y, xn, xb, aty, axb, ya, yp,
                      mx ? mx + (size_t)k * n : NULL, my ? my + (size_t)k * m : NULL);
    if (x_out) for (int64_t j = 0; j < n; ++j) x_out[j] = x[j] * s.cs[j];
    if (y_out) for (int64_t i = 0; i < m; ++i) y_out[i] = y[i] * s.rs[i];
    if (want_grad) {
        /* adjoints of the scaled state: x = s xhat, y = r yhat */
        double *gx = calloc((size_t)n + 1, sizeof(double)), *gy = calloc((size_t)m + 1, sizeof(double));
        double *gw = calloc((size_t)m + 1, sizeof(double)), *gxb = calloc((size_t)n + 1, sizeof(double));
        double *gz = calloc((size_t)n + 1, sizeof(double)), *agz = calloc((size_t)m + 1, sizeof(double));
        double *gc = calloc((size_t)n + 1, sizeof(double));
        double *grl = calloc((size_t)m + 1, sizeof(double)), *gru = calloc((size_t)m + 1, sizeof(double));
        for (int64_t j = 0; j < n; ++j) gx[j] = grad_x ? grad_x[j] * s.cs[j] : 0.0;
        for (int64_t i = 0; i < m; ++i) gy[i] = grad_y ? grad_y[i] * s.rs[i] : 0.0;
        for (int64_t k = iterations - 1; k >= 0; --k) {
            const unsigned char *kx = mx + (size_t)k * n, *ky = my + (size_t)k * m;
            for (int64_t i = 0; i < m; ++i) {
                gw[i] = ky[i] ? gy[i] : 0.0;
                if (ky[i] == 1) grl[i] -= sigma * gy[i];
                if (ky[i] == 2) gru[i] -= sigma * gy[i];
            }
            orc_spmv(m, n, s.rp, s.ci, s.v, sigma, gw, 0.0, gxb, 1);
            for (int64_t j = 0; j < n; ++j) {
                gz[j] = kx[j] ? gx[j] + 2.0 * gxb[j] : 0.0;
                gx[j] = gz[j] - gxb[j];
                gc[j] -= tau * gz[j];
            }
            orc_spmv(m, n, s.rp, s.ci, s.v, 1.0, gz, 0.0, agz, 0);
            for (int64_t i = 0; i < m; ++i) gy[i] = gw[i] - tau * agz[i];
        }
        /* c hat = s c, rl hat = r rl, ru hat = r ru */
        if (grad_c) for (int64_t j = 0; j < n; ++j) grad_c[j] = gc[j] * s.cs[j];
        if (grad_row_lower) for (int64_t i = 0; i < m; ++i) grad_row_lower[i] = grl[i] * s.rs[i];
        if (grad_row_upper) for (int64_t i = 0; i < m; ++i) grad_row_upper[i] = gru[i] * s.rs[i];
        free(gx); free(gy); free(gw); free(gxb); free(gz); free(agz); free(gc); free(grl); free(gru);
    }
    free(x); free(y); free(xn); free(xb); free(aty); free(axb); free(ya); free(yp); free(mx); free(my);
    scaled_free(&s);
    return 0;
}

/* ------------------------------------------------------------------ PDQP --
 * src/solver/qp/pdqp.cpp restated: solve_qp (:303-331) and solve_scaled_qp
 * (:20-300), with positive_semidefinite (continuous_common.hpp:242-275) and
 * objective_value (:284-299).  Q is n x n CSR stored symmetrically. */

/* continuous_common.hpp:227-240 csr_to_dense */
static double *dense_q(int64_t n, const int64_t *qrp, const int64_t *qci, const double *qv) {
    double *a = calloc((size_t)n * (size_t)n + 1, sizeof(double));
    if (!qrp) return a;
    for (int64_t r = 0; r < n; ++r)
        for (int64_t k = qrp[r]; k < qrp[r + 1]; ++k) a[(size_t)r * n + qci[k]] = qv[k];
    return a;
}

/* continuous_common.hpp:242-275: dense Cholesky with a semidefinite pivot rule */
static int q_positive_semidefinite(int64_t n, const int64_t *qrp, const int64_t *qci, const double *qv) {
    double *a = dense_q(n, qrp, qci, qv);
    const double tol = 1e-10;
    int ok = 1;
    for (int64_t k = 0; k < n && ok; ++k) {
        double pivot = a[(size_t)k * n + k];
        for (int64_t j = 0; j < k; ++j) pivot -= a[(size_t)k * n + j] * a[(size_t)k * n + j];
        if (pivot < -tol * smax(1.0, fabs(a[(size_t)k * n + k]))) { ok = 0; break; }
        if (pivot <= tol) {
            for (int64_t i = k + 1; i < n; ++i) {
                if (fabs(a[(size_t)i * n + k]) > tol) { ok = 0; break; }
                a[(size_t)i * n + k] = 0.0;
            }
            continue;
        }
        const double d = sqrt(pivot);
        a[(size_t)k * n + k] = d;
        for (int64_t i = k + 1; i < n; ++i) {
            double v = a[(size_t)i * n + k];
            for (int64_t j = 0; j < k; ++j) v -= a[(size_t)i * n + j] * a[(size_t)k * n + j];
            a[(size_t)i * n + k] = v / d;
        }
    }
    free(a);
    return ok;
}

/* Model.cpp:145-170: a square pattern matching the variables, entries stored
 * symmetrically with equal values. */
static int q_valid(int64_t n, const int64_t *qrp, const int64_t *qci, const double *qv) {
    if (orc_validate_pattern(n, n, n + 1, qrp, qrp[n], qci) != 0) return 0;
    for (int64_t k = 0; k < qrp[n]; ++k)
        if (!isfinite(qv[k])) return 0;
    for (int64_t r = 0; r < n; ++r)
        for (int64_t k = qrp[r]; k < qrp[r + 1]; ++k) {
            const int64_t c = qci[k];
            int found = 0;
            for (int64_t t = qrp[c]; t < qrp[c + 1]; ++t)
                if (qci[t] == r) { found = qv[t] == qv[k]; break; }
            if (!found) return 0;
        }
    return 1;
}

int orc_solve_qp(const orc_lp *lp, const int64_t *qrp, const int64_t *qci, const double *qv,
                 const double *lower_override, const double *upper_override, const double *warm_start,
                 const orc_options *options, double *x_out, double *y_out, orc_result *result) {
    result_init(result);
    const int64_t m = lp->m, n = lp->n;
    const double *lower = lower_override ? lower_override : lp->var_lower;
    const double *upper = upper_override ? upper_override : lp->var_upper;
    /* :307-315 copy_bounds and validate_continuous */
    for (int64_t j = 0; j < n; ++j)
        if (isnan(lower[j]) || isnan(upper[j]) || lower[j] > upper[j]) {
            result->status = ORC_INVALID_MODEL;
            return result->status;
        }
    for (int64_t r = 0; r < m; ++r)
        if (isnan(lp->row_lower[r]) || isnan(lp->row_upper[r]) || lp->row_lower[r] > lp->row_upper[r]) {
            result->status = ORC_INVALID_MODEL;
            return result->status;
        }
    if (orc_validate_pattern(m, n, m + 1, lp->row_offsets, lp->row_offsets[m], lp->column_indices) != 0 ||
        !qrp || !q_valid(n, qrp, qci, qv)) {
        result->status = ORC_INVALID_MODEL;
        return result->status;
    }
    if (qrp[n] == 0) {                       /* no quadratic values: an LP */
        result->status = ORC_UNSUPPORTED_MODEL;
        return result->status;
    }
    if (n == 0) { result->status = ORC_INVALID_MODEL; return result->status; }             /* :316-318 */
    if (!q_positive_semidefinite(n, qrp, qci, qv)) {                                        /* :319-321 */
        result->status = ORC_UNSUPPORTED_MODEL;
        return result->status;
    }
    if (!valid_options(options)) { result->status = ORC_INVALID_MODEL; return result->status; }   /* :322-324 */
    const double started = now_seconds();
    scaled_lp s;
    scaled_alloc(&s, lp);
    orc_scale_lp(lp, lower, upper, 0, s.v, s.c, s.rl, s.ru, s.l, s.u, s.rs, s.cs);
    if (warm_start)                                                                         /* :29-34 */
        for (int64_t j = 0; j < n; ++j)
            if (!isfinite(warm_start[j])) {
                scaled_free(&s);
                result->status = ORC_INVALID_MODEL;
                return result->status;
            }
    /* :36-45 Q scaled by s_row s_col */
    const int64_t qnnz = qrp[n];
    double *sq = malloc(((size_t)qnnz + 1) * sizeof(double));
    for (int64_t r = 0; r < n; ++r)
        for (int64_t k = qrp[r]; k < qrp[r + 1]; ++k) sq[k] = qv[k] * (s.cs[r] * s.cs[qci[k]]);
    work w;
    work_alloc(&w, m, n);
    double *qx = calloc((size_t)n + 1, sizeof(double));
    /* :75-85 */
    if (warm_start)
        for (int64_t j = 0; j < n; ++j) w.x[j] = warm_start[j] / s.cs[j];
    orc_project_box(n, w.x, s.l, s.u);
    /* :87-99 */
    const double norm = orc_operator_bound(m, n, s.rp, s.ci, s.v);
    double q_norm = 0.0;
    for (int64_t r = 0; r < n; ++r) {
        double row_sum = 0.0;
        for (int64_t k = qrp[r]; k < qrp[r + 1]; ++k) row_sum += fabs(sq[k]);
        q_norm = smax(q_norm, row_sum);
    }
    const double max_step = 0.99 / smax(1.0, norm * norm + q_norm);
    loop_state st;
    st.max_step = max_step;
    st.tau = st.sigma = max_step;
    st.prev_metric = INFINITY;
    st.last_metric = INFINITY;
    st.restart = 0;
    int status = ORC_NOT_SOLVED;
    int64_t iterations = 0;
    for (int64_t iter = 0; iter < options->iteration_limit; ++iter) {
        if (isfinite(options->time_limit_seconds) && now_seconds() - started >= options->time_limit_seconds) {
            status = ORC_TIME_LIMIT;
            break;
        }
        const double tau = st.tau, sigma = st.sigma;
        /* :115-122 */
        orc_spmv(m, n, s.rp, s.ci, s.v, 1.0, w.y, 0.0, w.aty, 1);
        orc_spmv(n, n, qrp, qci, sq, 1.0, w.x, 0.0, qx, 0);
        memcpy(w.grad, qx, (size_t)n * sizeof(double));
        orc_axpby(n, 1.0, s.c, 1.0, w.grad);
        orc_axpby(n, 1.0, w.aty, 1.0, w.grad);
        memcpy(w.x_next, w.x, (size_t)n * sizeof(double));
        orc_axpby(n, -tau, w.grad, 1.0, w.x_next);
        orc_project_box(n, w.x_next, s.l, s.u);
        /* :124-130 */
        if (st.restart) {
            memcpy(w.x_bar, w.x_next, (size_t)n * sizeof(double));
            st.restart = 0;
        } else {
            memcpy(w.x_bar, w.x, (size_t)n * sizeof(double));
            orc_axpby(n, 2.0, w.x_next, -1.0, w.x_bar);
        }
        /* :131-139 */
        orc_spmv(m, n, s.rp, s.ci, s.v, 1.0, w.x_bar, 0.0, w.ax_bar, 0);
        memcpy(w.y_arg, w.y, (size_t)m * sizeof(double));
        orc_axpby(m, sigma, w.ax_bar, 1.0, w.y_arg);
        memcpy(w.y_proj, w.y_arg, (size_t)m * sizeof(double));
        orc_axpby(m, 1.0 / sigma, w.y_proj, 0.0, w.y_proj);
        orc_project_box(m, w.y_proj, s.rl, s.ru);
        memcpy(w.y, w.y_arg, (size_t)m * sizeof(double));
        orc_axpby(m, -sigma, w.y_proj, 1.0, w.y);
        memcpy(w.x, w.x_next, (size_t)n * sizeof(double));
        iterations = iter + 1;
        /* :142-146 */
        orc_spmv(m, n, s.rp, s.ci, s.v, 1.0, w.x, 0.0, w.ax, 0);
        memcpy(w.pax, w.ax, (size_t)m * sizeof(double));
        orc_project_box(m, w.pax, s.rl, s.ru);
        orc_axpby(m, 1.0, w.pax, -1.0, w.ax);
        const double primal = orc_norm_inf(m, w.ax);
        /* :148-158 */
        orc_spmv(m, n, s.rp, s.ci, s.v, 1.0, w.y, 0.0, w.aty, 1);
        orc_spmv(n, n, qrp, qci, sq, 1.0, w.x, 0.0, qx, 0);
        memcpy(w.grad, qx, (size_t)n * sizeof(double));
        orc_axpby(n, 1.0, s.c, 1.0, w.grad);
        orc_axpby(n, 1.0, w.aty, 1.0, w.grad);
        memcpy(w.px, w.x, (size_t)n * sizeof(double));
        orc_axpby(n, -1.0, w.grad, 1.0, w.px);
        orc_project_box(n, w.px, s.l, s.u);
        memcpy(w.res, w.x, (size_t)n * sizeof(double));
        orc_axpby(n, -1.0, w.px, 1.0, w.res);
        const double metric = smax(primal, orc_norm_inf(n, w.res));
        st.last_metric = metric;
        step_rule(&st, iter, metric);                                                       /* :161-171 */
        if (metric <= smax(options->primal_tolerance, options->dual_tolerance)) {           /* :172-175 */
            status = ORC_OPTIMAL;
            break;
        }
    }
    result->iterations = iterations;
    result->tau = st.tau;
    result->sigma = st.sigma;
    result->last_metric = st.last_metric;
    /* :178-300 postsolve */
    double *xo = malloc(((size_t)n + 1) * sizeof(double));
    double *yo = malloc(((size_t)m + 1) * sizeof(double));
    unscale(&s, &w, xo, yo);
    double *qd = dense_q(n, qrp, qci, qv);
    /* objective_value: offset + c'x + 0.5 sum_r sum_c x_r q_rc x_c */
    double obj = lp->objective_offset + orc_dot(n, lp->objective, xo);
    double quad = 0.0;
    for (int64_t r = 0; r < n; ++r)
        for (int64_t c = 0; c < n; ++c) quad += xo[r] * qd[(size_t)r * n + c] * xo[c];
    obj += 0.5 * quad;
    result->objective_value = obj;
    if (!isfinite(obj)) status = ORC_NUMERICAL_FAILURE;
    double *row_values = calloc((size_t)m + 1, sizeof(double));
    for (int64_t r = 0; r < m; ++r)
        for (int64_t k = lp->row_offsets[r]; k < lp->row_offsets[r + 1]; ++k)
            row_values[r] += lp->values[k] * xo[lp->column_indices[k]];
    double pv = 0.0;
    for (int64_t r = 0; r < m; ++r) {
        if (isfinite(lp->row_lower[r])) pv = smax(pv, lp->row_lower[r] - row_values[r]);
        if (isfinite(lp->row_upper[r])) pv = smax(pv, row_values[r] - lp->row_upper[r]);
    }
    pv = smax(0.0, pv);
    double bv = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        if (isfinite(lower[j])) bv = smax(bv, lower[j] - xo[j]);
        if (isfinite(upper[j])) bv = smax(bv, xo[j] - upper[j]);
    }
    bv = smax(0.0, bv);
    result->primal_residual = smax(pv, bv);
    /* :216-232 gradient = c + Q x (dense rows) + A'y (row order) */
    double *grad = malloc(((size_t)n + 1) * sizeof(double));
    memcpy(grad, lp->objective, (size_t)n * sizeof(double));
    for (int64_t r = 0; r < n; ++r)
        for (int64_t c = 0; c < n; ++c) grad[r] += qd[(size_t)r * n + c] * xo[c];
    for (int64_t r = 0; r < m; ++r)
        for (int64_t k = lp->row_offsets[r]; k < lp->row_offsets[r + 1]; ++k)
            grad[lp->column_indices[k]] += yo[r] * lp->values[k];
    double sr = 0.0, cr = 0.0;
    const double active = smax(10.0 * options->primal_tolerance, 1e-8);
    for (int64_t j = 0; j < n; ++j) {
        const double lo = lower[j], hi = upper[j];
        const double t = xo[j] - grad[j];
        const double proj = (t < lo) ? lo : ((hi < t) ? hi : t);
        sr = smax(sr, fabs(xo[j] - proj));
        if (isfinite(lo) && xo[j] - lo <= active) cr = smax(cr, fabs((xo[j] - lo) * smax(0.0, grad[j])));
        if (isfinite(hi) && hi - xo[j] <= active) cr = smax(cr, fabs((hi - xo[j]) * smax(0.0, -grad[j])));
    }
    for (int64_t r = 0; r < m; ++r) {
        const double v = row_values[r], lo = lp->row_lower[r], hi = lp->row_upper[r];
        if (isfinite(lo)) cr = smax(cr, fabs((v - lo) * smax(0.0, -yo[r])));
        if (isfinite(hi)) cr = smax(cr, fabs((hi - v) * smax(0.0, yo[r])));
    }
    result->stationarity_residual = sr;
    result->complementarity_residual = cr;
    if (status == ORC_NOT_SOLVED) status = ORC_ITERATION_LIMIT;
    int finite = 1;
    for (int64_t j = 0; j < n; ++j) if (!isfinite(xo[j])) { finite = 0; break; }
    if (!finite) status = ORC_NUMERICAL_FAILURE;
    else if (status == ORC_OPTIMAL &&
             (result->primal_residual > options->primal_tolerance ||
              result->stationarity_residual > options->dual_tolerance))
        status = ORC_NUMERICAL_FAILURE;
    if (status == ORC_OPTIMAL) {
        result->absolute_gap = 0.0;
        result->relative_gap = 0.0;
    }
    result->status = status;
    result->solve_seconds = now_seconds() - started;
    if (x_out) memcpy(x_out, xo, (size_t)n * sizeof(double));
    if (y_out) memcpy(y_out, yo, (size_t)m * sizeof(double));
    free(grad); free(row_values); free(qd); free(xo); free(yo); free(qx); free(sq);
    work_free(&w);
    scaled_free(&s);
    return status;
}
