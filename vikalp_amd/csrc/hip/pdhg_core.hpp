// PDHG iteration core shared by the single-LP solver (pdhg.hip) and the
// batched persistent solver (batch.hip): device state, buffers, the fused
// epilogues and the step / restart / termination rules.  See pdhg.hip for
// the algorithm and its reference citations.
#pragma once

#include "common.hpp"
#include "vikalp_hip.h"

namespace vk {

constexpr int kDecisionPeriod = 50;   // pdhg.cpp:133

// Device-resident solver state: written by the one-block finalizer and
// check kernels (one thread), read by every kernel at its start.
struct DevState {
    long long k;            // completed iterations
    long long ks;           // iterations since the last restart
    long long ns;           // terms in the running sums
    long long ns_avg;       // terms of the average evaluated at the last check
    long long limit;
    long long period;       // adaptive restart-check period
    int status;             // VK_PDHG_*
    int policy;
    int restart_flag;       // reference: the next dual step uses xbar = x+
    int restart_apply;      // adaptive: the last check restarted
    int restart_to_avg;     // adaptive: ... to the average
    int final_from_avg;     // adaptive: converged on the average
    int restarts;
    double tau, sigma, inv_sigma, eta, omega, max_step;
    double prev_metric, metric, rp, rd;
    double metric_tol, rel_tol, norm_b, norm_c;
    double rel_p, rel_d, rel_g, rel_kkt, pobj, dobj;
    double cur_pr2, cur_dr2, cur_gap, cur_dx2, cur_dy2;
    double kkt_ref, kkt_prev;   // adaptive: weighted KKT at the restart / last check;
                                // Halpern: fixed-point residual at the restart / last check
    // Row part of the dual objective of the last dual iterate, sum of its
    // term magnitudes and count of -inf terms (primal infeasibility test).
    double d0_rows, d0_rows_abs, d0_rows_ninf;
    // Recentred solver: c'x of the base point, added to both objectives (the
    // partial sums hold those of the LP shifted to the base); 0 otherwise.
    double cx_base;
    // Primal-weight update at restarts (adaptive and Halpern): proportional
    // and integral gains on the log-balance error, and the error's sum.
    double pw_kp, pw_ki, pw_int;
};

// Partial-slot layout (slot 0 combines with max, the rest with +).
//   rows: 0 max|P(ax)-ax| (scaled)   1 sum (orig primal residual)^2
//         2 sum (scaled primal residual)^2   3 dual-objective row terms
//         4 count of -inf dual terms   5 sum (y - y_last)^2
//         6 (Halpern) sum (y+ - y_k)^2   7 sum |dual-objective row terms|
//   cols: 0 max|x-P(x-g)| (scaled)   1 sum (orig dual residual)^2
//         2 sum (scaled dual residual)^2   3 dual-objective column terms
//         4 count of -inf dual terms   5 c'x   6 sum (x - x_last)^2
//         7 (Halpern) sum (x+ - x_k)^2

template <class T>
struct Bufs {
    T *x[2];            // x_k in x[k & 1]
    T *y, *ax, *aty;    // current y, A x, A' y (in place)
    double *atyd;       // A' y before rounding to T (split primal pass; aliases
                        // aty when T is double)
    double *qx;         // Q x+ of a quadratic objective (PDQP); nullptr for an LP
    const T *c, *l, *u, *rl, *ru;
    const T *row_unscale, *col_scale;   // 1/r and s
    // Adaptive only.  The running sums of the iterates stay FP64 in every
    // storage precision: an FP32 sum of thousands of iterates would carry
    // its own rounding into the average far above FP32 iterate precision.
    double *xsum, *ysum;
    T *xlast, *ylast, *ax_avg, *aty_avg;
    // Halpern only: the Halpern iterate (xh, yh; its A x and A' y live in ax
    // and aty) and the anchor of the current restart (x0h, y0h, ax0h,
    // aty0h).  x[] and y then hold the PDHG outputs T(z), which the SpMVs
    // gather; ax_avg / aty_avg keep A x+ and A' y+ at restart checks.
    T *xh, *yh, *x0h, *y0h, *ax0h, *aty0h;
    // Recentred Halpern solver (null otherwise): every iterate above is a
    // correction to an FP64 base point (xbase, ybase), and c, l, u, rl, ru
    // hold the LP shifted to that base (c + A'ybase, l - xbase, u - xbase,
    // rl - A xbase, ru - A xbase).  The dual step acts on the total dual
    // ybase + y; cobj is the objective c itself (primal objective terms).
    const double *ybase;
    const T *cobj;
    // Halpern with a quadratic objective (FP64; null otherwise): Q x of the
    // Halpern iterate and of the anchor, carried by linearity like A'y, Q x+
    // kept at restart checks, and the absolute row sums of Q that set each
    // column's primal step (qp_column_step).
    double *qxh, *qx0h, *qx_avg;
    const double *qabs;
    double *part_rows, *part_cols, *part_erows, *part_ecols;
    DevState *st;
    int m, n;
    int grid_rows, grid_cols;
};

// Primal weight after a restart whose primal and dual moves since the last
// restart point were dx and dy (both > 0).  PDLP's smoothed update, w <-
// exp(log(dy/dx)/2 + log(w)/2), is a proportional controller on the
// log-balance error e = log w - log(dy/dx) with gain 1/2; this is its
// proportional-integral form, log w <- log w - kp e - ki sum(e), with
// kp = 1/2, ki = 0 by default (PID control of the primal weight: Lu, Peng
// and Yang, "cuPDLPx", 2025).
__device__ __forceinline__ double primal_weight_update(DevState &S, double dx, double dy) {
    const double target = log(dy / dx);
    S.pw_int += log(S.omega) - target;
    return exp(S.pw_kp * target + (1.0 - S.pw_kp) * log(S.omega) - S.pw_ki * S.pw_int);
}

// The recentred FP32 solver's FP64 base point (x = xb + correction, y = yb +
// correction) and the FP64 scaled LP the shifted one is rounded from.
template <class T>
struct RecentreBufs {
    double *xb, *yb;                 // base point (n, m)
    const double *c, *l, *u;         // scaled LP in FP64 (n)
    const double *rl, *ru;           // (m)
    double *ax, *aty;                // A xb, A' yb (FP64 sums)
    T *cs, *ls, *us, *rls, *rus;     // the shifted LP (aliases of Bufs c, l, u, rl, ru)
    // A quadratic objective (null for an LP): Q xb (FP64 sums) and the
    // objective's gradient at the base, c + Q xb (alias of Bufs cobj).
    double *qxb;
    T *cobj;
};

// The dual sign rule of the Moreau step (y+ = 0 inside [rl, ru], y+ <= 0 at
// rl, y+ >= 0 at ru): a total formed as ybase + v keeps it up to the
// rounding of v, which this removes (a wrong-signed dual on a one-sided row
// would make the dual objective -inf).
__device__ __forceinline__ double dual_sign_project(double y, double rl, double ru) {
    if (y > 0.0 && isinf(ru)) return 0.0;
    if (y < 0.0 && isinf(rl)) return 0.0;
    return y;
}

__device__ __forceinline__ double kkt_weighted(double w, double pr2, double dr2, double gap) {
    return sqrt(w * w * pr2 + dr2 / (w * w) + gap * gap);
}

// ------------------------------------------------------------------ epilogues

// Dual step on row i given acc = (A x+)_i.  Stream: non-temporal loads and
// stores for every vector but y, which the next kernel gathers.
template <class T, bool Stream = false>
struct DualEpi {
    const Bufs<T> &b;
    double sigma, inv_sigma;
    bool restart;
    bool sign_exact;
    struct Pre {
        T ax, y, lo, hi, unscale, ylast;
        double ysum;
    };
    __device__ Pre prefetch(int i) const {
        Pre p;
        p.ax = vec_load<Stream>(b.ax + i);
        p.y = vec_load<Stream>(b.y + i);
        p.lo = vec_load<Stream>(b.rl + i);
        p.hi = vec_load<Stream>(b.ru + i);
        p.unscale = vec_load<Stream>(b.row_unscale + i);
        p.ysum = b.ysum ? vec_load<Stream>(b.ysum + i) : 0.0;
        p.ylast = b.ylast ? vec_load<Stream>(b.ylast + i) : T(0);
        return p;
    }
    __device__ void row(int i, double acc, const Pre &q, Partials &p) const {
        const double axn = acc;
        const double ax_old = static_cast<double>(q.ax);
        const double axbar = restart ? axn : 2.0 * axn - ax_old;
        const double lo = static_cast<double>(q.lo), hi = static_cast<double>(q.hi);
        const double v = static_cast<double>(q.y) + sigma * axbar;
        double yn;
        if (sign_exact) {
            // Same Moreau step, written so that y+ is exactly 0 inside the
            // box and carries the sign of the active bound outside it (a
            // one-sided row never gets a dual of the wrong sign from
            // round-off, which would make the dual objective -inf).
            const double t = inv_sigma * v;
            yn = (t < lo) ? fmin(v - sigma * lo, 0.0) : ((hi < t) ? fmax(v - sigma * hi, 0.0) : 0.0);
        } else {
            const double proj = clamp_box(inv_sigma * v, lo, hi);   // pdhg.cpp:108-112
            yn = v - sigma * proj;
        }
        const T y_stored = static_cast<T>(yn);
        b.y[i] = y_stored;
        vec_store<Stream>(b.ax + i, static_cast<T>(axn));
        if (b.ysum) vec_store<Stream>(b.ysum + i, q.ysum + static_cast<double>(y_stored));
        const double d = clamp_box(axn, lo, hi) - axn;
        p.v[0] = dmax(p.v[0], fabs(d));
        const double du = d * static_cast<double>(q.unscale);
        p.v[1] += du * du;
        p.v[2] += d * d;
        if (yn > 0.0) {
            if (isinf(hi)) p.v[4] += 1.0; else { p.v[3] -= yn * hi; p.v[7] += fabs(yn * hi); }
        } else if (yn < 0.0) {
            if (isinf(lo)) p.v[4] += 1.0; else { p.v[3] -= yn * lo; p.v[7] += fabs(yn * lo); }
        }
        if (b.ylast) {
            const double dy = yn - static_cast<double>(q.ylast);
            p.v[5] += dy * dy;
        }
    }
};

// Restarted Halpern PDHG with reflection (Lu and Yang, "Restarted Halpern
// PDHG for linear programming", 2024): with T the PDHG map and z0 the point
// of the last restart, z_{k+1} = w1 (2 T(z_k) - z_k) + w0 z0 where, for the
// ks-th iteration since the restart, w1 = (ks+1)/(ks+2) and w0 = 1/(ks+2).
// The SpMVs run on T(z_k) = (x+, y+); A x_{k+1} and A' y_{k+1} follow by
// linearity from A x+, A x_k and A x0 (and likewise for A'), so an iteration
// still costs two SpMVs.  The KKT sums and the termination test are taken
// on T(z_k), the point returned.
struct HalpernWeights {
    double w1, w0;
};
__device__ __forceinline__ HalpernWeights halpern_weights(long long ks) {
    const double d = 1.0 / static_cast<double>(ks + 2);
    return HalpernWeights{static_cast<double>(ks + 1) * d, d};
}

// Whether iteration k of a Halpern or adaptive solve ends at a restart check.
__device__ __forceinline__ bool check_iteration(const DevState &s) { return (s.k + 1) % s.period == 0; }

// Halpern dual half on row i given acc = (A x+)_i: y+ = the dual step from
// (y_k, A x_k) with A xbar = 2 A x+ - A x_k; then the Halpern update of y
// and of A x.  Row slots as DualEpi on (A x+, y+), plus slot 6
// sum (y+ - y_k)^2 and, at checks, slot 5 sum (y+ - y_restart)^2.
// R: the recentred solver's form (Bufs::ybase; FP32 storage only).
template <class T, bool Stream = false, bool R = false>
struct HalpernDualEpi {
    const Bufs<T> &b;
    double sigma, inv_sigma;
    HalpernWeights w;
    bool check;
    struct Pre {
        T ax, yh, lo, hi, unscale, y0, ax0, ylast;
        double yb;
    };
    __device__ Pre prefetch(int i) const {
        Pre p;
        p.ax = vec_load<Stream>(b.ax + i);
        p.yh = vec_load<Stream>(b.yh + i);
        p.lo = vec_load<Stream>(b.rl + i);
        p.hi = vec_load<Stream>(b.ru + i);
        p.unscale = vec_load<Stream>(b.row_unscale + i);
        p.y0 = vec_load<Stream>(b.y0h + i);
        p.ax0 = vec_load<Stream>(b.ax0h + i);
        p.ylast = check ? vec_load<Stream>(b.ylast + i) : T(0);
        if constexpr (R) p.yb = vec_load<Stream>(b.ybase + i);
        else p.yb = 0.0;
        return p;
    }
    __device__ void row(int i, double acc, const Pre &q, Partials &p) const {
        const double axn = acc;
        const double ax_old = static_cast<double>(q.ax), y_old = static_cast<double>(q.yh);
        const double lo = static_cast<double>(q.lo), hi = static_cast<double>(q.hi);
        const double v = y_old + sigma * (2.0 * axn - ax_old);
        // Recentred: the step acts on the total dual ybase + v and the stored
        // y+ is its correction, both formed in FP64, so that only the
        // correction is rounded to T.  yt is the exact (sign-exact) total.
        const double vt = R ? v + q.yb : v;
        const double t = inv_sigma * vt;   // sign-exact Moreau step, as for the adaptive policy
        const double yt = (t < lo) ? fmin(vt - sigma * lo, 0.0) : ((hi < t) ? fmax(vt - sigma * hi, 0.0) : 0.0);
        const double yn = R ? yt - q.yb : yt;
        const T y_stored = static_cast<T>(yn);
        b.y[i] = y_stored;
        const double yp = static_cast<double>(y_stored);
        vec_store<Stream>(b.yh + i, static_cast<T>(w.w1 * (2.0 * yp - y_old) + w.w0 * static_cast<double>(q.y0)));
        vec_store<Stream>(b.ax + i, static_cast<T>(w.w1 * (2.0 * axn - ax_old) + w.w0 * static_cast<double>(q.ax0)));
        const double d = clamp_box(axn, lo, hi) - axn;
        p.v[0] = dmax(p.v[0], fabs(d));
        const double du = d * static_cast<double>(q.unscale);
        p.v[1] += du * du;
        p.v[2] += d * d;
        if (yt > 0.0) {
            if (isinf(hi)) p.v[4] += 1.0; else { p.v[3] -= yt * hi; p.v[7] += fabs(yt * hi); }
        } else if (yt < 0.0) {
            if (isinf(lo)) p.v[4] += 1.0; else { p.v[3] -= yt * lo; p.v[7] += fabs(yt * lo); }
        }
        const double dz = yp - y_old;
        p.v[6] += dz * dz;
        if (check) {
            vec_store<Stream>(b.ax_avg + i, static_cast<T>(axn));
            const double dy = yp - static_cast<double>(q.ylast);
            p.v[5] += dy * dy;
        }
    }
};

// The PDQP primal step before projection, x - tau (A'y + (c + Q x))
// (pdqp.cpp:115-122, the order of pstep_element's quadratic form).
__device__ __forceinline__ double qp_step(double x, double c, double qx, double aty, double tau) {
    return -tau * (aty + (c + qx)) + x;
}

// Primal step of column j under the Halpern policy with a quadratic
// objective: 1/tau_j = 1/tau + sum_k |q_jk|, tau = eta/w the LP step, sigma =
// eta w unchanged.  The linearised PDHG step is then a proximal point step in
// the metric [T^-1 - Q, -A'; -A, I/sigma] with T = diag(tau_j), positive
// definite because T^-1 - Q - sigma A'A = (diag(sum_k |q_jk|) - Q) + (I/tau
// - sigma A'A): a diagonally dominant matrix with a non-negative diagonal
// (PSD) plus (1 - eta^2 ||A||^2)/tau I (eta ||A|| = 0.99).  So the map is
// firmly nonexpansive in that metric and the reflected Halpern step stays
// nonexpansive.  Per column, unlike a scalar 1/(1/tau + ||Q||), the step
// follows the curvature each column has after the column scaling.
__device__ __forceinline__ double qp_column_step(double tau, double qabs) {
    return 1.0 / (1.0 / tau + qabs);
}

// Column terms of the KKT sums for (x_j, g_j = c_j + aty_j) (shared by the
// per-iteration and the average evaluation).  s = col_scale[j].
__device__ __forceinline__ void column_terms(double xj, double cj, double g, double lj, double uj,
                                             double s, bool have_last, double xlast, Partials &p) {
    const double d = xj - clamp_box(xj - g, lj, uj);
    p.v[0] = dmax(p.v[0], fabs(d));
    p.v[2] += d * d;
    const double dorig = s * (xj - clamp_box(xj - g / (s * s), lj, uj));
    p.v[1] += dorig * dorig;
    p.v[5] += cj * xj;
    if (g > 0.0) {
        if (isinf(lj)) p.v[4] += 1.0; else p.v[3] += g * lj;
    } else if (g < 0.0) {
        if (isinf(uj)) p.v[4] += 1.0; else p.v[3] += g * uj;
    }
    if (have_last) {
        const double dx = xj - xlast;
        p.v[6] += dx * dx;
    }
}

// A' y on column j, the residual terms, and (fused) the next primal step.
// Stream as for DualEpi; x+ (gathered next) keeps ordinary stores.
template <class T, bool Stream = false>
struct PrimalEpi {
    const Bufs<T> &b;
    const T *xcur;
    T *xnext;
    double tau;
    bool fuse;
    bool store_aty = true;   // false when aty already holds acc
    struct Pre {
        T x, c, l, u, s, xlast;
        double xsum, qx;
    };
    __device__ Pre prefetch(int j) const {
        Pre p;
        p.qx = b.qx ? vec_load<Stream>(b.qx + j) : 0.0;
        p.x = vec_load<Stream>(xcur + j);
        p.c = vec_load<Stream>(b.c + j);
        p.l = vec_load<Stream>(b.l + j);
        p.u = vec_load<Stream>(b.u + j);
        p.s = vec_load<Stream>(b.col_scale + j);
        p.xsum = (fuse && b.xsum) ? vec_load<Stream>(b.xsum + j) : 0.0;
        p.xlast = b.xlast ? vec_load<Stream>(b.xlast + j) : T(0);
        return p;
    }
    __device__ void row(int j, double acc, const Pre &q, Partials &p) const {
        if (store_aty) vec_store<Stream>(b.aty + j, static_cast<T>(acc));
        const double xj = static_cast<double>(q.x);
        const double cj = static_cast<double>(q.c);
        const double lj = static_cast<double>(q.l), uj = static_cast<double>(q.u);
        // pdhg.cpp:122-125 g = c + A'y; pdqp.cpp:148-152 g = A'y + (c + Q x).
        const double g = b.qx ? acc + (cj + q.qx) : cj + acc;
        column_terms(xj, cj, g, lj, uj, static_cast<double>(q.s), b.xlast != nullptr,
                     static_cast<double>(q.xlast), p);
        if (fuse) {
            // pdhg.cpp:93-96 (x - tau c) - tau A'y;  pdqp.cpp:120-122 x - tau g.
            const T xn = static_cast<T>(
                clamp_box(b.qx ? -tau * g + xj : (xj - tau * cj) - tau * acc, lj, uj));
            xnext[j] = xn;
            if (b.xsum) vec_store<Stream>(b.xsum + j, q.xsum + static_cast<double>(xn));
        }
    }
};

// Halpern primal half on column j given acc = (A' y+)_j: the KKT column
// terms of (x+, c + A' y+) (slot 6 at checks: sum (x+ - x_restart)^2), slot 7
// sum (x+ - x_k)^2, the Halpern update of x and of A' y, and (fused) the next
// PDHG primal step x+ = P_X((x_{k+1} - tau c) - tau A' y_{k+1}).
// R: the recentred solver's form (Bufs::cobj; FP32 storage only).  QP: a
// quadratic objective (Bufs::qx = Q x+ from k_qx; gradient A'y + (c + Q x),
// as pdqp.cpp:148-152; objectives with the 1/2 x'Qx terms; Q x of the
// Halpern iterate carried like A'y; FP64 only).
template <class T, bool Stream = false, bool R = false, bool QP = false>
struct HalpernPrimalEpi {
    const Bufs<T> &b;
    const T *xcur;   // x+ of this iteration
    T *xnext;        // x+ of the next one (fused step)
    double tau;
    bool fuse, check;
    HalpernWeights w;
    struct Pre {
        T xp, xh, aty, c, l, u, s, x0, aty0, xlast, cobj;
        double qx, qxh, qx0, qabs;
    };
    __device__ Pre prefetch(int j) const {
        Pre p;
        p.xp = vec_load<Stream>(xcur + j);
        p.xh = vec_load<Stream>(b.xh + j);
        p.aty = vec_load<Stream>(b.aty + j);
        p.c = vec_load<Stream>(b.c + j);
        p.l = vec_load<Stream>(b.l + j);
        p.u = vec_load<Stream>(b.u + j);
        p.s = vec_load<Stream>(b.col_scale + j);
        p.x0 = vec_load<Stream>(b.x0h + j);
        p.aty0 = vec_load<Stream>(b.aty0h + j);
        p.xlast = check ? vec_load<Stream>(b.xlast + j) : T(0);
        if constexpr (R) p.cobj = vec_load<Stream>(b.cobj + j);
        else p.cobj = T(0);
        if constexpr (QP) {
            p.qx = vec_load<Stream>(b.qx + j);
            p.qxh = vec_load<Stream>(b.qxh + j);
            p.qx0 = vec_load<Stream>(b.qx0h + j);
            p.qabs = vec_load<Stream>(b.qabs + j);
        } else {
            p.qx = p.qxh = p.qx0 = p.qabs = 0.0;
        }
        return p;
    }
    __device__ void row(int j, double acc, const Pre &q, Partials &p) const {
        const double xp = static_cast<double>(q.xp), x_old = static_cast<double>(q.xh);
        const double cj = static_cast<double>(q.c);
        const double lj = static_cast<double>(q.l), uj = static_cast<double>(q.u);
        // Recentred: c holds c + A'ybase, so the total gradient is c + acc
        // as before, while the primal objective terms take c itself.
        const double cobj = R ? static_cast<double>(q.cobj) : cj;
        if constexpr (QP) {
            column_terms(xp, cobj, acc + (cj + q.qx), lj, uj, static_cast<double>(q.s), check,
                         static_cast<double>(q.xlast), p);
            // 1/2 x'Qx: added to the primal objective and taken from the
            // (Wolfe) dual objective -1/2 x'Qx + min_X g'x - ...
            const double half = 0.5 * xp * q.qx;
            p.v[5] += half;
            p.v[3] -= half;
        } else {
            column_terms(xp, cobj, cj + acc, lj, uj, static_cast<double>(q.s), check, static_cast<double>(q.xlast), p);
        }
        const double dz = xp - x_old;
        p.v[7] += dz * dz;
        const T xh_new = static_cast<T>(w.w1 * (2.0 * xp - x_old) + w.w0 * static_cast<double>(q.x0));
        const T aty_new =
            static_cast<T>(w.w1 * (2.0 * acc - static_cast<double>(q.aty)) + w.w0 * static_cast<double>(q.aty0));
        vec_store<Stream>(b.xh + j, xh_new);
        vec_store<Stream>(b.aty + j, aty_new);
        if (check) vec_store<Stream>(b.aty_avg + j, static_cast<T>(acc));
        if constexpr (QP) {
            const double qxh_new = w.w1 * (2.0 * q.qx - q.qxh) + w.w0 * q.qx0;
            vec_store<Stream>(b.qxh + j, qxh_new);
            if (check) vec_store<Stream>(b.qx_avg + j, q.qx);
            if (fuse)
                xnext[j] = static_cast<T>(clamp_box(qp_step(static_cast<double>(xh_new), cj, qxh_new,
                                                            static_cast<double>(aty_new),
                                                            qp_column_step(tau, q.qabs)),
                                                    lj, uj));
        } else if (fuse) {
            xnext[j] = static_cast<T>(clamp_box((static_cast<double>(xh_new) - tau * cj) -
                                                    tau * static_cast<double>(aty_new),
                                                lj, uj));
        }
    }
};

// Average evaluation, rows: acc = (A xsum)_i, scaled by 1/ns.
template <class T>
struct EvalRowsEpi {
    const Bufs<T> &b;
    double inv_ns;
    struct Pre {
        double ysum;
        T lo, hi, unscale, ylast;
    };
    __device__ Pre prefetch(int i) const {
        return Pre{b.ysum[i], b.rl[i], b.ru[i], b.row_unscale[i], b.ylast[i]};
    }
    __device__ void row(int i, double acc, const Pre &q, Partials &p) const {
        const double axa = acc * inv_ns;
        b.ax_avg[i] = static_cast<T>(axa);
        const double ya = q.ysum * inv_ns;
        const double lo = static_cast<double>(q.lo), hi = static_cast<double>(q.hi);
        const double d = clamp_box(axa, lo, hi) - axa;
        p.v[0] = dmax(p.v[0], fabs(d));
        const double du = d * static_cast<double>(q.unscale);
        p.v[1] += du * du;
        p.v[2] += d * d;
        if (ya > 0.0) {
            if (isinf(hi)) p.v[4] += 1.0; else p.v[3] -= ya * hi;
        } else if (ya < 0.0) {
            if (isinf(lo)) p.v[4] += 1.0; else p.v[3] -= ya * lo;
        }
        const double dy = ya - static_cast<double>(q.ylast);
        p.v[5] += dy * dy;
    }
};

template <class T>
struct EvalColsEpi {
    const Bufs<T> &b;
    double inv_ns;
    struct Pre {
        double xsum;
        T c, l, u, s, xlast;
    };
    __device__ Pre prefetch(int j) const {
        return Pre{b.xsum[j], b.c[j], b.l[j], b.u[j], b.col_scale[j], b.xlast[j]};
    }
    __device__ void row(int j, double acc, const Pre &q, Partials &p) const {
        const double atya = acc * inv_ns;
        b.aty_avg[j] = static_cast<T>(atya);
        const double xa = q.xsum * inv_ns;
        const double cj = static_cast<double>(q.c);
        column_terms(xa, cj, cj + atya, static_cast<double>(q.l), static_cast<double>(q.u),
                     static_cast<double>(q.s), true, static_cast<double>(q.xlast), p);
    }
};

// ------------------------------------------------------------------ device helpers

__device__ __forceinline__ void relative_kkt(const DevState &s, const double *r, const double *c,
                                             double &rel_p, double &rel_d, double &rel_g,
                                             double &gap_abs, double &pobj, double &dobj) {
    rel_p = sqrt(r[1]) / (1.0 + s.norm_b);
    rel_d = sqrt(c[1]) / (1.0 + s.norm_c);
    // Both objectives of an LP shifted by a base point x_b drop c'x_b (the
    // dual one by shifting the bounds it takes), so the gap is taken on the
    // shifted sums and cx_base added back to each objective.
    pobj = s.cx_base + c[5];
    const bool neg_inf = (r[4] + c[4]) > 0.0;
    dobj = neg_inf ? -INFINITY : s.cx_base + (r[3] + c[3]);
    if (neg_inf) {
        rel_g = 1.0;
        gap_abs = INFINITY;
    } else {
        gap_abs = fabs(c[5] - (r[3] + c[3]));
        rel_g = gap_abs / (1.0 + fabs(pobj) + fabs(dobj));
    }
}

// Whether the primal kernel of iteration k may fuse the next primal step
// (tau is known to stay the same).
__device__ __forceinline__ bool may_fuse(const DevState &s) {
    if (s.policy == VK_PDHG_POLICY_REFERENCE) return !(s.k > 0 && s.k % kDecisionPeriod == 0);
    return (s.k + 1) % s.period != 0;
}

// Primal infeasibility test (Farkas).  For any y, with g = A'y and
// D0(y) = sum_j min(g_j l_j, g_j u_j) - sum_i (max(y_i,0) ru_i -
// max(-y_i,0) rl_i), a feasible LP has D0(y) <= 0 (take a feasible x:
// min_x g'x <= y'Ax <= max_t y't), so D0(y) > 0 with every term finite proves
// it infeasible.  On an infeasible LP the PDHG dual iterate grows along such
// a ray.  The finalizer records the row part of the last dual iterate
// (DevState::d0_rows*); farkas_column adds the column part of g = A'y+
// (partial slots 1 sum of terms, 2 count of -inf terms, 3 sum of term
// magnitudes); infeasibility_rule decides with a relative margin far above
// the rounding of FP32 or FP64 terms.  Scaling leaves D0 unchanged.
constexpr double kInfeasMargin = 1e-6;

__device__ __forceinline__ void farkas_column(double gj, double lj, double uj, Partials &p) {
    if (gj == 0.0) return;
    const double bound = gj > 0.0 ? lj : uj;
    if (isinf(bound)) {
        p.v[2] += 1.0;
    } else {
        p.v[1] += gj * bound;
        p.v[3] += fabs(gj * bound);
    }
}

__device__ inline void infeasibility_rule(DevState &S, const double *red) {
    if (S.status != VK_PDHG_RUNNING) return;
    const double d0 = S.d0_rows + red[1];
    const double scale = S.d0_rows_abs + red[3];
    if (S.d0_rows_ninf == 0.0 && red[2] == 0.0 && d0 > 0.0 && d0 > kInfeasMargin * scale)
        S.status = VK_PDHG_PRIMAL_INFEASIBLE;
}

// Halpern restart rule at the end of iteration S.k, from the reduced slots
// (before S.k and S.ks advance).  The fixed-point residual
// r = |z_k - T(z_k)| in the primal-weight norm (w |dx|^2 + |dy|^2 / w)^(1/2)
// is taken every iteration; the first one after a restart is the reference
// r0.  At a check: restart on r <= 0.2 r0, on r <= 0.8 r0 with r above the
// previous check's, or once the iterations since the restart reach 0.36 of
// all iterations; the primal weight then moves halfway (in log) to
// |y+ - y_r| / |x+ - x_r| since the last restart point, and the restart
// (k_halpern_apply) sets z = z0 = T(z_k).
__device__ inline void halpern_rule(DevState &S, const double *red_rows, const double *red_cols, bool converged) {
    const double r = sqrt(S.omega * red_cols[7] + red_rows[6] / S.omega);
    if (S.ks == 0) S.kkt_ref = r;
    S.restart_apply = 0;
    if (converged || (S.k + 1) % S.period != 0) return;
    const bool restart = r <= 0.2 * S.kkt_ref || (r <= 0.8 * S.kkt_ref && r > S.kkt_prev) ||
                         static_cast<double>(S.ks) >= 0.36 * static_cast<double>(S.k + 1);
    S.kkt_prev = r;
    if (!restart) return;
    const double dx = sqrt(red_cols[6]), dy = sqrt(red_rows[5]);
    if (dx > 1e-10 && dy > 1e-10) S.omega = primal_weight_update(S, dx, dy);
    S.tau = S.eta / S.omega;
    S.sigma = S.eta * S.omega;
    S.inv_sigma = 1.0 / S.sigma;
    S.kkt_prev = INFINITY;
    S.restart_apply = 1;
    S.restarts += 1;
}

// Step / restart / termination rule after one iteration, from the reduced
// dual (red_rows) and primal (red_cols) partial slots (pdhg.cpp:131-150 for
// the reference policy).  Runs in one thread.
__device__ inline void finalize_step(DevState &S, const double *red_rows, const double *red_cols) {
    const long long k = S.k;
    const double rp = red_rows[0], rd = red_cols[0];
    const double metric = dmax(rp, rd);   // pdhg.cpp:131
    double rel_p, rel_d, rel_g, gap_abs, pobj, dobj;
    relative_kkt(S, red_rows, red_cols, rel_p, rel_d, rel_g, gap_abs, pobj, dobj);
    const double rel_kkt = dmax(rel_p, dmax(rel_d, rel_g));
    S.rp = rp;
    S.rd = rd;
    S.metric = metric;
    S.rel_p = rel_p;
    S.rel_d = rel_d;
    S.rel_g = rel_g;
    S.rel_kkt = rel_kkt;
    S.pobj = pobj;
    S.dobj = dobj;
    S.cur_pr2 = red_rows[2];
    S.cur_dr2 = red_cols[2];
    S.cur_gap = gap_abs;
    S.cur_dx2 = red_cols[6];
    S.cur_dy2 = red_rows[5];
    S.d0_rows = red_rows[3];
    S.d0_rows_abs = red_rows[7];
    S.d0_rows_ninf = red_rows[4];
    if (S.policy == VK_PDHG_POLICY_REFERENCE) {
        // pdhg.cpp:133-145; the restart flag was consumed by this iteration's
        // dual step (pdhg.cpp:98-100).
        S.restart_flag = 0;
        if (k > 0 && k % kDecisionPeriod == 0) {
            if (metric > S.prev_metric * 1.25) {
                S.tau *= 0.5;
                S.sigma *= 0.5;
                S.restart_flag = 1;
            } else if (metric < S.prev_metric * 0.1) {
                const double grown = S.tau * 1.05;
                S.tau = (S.max_step < grown) ? S.max_step : grown;
                S.sigma = S.tau;
            }
            S.prev_metric = metric;
        } else if (k == 0) {
            S.prev_metric = metric;
        }
        S.inv_sigma = 1.0 / S.sigma;
    }
    const bool converged = (S.rel_tol > 0.0) ? (rel_kkt <= S.rel_tol) : (metric <= S.metric_tol);  // pdhg.cpp:147
    if (S.policy == VK_PDHG_POLICY_HALPERN) halpern_rule(S, red_rows, red_cols, converged);
    S.k = k + 1;
    S.ks = (S.policy == VK_PDHG_POLICY_HALPERN && S.restart_apply) ? 0 : S.ks + 1;
    S.ns += 1;
    if (converged) S.status = VK_PDHG_CONVERGED;
    else if (S.k >= S.limit) S.status = VK_PDHG_ITERATION_LIMIT;
}

// Adaptive restart rule on the evaluated average (er/ec: reduced partial
// slots of the average's rows and columns).  `init` records the start
// point's KKT error as the first restart reference.  Runs in one thread.
__device__ inline void check_step(DevState &S, const double *red_rows, const double *red_cols, bool init) {
    double rel_p, rel_d, rel_g, gap_abs, pobj, dobj;
    relative_kkt(S, red_rows, red_cols, rel_p, rel_d, rel_g, gap_abs, pobj, dobj);
    const double w = S.omega;
    const double kkt_avg = kkt_weighted(w, red_rows[2], red_cols[2], gap_abs);
    if (init) {
        S.kkt_ref = kkt_avg;
        S.kkt_prev = INFINITY;
        return;
    }
    const double rel_avg = dmax(rel_p, dmax(rel_d, rel_g));
    S.ns_avg = S.ns;
    S.restart_apply = 0;
    S.restart_to_avg = 0;
    if (S.rel_tol > 0.0 && rel_avg <= S.rel_tol) {
        S.status = VK_PDHG_CONVERGED;
        S.final_from_avg = 1;
        S.rel_p = rel_p;
        S.rel_d = rel_d;
        S.rel_g = rel_g;
        S.rel_kkt = rel_avg;
        S.pobj = pobj;
        S.dobj = dobj;
        return;
    }
    const double kkt_cur = kkt_weighted(w, S.cur_pr2, S.cur_dr2, S.cur_gap);
    const bool use_avg = kkt_avg < kkt_cur;
    const double kc = use_avg ? kkt_avg : kkt_cur;
    const bool restart = kc <= 0.2 * S.kkt_ref ||
                         (kc <= 0.8 * S.kkt_ref && kc > S.kkt_prev) ||
                         static_cast<double>(S.ks) >= 0.36 * static_cast<double>(S.k);
    S.kkt_prev = kc;
    if (!restart) return;
    const double dx = sqrt(use_avg ? red_cols[6] : S.cur_dx2);
    const double dy = sqrt(use_avg ? red_rows[5] : S.cur_dy2);
    if (dx > 1e-10 && dy > 1e-10) S.omega = primal_weight_update(S, dx, dy);
    S.tau = S.eta / S.omega;
    S.sigma = S.eta * S.omega;
    S.inv_sigma = 1.0 / S.sigma;
    S.kkt_ref = use_avg ? kkt_weighted(S.omega, red_rows[2], red_cols[2], gap_abs)
                        : kkt_weighted(S.omega, S.cur_pr2, S.cur_dr2, S.cur_gap);
    S.kkt_prev = INFINITY;
    S.ks = 0;
    S.ns = 0;
    S.restart_apply = 1;
    S.restart_to_avg = use_avg ? 1 : 0;
    S.restarts += 1;
}

// The post-check update, split by index kind: move to the average if chosen
// (or converged on it), record the restart point, clear the sums, and take
// the next primal step (columns) / refresh y and A x (rows).
template <class T>
__device__ __forceinline__ void apply_column(const Bufs<T> &b, const DevState &s, int j, T *xcur, T *xnext) {
    const bool copy_avg = s.final_from_avg || (s.restart_apply && s.restart_to_avg);
    const double inv = 1.0 / static_cast<double>(s.ns_avg > 0 ? s.ns_avg : 1);
    if (copy_avg) {
        xcur[j] = static_cast<T>(b.xsum[j] * inv);
        b.aty[j] = b.aty_avg[j];
    }
    if (s.restart_apply) {
        b.xlast[j] = xcur[j];
        b.xsum[j] = 0.0;
    }
    if (s.status == VK_PDHG_RUNNING) {
        const double xj = static_cast<double>(xcur[j]);
        const T xn = static_cast<T>(clamp_box((xj - s.tau * static_cast<double>(b.c[j])) -
                                                  s.tau * static_cast<double>(b.aty[j]),
                                              static_cast<double>(b.l[j]), static_cast<double>(b.u[j])));
        xnext[j] = xn;
        b.xsum[j] += static_cast<double>(xn);
    }
}

template <class T>
__device__ __forceinline__ void apply_row(const Bufs<T> &b, const DevState &s, int i) {
    const bool copy_avg = s.final_from_avg || (s.restart_apply && s.restart_to_avg);
    const double inv = 1.0 / static_cast<double>(s.ns_avg > 0 ? s.ns_avg : 1);
    if (copy_avg) {
        b.y[i] = static_cast<T>(b.ysum[i] * inv);
        b.ax[i] = b.ax_avg[i];
    }
    if (s.restart_apply) {
        b.ylast[i] = b.y[i];
        b.ysum[i] = 0.0;
    }
}

// Halpern policy after a restart check (the primal pass did not fuse the
// next step): on a restart z = z0 = T(z_k), i.e. x+ (xp) and the A' y+ kept
// in aty_avg, which also become the primal weight's restart point; then the
// next PDHG primal step x+ = P_X((xh - tau c) - tau A' yh) into xnext.
// QP: a quadratic objective (Bufs::qxh; FP64), whose Q x of the Halpern
// iterate and anchor restart at the Q x+ kept in qx_avg.
template <class T, bool QP = false>
__device__ __forceinline__ void halpern_apply_column(const Bufs<T> &b, const DevState &s, int j, const T *xp,
                                                     T *xnext) {
    T xh = b.xh[j], aty = b.aty[j];
    if (s.restart_apply) {
        xh = xp[j];
        aty = b.aty_avg[j];
        b.xh[j] = xh;
        b.aty[j] = aty;
        b.x0h[j] = xh;
        b.aty0h[j] = aty;
        b.xlast[j] = xh;
    }
    if constexpr (QP) {
        double qxh = b.qxh[j];
        if (s.restart_apply) {
            qxh = b.qx_avg[j];
            b.qxh[j] = qxh;
            b.qx0h[j] = qxh;
        }
        xnext[j] = static_cast<T>(clamp_box(qp_step(static_cast<double>(xh), static_cast<double>(b.c[j]), qxh,
                                                    static_cast<double>(aty), qp_column_step(s.tau, b.qabs[j])),
                                            static_cast<double>(b.l[j]), static_cast<double>(b.u[j])));
        return;
    }
    xnext[j] = static_cast<T>(clamp_box((static_cast<double>(xh) - s.tau * static_cast<double>(b.c[j])) -
                                            s.tau * static_cast<double>(aty),
                                        static_cast<double>(b.l[j]), static_cast<double>(b.u[j])));
}

// Row half of the restart: y = y0 = y+, A x = A x0 = the A x+ kept in ax_avg.
template <class T>
__device__ __forceinline__ void halpern_apply_row(const Bufs<T> &b, const DevState &s, int i) {
    if (!s.restart_apply) return;
    const T yh = b.y[i], ax = b.ax_avg[i];
    b.yh[i] = yh;
    b.ax[i] = ax;
    b.y0h[i] = yh;
    b.ax0h[i] = ax;
    b.ylast[i] = yh;
}

// One element of the standalone primal step x[k+1] = P_X(x[k] - tau (c + aty))
// (with Q: x[k] - tau (aty + (c + Q x[k])), pdqp.cpp:115-122; HalpernQP: the
// Halpern policy's column step qp_column_step).
template <class T, bool HalpernQP = false>
__device__ __forceinline__ void pstep_element(const Bufs<T> &b, const DevState &s, int j, const T *xcur, T *xnext) {
    const double xj = static_cast<double>(xcur[j]);
    const double cj = static_cast<double>(b.c[j]), aj = static_cast<double>(b.aty[j]);
    double z;
    if constexpr (HalpernQP) z = qp_step(xj, cj, b.qx[j], aj, qp_column_step(s.tau, b.qabs[j]));
    else z = b.qx ? -s.tau * (aj + (cj + b.qx[j])) + xj : (xj - s.tau * cj) - s.tau * aj;
    const T xn = static_cast<T>(clamp_box(z, static_cast<double>(b.l[j]), static_cast<double>(b.u[j])));
    xnext[j] = xn;
    if (b.xsum) b.xsum[j] += static_cast<double>(xn);
}

}  // namespace vk
