/*
 * dflock.c - TEST INFRASTRUCTURE ONLY: the FP64 CPU statement of the
 * differentiable (smooth) flocking model and of its adjoint, the reference
 * gradient for the GPU adjoint kernels.  No reference counterpart: the
 * reference has no differentiable path; the model is the smooth variant of
 * the flocking rule (m_flocking/agent.c:2-33: neighbours within a radius,
 * separation, alignment, cohesion) with every discontinuity removed.
 *
 * Model, per step, for every agent i and every other agent j with
 * r2 = |p_j - p_i|^2 < R^2 (d = p_j - p_i, u = r2 / R^2, w = (1 - u)^3,
 * q = r2 + h2):
 *     W_i += w,  A_i += w (v_j - v_i),  C_i += w d,  S_i -= w d / q
 *     a_i  = (ws S_i + wa A_i + wc C_i) / (W_i + 1)
 *     v_i' = v_i + dt (a_i - damping v_i),   p_i' = p_i + dt v_i'
 * w and its first two derivatives vanish at r = R, so the step is C^2 in
 * the positions and pairs entering or leaving the radius change nothing
 * abruptly.
 *
 * Loss: L = sum_i |p_T,i - target_i|^2 / (3 n).  The gradient is reverse
 * mode through every step with all states kept (no checkpointing); per pair
 * (i seer, j seen), with La the adjoint of a_i and D = W_i + 1:
 *     LS = ws La / D,  LA = wa La / D,  LC = wc La / D,  LW = -(La . a) / D
 *     Lw = LW + LA . (v_j - v_i) + LC . d - LS . d / q
 *     Ld = 2 w'(r2) Lw d + w LC - w LS / q + 2 w (d . LS) d / q^2
 *     Lp_j += Ld,  Lp_i -= Ld,  Lv_j += w LA,  Lv_i -= w LA
 * Pinned by central finite differences in FP64 and by the closed form of the
 * weight-free motion (tests/test_dflock_oracle.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "dflock.h"

/* Uniform grid of cell size R over the bounding box: agents of a cell are
 * order[start[c] .. start[c + 1]). */
typedef struct Grid {
    double lo[3];
    double cs;
    int count[3];
    unsigned *start;
    unsigned *order;
    int *cell;
} Grid;

static int axis_cell(double p, double lo, double cs, int count) {
    int c = (int)floor((p - lo) / cs);
    return c < 0 ? 0 : (c > count - 1 ? count - 1 : c);
}

static int grid_build(Grid *g, unsigned n, const double *p, double radius) {
    double hi[3];
    for (int k = 0; k < 3; ++k) {
        g->lo[k] = INFINITY;
        hi[k] = -INFINITY;
    }
    for (unsigned i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            g->lo[k] = fmin(g->lo[k], p[3 * i + k]);
            hi[k] = fmax(hi[k], p[3 * i + k]);
        }
    g->cs = radius;
    size_t cells = 1;
    for (int k = 0; k < 3; ++k) {
        double c = floor((hi[k] - g->lo[k]) / radius) + 1.0;
        g->count[k] = c > 4096.0 ? 4096 : (int)c;
        cells *= (size_t)g->count[k];
    }
    if (cells > ((size_t)1 << 28))
        return 0;
    g->start = calloc(cells + 1, sizeof(unsigned));
    g->order = malloc(sizeof(unsigned) * (n ? n : 1));
    g->cell = malloc(sizeof(int) * (n ? n : 1));
    if (!g->start || !g->order || !g->cell)
        return 0;
    for (unsigned i = 0; i < n; ++i) {
        int x = axis_cell(p[3 * i], g->lo[0], g->cs, g->count[0]);
        int y = axis_cell(p[3 * i + 1], g->lo[1], g->cs, g->count[1]);
        int z = axis_cell(p[3 * i + 2], g->lo[2], g->cs, g->count[2]);
        g->cell[i] = x + g->count[0] * (y + g->count[1] * z);
        g->start[g->cell[i] + 1]++;
    }
    for (size_t c = 0; c < cells; ++c)
        g->start[c + 1] += g->start[c];
    unsigned *fill = malloc(sizeof(unsigned) * cells);
    if (!fill)
        return 0;
    memcpy(fill, g->start, sizeof(unsigned) * cells);
    for (unsigned i = 0; i < n; ++i)
        g->order[fill[g->cell[i]]++] = i;
    free(fill);
    return 1;
}

static void grid_free(Grid *g) {
    free(g->start);
    free(g->order);
    free(g->cell);
}

/* Runs the body for every j != i with r2 < R^2 (d and r2 in scope); the cells are R
 * wide, so the 27 cells around i's cell hold all of them. */
#define FOR_NEIGHBOURS(g, p, i, R2, ...)                                                               \
    do {                                                                                               \
        const int cx_ = (g)->cell[i] % (g)->count[0];                                                  \
        const int cy_ = ((g)->cell[i] / (g)->count[0]) % (g)->count[1];                                \
        const int cz_ = (g)->cell[i] / ((g)->count[0] * (g)->count[1]);                                \
        for (int x_ = cx_ - 1; x_ <= cx_ + 1; ++x_)                                                    \
            for (int y_ = cy_ - 1; y_ <= cy_ + 1; ++y_)                                                \
                for (int z_ = cz_ - 1; z_ <= cz_ + 1; ++z_) {                                          \
                    if (x_ < 0 || y_ < 0 || z_ < 0 || x_ >= (g)->count[0] || y_ >= (g)->count[1] ||    \
                        z_ >= (g)->count[2])                                                           \
                        continue;                                                                      \
                    const int c_ = x_ + (g)->count[0] * (y_ + (g)->count[1] * z_);                     \
                    for (unsigned s_ = (g)->start[c_]; s_ < (g)->start[c_ + 1]; ++s_) {                \
                        const unsigned j = (g)->order[s_];                                             \
                        if (j == (i))                                                                  \
                            continue;                                                                  \
                        double d[3] = {(p)[3 * j] - (p)[3 * (i)], (p)[3 * j + 1] - (p)[3 * (i) + 1],   \
                                       (p)[3 * j + 2] - (p)[3 * (i) + 2]};                              \
                        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];                     \
                        if (!(r2 < (R2)))                                                              \
                            continue;                                                                  \
                        __VA_ARGS__                                                                    \
                    }                                                                                  \
                }                                                                                      \
    } while (0)

/* Accumulators and acceleration of every agent; acc receives a, sums
 * (optional) S, A, C (9 per agent) and wsum W. */
static int accumulate(unsigned n, const double *p, const double *v, const double *wts, const DflockParams *prm,
                      double *acc, double *sums, double *wsum, Grid *g) {
    if (!grid_build(g, n, p, prm->radius))
        return 0;
    const double R2 = prm->radius * prm->radius;
#pragma omp parallel for schedule(dynamic, 256)
    for (unsigned i = 0; i < n; ++i) {
        double W = 0.0, A[3] = {0, 0, 0}, C[3] = {0, 0, 0}, S[3] = {0, 0, 0};
        FOR_NEIGHBOURS(g, p, i, R2, {
            const double t = 1.0 - r2 / R2;
            const double w = t * t * t;
            const double q = r2 + prm->h2;
            W += w;
            for (int k = 0; k < 3; ++k) {
                A[k] += w * (v[3 * j + k] - v[3 * i + k]);
                C[k] += w * d[k];
                S[k] -= w * d[k] / q;
            }
        });
        const double D = W + 1.0;
        for (int k = 0; k < 3; ++k) {
            acc[3 * i + k] = (wts[0] * S[k] + wts[1] * A[k] + wts[2] * C[k]) / D;
            if (sums) {
                sums[9 * i + k] = S[k];
                sums[9 * i + 3 + k] = A[k];
                sums[9 * i + 6 + k] = C[k];
            }
        }
        if (wsum)
            wsum[i] = W;
    }
    return 1;
}

static int step(unsigned n, const double *p, const double *v, const double *wts, const DflockParams *prm,
                double *p_out, double *v_out) {
    Grid g;
    double *acc = malloc(sizeof(double) * 3 * (n ? n : 1));
    int ok = acc && accumulate(n, p, v, wts, prm, acc, NULL, NULL, &g);
    if (ok)
        for (unsigned i = 0; i < 3 * n; ++i) {
            v_out[i] = v[i] + prm->dt * (acc[i] - prm->damping * v[i]);
            p_out[i] = p[i] + prm->dt * v_out[i];
        }
    grid_free(&g);
    free(acc);
    return ok;
}

int oracle_dflock_forward(unsigned n, const double *p0, const double *v0, const double *weights,
                          const DflockParams *prm, unsigned steps, double *p_out, double *v_out) {
    double *p = malloc(sizeof(double) * 3 * (n ? n : 1)), *v = malloc(sizeof(double) * 3 * (n ? n : 1));
    if (!p || !v)
        return 0;
    memcpy(p, p0, sizeof(double) * 3 * n);
    memcpy(v, v0, sizeof(double) * 3 * n);
    int ok = 1;
    for (unsigned t = 0; ok && t < steps; ++t) {
        ok = step(n, p, v, weights, prm, p_out, v_out);
        memcpy(p, p_out, sizeof(double) * 3 * n);
        memcpy(v, v_out, sizeof(double) * 3 * n);
    }
    memcpy(p_out, p, sizeof(double) * 3 * n);
    memcpy(v_out, v, sizeof(double) * 3 * n);
    free(p);
    free(v);
    return ok;
}

/* states_p / states_v (optional): the trajectory to differentiate along,
 * steps + 1 states of n x 3; without them the forward model makes it. */
static int backprop(unsigned n, const double *p0, const double *v0, const double *weights, const DflockParams *prm,
                    unsigned steps, const double *target, double *loss, const double *gp_final,
                    const double *gv_final, double *g_p0, double *g_v0, double *g_weights,
                    const double *states_p, const double *states_v) {
    const size_t m = 3 * (size_t)(n ? n : 1);
    double *ps = malloc(sizeof(double) * m * (steps + 1)), *vs = malloc(sizeof(double) * m * (steps + 1));
    double *acc = malloc(sizeof(double) * m), *sums = malloc(sizeof(double) * 3 * m), *wsum = malloc(sizeof(double) * m);
    double *Lp = malloc(sizeof(double) * m), *Lv = malloc(sizeof(double) * m), *nLp = malloc(sizeof(double) * m),
           *nLv = malloc(sizeof(double) * m), *adj = malloc(sizeof(double) * 10 * m / 3 + 10),
           *gpart = malloc(sizeof(double) * m);
    int ok = ps && vs && acc && sums && wsum && Lp && Lv && nLp && nLv && adj && gpart;
    if (ok && states_p) {
        for (unsigned t = 0; t <= steps; ++t) {
            memcpy(ps + m * t, states_p + (size_t)3 * n * t, sizeof(double) * 3 * n);
            memcpy(vs + m * t, states_v + (size_t)3 * n * t, sizeof(double) * 3 * n);
        }
    } else if (ok) {
        memcpy(ps, p0, sizeof(double) * 3 * n);
        memcpy(vs, v0, sizeof(double) * 3 * n);
        for (unsigned t = 0; ok && t < steps; ++t)
            ok = step(n, ps + m * t, vs + m * t, weights, prm, ps + m * (t + 1), vs + m * (t + 1));
    }
    double L = 0.0;
    for (unsigned i = 0; ok && i < 3 * n; ++i) {
        if (target) {
            const double e = ps[m * steps + i] - target[i];
            L += e * e;
            Lp[i] = 2.0 * e / (3.0 * n);
            Lv[i] = 0.0;
        } else {
            Lp[i] = gp_final[i];
            Lv[i] = gv_final ? gv_final[i] : 0.0;
        }
    }
    if (loss)
        *loss = ok ? L / (3.0 * n) : 0.0;
    double gw[3] = {0.0, 0.0, 0.0};
    const double R2 = prm->radius * prm->radius, dt = prm->dt;
    for (unsigned tt = steps; ok && tt-- > 0;) {
        const double *p = ps + m * tt, *v = vs + m * tt;
        Grid g;
        ok = accumulate(n, p, v, weights, prm, acc, sums, wsum, &g);
        if (!ok)
            break;
        /* per agent: direct terms and the adjoints of its accumulators */
        for (unsigned i = 0; i < n; ++i) {
            const double D = wsum[i] + 1.0;
            double La[3], dot = 0.0;
            for (int k = 0; k < 3; ++k) {
                const double Lvt = Lv[3 * i + k] + dt * Lp[3 * i + k];
                nLp[3 * i + k] = Lp[3 * i + k];
                nLv[3 * i + k] = (1.0 - dt * prm->damping) * Lvt;
                La[k] = dt * Lvt;
                dot += La[k] * acc[3 * i + k];
            }
            double *a = adj + 10 * (size_t)i;   /* LS[3] LA[3] LC[3] LW */
            for (int k = 0; k < 3; ++k) {
                a[k] = weights[0] * La[k] / D;
                a[3 + k] = weights[1] * La[k] / D;
                a[6 + k] = weights[2] * La[k] / D;
            }
            a[9] = -dot / D;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
            for (int k = 0; k < 3; ++k) {
                s0 += La[k] * sums[9 * i + k] / D;
                s1 += La[k] * sums[9 * i + 3 + k] / D;
                s2 += La[k] * sums[9 * i + 6 + k] / D;
            }
            gpart[3 * i] = s0;
            gpart[3 * i + 1] = s1;
            gpart[3 * i + 2] = s2;
        }
        for (unsigned i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k)
                gw[k] += gpart[3 * i + k];
        /* pair terms, gathered per agent k over its symmetric neighbour set:
         * as seen in (m, k) and as seer in (k, m) */
#pragma omp parallel for schedule(dynamic, 256)
        for (unsigned kk = 0; kk < n; ++kk) {
            double gp[3] = {0, 0, 0}, gv[3] = {0, 0, 0};
            const double *ak = adj + 10 * (size_t)kk;
            FOR_NEIGHBOURS(&g, p, kk, R2, {
                /* d = p_j - p_kk with j the neighbour */
                const double *am = adj + 10 * (size_t)j;
                const double t = 1.0 - r2 / R2;
                const double w = t * t * t, dw = -3.0 * t * t / R2, q = r2 + prm->h2;
                /* seer kk, seen j: d */
                double dLS = 0.0, dLC = 0.0, vLA = 0.0;
                for (int c = 0; c < 3; ++c) {
                    dLS += d[c] * ak[c];
                    dLC += d[c] * ak[6 + c];
                    vLA += (v[3 * j + c] - v[3 * kk + c]) * ak[3 + c];
                }
                double Lw = ak[9] + vLA + dLC - dLS / q;
                for (int c = 0; c < 3; ++c) {
                    const double Ld = 2.0 * dw * Lw * d[c] + w * ak[6 + c] - w * ak[c] / q +
                                      2.0 * w * dLS * d[c] / (q * q);
                    gp[c] -= Ld;
                    gv[c] -= w * ak[3 + c];
                }
                /* seer j, seen kk: d' = -d */
                dLS = dLC = vLA = 0.0;
                for (int c = 0; c < 3; ++c) {
                    dLS -= d[c] * am[c];
                    dLC -= d[c] * am[6 + c];
                    vLA += (v[3 * kk + c] - v[3 * j + c]) * am[3 + c];
                }
                Lw = am[9] + vLA + dLC - dLS / q;
                for (int c = 0; c < 3; ++c) {
                    const double Ld = 2.0 * dw * Lw * (-d[c]) + w * am[6 + c] - w * am[c] / q +
                                      2.0 * w * dLS * (-d[c]) / (q * q);
                    gp[c] += Ld;
                    gv[c] += w * am[3 + c];
                }
            });
            for (int c = 0; c < 3; ++c) {
                nLp[3 * kk + c] += gp[c];
                nLv[3 * kk + c] += gv[c];
            }
        }
        grid_free(&g);
        double *tmp = Lp;
        Lp = nLp;
        nLp = tmp;
        tmp = Lv;
        Lv = nLv;
        nLv = tmp;
    }
    if (ok) {
        memcpy(g_p0, Lp, sizeof(double) * 3 * n);
        memcpy(g_v0, Lv, sizeof(double) * 3 * n);
        for (int k = 0; k < 3; ++k)
            g_weights[k] = gw[k];
    }
    free(ps);
    free(vs);
    free(acc);
    free(sums);
    free(wsum);
    free(Lp);
    free(Lv);
    free(nLp);
    free(nLv);
    free(adj);
    free(gpart);
    return ok;
}

int oracle_dflock_gradient(unsigned n, const double *p0, const double *v0, const double *weights,
                           const DflockParams *prm, unsigned steps, const double *target, double *loss,
                           double *g_p0, double *g_v0, double *g_weights) {
    return backprop(n, p0, v0, weights, prm, steps, target, loss, NULL, NULL, g_p0, g_v0, g_weights, NULL, NULL);
}

int oracle_dflock_backprop(unsigned n, const double *p0, const double *v0, const double *weights,
                           const DflockParams *prm, unsigned steps, const double *gp_final, const double *gv_final,
                           double *g_p0, double *g_v0, double *g_weights) {
    return backprop(n, p0, v0, weights, prm, steps, NULL, NULL, gp_final, gv_final, g_p0, g_v0, g_weights, NULL,
                    NULL);
}

int oracle_dflock_backprop_along(unsigned n, const double *states_p, const double *states_v, const double *weights,
                                 const DflockParams *prm, unsigned steps, const double *gp_final,
                                 const double *gv_final, double *g_p0, double *g_v0, double *g_weights) {
    return backprop(n, states_p, states_v, weights, prm, steps, NULL, NULL, gp_final, gv_final, g_p0, g_v0,
                    g_weights, states_p, states_v);
}
