/*
 * dflock.hip - differentiable flocking: hand-written forward and adjoint
 * kernels for the smooth flocking model, with checkpointed recomputation.
 * No reference counterpart (the reference has no differentiable path).
 *
 * Model, per step, for every agent i and every other agent j with
 * r2 = |p_j - p_i|^2 < R^2 (d = p_j - p_i, t = 1 - r2 / R^2, w = t^3,
 * q = r2 + h2):
 *     W_i += w,  A_i += w (v_j - v_i),  C_i += w d,  S_i -= w d / q
 *     a_i  = (ws S_i + wa A_i + wc C_i) / (W_i + 1)
 *     v_i' = v_i + dt (a_i - damping v_i),   p_i' = p_i + dt v_i'
 * The cutoff weight w and its first two derivatives vanish at r = R, so
 * gradients exist everywhere.  The same model in Tay's model dialect is
 * smooth_see / smooth_act (agent_src/flocking.c).
 *
 * Storage: agents keep their identity order; positions and velocities are
 * structure-of-arrays words 2..4 and 5..7 of an 8-word column set, so the
 * Z-grid build kernels (grid_build.h) sort (key, agent) pairs and fill the
 * cell table without moving the agents.  Threads walk agents in key order
 * for locality.
 *
 * Forward: per step, grid build + one fused see/act kernel (eight lanes per
 * agent, their partial sums combined by a butterfly); the state at
 * every checkpoint_every-th step is kept.  Backward: per segment between
 * checkpoints, the forward is recomputed from its checkpoint recording every
 * step's state, sums (S, A, C, W) and neighbour lists (at most K per agent,
 * K = the largest neighbour count the forward saw), then the steps are
 * reversed:
 *   local   per agent: adjoints of the step's outputs and of its sums
 *           (LS = ws La / D, LA = wa La / D, LC = wc La / D,
 *            LW = -(La . a) / D with La = dt (Lv' + dt Lp'), D = W + 1),
 *           and its share of the weight gradient La . (S, A, C) / D
 *   gather  per agent k over its neighbours m: the pair terms with k as seer
 *           and as seen, combined so that their cancelling parts are never
 *           formed.  The neighbour relation is symmetric (the same rounded
 *           r2 decides both directions), so every pair term is gathered by
 *           the agent it updates: no atomics, and the result is the same on
 *           every run.
 * The weight gradient is reduced per step in double, in a fixed order.
 */
#include "internal.h"

#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

using namespace tay;

#include "grid_build.h"
#include "tay_dflock.h"

namespace {

constexpr unsigned kWords = 8;      /* tag (2, unused), p (3), v (3) */
constexpr unsigned kReduceBlock = 256;
constexpr unsigned kReduceParts = 64;     /* blocks of the first reduction stage */

struct Prm {
    float radius, r2max, inv_r2, h2, dt, damping;
    float ws, wa, wc;
};

/* Cells along one axis that can hold a point within r of p, with the
 * see kernel's rounding margin. */
__device__ __forceinline__ void axis_span(float p, float r, const TayGridGeom *g, int k, int *lo, int *hi) {
    const float m = (fabsf(p) + r) * 0x1p-20f;
    *lo = cell_axis((p - r) - m, g->origin[k], g->cell_size[k], g->count[k]);
    *hi = cell_axis((p + r) + m, g->origin[k], g->cell_size[k], g->count[k]);
}

__device__ __forceinline__ unsigned wave_max_u(unsigned v) {
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned u = __shfl_xor(v, o);
        v = u > v ? u : v;
    }
    return v;
}

/* Pair weight terms. */
struct PairW {
    float w, dw, q;
};

__device__ __forceinline__ PairW pair_weight(float r2, const Prm &c) {
    const float t = 1.0f - r2 * c.inv_r2;
    PairW o;
    o.w = t * t * t;
    o.dw = -3.0f * t * t * c.inv_r2;
    o.q = r2 + c.h2;
    return o;
}

/* Lanes per agent in the step and gather kernels: the candidates of an agent
 * are dealt to LANES lanes (each takes every LANES-th agent of every cell,
 * cells in x, y, z order) and their partial sums are combined by a butterfly
 * over the group, which gives every lane the same, run-to-run identical
 * totals.  Few agents (tens of thousands) would otherwise leave one
 * wavefront per SIMD and no latency hidden. */
constexpr unsigned LANES = 8;

__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (unsigned o = 1; o < LANES; o <<= 1)
        v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ unsigned group_exclusive_count(unsigned v, unsigned sub) {
    unsigned inc = v;
#pragma unroll
    for (unsigned o = 1; o < LANES; o <<= 1) {
        const unsigned u = __shfl_up(inc, o, LANES);
        inc += sub >= o ? u : 0u;
    }
    return inc - v;
}

/* One step for the agents in key order: neighbour sums, acceleration, new
 * state.  RECORD also writes the sums and, in the same sweep, each lane's
 * neighbours (their key-order slots) to its own list of KL entries with its
 * count, everything indexed by the agent's key-order slot t; entry e of
 * lane l is word (t * KL + e) * LANES + l, so the lanes of an agent that
 * read their e-th entries together read one contiguous run.  Block b keeps
 * running statistics in stats[4b..4b+3] across the steps of a run: the
 * largest neighbour count of an agent, the largest of one lane (which sizes
 * the lists of the recording recomputation) and the 64-bit sum of every
 * agent's neighbour count.  Each block updates only its own slot, after the
 * previous step's launch has finished with it, so no atomics are needed
 * (one slot per wavefront on a single address serialised the whole grid);
 * k_df_stats_reduce folds the slots when a caller asks. */
template <bool RECORD>
__global__ __launch_bounds__(256) void k_df_step(const float *__restrict__ cols, unsigned n,
                                                 const unsigned *__restrict__ ids, const float *__restrict__ srt,
                                                 const TayCellRange *__restrict__ ranges,
                                                 const TayGridGeom *__restrict__ g, Prm c,
                                                 float *__restrict__ out, unsigned *__restrict__ nbr, unsigned KL,
                                                 unsigned *__restrict__ counts, float *__restrict__ sums,
                                                 unsigned *__restrict__ stats) {
    const unsigned t = (blockIdx.x * 256u + threadIdx.x) / LANES;
    const unsigned sub = threadIdx.x % LANES;
    const bool live = t < n;
    const unsigned i = live ? ids[t] : 0u;
    const float *p = cols + 2u * n, *v = cols + 5u * n;
    const float pi[3] = {p[i], p[n + i], p[2u * n + i]};
    const float vi[3] = {v[i], v[n + i], v[2u * n + i]};
    /* candidates come from the key-ordered copy: consecutive slots of a
     * cell are consecutive words */
    const float *ps = srt, *vs = srt + 3u * n;
    int lo[3], hi[3];
    for (int k = 0; k < 3; ++k)
        axis_span(pi[k], c.radius, g, k, &lo[k], &hi[k]);
    const unsigned cx = (unsigned)g->count[0], cxy = cx * (unsigned)g->count[1];
    /* visits this lane's candidates j of agent i with r2 < R^2 */
    auto sweep = [&](auto &&visit) {
        if (!live)
            return;
        for (int x = lo[0]; x <= hi[0]; ++x)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int z = lo[2]; z <= hi[2]; ++z) {
                    const TayCellRange r = ranges[(unsigned)x + (unsigned)y * cx + (unsigned)z * cxy];
                    for (unsigned s = r.beg + sub; s < r.end; s += LANES) {
                        if (s == t)
                            continue;
                        const float d[3] = {ps[s] - pi[0], ps[n + s] - pi[1], ps[2u * n + s] - pi[2]};
                        const float r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                        if (r2 < c.r2max)
                            visit(s, d, r2);
                    }
                }
    };
    float W = 0.0f, A[3] = {0.0f, 0.0f, 0.0f}, C[3] = {0.0f, 0.0f, 0.0f}, S[3] = {0.0f, 0.0f, 0.0f};
    unsigned cnt = 0;
    unsigned *list = RECORD ? nbr + (size_t)t * KL * LANES + sub : nullptr;
    sweep([&](unsigned s, const float *d, float r2) {
        if (RECORD && cnt < KL)
            list[(size_t)cnt * LANES] = s;
        const PairW pw = pair_weight(r2, c);
        W += pw.w;
        const float vj[3] = {vs[s], vs[n + s], vs[2u * n + s]};
        for (int k = 0; k < 3; ++k) {
            A[k] += pw.w * (vj[k] - vi[k]);
            C[k] += pw.w * d[k];
            S[k] -= pw.w * d[k] / pw.q;
        }
        ++cnt;
    });
    W = group_sum(W);
    for (int k = 0; k < 3; ++k) {
        A[k] = group_sum(A[k]);
        C[k] = group_sum(C[k]);
        S[k] = group_sum(S[k]);
    }
    const unsigned before = group_exclusive_count(cnt, sub);
    const unsigned total = __shfl(before + cnt, LANES - 1u, LANES);
    if (RECORD && live)
        counts[(size_t)t * LANES + sub] = cnt;
    if (live && sub == 0) {
        const float D = W + 1.0f;
        for (int k = 0; k < 3; ++k) {
            const float a = (c.ws * S[k] + c.wa * A[k] + c.wc * C[k]) / D;
            const float vn = vi[k] + c.dt * (a - c.damping * vi[k]);
            out[(5u + k) * n + i] = vn;
            out[(2u + k) * n + i] = pi[k] + c.dt * vn;
            if (RECORD) {
                sums[(size_t)k * n + t] = S[k];
                sums[(size_t)(3 + k) * n + t] = A[k];
                sums[(size_t)(6 + k) * n + t] = C[k];
            }
        }
        if (RECORD)
            sums[(size_t)9 * n + t] = W;
    }
    const unsigned m = wave_max_u(live ? total : 0u);
    const unsigned ml = wave_max_u(live ? cnt : 0u);
    /* each agent's count once (its lanes hold partial counts) */
    unsigned own = live ? cnt : 0u;
    for (int o = 32; o > 0; o >>= 1)
        own += __shfl_xor(own, o);
    __shared__ unsigned wave_stats[256 / 64][3];
    const unsigned w = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0u) {
        wave_stats[w][0] = m;
        wave_stats[w][1] = ml;
        wave_stats[w][2] = own;
    }
    __syncthreads();
    if (threadIdx.x == 0u) {
        unsigned bm = 0, bml = 0;
        unsigned long long bsum = 0;
        for (unsigned k = 0; k < 256 / 64; ++k) {
            bm = wave_stats[k][0] > bm ? wave_stats[k][0] : bm;
            bml = wave_stats[k][1] > bml ? wave_stats[k][1] : bml;
            bsum += wave_stats[k][2];
        }
        unsigned *slot = stats + 4ull * blockIdx.x;
        slot[0] = slot[0] > bm ? slot[0] : bm;
        slot[1] = slot[1] > bml ? slot[1] : bml;
        *(unsigned long long *)(slot + 2) += bsum;
    }
}

/* Folds the per-block statistics of k_df_step: out[0] = largest neighbour
 * count of an agent, out[1] = largest of one lane, out[2..3] = the 64-bit
 * pair total.  One block of 256 threads. */
__global__ __launch_bounds__(256) void k_df_stats_reduce(const unsigned *__restrict__ stats, unsigned blocks,
                                                         unsigned *__restrict__ out) {
    unsigned m = 0, ml = 0;
    unsigned long long sum = 0;
    for (unsigned b = threadIdx.x; b < blocks; b += 256u) {
        const unsigned *slot = stats + 4ull * b;
        m = slot[0] > m ? slot[0] : m;
        ml = slot[1] > ml ? slot[1] : ml;
        sum += *(const unsigned long long *)(slot + 2);
    }
    __shared__ unsigned sm[256], sml[256];
    __shared__ unsigned long long ssum[256];
    sm[threadIdx.x] = m;
    sml[threadIdx.x] = ml;
    ssum[threadIdx.x] = sum;
    __syncthreads();
    for (unsigned h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            sm[threadIdx.x] = sm[threadIdx.x + h] > sm[threadIdx.x] ? sm[threadIdx.x + h] : sm[threadIdx.x];
            sml[threadIdx.x] = sml[threadIdx.x + h] > sml[threadIdx.x] ? sml[threadIdx.x + h] : sml[threadIdx.x];
            ssum[threadIdx.x] += ssum[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0u) {
        out[0] = sm[0];
        out[1] = sml[0];
        *(unsigned long long *)(out + 2) = ssum[0];
    }
}

/* p and v (words 2..7) of the agents in key order: srt[k * n + s] =
 * cols[(2 + k) * n + ids[s]]. */
__global__ __launch_bounds__(256) void k_df_sorted_copy(const float *__restrict__ cols, unsigned n,
                                                        const unsigned *__restrict__ ids, float *__restrict__ srt) {
    const unsigned s = blockIdx.x * 256u + threadIdx.x;
    if (s >= n)
        return;
    const unsigned j = ids[s];
    float w[6];
#pragma unroll
    for (unsigned k = 0; k < 6u; ++k)
        w[k] = cols[(2u + k) * n + j];
#pragma unroll
    for (unsigned k = 0; k < 6u; ++k)
        srt[k * n + s] = w[k];
}

/* Per agent: the direct adjoint terms of the step, the adjoints of its sums
 * (adj: LS[3], LA[3], LC[3], LW and two pad words, 48 bytes per agent, so a
 * neighbour's are three 16-byte loads) and its share of the weight
 * gradient. */
__global__ __launch_bounds__(256) void k_df_adj_local(unsigned n, const unsigned *__restrict__ ids,
                                                      const float *__restrict__ sums, Prm c,
                                                      const float *__restrict__ Lp, const float *__restrict__ Lv,
                                                      float *__restrict__ nLp, float *__restrict__ nLv,
                                                      float *__restrict__ adj, double *__restrict__ gpart) {
    const unsigned t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n)
        return;
    const unsigned i = ids[t];
    const float W = sums[(size_t)9 * n + t];
    const float D = W + 1.0f;
    float La[3], S[3], A[3], C[3];
    float dot = 0.0f;
    for (int k = 0; k < 3; ++k) {
        S[k] = sums[(size_t)k * n + t];
        A[k] = sums[(size_t)(3 + k) * n + t];
        C[k] = sums[(size_t)(6 + k) * n + t];
        const float a = (c.ws * S[k] + c.wa * A[k] + c.wc * C[k]) / D;
        const float lvt = Lv[(size_t)k * n + i] + c.dt * Lp[(size_t)k * n + i];
        nLp[(size_t)k * n + i] = Lp[(size_t)k * n + i];
        nLv[(size_t)k * n + i] = (1.0f - c.dt * c.damping) * lvt;
        La[k] = c.dt * lvt;
        dot += La[k] * a;
    }
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    float a[12];
    for (int k = 0; k < 3; ++k) {
        a[k] = c.ws * La[k] / D;
        a[3 + k] = c.wa * La[k] / D;
        a[6 + k] = c.wc * La[k] / D;
        g0 += (double)(La[k] * S[k] / D);
        g1 += (double)(La[k] * A[k] / D);
        g2 += (double)(La[k] * C[k] / D);
    }
    a[9] = -dot / D;
    a[10] = a[11] = 0.0f;
    float4 *out = (float4 *)(adj + (size_t)12 * t);
    for (int q = 0; q < 3; ++q)
        out[q] = make_float4(a[4 * q], a[4 * q + 1], a[4 * q + 2], a[4 * q + 3]);
    gpart[t] = g0;
    gpart[(size_t)n + t] = g1;
    gpart[(size_t)2 * n + t] = g2;
}

/* Pair term of agent k's adjoints from one neighbour m (d = p_m - p_k,
 * dv = v_m - v_k): the adjoint with k as the seen agent of m's sums minus
 * the adjoint with k as the seer of its own.  Written out, the parts
 * w LC - w LS / q + 2 w (d . LS) d / q^2 of the two directions cancel up to
 * the difference of the two agents' sum adjoints, so the kernel forms that
 * difference (da = a_m - a_k) first and never sums the large cancelling
 * terms:
 *   Lp_k += -2 w' [LW_m + LW_k - dv . dLA - d . dLC + d . dLS / q] d
 *           + w dLC - w dLS / q + 2 w (d . dLS) d / q^2
 *   Lv_k += w dLA */
__device__ __forceinline__ void pair_gather(const float *d, const PairW &pw, const float *ak, const float *am,
                                            const float *dv, float *gp, float *gv) {
    float da[9];
    for (int q = 0; q < 9; ++q)
        da[q] = am[q] - ak[q];
    const float dS = d[0] * da[0] + d[1] * da[1] + d[2] * da[2];
    const float vA = dv[0] * da[3] + dv[1] * da[4] + dv[2] * da[5];
    const float dC = d[0] * da[6] + d[1] * da[7] + d[2] * da[8];
    const float Lw = (am[9] + ak[9]) - vA - dC + dS / pw.q;
    for (int k = 0; k < 3; ++k) {
        gp[k] += -2.0f * pw.dw * Lw * d[k] + pw.w * da[6 + k] - pw.w * da[k] / pw.q +
                 2.0f * pw.w * dS * d[k] / (pw.q * pw.q);
        gv[k] += pw.w * da[3 + k];
    }
}

__device__ __forceinline__ void load_adj(const float *__restrict__ adj, unsigned i, float *a) {
    const float4 *q = (const float4 *)(adj + (size_t)12 * i);
    for (int k = 0; k < 3; ++k) {
        const float4 v = q[k];
        a[4 * k] = v.x;
        a[4 * k + 1] = v.y;
        a[4 * k + 2] = v.z;
        a[4 * k + 3] = v.w;
    }
}

__global__ __launch_bounds__(256) void k_df_adj_gather(unsigned n, const unsigned *__restrict__ ids,
                                                       const float *__restrict__ srt,
                                                       const unsigned *__restrict__ nbr, unsigned KL,
                                                       const unsigned *__restrict__ counts,
                                                       const float *__restrict__ adj, Prm c,
                                                       float *__restrict__ nLp, float *__restrict__ nLv) {
    const unsigned t = (blockIdx.x * 256u + threadIdx.x) / LANES;
    const unsigned sub = threadIdx.x % LANES;
    const bool live = t < n;
    const unsigned k = live ? ids[t] : 0u;
    const unsigned tk = live ? t : 0u;
    /* everything of the step in key order: p and v from its key-ordered
     * copy, sum adjoints and neighbour lists by key-order slot */
    const float *p = srt, *v = srt + 3u * n;
    const float pk[3] = {p[tk], p[n + tk], p[2u * n + tk]};
    const float vk[3] = {v[tk], v[n + tk], v[2u * n + tk]};
    float ak[12];
    load_adj(adj, tk, ak);
    float gp[3] = {0.0f, 0.0f, 0.0f}, gv[3] = {0.0f, 0.0f, 0.0f};
    /* lane sub takes the neighbours lane sub of the step recorded */
    const unsigned lc = live ? counts[(size_t)t * LANES + sub] : 0u;
    const unsigned cnt = lc < KL ? lc : KL;
    const unsigned *list = nbr + (size_t)tk * KL * LANES + sub;
    for (unsigned e = 0; e < cnt; ++e) {
        const unsigned m = list[(size_t)e * LANES];
        const float d[3] = {p[m] - pk[0], p[n + m] - pk[1], p[2u * n + m] - pk[2]};
        const float r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const PairW pw = pair_weight(r2, c);
        const float dv[3] = {v[m] - vk[0], v[n + m] - vk[1], v[2u * n + m] - vk[2]};
        float am[12];
        load_adj(adj, m, am);
        pair_gather(d, pw, ak, am, dv, gp, gv);
    }
    for (int q = 0; q < 3; ++q) {
        gp[q] = group_sum(gp[q]);
        gv[q] = group_sum(gv[q]);
    }
    if (live && sub == 0)
        for (int q = 0; q < 3; ++q) {
            nLp[(size_t)q * n + k] += gp[q];
            nLv[(size_t)q * n + k] += gv[q];
        }
}

/* One block: gw[r] += sum over agents of gpart[r], in a fixed order. */
/* Block b: part[r * kReduceParts + b] = sum of gpart[r] over its contiguous
 * share of the agents, in a fixed order (strided per thread, then a tree). */
__global__ __launch_bounds__(kReduceBlock) void k_df_reduce(unsigned n, const double *__restrict__ gpart,
                                                            double *__restrict__ part) {
    __shared__ double s[kReduceBlock];
    const unsigned per = (n + kReduceParts - 1) / kReduceParts;
    const unsigned beg = blockIdx.x * per, end = beg + per < n ? beg + per : n;
    for (unsigned r = 0; r < 3; ++r) {
        double acc = 0.0;
        for (unsigned i = beg + threadIdx.x; i < end; i += kReduceBlock)
            acc += gpart[(size_t)r * n + i];
        s[threadIdx.x] = acc;
        __syncthreads();
        for (unsigned h = kReduceBlock / 2; h > 0; h >>= 1) {
            if (threadIdx.x < h)
                s[threadIdx.x] += s[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            part[r * kReduceParts + blockIdx.x] = s[0];
        __syncthreads();
    }
}

/* gw[r] += the kReduceParts block sums of k_df_reduce, in block order. */
__global__ __launch_bounds__(64) void k_df_reduce_parts(const double *__restrict__ part, double *__restrict__ gw) {
    if (threadIdx.x < 3) {
        double acc = 0.0;
        for (unsigned b = 0; b < kReduceParts; ++b)
            acc += part[threadIdx.x * kReduceParts + b];
        gw[threadIdx.x] += acc;
    }
}

/* n x 3 host (AoS) <-> 3 rows of n (SoA) at rows[row0 .. row0 + 2]. */
__global__ __launch_bounds__(256) void k_df_unpack(const float *__restrict__ aos, unsigned n,
                                                   float *__restrict__ rows) {
    const unsigned i = blockIdx.x * 256u + threadIdx.x;
    if (i < n)
        for (int k = 0; k < 3; ++k)
            rows[(size_t)k * n + i] = aos[3u * i + k];
}

__global__ __launch_bounds__(256) void k_df_pack(const float *__restrict__ rows, unsigned n,
                                                 float *__restrict__ aos) {
    const unsigned i = blockIdx.x * 256u + threadIdx.x;
    if (i < n)
        for (int k = 0; k < 3; ++k)
            aos[3u * i + k] = rows[(size_t)k * n + i];
}

}  // namespace

struct TayDflock {
    unsigned n;
    TayDflockParams params;
    unsigned every;
    unsigned long long max_cells;
    float weights[3];
    unsigned steps;             /* of the last forward */
    unsigned KL;                /* per-lane neighbour-list capacity of the recorded segment */

    float *cols[2];             /* 8 words per agent */
    float *srt;                 /* p and v in key order, 6 words per agent */
    std::vector<float *> ckpt;  /* 8 words per agent, one per checkpoint */
    /* grid */
    float *box_partials;
    unsigned box_blocks;
    TayGridGeom *geom;
    unsigned *keys[2], *vals[2], *block_hist, *scan_sums;
    TayCellRange *ranges;
    unsigned *max_count;        /* folded statistics: two maxima and the 64-bit pair total */
    unsigned *stats[2];         /* per-block statistics of the forward [0] and of the recording [1] */
    unsigned stats_blocks;
    /* recorded segment */
    float *seg_states;          /* every x 8 words per agent */
    unsigned *seg_ids;          /* every x n: agents in key order */
    float *seg_sums;            /* every x 10 rows, key order */
    float *seg_srt;             /* every x 6 rows: p and v in key order */
    unsigned *seg_counts;       /* every x n x LANES */
    unsigned *seg_nbr;          /* every x n x LANES x KL */
    size_t seg_nbr_cap;
    /* adjoint */
    float *Lp[2], *Lv[2], *adj;
    double *gpart, *gw, *gparts;
    float *io;                  /* n x 3 host transfer staging */
    double forward_ms, backward_ms;
};

namespace {

Prm make_prm(const TayDflock *f) {
    Prm c;
    c.radius = f->params.radius;
    c.r2max = f->params.radius * f->params.radius;
    c.inv_r2 = 1.0f / c.r2max;
    c.h2 = f->params.h2;
    c.dt = f->params.dt;
    c.damping = f->params.damping;
    c.ws = f->weights[0];
    c.wa = f->weights[1];
    c.wc = f->weights[2];
    return c;
}

/* Grid of the positions in cols: keys[0] / vals[0] hold (key, agent) in key
 * order afterwards, ranges the cells. */
void df_sort(TayDflock *f, const float *cols, hipStream_t s) {
    const unsigned n = f->n;
    const unsigned *c = (const unsigned *)cols;
    const unsigned box_blocks = blocks_for(n, kBuildBlock * 4, f->box_blocks);
    const float part = f->params.cell_size;
    k_box_partial<<<box_blocks, kBuildBlock, 0, s>>>(c, n, n, 3, 1, f->box_partials);
    k_geometry<<<1, 64, 0, s>>>(f->box_partials, box_blocks, 3, part, part, part, f->max_cells,
                                TAY_GPU_DEFAULT_ORDER_BITS, 0, 0, 0, 0, 0, 0, 0, f->geom);
    const unsigned agent_blocks = blocks_for(n, kBuildBlock, 0xFFFFFFFFu);
    k_keys<<<agent_blocks, kBuildBlock, 0, s>>>(c, n, n, 3, f->geom, f->keys[0]);
    const unsigned nblocks = (n + kRadixTile - 1) / kRadixTile;
    const unsigned total = kRadix * nblocks;
    const unsigned segs = total < kScanBlocks * kBuildBlock ? (total + kBuildBlock - 1) / kBuildBlock : kScanBlocks;
    for (unsigned pass = 0; pass < kRadixPasses; ++pass) {
        const unsigned in = pass & 1u, out = in ^ 1u;
        k_radix_hist<<<nblocks, kRadixBlock, 0, s>>>(f->keys[in], n, pass * kRadixBits, f->block_hist, nblocks,
                                                     nullptr, nullptr);
        k_scan_reduce<<<segs, kBuildBlock, 0, s>>>(f->block_hist, total, f->scan_sums, nullptr, 0u);
        k_scan_down<<<segs, kBuildBlock, 0, s>>>(f->block_hist, total, f->scan_sums, nullptr, 0u);
        k_radix_scatter<<<nblocks, kRadixBlock, 0, s>>>(f->keys[in], f->vals[in], f->keys[out], f->vals[out], n,
                                                        pass * kRadixBits, f->block_hist, nblocks, pass == 0, nullptr,
                                                        nullptr);
    }
    k_ranges_clear<<<blocks_for(f->max_cells, kBuildBlock, 4096), kBuildBlock, 0, s>>>(f->ranges, f->geom);
    k_ranges<<<agent_blocks, kBuildBlock, 0, s>>>(f->keys[0], n, 3, f->geom, f->ranges);
}

size_t state_bytes(const TayDflock *f) {
    return (size_t)kWords * f->n * sizeof(float);
}

bool timed_begin(hipEvent_t *e, hipStream_t s) {
    return hip_ok(hipEventCreate(&e[0]), "hipEventCreate") && hip_ok(hipEventCreate(&e[1]), "hipEventCreate") &&
           hip_ok(hipEventRecord(e[0], s), "hipEventRecord");
}

bool timed_end(hipEvent_t *e, hipStream_t s, double *ms) {
    float t = 0.0f;
    bool ok = hip_ok(hipEventRecord(e[1], s), "hipEventRecord") &&
              hip_ok(hipEventSynchronize(e[1]), "hipEventSynchronize") &&
              hip_ok(hipEventElapsedTime(&t, e[0], e[1]), "hipEventElapsedTime");
    (void)hipEventDestroy(e[0]);
    (void)hipEventDestroy(e[1]);
    *ms = t;
    return ok;
}

}  // namespace

TAY_API void tay_dflock_destroy(TayDflock *f) {
    if (!f)
        return;
    (void)tay_gpu_synchronize();
    for (float *c : f->ckpt)
        device_free(c);
    void *bufs[] = {f->cols[0], f->cols[1], f->box_partials, f->geom, f->keys[0], f->keys[1], f->vals[0],
                    f->vals[1], f->block_hist, f->scan_sums, f->ranges, f->max_count, f->stats[0], f->stats[1],
                    f->seg_states, f->seg_ids,
                    f->seg_sums, f->seg_counts, f->seg_nbr, f->Lp[0], f->Lp[1], f->Lv[0], f->Lv[1], f->adj,
                    f->gpart, f->gw, f->gparts, f->io, f->srt, f->seg_srt};
    for (void *p : bufs)
        device_free(p);
    delete f;
}

TAY_API TayDflock *tay_dflock_create(unsigned n, const TayDflockParams *params, unsigned checkpoint_every,
                                     unsigned long long cell_table_bytes) {
    if (n == 0 || n > (1u << 26)) {
        set_error("differentiable flock of %u agents: needs 1 to 2^26", n);
        return nullptr;
    }
    if (!params || !(params->radius > 0.0f) || !(params->cell_size > 0.0f) || !(params->dt > 0.0f) ||
        !(params->h2 > 0.0f) || !(params->damping >= 0.0f)) {
        set_error("parameters: radius, cell_size, dt and h2 must be positive and damping non-negative");
        return nullptr;
    }
    if (checkpoint_every == 0) {
        set_error("checkpoint_every must be at least 1");
        return nullptr;
    }
    const unsigned long long max_cells = cell_table_bytes / sizeof(TayCellRange);
    if (max_cells == 0) {
        set_error("cell table needs at least %zu bytes", sizeof(TayCellRange));
        return nullptr;
    }
    if (!stream())
        return nullptr;
    TayDflock *f = new TayDflock();   /* value-initialised: every pointer null */
    f->n = n;
    f->params = *params;
    f->every = checkpoint_every;
    f->stats_blocks = (unsigned)(((unsigned long long)n * LANES + 255u) / 256u);
    f->max_cells = max_cells;
    f->box_blocks = blocks_for(n, kBuildBlock * 4, 1024);
    const unsigned nblocks = (n + kRadixTile - 1) / kRadixTile;
    const size_t rows3 = 3ull * n * sizeof(float);
    bool ok = device_alloc((void **)&f->cols[0], state_bytes(f), "agents") &&
              device_alloc((void **)&f->cols[1], state_bytes(f), "agents") &&
              hip_ok(hipMemsetAsync(f->cols[0], 0, state_bytes(f), stream()), "hipMemsetAsync") &&
              hip_ok(hipMemsetAsync(f->cols[1], 0, state_bytes(f), stream()), "hipMemsetAsync") &&
              device_alloc((void **)&f->box_partials, sizeof(float) * kBoxPartial * f->box_blocks, "box partials") &&
              device_alloc((void **)&f->geom, sizeof(TayGridGeom), "geometry") &&
              device_alloc((void **)&f->keys[0], 4ull * n, "keys") && device_alloc((void **)&f->keys[1], 4ull * n, "keys") &&
              device_alloc((void **)&f->vals[0], 4ull * n, "agents by key") &&
              device_alloc((void **)&f->vals[1], 4ull * n, "agents by key") &&
              device_alloc((void **)&f->block_hist, 4ull * kRadix * nblocks, "radix histogram") &&
              device_alloc((void **)&f->scan_sums, 4ull * kScanBlocks, "scan sums") &&
              device_alloc((void **)&f->ranges, sizeof(TayCellRange) * max_cells, "cell table") &&
              device_alloc((void **)&f->max_count, 16, "neighbour maxima and pair total") &&
              device_alloc((void **)&f->stats[0], 16ull * f->stats_blocks, "forward statistics") &&
              device_alloc((void **)&f->stats[1], 16ull * f->stats_blocks, "recording statistics") &&
              device_alloc((void **)&f->seg_states, state_bytes(f) * checkpoint_every, "segment states") &&
              device_alloc((void **)&f->seg_ids, 4ull * n * checkpoint_every, "segment orders") &&
              device_alloc((void **)&f->seg_sums, 10ull * n * 4 * checkpoint_every, "segment sums") &&
              device_alloc((void **)&f->seg_srt, 6ull * n * 4 * checkpoint_every, "segment key-ordered states") &&
              device_alloc((void **)&f->seg_counts, 4ull * LANES * n * checkpoint_every, "segment counts") &&
              device_alloc((void **)&f->Lp[0], rows3, "adjoint") && device_alloc((void **)&f->Lp[1], rows3, "adjoint") &&
              device_alloc((void **)&f->Lv[0], rows3, "adjoint") && device_alloc((void **)&f->Lv[1], rows3, "adjoint") &&
              device_alloc((void **)&f->adj, 12ull * n * 4, "sum adjoints") &&
              device_alloc((void **)&f->gpart, 3ull * n * 8, "weight gradient parts") &&
              device_alloc((void **)&f->gw, 3 * 8, "weight gradient") &&
              device_alloc((void **)&f->srt, 6ull * n * 4, "key-ordered positions and velocities") &&
              device_alloc((void **)&f->gparts, 3ull * kReduceParts * 8, "weight gradient block sums") &&
              device_alloc((void **)&f->io, rows3, "transfer staging") &&
              hip_ok(hipMemsetAsync(f->geom, 0, sizeof(TayGridGeom), stream()), "hipMemsetAsync");
    if (!ok) {
        tay_dflock_destroy(f);
        return nullptr;
    }
    return f;
}

/* Folds the forward's (which = 0) or the recording's (1) per-block
 * statistics into max_count and copies them to out[4] (synchronises). */
static bool read_stats(TayDflock *f, int which, unsigned *out, hipStream_t s) {
    k_df_stats_reduce<<<1, 256, 0, s>>>(f->stats[which], f->stats_blocks, f->max_count);
    return launch_ok("statistics") &&
           hip_ok(hipMemcpyAsync(out, f->max_count, 16, hipMemcpyDeviceToHost, s), "read neighbour statistics") &&
           hip_ok(hipStreamSynchronize(s), "read neighbour statistics");
}

TAY_API int tay_dflock_forward(TayDflock *f, const float *p0, const float *v0, const float *weights,
                               unsigned steps, float *p_final, float *v_final) {
    hipStream_t s = stream();
    if (!s)
        return 0;
    const unsigned n = f->n, blocks = (n + 255u) / 256u;
    const unsigned lane_blocks = (unsigned)(((unsigned long long)n * LANES + 255u) / 256u);
    for (int k = 0; k < 3; ++k)
        f->weights[k] = weights[k];
    const unsigned segments = (steps + f->every - 1) / f->every;
    while (f->ckpt.size() < segments) {
        float *c = nullptr;
        if (!device_alloc((void **)&c, state_bytes(f), "checkpoint"))
            return 0;
        f->ckpt.push_back(c);
    }
    const size_t rows3 = 3ull * n * sizeof(float);
    bool ok = hip_ok(hipMemcpyAsync(f->io, p0, rows3, hipMemcpyHostToDevice, s), "upload p0");
    if (ok)
        k_df_unpack<<<blocks, 256, 0, s>>>(f->io, n, f->cols[0] + 2ull * n);
    ok = ok && hip_ok(hipMemcpyAsync(f->io, v0, rows3, hipMemcpyHostToDevice, s), "upload v0");
    if (ok)
        k_df_unpack<<<blocks, 256, 0, s>>>(f->io, n, f->cols[0] + 5ull * n);
    ok = ok && hip_ok(hipMemsetAsync(f->stats[0], 0, 16ull * f->stats_blocks, s), "hipMemsetAsync") &&
         hip_ok(hipStreamSynchronize(s), "upload initial state");
    if (!ok)
        return 0;
    hipEvent_t ev[2];
    if (!timed_begin(ev, s))
        return 0;
    const Prm c = make_prm(f);
    int cur = 0;
    for (unsigned t = 0; t < steps; ++t) {
        if (t % f->every == 0)
            ok = ok && hip_ok(hipMemcpyAsync(f->ckpt[t / f->every], f->cols[cur], state_bytes(f),
                                             hipMemcpyDeviceToDevice, s), "checkpoint");
        df_sort(f, f->cols[cur], s);
        k_df_sorted_copy<<<blocks, 256, 0, s>>>(f->cols[cur], n, f->vals[0], f->srt);
        k_df_step<false><<<lane_blocks, 256, 0, s>>>(f->cols[cur], n, f->vals[0], f->srt, f->ranges, f->geom, c,
                                                f->cols[cur ^ 1], nullptr, 0, nullptr, nullptr, f->stats[0]);
        cur ^= 1;
    }
    ok = ok && launch_ok("differentiable forward") && timed_end(ev, s, &f->forward_ms);
    f->steps = ok ? steps : 0;
    if (ok && cur) {
        ok = hip_ok(hipMemcpyAsync(f->cols[0], f->cols[1], state_bytes(f), hipMemcpyDeviceToDevice, s), "copy");
    }
    for (int which = 0; ok && which < 2; ++which) {
        float *dst = which == 0 ? p_final : v_final;
        if (!dst)
            continue;
        k_df_pack<<<blocks, 256, 0, s>>>(f->cols[0] + (which == 0 ? 2ull : 5ull) * n, n, f->io);
        ok = hip_ok(hipMemcpyAsync(dst, f->io, rows3, hipMemcpyDeviceToHost, s), "fetch state") &&
             hip_ok(hipStreamSynchronize(s), "fetch state");
    }
    return ok;
}

TAY_API int tay_dflock_backward(TayDflock *f, const float *grad_p_final, const float *grad_v_final,
                                float *grad_p0, float *grad_v0, float *grad_weights) {
    hipStream_t s = stream();
    if (!s)
        return 0;
    if (!f->steps) {
        set_error("backward needs a forward run first");
        return 0;
    }
    const unsigned n = f->n, blocks = (n + 255u) / 256u;
    const unsigned lane_blocks = (unsigned)(((unsigned long long)n * LANES + 255u) / 256u);
    const size_t rows3 = 3ull * n * sizeof(float);
    unsigned maxc[4] = {0u, 0u, 0u, 0u};
    bool ok = read_stats(f, 0, maxc, s);
    if (!ok)
        return 0;
    const unsigned KL = maxc[1] ? maxc[1] : 1u;
    const size_t need = (size_t)KL * LANES * n * f->every;
    if (need > f->seg_nbr_cap) {
        device_free(f->seg_nbr);
        f->seg_nbr = nullptr;
        f->seg_nbr_cap = 0;
        if (!device_alloc((void **)&f->seg_nbr, 4 * need, "neighbour lists"))
            return 0;
        f->seg_nbr_cap = need;
    }
    f->KL = KL;
    ok = hip_ok(hipMemcpyAsync(f->io, grad_p_final, rows3, hipMemcpyHostToDevice, s), "upload gradient");
    if (ok)
        k_df_unpack<<<blocks, 256, 0, s>>>(f->io, n, f->Lp[0]);
    if (grad_v_final) {
        ok = ok && hip_ok(hipMemcpyAsync(f->io, grad_v_final, rows3, hipMemcpyHostToDevice, s), "upload gradient");
        if (ok)
            k_df_unpack<<<blocks, 256, 0, s>>>(f->io, n, f->Lv[0]);
    } else {
        ok = ok && hip_ok(hipMemsetAsync(f->Lv[0], 0, rows3, s), "hipMemsetAsync");
    }
    ok = ok && hip_ok(hipMemsetAsync(f->gw, 0, 24, s), "hipMemsetAsync") &&
         hip_ok(hipMemsetAsync(f->stats[1], 0, 16ull * f->stats_blocks, s), "hipMemsetAsync") &&
         hip_ok(hipStreamSynchronize(s), "upload gradient");
    if (!ok)
        return 0;
    hipEvent_t ev[2];
    if (!timed_begin(ev, s))
        return 0;
    const Prm c = make_prm(f);
    int L = 0;
    const unsigned segments = (f->steps + f->every - 1) / f->every;
    for (unsigned seg = segments; ok && seg-- > 0;) {
        const unsigned t0 = seg * f->every;
        const unsigned t1 = t0 + f->every < f->steps ? t0 + f->every : f->steps;
        /* recompute the segment from its checkpoint, recording every step */
        ok = hip_ok(hipMemcpyAsync(f->seg_states, f->ckpt[seg], state_bytes(f), hipMemcpyDeviceToDevice, s),
                    "restore checkpoint");
        for (unsigned t = t0; ok && t < t1; ++t) {
            const unsigned r = t - t0;
            float *x = f->seg_states + (size_t)r * kWords * n;
            df_sort(f, x, s);
            ok = hip_ok(hipMemcpyAsync(f->seg_ids + (size_t)r * n, f->vals[0], 4ull * n, hipMemcpyDeviceToDevice, s),
                        "record order");
            float *next = t + 1 < t1 ? x + (size_t)kWords * n : f->cols[1];
            float *srt = f->seg_srt + (size_t)r * 6 * n;
            k_df_sorted_copy<<<blocks, 256, 0, s>>>(x, n, f->vals[0], srt);
            k_df_step<true><<<lane_blocks, 256, 0, s>>>(x, n, f->vals[0], srt, f->ranges, f->geom, c, next,
                                                   f->seg_nbr + (size_t)r * KL * LANES * n, KL,
                                                   f->seg_counts + (size_t)r * LANES * n,
                                                   f->seg_sums + (size_t)r * 10 * n, f->stats[1]);
        }
        /* reverse the segment's steps */
        for (unsigned t = t1; ok && t-- > t0;) {
            const unsigned r = t - t0;
            const unsigned *ids = f->seg_ids + (size_t)r * n;
            k_df_adj_local<<<blocks, 256, 0, s>>>(n, ids, f->seg_sums + (size_t)r * 10 * n, c, f->Lp[L], f->Lv[L],
                                                  f->Lp[L ^ 1], f->Lv[L ^ 1], f->adj, f->gpart);
            k_df_reduce<<<kReduceParts, kReduceBlock, 0, s>>>(n, f->gpart, f->gparts);
            k_df_reduce_parts<<<1, 64, 0, s>>>(f->gparts, f->gw);
            k_df_adj_gather<<<lane_blocks, 256, 0, s>>>(n, ids, f->seg_srt + (size_t)r * 6 * n,
                                                   f->seg_nbr + (size_t)r * KL * LANES * n, KL,
                                                   f->seg_counts + (size_t)r * LANES * n, f->adj, c, f->Lp[L ^ 1],
                                                   f->Lv[L ^ 1]);
            L ^= 1;
        }
        ok = ok && launch_ok("differentiable backward");
    }
    ok = ok && timed_end(ev, s, &f->backward_ms);
    unsigned recount[4] = {0u, 0u, 0u, 0u};
    ok = ok && read_stats(f, 1, recount, s);
    if (ok && recount[1] > KL) {
        set_error("recomputed per-lane neighbour count %u exceeds the forward's %u", recount[1], KL);
        return 0;
    }
    double gw[3] = {0.0, 0.0, 0.0};
    for (int which = 0; ok && which < 2; ++which) {
        float *dst = which == 0 ? grad_p0 : grad_v0;
        if (!dst)
            continue;
        k_df_pack<<<blocks, 256, 0, s>>>(which == 0 ? f->Lp[L] : f->Lv[L], n, f->io);
        ok = hip_ok(hipMemcpyAsync(dst, f->io, rows3, hipMemcpyDeviceToHost, s), "fetch gradient") &&
             hip_ok(hipStreamSynchronize(s), "fetch gradient");
    }
    ok = ok && hip_ok(hipMemcpyAsync(gw, f->gw, 24, hipMemcpyDeviceToHost, s), "fetch weight gradient") &&
         hip_ok(hipStreamSynchronize(s), "fetch weight gradient");
    if (ok && grad_weights)
        for (int k = 0; k < 3; ++k)
            grad_weights[k] = (float)gw[k];
    return ok;
}

TAY_API int tay_dflock_times(TayDflock *f, double *forward_ms, double *backward_ms) {
    if (forward_ms)
        *forward_ms = f->forward_ms;
    if (backward_ms)
        *backward_ms = f->backward_ms;
    return 1;
}

TAY_API int tay_dflock_max_neighbours(TayDflock *f, unsigned *count) {
    hipStream_t s = stream();
    unsigned st[4];
    if (!s || !read_stats(f, 0, st, s))
        return 0;
    *count = st[0];
    return 1;
}

TAY_API int tay_dflock_pair_count(TayDflock *f, unsigned long long *pairs) {
    hipStream_t s = stream();
    unsigned st[4];
    if (!s || !read_stats(f, 0, st, s))
        return 0;
    std::memcpy(pairs, st + 2, sizeof(*pairs));
    return 1;
}
