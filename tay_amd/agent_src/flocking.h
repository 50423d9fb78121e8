/* Flocking (boids) models in Tay's model dialect.
 *
 * Boid / boid_see / boid_act: the flocking model of the reference's
 * m_flocking plug-in (m_flocking/agent.h:2-13, m_flocking/agent.c:2-79), same
 * record layout.  see: within Euclidean distance r and in front of the boid
 * (dot(dir, d) > 0) accumulate the neighbour's heading (alignment), the
 * offset (cohesion) and, closer than separation_r, a push away whose length
 * is separation_r - distance.  act: pull towards the origin beyond radius
 * 400, steer with weights 0.02 / 0.02 / 0.001, renormalise the heading and
 * advance by speed.
 *
 * FlockAgent / flock_see / flock_act: the weighted, integrated variant used
 * for the large runs.  The see pass is the same neighbourhood rule on the
 * velocity instead of the heading; the act pass forms
 *   acc = w_sep * separation / n_sep + w_ali * alignment / n + w_coh * cohesion / n
 * then v += acc * dt, clamps |v| to max_speed and moves p += v * dt. */

typedef struct __attribute__((packed)) Boid {
    TayAgentTag tag;
    float4 p;
    float4 dir;
    unsigned color;
    float speed;
    float4 separation;
    float4 alignment;
    float4 cohesion;
    int cohesion_count;
    int separation_count;
} Boid;

typedef struct __attribute__((packed)) BoidSeeContext {
    float r;
    float separation_r;
} BoidSeeContext;

typedef struct __attribute__((packed)) BoidActContext {
    int dummy;
} BoidActContext;

typedef struct __attribute__((packed)) FlockAgent {
    TayAgentTag tag;
    float4 p;
    float4 v;
    float4 separation;
    float4 alignment;
    float4 cohesion;
    int separation_count;
    int cohesion_count;
    int result_index;
} FlockAgent;

typedef struct __attribute__((packed)) FlockSeeContext {
    float r;
    float separation_r;
} FlockSeeContext;

typedef struct __attribute__((packed)) FlockActContext {
    float separation_w;
    float alignment_w;
    float cohesion_w;
    float dt;
    float max_speed;
} FlockActContext;

/* SmoothAgent / smooth_see / smooth_act: the differentiable variant
 * (include/tay_dflock.h): cubic cutoff weight w = (1 - r^2 / R^2)^3 for
 * r^2 < R^2, softened separation, no front filter and no speed clamp.
 *   see: weight_sum += w, alignment += w (v_b - v_a), cohesion += w d,
 *        separation -= w d / (r^2 + h2)                      (d = p_b - p_a)
 *   act: a = (ws S + wa A + wc C) / (weight_sum + 1),
 *        v += dt (a - damping v), p += dt v, clear the sums */
typedef struct __attribute__((packed)) SmoothAgent {
    TayAgentTag tag;
    float4 p;
    float4 v;
    float4 separation;
    float4 alignment;
    float4 cohesion;
    float weight_sum;
    int result_index;
} SmoothAgent;

typedef struct __attribute__((packed)) SmoothSeeContext {
    float r2;           /* R^2 */
    float inv_r2;       /* 1 / R^2 */
    float h2;
} SmoothSeeContext;

typedef struct __attribute__((packed)) SmoothActContext {
    float separation_w;
    float alignment_w;
    float cohesion_w;
    float dt;
    float damping;
} SmoothActContext;

void boid_see(global Boid *a, global Boid *b, constant BoidSeeContext *c);
void boid_act(global Boid *a, constant BoidActContext *c);
void flock_see(global FlockAgent *a, global FlockAgent *b, constant FlockSeeContext *c);
void flock_act(global FlockAgent *a, constant FlockActContext *c);
/* For a split see pass (tay_set_see_split): flock_see only adds to the
 * sums and counts, so a partial of a seer starts with them emptied
 * (flock_see_clear) and partials merge by adding them (flock_see_combine,
 * a += b). */
void flock_see_clear(global FlockAgent *a);
void flock_see_combine(global FlockAgent *a, global FlockAgent *b);
void smooth_see(global SmoothAgent *a, global SmoothAgent *b, constant SmoothSeeContext *c);
void smooth_act(global SmoothAgent *a, constant SmoothActContext *c);
