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

void boid_see(global Boid *a, global Boid *b, constant BoidSeeContext *c);
void boid_act(global Boid *a, constant BoidActContext *c);
void flock_see(global FlockAgent *a, global FlockAgent *b, constant FlockSeeContext *c);
void flock_act(global FlockAgent *a, constant FlockActContext *c);
