/* Flocking models; behaviour described in flocking.h. */

void boid_see(global Boid *a, global Boid *b, constant BoidSeeContext *c) {
    float4 d = float4_sub(float4_agent_position(b), float4_agent_position(a));
    float dl = float4_length(d);
    if (dl >= c->r)
        return;
    if (float4_dot(a->dir, d) > 0.0f) {
        a->alignment = float4_add(a->alignment, b->dir);
        if (dl < c->separation_r) {
            a->separation = float4_sub(a->separation, float4_normalize_to(d, c->separation_r - dl));
            a->separation_count += 1;
        }
        a->cohesion = float4_add(a->cohesion, d);
        a->cohesion_count += 1;
    }
}

void boid_act(global Boid *a, constant BoidActContext *c) {
    float4 p = float4_agent_position(a);
    float4 acc = float4_null();

    /* weak pull back towards the origin outside radius 400 */
    float dist = float4_length(p);
    if (dist > 400.0f)
        acc = float4_sub(acc, float4_mul_scalar(p, (dist - 400.0f) * 0.00001f / dist));

    if (a->cohesion_count)
        acc = float4_add(acc, float4_mul_scalar(a->alignment, 0.02f / (float)a->cohesion_count));
    if (a->separation_count)
        acc = float4_add(acc, float4_mul_scalar(a->separation, 0.02f / (float)a->separation_count));
    if (a->cohesion_count)
        acc = float4_add(acc, float4_mul_scalar(a->cohesion, 0.001f / (float)a->cohesion_count));

    a->dir = float4_normalize(float4_add(a->dir, acc));
    float4_agent_position(a) = float4_add(p, float4_mul_scalar(a->dir, a->speed));

    a->separation = float4_null();
    a->alignment = float4_null();
    a->cohesion = float4_null();
    a->cohesion_count = 0;
    a->separation_count = 0;
}

void flock_see(global FlockAgent *a, global FlockAgent *b, constant FlockSeeContext *c) {
    float4 d = float4_sub(b->p, a->p);
    float dl = float4_length(d);
    if (dl >= c->r)
        return;
    if (float4_dot(a->v, d) > 0.0f) {
        a->alignment = float4_add(a->alignment, b->v);
        if (dl < c->separation_r) {
            a->separation = float4_sub(a->separation, float4_normalize_to(d, c->separation_r - dl));
            a->separation_count += 1;
        }
        a->cohesion = float4_add(a->cohesion, d);
        a->cohesion_count += 1;
    }
}

void flock_see_clear(global FlockAgent *a) {
    a->separation = float4_null();
    a->alignment = float4_null();
    a->cohesion = float4_null();
    a->separation_count = 0;
    a->cohesion_count = 0;
}

void flock_see_combine(global FlockAgent *a, global FlockAgent *b) {
    a->separation = float4_add(a->separation, b->separation);
    a->alignment = float4_add(a->alignment, b->alignment);
    a->cohesion = float4_add(a->cohesion, b->cohesion);
    a->separation_count += b->separation_count;
    a->cohesion_count += b->cohesion_count;
}

void flock_act(global FlockAgent *a, constant FlockActContext *c) {
    float4 acc = float4_null();
    if (a->separation_count)
        acc = float4_add(acc, float4_mul_scalar(a->separation, c->separation_w / (float)a->separation_count));
    if (a->cohesion_count) {
        acc = float4_add(acc, float4_mul_scalar(a->alignment, c->alignment_w / (float)a->cohesion_count));
        acc = float4_add(acc, float4_mul_scalar(a->cohesion, c->cohesion_w / (float)a->cohesion_count));
    }

    float4 v = float4_add(a->v, float4_mul_scalar(acc, c->dt));
    float speed = float4_length(v);
    if (speed > c->max_speed)
        v = float4_mul_scalar(v, c->max_speed / speed);
    a->v = v;
    a->p = float4_add(a->p, float4_mul_scalar(v, c->dt));

    a->separation = float4_null();
    a->alignment = float4_null();
    a->cohesion = float4_null();
    a->separation_count = 0;
    a->cohesion_count = 0;
}

void smooth_see(global SmoothAgent *a, global SmoothAgent *b, constant SmoothSeeContext *c) {
    float dx = b->p.x - a->p.x, dy = b->p.y - a->p.y, dz = b->p.z - a->p.z;
    float r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < c->r2))
        return;
    float t = 1.0f - r2 * c->inv_r2;
    float w = t * t * t;
    float q = r2 + c->h2;
    a->weight_sum += w;
    a->alignment.x += w * (b->v.x - a->v.x);
    a->cohesion.x += w * dx;
    a->separation.x -= w * dx / q;
    a->alignment.y += w * (b->v.y - a->v.y);
    a->cohesion.y += w * dy;
    a->separation.y -= w * dy / q;
    a->alignment.z += w * (b->v.z - a->v.z);
    a->cohesion.z += w * dz;
    a->separation.z -= w * dz / q;
}

void smooth_act(global SmoothAgent *a, constant SmoothActContext *c) {
    float D = a->weight_sum + 1.0f;
    float ax = (c->separation_w * a->separation.x + c->alignment_w * a->alignment.x + c->cohesion_w * a->cohesion.x) / D;
    float ay = (c->separation_w * a->separation.y + c->alignment_w * a->alignment.y + c->cohesion_w * a->cohesion.y) / D;
    float az = (c->separation_w * a->separation.z + c->alignment_w * a->alignment.z + c->cohesion_w * a->cohesion.z) / D;
    a->v.x = a->v.x + c->dt * (ax - c->damping * a->v.x);
    a->v.y = a->v.y + c->dt * (ay - c->damping * a->v.y);
    a->v.z = a->v.z + c->dt * (az - c->damping * a->v.z);
    a->p.x = a->p.x + c->dt * a->v.x;
    a->p.y = a->p.y + c->dt * a->v.y;
    a->p.z = a->p.z + c->dt * a->v.z;
    a->separation = float4_null();
    a->alignment = float4_null();
    a->cohesion = float4_null();
    a->weight_sum = 0.0f;
}
