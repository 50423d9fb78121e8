/* Test models; behaviour described in test_agent.h. */

void agent_see(global Agent *a, global Agent *b, constant SeeContext *c) {
    float4 offset = float4_sub(float4_agent_position(b), float4_agent_position(a));
    a->b_buffer = float4_add(a->b_buffer, offset);
    a->b_buffer_count += 1;
}

/* reverse the velocity component of any axis that left [lo, hi] */
static void reflect_point(global float4 *p, global float4 *v, float4 lo, float4 hi) {
    if (p->x < lo.x) { p->x = lo.x; v->x = -v->x; }
    if (p->y < lo.y) { p->y = lo.y; v->y = -v->y; }
    if (p->z < lo.z) { p->z = lo.z; v->z = -v->z; }
    if (p->x > hi.x) { p->x = hi.x; v->x = -v->x; }
    if (p->y > hi.y) { p->y = hi.y; v->y = -v->y; }
    if (p->z > hi.z) { p->z = hi.z; v->z = -v->z; }
}

void agent_act(global Agent *a, constant ActContext *c) {
    if (a->b_buffer_count != 0) {
        float4 mean = float4_div_scalar(a->b_buffer, (float)a->b_buffer_count);
        a->f_buffer = float4_add(a->f_buffer, mean);
        a->b_buffer = float4_null();
        a->b_buffer_count = 0;
    }
    float4_agent_position(a) = float4_add(float4_agent_position(a), a->v);
    reflect_point(&a->p, &a->v, c->min, c->max);
}

void box_agent_see(global BoxAgent *a, global BoxAgent *b, constant SeeContext *c) {
    float4 offset = float4_sub(float4_agent_min(b), float4_agent_min(a));
    a->b_buffer = float4_add(a->b_buffer, offset);
    a->b_buffer_count += 1;
}

void box_agent_act(global BoxAgent *a, constant ActContext *c) {
    if (a->b_buffer_count != 0) {
        float4 mean = float4_div_scalar(a->b_buffer, (float)a->b_buffer_count);
        a->f_buffer = float4_add(a->f_buffer, mean);
        a->b_buffer = float4_null();
        a->b_buffer_count = 0;
    }
    float4 size = float4_sub(a->max, a->min);
    a->min = float4_add(a->min, a->v);
    a->max = float4_add(a->max, a->v);
    /* keep the box inside [min, max], preserving its size */
    if (a->min.x < c->min.x) { a->min.x = c->min.x; a->max.x = c->min.x + size.x; a->v.x = -a->v.x; }
    if (a->min.y < c->min.y) { a->min.y = c->min.y; a->max.y = c->min.y + size.y; a->v.y = -a->v.y; }
    if (a->min.z < c->min.z) { a->min.z = c->min.z; a->max.z = c->min.z + size.z; a->v.z = -a->v.z; }
    if (a->max.x > c->max.x) { a->max.x = c->max.x; a->min.x = c->max.x - size.x; a->v.x = -a->v.x; }
    if (a->max.y > c->max.y) { a->max.y = c->max.y; a->min.y = c->max.y - size.y; a->v.y = -a->v.y; }
    if (a->max.z > c->max.z) { a->max.z = c->max.z; a->min.z = c->max.z - size.z; a->v.z = -a->v.z; }
}

void point_box_see(global Agent *a, global BoxAgent *b, constant SeeContext *c) {
    a->b_buffer.x += b->min.x - a->p.x;
    a->b_buffer.y += b->min.y - a->p.y;
    a->b_buffer.z += b->min.z - a->p.z;
    a->b_buffer_count += 1;
}

void box_point_see(global BoxAgent *a, global Agent *b, constant SeeContext *c) {
    a->b_buffer.x += b->p.x - a->min.x;
    a->b_buffer.y += b->p.y - a->min.y;
    a->b_buffer.z += b->p.z - a->min.z;
    a->b_buffer_count += 1;
}

void plain_see(global PlainAgent *a, global PlainAgent *b, constant PlainSeeContext *c) {
    float dx = b->p.x - a->p.x;
    float dy = b->p.y - a->p.y;
    float dz = b->p.z - a->p.z;
    if (dx * dx + dy * dy + dz * dz <= c->r2) {
        a->nsum.x += b->p.x;
        a->nsum.y += b->p.y;
        a->nsum.z += b->p.z;
        a->ncount += 1u;
    }
}

void plain_act(global PlainAgent *a, constant ActContext *c) {
    a->last_nsum = a->nsum;
    a->last_ncount = a->ncount;
    a->nsum = float4_null();
    a->ncount = 0u;
    a->p.x += a->v.x;
    a->p.y += a->v.y;
    a->p.z += a->v.z;
    reflect_point(&a->p, &a->v, c->min, c->max);
}

void plain_see_clear(global PlainAgent *a) {
    a->nsum.x = 0.0f;
    a->nsum.y = 0.0f;
    a->nsum.z = 0.0f;
    a->ncount = 0u;
}

void plain_see_combine(global PlainAgent *a, global PlainAgent *b) {
    a->nsum.x += b->nsum.x;
    a->nsum.y += b->nsum.y;
    a->nsum.z += b->nsum.z;
    a->ncount += b->ncount;
}

void agent_see_clear(global Agent *a) {
    a->b_buffer = float4_null();
    a->b_buffer_count = 0;
}

void agent_see_combine(global Agent *a, global Agent *b) {
    a->b_buffer = float4_add(a->b_buffer, b->b_buffer);
    a->b_buffer_count += b->b_buffer_count;
}

void box_agent_see_clear(global BoxAgent *a) {
    a->b_buffer = float4_null();
    a->b_buffer_count = 0;
}

void box_agent_see_combine(global BoxAgent *a, global BoxAgent *b) {
    a->b_buffer = float4_add(a->b_buffer, b->b_buffer);
    a->b_buffer_count += b->b_buffer_count;
}
