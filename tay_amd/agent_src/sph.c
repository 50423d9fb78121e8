/* SPH model; behaviour described in sph.h. */

void sph_density(global SphParticle *a, global SphParticle *b, constant SphContext *c) {
    float dx = b->p.x - a->p.x;
    float dy = b->p.y - a->p.y;
    float dz = b->p.z - a->p.z;
    float q = c->h2 - (dx * dx + dy * dy + dz * dz);
    if (q > 0.0f)
        a->density += c->poly6 * q * q * q;
}

void sph_pressure(global SphParticle *a, constant SphContext *c) {
    float x = a->density / c->density;
    float x2 = x * x;
    float x4 = x2 * x2;
    a->pressure = c->K * (x4 * x2 * x - 1.0f);
}

void sph_forces(global SphParticle *a, global SphParticle *b, constant SphContext *c) {
    float4 r = float4_sub(b->p, a->p);
    float rl = float4_length(r);
    if (rl >= c->h)
        return;

    float w = c->h - rl;
    if (rl > 0.00001f) {
        float grad = c->spiky * w * w;
        float s = grad * (a->pressure + b->pressure) / (2.0f * b->density) / rl;
        a->pressure_accum = float4_add(a->pressure_accum, float4_mul_scalar(r, s));
    }
    float4 dv = float4_sub(b->v, a->v);
    a->viscosity_accum = float4_add(a->viscosity_accum, float4_mul_scalar(dv, c->viscosity * w / b->density));
}

void sph_density_clear(global SphParticle *a) {
    a->density = 0.0f;
}

void sph_density_combine(global SphParticle *a, global SphParticle *b) {
    a->density += b->density;
}

void sph_forces_clear(global SphParticle *a) {
    a->pressure_accum.x = 0.0f;
    a->pressure_accum.y = 0.0f;
    a->pressure_accum.z = 0.0f;
    a->viscosity_accum.x = 0.0f;
    a->viscosity_accum.y = 0.0f;
    a->viscosity_accum.z = 0.0f;
}

void sph_forces_combine(global SphParticle *a, global SphParticle *b) {
    a->pressure_accum.x += b->pressure_accum.x;
    a->pressure_accum.y += b->pressure_accum.y;
    a->pressure_accum.z += b->pressure_accum.z;
    a->viscosity_accum.x += b->viscosity_accum.x;
    a->viscosity_accum.y += b->viscosity_accum.y;
    a->viscosity_accum.z += b->viscosity_accum.z;
}

void sph_integrate(global SphParticle *a, constant SphContext *c) {
    float4 f = float4_add(a->pressure_accum, float4_mul_scalar(a->viscosity_accum, c->dynamic_viscosity));
    float4 acc = float4_div_scalar(f, a->density);
    acc.z -= 9.81f;

    a->v = float4_add(a->v, float4_mul_scalar(acc, c->dt));
    a->vh = a->v;
    a->p = float4_add(a->p, float4_mul_scalar(a->v, c->dt));

    const float damp = -0.5f;
    if (a->p.x < c->min.x) { a->p.x = c->min.x; a->v.x *= damp; a->vh.x *= damp; }
    if (a->p.y < c->min.y) { a->p.y = c->min.y; a->v.y *= damp; a->vh.y *= damp; }
    if (a->p.z < c->min.z) { a->p.z = c->min.z; a->v.z *= damp; a->vh.z *= damp; }
    if (a->p.x > c->max.x) { a->p.x = c->max.x; a->v.x *= damp; a->vh.x *= damp; }
    if (a->p.y > c->max.y) { a->p.y = c->max.y; a->v.y *= damp; a->vh.y *= damp; }
    if (a->p.z > c->max.z) { a->p.z = c->max.z; a->v.z *= damp; a->vh.z *= damp; }

    a->density = 0.0f;
    a->pressure_accum = float4_null();
    a->viscosity_accum = float4_null();
}

void sph_particle_density(global SphParticle *a, global SphParticle *b, constant SphContext *c) {
    sph_density(a, b, c);
}

void sph_force_terms(global SphParticle *a, global SphParticle *b, constant SphContext *c) {
    sph_forces(a, b, c);
}

void sph_particle_pressure(global SphParticle *a, constant SphContext *c) {
    double ratio = (double)(a->density / c->density);
    a->pressure = c->K * ((float)pow(ratio, 7.0) - 1.0f);
}

/* one axis of the walls: a coordinate below lo, then one above hi, is put
 * on the wall and both velocities of the axis are scaled by damp (axes are
 * independent, so doing them one after another keeps every result) */
static void sph_wall(global float *p, global float *v, global float *vh, float lo, float hi, float damp) {
    if (*p < lo) { *p = lo; *v = damp * *v; *vh = damp * *vh; }
    if (*p > hi) { *p = hi; *v = damp * *v; *vh = damp * *vh; }
}

void sph_particle_leapfrog(global SphParticle *a, constant SphContext *c) {
    float4 f = float4_add(a->pressure_accum, float4_mul_scalar(a->viscosity_accum, c->dynamic_viscosity));
    float4 acc = float4_div_scalar(f, a->density);
    acc.z -= 9.81f;

    float4 vh = float4_add(a->vh, float4_mul_scalar(acc, c->dt));
    a->vh = vh;
    a->v = float4_add(vh, float4_mul_scalar(acc, c->dt * 0.5f));
    a->p = float4_add(a->p, float4_mul_scalar(vh, c->dt));

    const float damp = -0.5f;
    sph_wall(&a->p.x, &a->v.x, &a->vh.x, c->min.x, c->max.x, damp);
    sph_wall(&a->p.y, &a->v.y, &a->vh.y, c->min.y, c->max.y, damp);
    sph_wall(&a->p.z, &a->v.z, &a->vh.z, c->min.z, c->max.z, damp);

    sph_particle_reset(a);
}

void sph_particle_reset(global SphParticle *a) {
    a->density = 0.0f;
    a->pressure_accum.x = 0.0f;
    a->pressure_accum.y = 0.0f;
    a->pressure_accum.z = 0.0f;
    a->viscosity_accum.x = 0.0f;
    a->viscosity_accum.y = 0.0f;
    a->viscosity_accum.z = 0.0f;
}
