/* Smoothed-particle hydrodynamics model in Tay's model dialect, after the
 * reference's m_sph plug-in (m_sph/agent.h:2-32, m_sph/agent.c): same
 * particle record and context layout, Mueller-Charypar-Gross kernels.
 *
 *   sph_density   (see, self included)  density += poly6 * (h^2 - r^2)^3 for r^2 < h^2
 *   sph_pressure  (act)                 Tait equation, gamma 7:
 *                                       pressure = K * ((density / rest_density)^7 - 1)
 *   sph_forces    (see, self excluded)  spiky pressure gradient and viscosity
 *                                       Laplacian terms for r < h
 *   sph_integrate (act)                 symplectic Euler: v += a dt, p += v dt,
 *                                       gravity -9.81 on z, clamp to the walls
 *                                       with damping -0.5, clear the sums
 *
 * The seventh power is three multiplies (x^2, x^4, x^7 = x^4 x^2 x), not
 * libm pow, so host and device agree to the last bit. */

typedef struct __attribute__((packed)) SphParticle {
    TayAgentTag tag;
    float4 p;
    float4 pressure_accum;
    float4 viscosity_accum;
    float4 vh;
    float4 v;
    float density;
    float pressure;
} SphParticle;

typedef struct __attribute__((packed)) SphContext {
    float h;
    float dynamic_viscosity;
    float dt;
    float m;
    float surface_tension;
    float surface_tension_threshold;
    float K;
    float density;

    float4 min;
    float4 max;

    float h2;
    float poly6;
    float poly6_gradient;
    float poly6_laplacian;
    float spiky;
    float viscosity;
} SphContext;

void sph_density(global SphParticle *a, global SphParticle *b, constant SphContext *c);
void sph_pressure(global SphParticle *a, constant SphContext *c);
void sph_forces(global SphParticle *a, global SphParticle *b, constant SphContext *c);
void sph_integrate(global SphParticle *a, constant SphContext *c);
/* For split see passes (tay_set_see_split): sph_density only adds to the
 * density and sph_forces only to the x, y, z of the two accumulators (it
 * reads the seer's pressure, which a partial keeps), so a partial starts
 * with those emptied (*_clear) and partials merge by adding them
 * (*_combine, a += b); the accumulators' w is left alone. */
void sph_density_clear(global SphParticle *a);
void sph_density_combine(global SphParticle *a, global SphParticle *b);
void sph_forces_clear(global SphParticle *a);
void sph_forces_combine(global SphParticle *a, global SphParticle *b);

/* The reference plug-in's own pass functions, under its names
 * (m_sph/agent.h:34-38, m_sph/main.c:70-74):
 *   sph_particle_density   = sph_density
 *   sph_force_terms        = sph_forces
 *   sph_particle_pressure  Tait equation with the double-precision pow of
 *                          the density ratio, rounded to float
 *   sph_particle_leapfrog  leapfrog: vh += a dt, v = vh + a dt/2, p += vh dt,
 *                          walls as sph_integrate, then sph_particle_reset
 *   sph_particle_reset     clears density and the xyz of both accumulators */
void sph_particle_density(global SphParticle *a, global SphParticle *b, constant SphContext *c);
void sph_particle_pressure(global SphParticle *a, constant SphContext *c);
void sph_force_terms(global SphParticle *a, global SphParticle *b, constant SphContext *c);
void sph_particle_leapfrog(global SphParticle *a, constant SphContext *c);
void sph_particle_reset(global SphParticle *a);
