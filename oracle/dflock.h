/* dflock.h - TEST INFRASTRUCTURE ONLY: FP64 smooth flocking and its
 * gradient (dflock.c).  Positions and velocities are n x 3 doubles. */
#ifndef ORACLE_DFLOCK_H
#define ORACLE_DFLOCK_H

typedef struct DflockParams {
    double radius;      /* smooth cutoff R */
    double h2;          /* separation softening, q = r^2 + h2 */
    double dt;
    double damping;
} DflockParams;

/* Final state after `steps` steps; returns 1 on success. */
int oracle_dflock_forward(unsigned n, const double *p0, const double *v0, const double *weights,
                          const DflockParams *params, unsigned steps, double *p_out, double *v_out);

/* Loss sum |p_T - target|^2 / (3 n) and its gradient with respect to the
 * initial positions, the initial velocities and the weights
 * (separation, alignment, cohesion). */
int oracle_dflock_gradient(unsigned n, const double *p0, const double *v0, const double *weights,
                           const DflockParams *params, unsigned steps, const double *target, double *loss,
                           double *g_p0, double *g_v0, double *g_weights);

/* The same gradient for any loss, given its gradient with respect to the
 * final positions and velocities (gv_final may be NULL for zero). */
int oracle_dflock_backprop(unsigned n, const double *p0, const double *v0, const double *weights,
                           const DflockParams *params, unsigned steps, const double *gp_final, const double *gv_final,
                           double *g_p0, double *g_v0, double *g_weights);

/* The same, differentiating along a given trajectory (steps + 1 states of
 * n x 3 positions and velocities, state 0 first) instead of the one this
 * model computes: the FP64 adjoint of another implementation's trajectory. */
int oracle_dflock_backprop_along(unsigned n, const double *states_p, const double *states_v, const double *weights,
                                 const DflockParams *params, unsigned steps, const double *gp_final,
                                 const double *gv_final, double *g_p0, double *g_v0, double *g_weights);

#endif
