/*
 * tay_dflock.h - differentiable flocking on the GPU (libtay_amd.so): a
 * forward/backward pair over the smooth flocking model with hand-written
 * adjoint kernels and checkpointed recomputation.
 *
 * No reference counterpart: the reference has no differentiable path.  The
 * model is the smooth variant of its flocking rule (m_flocking/agent.c:2-33)
 * with a cubic cutoff weight w = (1 - r^2/R^2)^3 in place of the hard radius
 * test, a softened separation term and no speed clamp, so every step is
 * twice continuously differentiable:
 *     W_i = sum_j w,  A_i = sum_j w (v_j - v_i),  C_i = sum_j w d,
 *     S_i = -sum_j w d / (r^2 + h2)         (d = p_j - p_i, r < R, j != i)
 *     a_i = (ws S_i + wa A_i + wc C_i) / (W_i + 1)
 *     v_i' = v_i + dt (a_i - damping v_i),  p_i' = p_i + dt v_i'
 * The same model in Tay's model dialect is smooth_see / smooth_act
 * (tay_amd/agent_src/flocking.c), so it also runs through tay.h.
 *
 * Arrays are n x 3 floats (x, y, z per agent) in host memory; every call
 * returns 1 on success and 0 with tay_gpu_last_error() set.
 */
#ifndef tay_dflock_h
#define tay_dflock_h

#ifdef __cplusplus
extern "C" {
#endif

typedef struct TayDflock TayDflock;

typedef struct TayDflockParams {
    float radius;       /* R, the smooth cutoff */
    float h2;           /* separation softening */
    float dt;
    float damping;
    float cell_size;    /* Z-grid cell size of the neighbour search */
} TayDflockParams;

/* n agents; a checkpoint every checkpoint_every steps; cell_table_bytes
 * bounds the neighbour-search grid (8 bytes per cell). */
TayDflock *tay_dflock_create(unsigned n, const TayDflockParams *params, unsigned checkpoint_every,
                             unsigned long long cell_table_bytes);
void tay_dflock_destroy(TayDflock *flock);

/* Runs `steps` steps from (p0, v0) with weights (separation, alignment,
 * cohesion), keeping the checkpoints the backward pass needs; p_final and
 * v_final (each may be NULL) receive the final state. */
int tay_dflock_forward(TayDflock *flock, const float *p0, const float *v0, const float *weights, unsigned steps,
                       float *p_final, float *v_final);

/* Gradient of a scalar loss of the final state, given its gradient with
 * respect to the final positions and velocities (grad_v_final may be NULL
 * for zero), with respect to the initial positions, the initial velocities
 * and the three weights (outputs may be NULL).  Recomputes each segment
 * between checkpoints from its checkpoint. */
int tay_dflock_backward(TayDflock *flock, const float *grad_p_final, const float *grad_v_final, float *grad_p0,
                        float *grad_v0, float *grad_weights);

/* Device time of the last forward and backward call, in milliseconds. */
int tay_dflock_times(TayDflock *flock, double *forward_ms, double *backward_ms);

/* Largest neighbour count of any agent in any step of the last call. */
int tay_dflock_max_neighbours(TayDflock *flock, unsigned *count);

/* Neighbour pairs (r^2 < R^2, each agent's own count, self excluded) summed
 * over the steps of the last forward call: the pair interactions of the
 * forward pass. */
int tay_dflock_pair_count(TayDflock *flock, unsigned long long *pairs);

#ifdef __cplusplus
}
#endif

#endif
