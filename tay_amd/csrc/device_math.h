/* Device math shared by the library's kernels and the model dialect of the
 * JIT programs.
 *
 * tay_sqrt_rn: the correctly rounded float square root, bit-identical to
 * the IEEE operation (and so to the host's sqrtf and to (float)sqrt(double)).
 * For inputs in [2^-96, +inf) the hardware square root (v_sqrt_f32, within
 * one ulp) is corrected to the nearest float by the signs of the exact
 * residuals x - s'*s of its two neighbours s' (fma, no rounding of the
 * residual): x <= s_down*s means sqrt(x) lies below the midpoint of s_down
 * and s, x > s_up*s means it lies above the midpoint of s and s_up.  Every
 * other input (zero, denormals and tiny normals, whose residuals would lose
 * bits, infinities, negatives, NaN) takes the compiler's full IEEE square
 * root.  The identity with the IEEE result is checked over all 2^32 inputs
 * by tay_gpu_sqrt_selftest. */
#ifndef TAY_DEVICE_MATH_H
#define TAY_DEVICE_MATH_H

__device__ __forceinline__ float tay_sqrt_rn(float x) {
    /* the corrected hardware root is computed for every input (straight-line
     * code, so two calls on one value share it after inlining); lanes with
     * an input outside [2^-96, +inf) branch to the full sequence, which a
     * wavefront skips when none has one */
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u);
    const float su = __uint_as_float(__float_as_uint(s) + 1u);
    float r = __builtin_fmaf(-sd, s, x) <= 0.0f ? sd : s;
    r = __builtin_fmaf(-su, s, x) > 0.0f ? su : r;
    const bool fast = x >= 0x1p-96f && x < __builtin_inff();
    if (__builtin_expect(!fast, 0))
        r = __builtin_sqrtf(x);
    return r;
}

#endif
