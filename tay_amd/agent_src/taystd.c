/* Definitions for taystd.h.  All helpers work on x, y, z; the unused fourth
 * lane w of a result is that of the first operand (0 for float4_null), so a
 * field updated in place, a->f = float4_add(a->f, ...), keeps its w and the
 * lane is never rewritten.
 * Lengths use sqrt of a float sum of squares; on the host that is the double
 * sqrt rounded to float, on the device the float sqrt overload.  Both are
 * correctly rounded, so both give the same float. */

float4 float4_null(void) {
    float4 r;
    r.x = 0.0f; r.y = 0.0f; r.z = 0.0f; r.w = 0.0f;
    return r;
}

float4 float4_make(float x, float y, float z, float w) {
    float4 r;
    r.x = x; r.y = y; r.z = z; r.w = w;
    return r;
}

float4 float4_add(float4 a, float4 b) {
    float4 r;
    r.x = a.x + b.x; r.y = a.y + b.y; r.z = a.z + b.z; r.w = a.w;
    return r;
}

float4 float4_sub(float4 a, float4 b) {
    float4 r;
    r.x = a.x - b.x; r.y = a.y - b.y; r.z = a.z - b.z; r.w = a.w;
    return r;
}

float4 float4_mul_scalar(float4 a, float s) {
    float4 r;
    r.x = a.x * s; r.y = a.y * s; r.z = a.z * s; r.w = a.w;
    return r;
}

float4 float4_div_scalar(float4 a, float s) {
    float4 r;
    r.x = a.x / s; r.y = a.y / s; r.z = a.z / s; r.w = a.w;
    return r;
}

float float4_dot(float4 a, float4 b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

float float4_length(float4 a) {
    return (float)sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
}

/* unit vector along a; (1,0,0) when a is (near) zero */
float4 float4_normalize(float4 a) {
    float l = float4_length(a);
    if (l < 0.000001f)
        return float4_make(1.0f, 0.0f, 0.0f, a.w);
    return float4_make(a.x / l, a.y / l, a.z / l, a.w);
}

/* a rescaled to length b; a's xyz unchanged when (near) zero */
float4 float4_normalize_to(float4 a, float b) {
    float l = float4_length(a);
    /* x * 1.0f is x exactly (zeros, denormals, infinities and NaN included),
     * so the (near) zero case needs no branch of its own */
    float c = l < 0.000001f ? 1.0f : b / l;
    return float4_make(a.x * c, a.y * c, a.z * c, a.w);
}
