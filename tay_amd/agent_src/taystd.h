/* Vector helpers for agent models, written in Tay's model dialect: plain C
 * that also compiles as device code once the host supplies float4/float3 and
 * the global/constant qualifiers.  Same helper set as the reference models use
 * (test/taystd.h, m_flocking/taystd.c).  Only x, y, z take part; every
 * helper returns w = 0. */

float4 float4_null(void);
float4 float4_make(float x, float y, float z, float w);
float4 float4_add(float4 a, float4 b);
float4 float4_sub(float4 a, float4 b);
float4 float4_mul_scalar(float4 a, float s);
float4 float4_div_scalar(float4 a, float s);
float float4_dot(float4 a, float4 b);
float float4_length(float4 a);
float4 float4_normalize(float4 a);
float4 float4_normalize_to(float4 a, float b);
