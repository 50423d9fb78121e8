/* Test models in Tay's model dialect.
 *
 * Agent / agent_see / agent_act: the model of Tay's own test harness
 * (test/agent.h:2-10, test/agent.c:2-55).  The record layout is the
 * reference's, byte for byte, so results can be compared agent by agent.
 *   see: b_buffer += seen.p - seer.p, b_buffer_count += 1
 *   act: f_buffer += b_buffer / b_buffer_count (when nonzero), clear the
 *        buffers, p += v, clamp p to [min, max] per axis reversing v there.
 *
 * BoxAgent / box_agent_see / box_agent_act: the non-point variant
 * (test/agent.h:19-27, test/agent.c:57-119) with min/max corners.
 *
 * point_box_see / box_point_see: a point group seeing a box group and the
 * reverse, which the reference's harness has no model for
 * (test/test_combo.c:11-17): the seer's b_buffer gains the seen min corner
 * (or position) minus its own position (or min corner), b_buffer_count
 * counts the pairs; the act passes are agent_act and box_agent_act.
 *
 * PlainAgent / plain_see / plain_act: the plain 3D point workload: count
 * the neighbours within Euclidean distance r (d^2 <= r^2, self excluded by
 * the pass) and sum their positions; the act pass keeps the finished step's
 * count and sum in last_ncount/last_nsum, clears the accumulators, moves by v
 * and reflects at the walls. */

typedef struct __attribute__((packed)) Agent {
    TayAgentTag tag;
    float4 p;
    float4 v;
    float4 b_buffer;
    int b_buffer_count;
    float4 f_buffer;
    int result_index;
} Agent;

typedef struct __attribute__((packed)) BoxAgent {
    TayAgentTag tag;
    float4 min;
    float4 max;
    float4 v;
    float4 b_buffer;
    int b_buffer_count;
    float4 f_buffer;
    int result_index;
} BoxAgent;

typedef struct __attribute__((packed)) PlainAgent {
    TayAgentTag tag;
    float4 p;
    float4 v;
    float4 nsum;
    unsigned ncount;
    float4 last_nsum;
    unsigned last_ncount;
    int result_index;
} PlainAgent;

typedef struct __attribute__((packed)) ActContext {
    float4 min;
    float4 max;
} ActContext;

typedef struct __attribute__((packed)) SeeContext {
    float4 radii;
} SeeContext;

typedef struct __attribute__((packed)) PlainSeeContext {
    float r2;   /* squared Euclidean radius */
} PlainSeeContext;

void agent_see(global Agent *a, global Agent *b, constant SeeContext *c);
void agent_act(global Agent *a, constant ActContext *c);
void box_agent_see(global BoxAgent *a, global BoxAgent *b, constant SeeContext *c);
void box_agent_act(global BoxAgent *a, constant ActContext *c);
void point_box_see(global Agent *a, global BoxAgent *b, constant SeeContext *c);
void box_point_see(global BoxAgent *a, global Agent *b, constant SeeContext *c);
void plain_see(global PlainAgent *a, global PlainAgent *b, constant PlainSeeContext *c);
void plain_act(global PlainAgent *a, constant ActContext *c);
/* For split see passes (tay_set_see_split): each see function above only
 * adds to its seer's buffer and count, so a partial starts with them
 * emptied (*_clear) and partials merge by adding them (*_combine, a += b). */
void plain_see_clear(global PlainAgent *a);
void plain_see_combine(global PlainAgent *a, global PlainAgent *b);
void agent_see_clear(global Agent *a);
void agent_see_combine(global Agent *a, global Agent *b);
void box_agent_see_clear(global BoxAgent *a);
void box_agent_see_combine(global BoxAgent *a, global BoxAgent *b);
