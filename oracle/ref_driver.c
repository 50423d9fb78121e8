/*
 * ref_driver.c - TEST INFRASTRUCTURE ONLY (built by `make ref`, this
 * container only; the reference tree does not exist on the GPU hosts).
 *
 * Runs the reference's own CPU structures on one or two groups of Tay
 * test-model agents (test/agent.h: Agent for point groups, BoxAgent for
 * non-point groups) and dumps what the golden fixtures hold.  All arithmetic
 * on the path is the reference's: sort = cpu_grid_sort / cpu_z_grid_sort,
 * seen sweeps = cpu_grid_see_seen / cpu_z_grid_see_seen /
 * cpu_simple_see_seen, pairing = the space_see_* loop tay/state_compile.c
 * selects (tay/state_compile.c:109-120), models = test/agent.c.  Only the
 * outer loop of tay_run (tay/state.c:221-306) is restated, single-threaded.
 *
 * Passes: one group sees itself and acts (test/test_basic.c,
 * test/test_nonpoint.c); two groups see themselves and each other in the
 * order of test/test_combo.c:114-119 and then act.  The reference has no
 * model for a point group seeing a box group (test/test_combo.c:11-17);
 * point_box_see / box_point_see below add the seen min corner minus the
 * seer position (or the reverse) to the seer's buffer, the same arithmetic
 * as the GPU models of the same names.
 *
 * usage: ref_driver STEPS RADIUS SHARED_MB OUT  KIND STRUCT N PART IN  [KIND STRUCT N PART IN]
 *   KIND    point | box
 *   STRUCT  simple | grid | zgrid
 *   IN      N records of sizeof(Agent) or sizeof(BoxAgent) bytes
 *   OUT     for every step and group: the records after the sort, then for
 *           every group after the see passes, then after the act passes;
 *           then the geometry (origin, cell sizes, cell counts) of every
 *           step's sort of every group
 */
#include "state.h"
#include "space.h"
#include "thread.h"
#include "agent_host.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void ref_grid_see(TayPass *pass, TayThreadContext *ctx);
void ref_z_grid_see(TayPass *pass, TayThreadContext *ctx);

static void point_box_see(void *seer, void *seen, void *c) {
    Agent *a = (Agent *)seer;
    BoxAgent *b = (BoxAgent *)seen;
    a->b_buffer.x += b->min.x - a->p.x;
    a->b_buffer.y += b->min.y - a->p.y;
    a->b_buffer.z += b->min.z - a->p.z;
    a->b_buffer_count++;
}

static void box_point_see(void *seer, void *seen, void *c) {
    BoxAgent *a = (BoxAgent *)seer;
    Agent *b = (Agent *)seen;
    a->b_buffer.x += b->p.x - a->min.x;
    a->b_buffer.y += b->p.y - a->min.y;
    a->b_buffer.z += b->p.z - a->min.z;
    a->b_buffer_count++;
}

static void write_records(FILE *f, TayGroup *g) {
    fwrite(g->storage, (size_t)g->agent_size, g->space.count, f);
}

static TaySpaceType parse_type(const char *s) {
    return !strcmp(s, "grid") ? TAY_CPU_GRID : !strcmp(s, "zgrid") ? TAY_CPU_Z_GRID : TAY_CPU_SIMPLE;
}

static TAY_SEE_FUNC see_func(int seer_point, int seen_point) {
    if (seer_point)
        return seen_point ? (TAY_SEE_FUNC)agent_see : point_box_see;
    return seen_point ? box_point_see : (TAY_SEE_FUNC)box_agent_see;
}

/* tay/state_compile.c:109-120 */
static PAIRING_FUNC pairing(TayGroup *seer, TayGroup *seen, int self_see) {
    if (seer == seen) {
        if (self_see)
            return seer->is_point ? space_see_point_point_self_see : space_see_nonpoint_nonpoint_self_see;
        return seer->is_point ? space_see_point_point : space_see_nonpoint_nonpoint;
    }
    if (seer->is_point == seen->is_point)
        return seer->is_point ? space_see_point_point_self_see : space_see_nonpoint_nonpoint_self_see;
    return seer->is_point ? space_see_point_nonpoint : space_see_nonpoint_point;
}

int main(int argc, char **argv) {
    if (argc != 10 && argc != 15) {
        fprintf(stderr, "usage: %s STEPS RADIUS SHARED_MB OUT KIND STRUCT N PART IN [KIND STRUCT N PART IN]\n",
                argv[0]);
        return 2;
    }
    int steps = atoi(argv[1]);
    float radius = (float)atof(argv[2]);
    int shared_mb = atoi(argv[3]);
    const char *out_path = argv[4];
    int groups_count = argc == 10 ? 1 : 2;

    TayGroup groups[2];
    memset(groups, 0, sizeof(groups));
    for (int gi = 0; gi < groups_count; ++gi) {
        char **a = argv + 5 + 5 * gi;
        TayGroup *g = groups + gi;
        int point = strcmp(a[0], "box") != 0;
        unsigned n = (unsigned)atoi(a[2]);
        float part = (float)atof(a[3]);
        g->agent_size = point ? (int)sizeof(Agent) : (int)sizeof(BoxAgent);
        g->capacity = (int)n;
        g->is_point = point;
        g->id = (unsigned)gi;
        g->agent_storage[0] = g->storage = calloc(n ? n : 1, (size_t)g->agent_size);
        g->agent_storage[1] = g->sort_storage = calloc(n ? n : 1, (size_t)g->agent_size);
        g->space.type = parse_type(a[1]);
        g->space.dims = 3;
        g->space.min_part_sizes = (float4){{{part, part, part, part}}};
        g->space.shared_size = shared_mb * TAY_MB;
        g->space.shared = calloc(1, (size_t)g->space.shared_size);
        g->space.count = n;
        FILE *in = fopen(a[4], "rb");
        if (!in || fread(g->storage, (size_t)g->agent_size, n, in) != n) {
            fprintf(stderr, "cannot read %s\n", a[4]);
            return 1;
        }
        fclose(in);
        if (g->space.type == TAY_CPU_GRID)
            cpu_grid_on_simulation_start(&g->space);
        else if (g->space.type == TAY_CPU_Z_GRID)
            cpu_z_grid_on_simulation_start(&g->space);
    }

    SeeContext see_context;
    see_context.radii = (float4){{{radius, radius, radius, 0.0f}}};
    ActContext act_context;
    act_context.min = (float4){{{0.0f, 0.0f, 0.0f, 0.0f}}};
    act_context.max = (float4){{{1000.0f, 1000.0f, 1000.0f, 0.0f}}};

    /* test/test_combo.c:114-119: a sees a, a sees b, b sees b, b sees a */
    static const int order[4][2] = {{0, 0}, {0, 1}, {1, 1}, {1, 0}};
    TayPass passes[4];
    int passes_count = groups_count == 1 ? 1 : 4;
    for (int pi = 0; pi < passes_count; ++pi) {
        TayPass *p = passes + pi;
        memset(p, 0, sizeof(*p));
        p->type = TAY_PASS_SEE;
        p->seer_group = groups + order[pi][0];
        p->seen_group = groups + order[pi][1];
        p->see = see_func(p->seer_group->is_point, p->seen_group->is_point);
        p->radii = (float4){{{radius, radius, radius, radius}}};
        p->self_see = 0;
        p->context = &see_context;
        p->pairing_func = pairing(p->seer_group, p->seen_group, 0);
        TaySpaceType t = p->seen_group->space.type;
        p->seen_func = t == TAY_CPU_GRID ? cpu_grid_see_seen : t == TAY_CPU_Z_GRID ? cpu_z_grid_see_seen
                                                                                     : cpu_simple_see_seen;
    }

    TayThreadContext ctx;
    memset(&ctx, 0, sizeof(ctx));

    FILE *out = fopen(out_path, "wb");
    if (!out)
        return 1;
    float *geometry = calloc((size_t)(steps ? steps : 1) * 2, 12 * sizeof(float));
    for (int step = 0; step < steps; ++step) {
        for (int gi = 0; gi < groups_count; ++gi) {
            TayGroup *g = groups + gi;
            if (g->space.type == TAY_CPU_GRID)
                cpu_grid_sort(g);
            else if (g->space.type == TAY_CPU_Z_GRID)
                cpu_z_grid_sort(g);
            write_records(out, g);
        }

        for (int pi = 0; pi < passes_count; ++pi) {
            TayPass *p = passes + pi;
            ctx.context = p->context;
            TaySpaceType t = p->seer_group->space.type;
            if (t == TAY_CPU_GRID)
                ref_grid_see(p, &ctx);
            else if (t == TAY_CPU_Z_GRID)
                ref_z_grid_see(p, &ctx);
            else {
                /* tay/cpu_simple.c:18-38, one thread: every seer, an all-zero box */
                AgentsSlice all = {p->seer_group->storage, (unsigned)p->seer_group->agent_size, 0,
                                   p->seer_group->space.count};
                p->seen_func(p, all, (Box){0}, 3, &ctx);
            }
        }
        for (int gi = 0; gi < groups_count; ++gi)
            write_records(out, groups + gi);

        for (int gi = 0; gi < groups_count; ++gi) {
            TayGroup *g = groups + gi;
            for (unsigned i = 0; i < g->space.count; ++i) {
                void *agent = g->storage + (size_t)g->agent_size * i;
                if (g->is_point)
                    agent_act((Agent *)agent, &act_context);
                else
                    box_agent_act((BoxAgent *)agent, &act_context);
            }
            write_records(out, g);
        }

        for (int gi = 0; gi < groups_count; ++gi) {
            TayGroup *g = groups + gi;
            float *geo = geometry + 12 * ((size_t)step * groups_count + gi);
            if (g->space.type == TAY_CPU_GRID) {
                memcpy(geo, &g->space.cpu_grid.origin, 16);
                memcpy(geo + 4, &g->space.cpu_grid.cell_sizes, 16);
                memcpy(geo + 8, &g->space.cpu_grid.cell_counts, 16);
                cpu_grid_unsort(g);
            } else if (g->space.type == TAY_CPU_Z_GRID) {
                memcpy(geo, &g->space.cpu_z_grid.origin, 16);
                memcpy(geo + 4, &g->space.cpu_z_grid.cell_sizes, 16);
                memcpy(geo + 8, &g->space.cpu_z_grid.cell_counts, 16);
                cpu_z_grid_unsort(g);
            }
        }
    }
    fwrite(geometry, sizeof(float), 12 * (size_t)steps * groups_count, out);
    fclose(out);
    return 0;
}
