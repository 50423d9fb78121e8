/*
 * ref_driver.c - TEST INFRASTRUCTURE ONLY (built by `make ref`, this
 * container only; the reference tree does not exist on the GPU hosts).
 *
 * Runs the reference's own CPU structures on an input file of Tay test-model
 * agents (test/agent.h layout) and dumps what the golden fixtures hold.  All
 * arithmetic is the reference's: sort = cpu_grid_sort / cpu_z_grid_sort,
 * seen sweeps = cpu_grid_see_seen / cpu_z_grid_see_seen / cpu_simple_see_seen,
 * pairing = space_see_point_point, model = test/agent.c.  Only the outer loop
 * of tay_run (tay/state.c:221-306) is restated here, single-threaded.
 *
 * usage: ref_driver STRUCT N STEPS RADIUS PART SHARED_MB IN OUT
 *   STRUCT  simple | grid | zgrid
 *   IN      N records of sizeof(Agent) bytes
 *   OUT     for every step: the records after the sort, after the see pass
 *           and after the act pass, then the geometry (origin, cell sizes,
 *           cell counts) of every step's sort
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

static void write_records(FILE *f, TayGroup *g) {
    fwrite(g->storage, (size_t)g->agent_size, g->space.count, f);
}

int main(int argc, char **argv) {
    if (argc != 9) {
        fprintf(stderr, "usage: %s simple|grid|zgrid N STEPS RADIUS PART SHARED_MB IN OUT\n", argv[0]);
        return 2;
    }
    const char *structure = argv[1];
    unsigned n = (unsigned)atoi(argv[2]);
    int steps = atoi(argv[3]);
    float radius = (float)atof(argv[4]);
    float part = (float)atof(argv[5]);
    int shared_mb = atoi(argv[6]);

    TaySpaceType type = !strcmp(structure, "grid") ? TAY_CPU_GRID
                      : !strcmp(structure, "zgrid") ? TAY_CPU_Z_GRID : TAY_CPU_SIMPLE;

    TayGroup group;
    memset(&group, 0, sizeof(group));
    group.agent_size = sizeof(Agent);
    group.capacity = (int)n;
    group.is_point = 1;
    group.agent_storage[0] = group.storage = calloc(n, sizeof(Agent));
    group.agent_storage[1] = group.sort_storage = calloc(n, sizeof(Agent));
    group.space.type = type;
    group.space.dims = 3;
    group.space.min_part_sizes = (float4){{{part, part, part, part}}};
    group.space.shared_size = shared_mb * TAY_MB;
    group.space.shared = calloc(1, (size_t)group.space.shared_size);
    group.space.count = n;

    FILE *in = fopen(argv[7], "rb");
    if (!in || fread(group.storage, sizeof(Agent), n, in) != n) {
        fprintf(stderr, "cannot read %s\n", argv[7]);
        return 1;
    }
    fclose(in);

    SeeContext see_context;
    see_context.radii = (float4){{{radius, radius, radius, 0.0f}}};
    ActContext act_context;
    act_context.min = (float4){{{0.0f, 0.0f, 0.0f, 0.0f}}};
    act_context.max = (float4){{{1000.0f, 1000.0f, 1000.0f, 0.0f}}};

    TayPass see;
    memset(&see, 0, sizeof(see));
    see.type = TAY_PASS_SEE;
    see.seer_group = &group;
    see.seen_group = &group;
    see.see = (TAY_SEE_FUNC)agent_see;
    see.radii = see_context.radii;
    see.self_see = 0;
    see.context = &see_context;
    see.pairing_func = space_see_point_point;   /* tay/state_compile.c:109-113 */
    see.seen_func = type == TAY_CPU_GRID ? cpu_grid_see_seen
                  : type == TAY_CPU_Z_GRID ? cpu_z_grid_see_seen : cpu_simple_see_seen;

    TayThreadContext ctx;
    memset(&ctx, 0, sizeof(ctx));
    ctx.context = &see_context;

    if (type == TAY_CPU_GRID)
        cpu_grid_on_simulation_start(&group.space);
    else if (type == TAY_CPU_Z_GRID)
        cpu_z_grid_on_simulation_start(&group.space);

    FILE *out = fopen(argv[8], "wb");
    if (!out)
        return 1;
    float geometry[12 * 64];
    for (int step = 0; step < steps; ++step) {
        if (type == TAY_CPU_GRID)
            cpu_grid_sort(&group);
        else if (type == TAY_CPU_Z_GRID)
            cpu_z_grid_sort(&group);
        write_records(out, &group);

        if (type == TAY_CPU_GRID)
            ref_grid_see(&see, &ctx);
        else if (type == TAY_CPU_Z_GRID)
            ref_z_grid_see(&see, &ctx);
        else {
            AgentsSlice all = {group.storage, group.agent_size, 0, group.space.count};
            cpu_simple_see_seen(&see, all, (Box){0}, 3, &ctx);
        }
        write_records(out, &group);

        for (unsigned i = 0; i < group.space.count; ++i)
            agent_act((Agent *)(group.storage + group.agent_size * i), &act_context);
        write_records(out, &group);

        float *geo = geometry + 12 * (step < 64 ? step : 63);
        const float4 *origin = type == TAY_CPU_GRID ? &group.space.cpu_grid.origin : &group.space.cpu_z_grid.origin;
        const float4 *sizes = type == TAY_CPU_GRID ? &group.space.cpu_grid.cell_sizes : &group.space.cpu_z_grid.cell_sizes;
        const int4 *counts = type == TAY_CPU_GRID ? &group.space.cpu_grid.cell_counts : &group.space.cpu_z_grid.cell_counts;
        memcpy(geo, origin, 16);
        memcpy(geo + 4, sizes, 16);
        memcpy(geo + 8, counts, 16);

        if (type == TAY_CPU_GRID)
            cpu_grid_unsort(&group);
        else if (type == TAY_CPU_Z_GRID)
            cpu_z_grid_unsort(&group);
    }
    fwrite(geometry, sizeof(float), 12 * (size_t)(steps < 64 ? steps : 64), out);
    fclose(out);
    return 0;
}
