/*
 * flock_host.c - a plain C host of the tay.h API (include/tay.h), built
 * against libtay_amd.so the way a reference host links Tay: no Python and
 * no library internals, only the reference's public calls
 * (tay/tay.h:87-117) in the order its own harness makes them
 * (test/test_basic.c:34-84: state, sources, group, space, passes, agents,
 * start, run, read back).
 *
 *   flock_host <agent_src dir> <space> <steps> <in.bin> <out.bin>
 *
 * in.bin holds FlockAgent records (tay_amd/agent_src/flocking.h) back to
 * back; the program runs the flocking model (flock_see radius 25,
 * separation radius 12.5; flock_act weights 1.5 / 1.0 / 1.0, dt 0.1, max
 * speed 2) for <steps> steps in a space of type <space> (z_grid:
 * TAY_GPU_Z_GRID, simple: TAY_GPU_SIMPLE, cpu_grid: TAY_CPU_GRID; cell size
 * 25, 250 MB) and writes the group's agents, in storage order,
 * to out.bin.  Exit status 0 on success; on a Tay error the message goes
 * to stderr and the status is 2.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tay.h"
#include "tay_ext.h"

#define global
#define constant
#include "taystd.h"
#include "flocking.h"

static const char *const kSources[] = {"taystd.h", "test_agent.h", "flocking.h", "sph.h",
                                       "taystd.c", "test_agent.c", "flocking.c", "sph.c"};

static char *read_text(const char *dir, const char *name) {
    char path[4096];
    snprintf(path, sizeof(path), "%s/%s", dir, name);
    FILE *f = fopen(path, "rb");
    if (!f)
        return NULL;
    fseek(f, 0, SEEK_END);
    const long size = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *text = (char *)malloc((size_t)size + 1);
    if (text && fread(text, 1, (size_t)size, f) != (size_t)size) {
        free(text);
        text = NULL;
    }
    if (text)
        text[size] = '\0';
    fclose(f);
    return text;
}

static int fail(TayState *tay, const char *what) {
    fprintf(stderr, "flock_host: %s failed: error %d %s\n", what, (int)tay_get_error(tay),
            tay_get_error_message(tay));
    return 2;
}

int main(int argc, char **argv) {
    if (argc != 6) {
        fprintf(stderr, "usage: %s <agent_src dir> <z_grid|simple|cpu_grid> <steps> <in.bin> <out.bin>\n",
                argv[0]);
        return 1;
    }
    TaySpaceType space_type;
    if (strcmp(argv[2], "z_grid") == 0)
        space_type = TAY_GPU_Z_GRID;
    else if (strcmp(argv[2], "simple") == 0)
        space_type = TAY_GPU_SIMPLE;
    else if (strcmp(argv[2], "cpu_grid") == 0)
        space_type = TAY_CPU_GRID;
    else {
        fprintf(stderr, "flock_host: unknown space %s\n", argv[2]);
        return 1;
    }
    const int steps = atoi(argv[3]);

    FILE *in = fopen(argv[4], "rb");
    if (!in) {
        fprintf(stderr, "flock_host: cannot open %s\n", argv[4]);
        return 1;
    }
    fseek(in, 0, SEEK_END);
    const long bytes = ftell(in);
    fseek(in, 0, SEEK_SET);
    const int n = (int)(bytes / (long)sizeof(FlockAgent));
    FlockAgent *agents = (FlockAgent *)malloc(sizeof(FlockAgent) * (size_t)(n > 0 ? n : 1));
    if (!agents || fread(agents, sizeof(FlockAgent), (size_t)n, in) != (size_t)n) {
        fprintf(stderr, "flock_host: cannot read %d agents\n", n);
        return 1;
    }
    fclose(in);

    TayState *tay = tay_create_state();
    for (size_t i = 0; i < sizeof(kSources) / sizeof(kSources[0]); ++i) {
        char *text = read_text(argv[1], kSources[i]);
        if (!text) {
            fprintf(stderr, "flock_host: cannot read %s/%s\n", argv[1], kSources[i]);
            return 1;
        }
        ocl_add_source(tay, text);   /* Tay copies the text */
        free(text);
    }

    TayGroup *group = tay_add_group(tay, sizeof(FlockAgent), n > 0 ? n : 1, TAY_TRUE);
    float4 part = {.arr = {25.0f, 25.0f, 25.0f, 25.0f}};
    tay_configure_space(tay, group, space_type, 3, part, 250);

    FlockSeeContext see_ctx = {25.0f, 12.5f};
    FlockActContext act_ctx = {1.5f, 1.0f, 1.0f, 0.1f, 2.0f};
    float4 radii = {.arr = {25.0f, 25.0f, 25.0f, 25.0f}};
    /* the device compiles flock_see / flock_act from the sources: no host
     * callbacks are needed */
    tay_add_see(tay, group, group, NULL, "flock_see", radii, TAY_FALSE, &see_ctx, sizeof(see_ctx));
    tay_add_act(tay, group, NULL, "flock_act", &act_ctx, sizeof(act_ctx));

    for (int i = 0; i < n; ++i) {
        FlockAgent *a = (FlockAgent *)tay_get_available_agent(tay, group);
        if (!a)
            return fail(tay, "tay_get_available_agent");
        *a = agents[i];
        tay_commit_available_agent(tay, group);
    }

    tay_simulation_start(tay);
    if (tay_get_error(tay) != TAY_ERROR_NONE)
        return fail(tay, "tay_simulation_start");
    if (tay_run(tay, steps) != steps)
        return fail(tay, "tay_run");

    for (int i = 0; i < n; ++i)
        agents[i] = *(FlockAgent *)tay_get_agent(tay, group, i);
    FILE *out = fopen(argv[5], "wb");
    if (!out || fwrite(agents, sizeof(FlockAgent), (size_t)n, out) != (size_t)n) {
        fprintf(stderr, "flock_host: cannot write %s\n", argv[5]);
        return 1;
    }
    fclose(out);
    printf("flock_host: %d agents, %d steps, %.3f ms per step\n", n, steps,
           tay_get_ms_per_step_for_last_run(tay));

    tay_simulation_end(tay);
    tay_destroy_state(tay);
    free(agents);
    return 0;
}
