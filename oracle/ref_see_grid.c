/*
 * ref_see_grid.c - TEST INFRASTRUCTURE ONLY (built by `make ref`, this
 * container only).  Compiles the reference's tay/cpu_grid.c in place and adds
 * a single-threaded walk over the seers that restates its threaded seer
 * driver (tay/cpu_grid.c:286-323): one seer run per occupied cell, the cell's
 * box grown by the radii as the seer box, the reference's own seen sweep.
 */
#include "cpu_grid.c"

void ref_grid_see(TayPass *pass, TayThreadContext *ctx) {
    TayGroup *seer_group = pass->seer_group;
    CpuGrid *grid = &seer_group->space.cpu_grid;
    int dims = seer_group->space.dims;
    unsigned i = 0, end = seer_group->space.count;
    while (i < end) {
        TayAgentTag *seer = (TayAgentTag *)(seer_group->storage + seer_group->agent_size * i);
        GridCell *cell = grid->cells + seer->part_i;
        Box box;
        for (int k = 0; k < dims; ++k) {
            float lo = grid->origin.arr[k] + cell->indices.arr[k] * grid->cell_sizes.arr[k];
            box.min.arr[k] = lo - pass->radii.arr[k];
            box.max.arr[k] = lo + grid->cell_sizes.arr[k] + pass->radii.arr[k];
        }
        unsigned cell_end = cell->first_agent_i + cell->count;
        AgentsSlice seer_slice = {seer_group->storage, seer_group->agent_size, i, cell_end};
        pass->seen_func(pass, seer_slice, box, dims, ctx);
        i = cell_end;
    }
}
