/*
 * jit_device.h - device templates for the generated see and act kernels.
 *
 * The build embeds this text into every hipRTC program after the model
 * sources; each pass then instantiates one template with a call to its model
 * function.  Agent records are rebuilt in registers from the
 * structure-of-arrays columns, the model function runs on those private
 * copies, and only the words it changed are stored back, so loads of columns
 * a model never reads are removed by the compiler.
 *
 * See kernel (tay_see_tiles): one wavefront owns a tile of 64 consecutive
 * seers of the sorted storage.  The tile's bounding box, grown by the radii
 * and a rounding margin, selects the seen cells; the wavefront walks them x
 * outermost, then y, then z, and each cell's agents in storage order, so every
 * seer meets its pairs in the order of the reference's grid sweep
 * (tay/cpu_grid.c:222-261) and accumulates exactly the same float sums.
 * Every lane applies the reference's per-axis box test (tay/space.c:62-104)
 * to its own seer.
 *
 * Batched kernels (tay_see_batch, tay_act_batch) run the same tile body for
 * many independent instances sharing one set of columns: a persistent see
 * grid takes tiles of all instances from a work queue, each tile with its
 * instance's grid, radii and context.  A seer's pairs, and their order, do
 * not depend on which tiles run together, so every instance computes exactly
 * what it computes alone.
 */

template <typename F> struct tay_see_traits;
template <typename A, typename B, typename C> struct tay_see_traits<void(A *, B *, C *)> {
    typedef A seer;
    typedef B seen;
    typedef C context;
};

template <typename F> struct tay_act_traits;
template <typename A, typename C> struct tay_act_traits<void(A *, C *)> {
    typedef A agent;
    typedef C context;
};

enum { TAY_NARROW_PP = 0, TAY_NARROW_PN = 1, TAY_NARROW_NP = 2, TAY_NARROW_NN = 3 };

template <class C> struct tay_is_void { static constexpr bool value = false; };
template <> struct tay_is_void<void> { static constexpr bool value = true; };
template <class C, unsigned WORDS> struct tay_ctx_fits {
    static constexpr bool value = WORDS > 0u && WORDS <= 64u && sizeof(C) <= WORDS * 4u;
};
template <unsigned WORDS> struct tay_ctx_fits<void, WORDS> { static constexpr bool value = false; };

/* A pass context of up to 64 words is copied once per kernel into
 * registers with aligned uniform loads.  Model contexts are packed structs
 * (alignment 1), so reading their fields from memory inside the pair loop
 * would cost a vector load per field per pair. */
template <class C, unsigned WORDS, bool COPY = tay_ctx_fits<C, WORDS>::value> struct tay_ctx_holder {
    C *p;
    __device__ __forceinline__ explicit tay_ctx_holder(const void *q) : p((C *)q) {}
    __device__ __forceinline__ C *get() { return p; }
};
template <class C, unsigned WORDS> struct tay_ctx_holder<C, WORDS, true> {
    union {
        C ctx;
        unsigned w[WORDS];
    } u;
    __device__ __forceinline__ explicit tay_ctx_holder(const void *q) {
        const unsigned *words = (const unsigned *)q;
#pragma unroll
        for (unsigned k = 0; k < WORDS; ++k)
            u.w[k] = words[k];
    }
    __device__ __forceinline__ C *get() { return &u.ctx; }
};

#define TAY_SEE_BLOCK 256u

/* Candidates per run of a chunk in the balanced see order (a power of two,
 * 1 to 64).  Measured on the 2M-particle SPH dam break (force pass, first
 * steps): 1.17 ms per pass in the sweep order; runs of 1, 4, 16, 32:
 * 0.99, 0.98, 1.17, 1.12 ms. */
#ifndef TAY_BALANCE_RUN
#define TAY_BALANCE_RUN 4u
#endif

__device__ __forceinline__ float tay_wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = fminf(v, __shfl_xor(v, o));
    return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(v)));
}

__device__ __forceinline__ float tay_wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = fmaxf(v, __shfl_xor(v, o));
    return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(v)));
}

/* A wave-uniform value moved to a scalar register: loads from a uniform
 * address after a store of the same kernel are vector loads, and values
 * derived from them would otherwise stay in vector registers. */
__device__ __forceinline__ float tay_uniform(float v) {
    return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(v)));
}
__device__ __forceinline__ int tay_uniform(int v) {
    return (int)__builtin_amdgcn_readfirstlane((unsigned)v);
}
__device__ __forceinline__ unsigned tay_uniform(unsigned v) {
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ unsigned tay_wave_sum(unsigned v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    return v;
}

/* Inclusive prefix sum over the wavefront's lanes. */
__device__ __forceinline__ unsigned tay_wave_incl_sum(unsigned v) {
    const unsigned lane = threadIdx.x & 63u;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned t = __shfl_up(v, o);
        v += lane >= (unsigned)o ? t : 0u;
    }
    return v;
}

/* Cell indices of an agent's tag: part_i holds the Morton key of the cell
 * the sort stored it in (3D 10 bits per axis interleaved x, y, z from bit
 * 0; 2D 16 bits, x, y; 1D the index itself). */
__device__ __forceinline__ unsigned tay_compact3(unsigned v) {
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030C30C3u;
    v = (v | (v >> 4)) & 0x0300F00Fu;
    v = (v | (v >> 8)) & 0xFF0000FFu;
    return (v | (v >> 16)) & 0x000003FFu;
}
__device__ __forceinline__ unsigned tay_compact2(unsigned v) {
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0F0F0F0Fu;
    v = (v | (v >> 4)) & 0x00FF00FFu;
    return (v | (v >> 8)) & 0x0000FFFFu;
}
__device__ __forceinline__ void tay_cell_of_tag(unsigned part_i, int dims, int *c) {
    if (dims == 3) {
        c[0] = (int)tay_compact3(part_i);
        c[1] = (int)tay_compact3(part_i >> 1);
        c[2] = (int)tay_compact3(part_i >> 2);
    } else if (dims == 2) {
        c[0] = (int)tay_compact2(part_i);
        c[1] = (int)tay_compact2(part_i >> 1);
        c[2] = 0;
    } else {
        c[0] = (int)part_i;
        c[1] = c[2] = 0;
    }
}

/* Logical block order that hands each XCD one contiguous run of tiles: the
 * hardware deals consecutive workgroups round-robin over the 8 XCDs, each
 * with its own L2, and neighbouring tiles read the same seen cells. */
__device__ __forceinline__ unsigned tay_xcd_block(void) {
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const unsigned q = nb >> 3, r = nb & 7u, xcd = b & 7u, k = b >> 3;
    return xcd * q + (xcd < r ? xcd : r) + k;
}

/* Cell index along one axis, (int)floorf((p - origin) / size) as in
 * tay/cpu_grid.c:18-41, clamped to [0, count - 1]. */
__device__ __forceinline__ int tay_cell_axis(float p, float origin, float size, int count) {
    int c = (int)floorf((p - origin) / size);
    c = c < 0 ? 0 : c;
    return c > count - 1 ? count - 1 : c;
}

/* The reference's per-axis tests (tay/space.c:62-104), evaluated without
 * short-circuit branches.  For point pairs the reference skips when
 * d < -r || d > r; !(|d| > r) is the same predicate for every d and r,
 * NaN and negative radii included, and costs one compare. */
template <int NARROW, int DIMS>
__device__ __forceinline__ bool tay_narrow(const float *a_lo, const float *a_hi, const float *b_lo, const float *b_hi,
                                           const float *r) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < DIMS; ++k) {
        if (NARROW == TAY_NARROW_PP) {
            ok = ok & !(fabsf(a_lo[k] - b_lo[k]) > r[k]);
        } else if (NARROW == TAY_NARROW_PN) {
            ok = ok & !(a_lo[k] - b_hi[k] > r[k]) & !(b_lo[k] - a_lo[k] > r[k]);
        } else if (NARROW == TAY_NARROW_NP) {
            ok = ok & !(a_lo[k] - b_lo[k] > r[k]) & !(b_lo[k] - a_hi[k] > r[k]);
        } else {
            ok = ok & !(a_lo[k] - b_hi[k] > r[k]) & !(b_lo[k] - a_hi[k] > r[k]);
        }
    }
    return ok;
}

/* Orders this wavefront's LDS accesses: everything before is complete and
 * visible to every lane of the wavefront before anything after starts. */
__device__ __forceinline__ void tay_wave_sync(void) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* bits * 2 + !(m > r): the mask of a chunk is built from its top bit down,
 * one compare into VCC and one add with that carry per candidate. */
__device__ __forceinline__ unsigned tay_shift_in_ngt(unsigned bits, float m, float r) {
    asm("v_cmp_nlt_f32_e32 vcc, %2, %1\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc"
        : "+v"(bits)
        : "v"(m), "s"(r)
        : "vcc");
    return bits;
}

/* Wavefront maximum of a value below 128, from the top bit down: one
 * compare and one ballot per bit, no cross-lane data movement.  Every lane
 * of the wavefront must be active. */
__device__ __forceinline__ unsigned tay_wave_max_u7(unsigned v) {
    unsigned r = 0;
#pragma unroll
    for (int b = 6; b >= 0; --b) {
        const unsigned t = r | (1u << b);
        if (__ballot(v >= t) != 0ull)
            r = t;
    }
    return r;
}

__device__ __forceinline__ unsigned tay_wave_max_u(unsigned v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned u = __shfl_xor(v, o);
        v = u > v ? u : v;
    }
    return __builtin_amdgcn_readfirstlane(v);
}

/* One tile's view of the storage it reads and writes: the plain kernel has
 * one view for all tiles, a batched kernel one per instance. */
struct TaySeeView {
    const unsigned *seer_cols;
    unsigned *seer_out;
    unsigned seer_cap;
    unsigned seer_end;                  /* one past the last seer slot of the tile's group */
    const unsigned *seen_cols;
    unsigned seen_cap;
    unsigned seen_beg, seen_end;        /* the seen slots of a Simple seen space */
    const TayCellRange *ranges;
    const TayGridGeom *geom;
    const void *context;
    float radii[4];
};

/* The seen candidates of a tile are consumed in chunks of 64.  Filter: each
 * chunk's boxes are staged in LDS and every lane box-tests all of them
 * (broadcast reads), setting bit c of its 64-bit mask for each passing
 * candidate c.  Compact: the wavefront then loops over the set bits, each
 * lane taking its lowest remaining bit, so the model body runs with nearly
 * all lanes busy instead of once per candidate under a sparse lane mask.
 * Bits are taken in ascending order and chunks follow the candidate order,
 * so every seer still meets its pairs in the reference's sweep order.
 *
 * i0 is the tile's first slot and `seers` (64, 32 or 16) the number of
 * seers it holds on lanes 0 to seers - 1 (fewer than 64 for small states,
 * whose 64-seer tiles would leave most SIMDs idle; the other lanes still
 * stage and filter candidates but own no seer); `width` (a power of two up
 * to `seers`) is the size of the lane groups a tile starts with; a group of at most lane_walk
 * lanes whose box is still too wide walks per lane instead of splitting
 * further (0: always split).  Returns the pairs this lane passed to the
 * model when COUNT is set.
 *
 * BALANCED (the balanced see order): a swept group takes its cells 64 at a
 * time; the M candidates of those cells, in sweep order, are cut into
 * P = 64 / TAY_BALANCE_RUN parts of L = ceil(M / P), and chunk k holds run k
 * (TAY_BALANCE_RUN consecutive candidates) of every part: slot
 * j * TAY_BALANCE_RUN + t holds sweep position j * L + k * TAY_BALANCE_RUN + t,
 * slots past the end of their part left empty.  A run of consecutive
 * candidates lies on one side of the group, where some seers pass most of
 * them and others none; a chunk made of runs from every part of the
 * neighbourhood spreads each seer's pairs over the chunks more evenly, so
 * fewer compact rounds carry the same pairs, while runs keep the loads of
 * a round on few cache lines.  Each seer meets the same pairs in that
 * fixed order instead of the sweep order; TAY_BALANCE_RUN 64 (or M <= 64)
 * is the sweep order.
 *
 * SPLIT (a split see pass, 2, 4 or 8): the tile holds 64 / SPLIT seers and
 * lane s * SPLIT + part works for seer s: part 0 on the seer's record, the
 * others on partial copies emptied by `clear`.  A swept group fills its
 * chunks 64 candidates at a time in sweep order and lane `part` takes
 * slots part, part + SPLIT, ..., so the candidate at sweep position pos
 * goes to partial pos mod SPLIT; a walking group hands each seer's cell
 * rows (x, y) to its lanes by row index mod SPLIT.  At the end of the tile
 * the partials merge into part 0 by `combine` over lane shuffles, in a
 * fixed tree (0 + 1), (2 + 3), ..., then pairs of those. */
template <class T, unsigned SEER_WORDS, unsigned SEEN_WORDS, int NARROW, bool SKIP_SELF, bool SEEN_GRID, int DIMS,
          bool COUNT, unsigned CTX_WORDS, bool ROW_WALK, bool BALANCED, unsigned SPLIT, class Call, class Clear,
          class Combine>
__device__ __forceinline__ unsigned tay_see_tile(const TaySeeView &v, unsigned i0, unsigned seers, unsigned width0,
                                                unsigned lane_walk,
                                                Call call, Clear clear, Combine combine) {
    static_assert(SPLIT == 1u || SPLIT == 2u || SPLIT == 4u || SPLIT == 8u, "lanes per seer: 1, 2, 4 or 8");
    static_assert(SPLIT == 1u || !BALANCED, "a split pass runs in the sweep order");
    static_assert(sizeof(typename T::seer) <= SEER_WORDS * 4u, "seer record larger than the group's agent size");
    static_assert(sizeof(typename T::seen) <= SEEN_WORDS * 4u, "seen record larger than the group's agent size");
    constexpr bool SEER_BOX = NARROW == TAY_NARROW_NP || NARROW == TAY_NARROW_NN;
    constexpr bool SEEN_BOX = NARROW == TAY_NARROW_PN || NARROW == TAY_NARROW_NN;
    constexpr unsigned WAVES = TAY_SEE_BLOCK / 64u;

    __shared__ __align__(16) float chunk_lo[WAVES][64][4];
    __shared__ float chunk_hi[SEEN_BOX ? WAVES : 1u][64][4];
    __shared__ __align__(8) unsigned cell_ranges[SEEN_GRID ? WAVES : 1u][64][2];
    __shared__ unsigned lane_cell[BALANCED ? WAVES : 1u][64];   /* balanced order: each lane's cell in the segment */

    const unsigned lane = threadIdx.x & 63u;
    const unsigned wave = threadIdx.x >> 6;
    /* lane = seer slot x SPLIT + part: a split pass gives each seer SPLIT
     * lanes, part 0 the seer itself, the others partials of it */
    const unsigned sl = lane / SPLIT, part = lane % SPLIT;
    const unsigned i = i0 + sl;
    const bool active = sl < seers && i < v.seer_end;
    const unsigned ia = active ? i : i0;

    union {
        typename T::seer agent;
        unsigned w[SEER_WORDS];
    } A;
    unsigned orig[SEER_WORDS];
#pragma unroll
    for (unsigned k = 0; k < SEER_WORDS; ++k)
        orig[k] = A.w[k] = v.seer_cols[k * v.seer_cap + ia];
    if (SPLIT > 1u && part != 0u)
        clear(&A.agent);    /* a partial starts with the sums emptied */

    /* position words: 2.. (min corner), 6.. (max corner of a box agent) */
    float a_lo[3] = {0.0f, 0.0f, 0.0f}, a_hi[3] = {0.0f, 0.0f, 0.0f}, r[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < DIMS; ++k) {
        a_lo[k] = __uint_as_float(A.w[2 + k]);
        a_hi[k] = SEER_BOX ? __uint_as_float(A.w[6 + k]) : a_lo[k];
        r[k] = v.radii[k];
    }
    tay_ctx_holder<typename T::context, CTX_WORDS> ctx_holder(v.context);
    typename T::context *ctx = ctx_holder.get();
    const char *seen_col[SEEN_WORDS];
#pragma unroll
    for (unsigned k = 0; k < SEEN_WORDS; ++k)
        seen_col[k] = (const char *)(v.seen_cols + (unsigned long long)k * v.seen_cap);
    unsigned pairs = 0;
    bool member = active;   /* lane belongs to the lane group being processed */
    unsigned fill = 0;      /* slots used in the current chunk (uniform) */
    unsigned cand = 0;      /* lane k: storage index of the chunk's slot k */
    unsigned self_slot = 64u; /* the chunk slot holding this lane's own seer, 64: none */
    unsigned long long chunk_valid = ~0ull;  /* balanced order: the chunk's occupied slots */

    /* Point pairs with one radius on every axis: !(|d_k| > r) for all k is
     * !(max_k |d_k| > r) (the maximum ignores a NaN difference exactly as
     * the per-axis test passes it), one max3 and one compare per candidate. */
    bool iso = NARROW == TAY_NARROW_PP;
#pragma unroll
    for (int k = 1; k < DIMS; ++k)
        iso = iso && r[k] == r[0];
    const float r0 = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(r[0])));

    /* box-test candidates [c0, c1) of the staged chunk into a 32-bit mask,
     * built from its top bit down (candidate c1 - 1 first) */
    auto filter = [&](unsigned c0, unsigned c1) {
        unsigned bits = 0;
        if (iso) {
#pragma unroll 4
            for (unsigned c = c1; c-- > c0;) {
                const float *b = chunk_lo[wave][c];
                float m = fabsf(a_lo[0] - b[0]);
#pragma unroll
                for (int k = 1; k < DIMS; ++k)
                    m = fmaxf(m, fabsf(a_lo[k] - b[k]));
                bits = tay_shift_in_ngt(bits, m, r0);
            }
        } else {
#pragma unroll 4
            for (unsigned c = c1; c-- > c0;) {
                const float *b_lo = chunk_lo[wave][c];
                const float *b_hi = SEEN_BOX ? chunk_hi[wave][c] : b_lo;
                const bool ok = tay_narrow<NARROW, DIMS>(a_lo, a_hi, b_lo, b_hi, r);
                bits = bits + bits + (unsigned)ok;
            }
        }
        return bits;
    };

    /* split pass: box-test this lane's slots part, part + SPLIT, ... of the
     * staged chunk into bit t = slot / SPLIT, from the top bit down */
    auto filter_split = [&]() {
        const unsigned nt = (fill + SPLIT - 1u) / SPLIT;
        unsigned bits = 0;
        if (iso) {
            for (unsigned t = nt; t-- > 0u;) {
                const float *b = chunk_lo[wave][t * SPLIT + part];
                float m = fabsf(a_lo[0] - b[0]);
#pragma unroll
                for (int k = 1; k < DIMS; ++k)
                    m = fmaxf(m, fabsf(a_lo[k] - b[k]));
                bits = tay_shift_in_ngt(bits, m, r0);
            }
        } else {
            for (unsigned t = nt; t-- > 0u;) {
                const unsigned c = t * SPLIT + part;
                const float *b_lo = chunk_lo[wave][c];
                const float *b_hi = SEEN_BOX ? chunk_hi[wave][c] : b_lo;
                bits = bits + bits + (unsigned)tay_narrow<NARROW, DIMS>(a_lo, a_hi, b_lo, b_hi, r);
            }
        }
        /* slot (nt - 1) * SPLIT + part lies past the chunk on some lanes */
        if (nt != 0u && (nt - 1u) * SPLIT + part >= fill)
            bits &= ~(1u << (nt - 1u));
        return bits;
    };

    auto process = [&]() {
        if (lane < fill) {
            const unsigned off = cand * 4u;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                chunk_lo[wave][lane][k] =
                    k < DIMS ? __uint_as_float(*(const unsigned *)(seen_col[2 + k] + off)) : 0.0f;
                if (SEEN_BOX)
                    chunk_hi[wave][lane][k] =
                        k < DIMS ? __uint_as_float(*(const unsigned *)(seen_col[6 + k] + off))
                                 : 0.0f;
            }
            /* the fourth word carries the candidate's byte offset in the
             * columns, so the compact loop reads position and offset as
             * one 16-byte record */
            chunk_lo[wave][lane][3] = __uint_as_float(off);
        }
        tay_wave_sync();

        auto visit = [&](unsigned c) {
            const float4 rec = *(const float4 *)chunk_lo[wave][c];
            union {
                typename T::seen agent;
                unsigned w[SEEN_WORDS];
            } B;
            const unsigned off = __float_as_uint(rec.w);
#pragma unroll
            for (unsigned k = 0; k < SEEN_WORDS; ++k)
                B.w[k] = *(const unsigned *)(seen_col[k] + off);
            const float pos[3] = {rec.x, rec.y, rec.z};
#pragma unroll
            for (int k = 0; k < DIMS; ++k)
                B.w[2 + k] = __float_as_uint(pos[k]);
            call(&A.agent, &B.agent, ctx);
        };
        if constexpr (SPLIT > 1u) {
            /* this lane's slots part, part + SPLIT, ...: bit t is slot
             * t * SPLIT + part, taken in ascending order */
            unsigned m = filter_split();
            if (SKIP_SELF && self_slot < 64u && self_slot % SPLIT == part)
                m &= ~(1u << (self_slot / SPLIT));
            self_slot = 64u;
            m = member ? m : 0u;
            if (COUNT)
                pairs += (unsigned)__popc(m);
            const unsigned rounds = tay_wave_max_u7((unsigned)__popc(m));
            for (unsigned round = 0; round < rounds; ++round) {
                const bool has = m != 0u;
                const unsigned c = ((unsigned)__builtin_ctz(m) & 31u) * SPLIT + part;
                m &= m - 1u;
                if (has)
                    visit(c);
            }
            tay_wave_sync();
            fill = 0;
            return;
        }
        const unsigned lo_bits = filter(0u, fill < 32u ? fill : 32u);
        const unsigned hi_bits = fill > 32u ? filter(32u, fill) : 0u;
        unsigned long long mask = ((unsigned long long)hi_bits << 32) | lo_bits;
        if (SKIP_SELF && self_slot < 64u)
            mask &= ~(1ull << self_slot);
        self_slot = 64u;
        if (BALANCED)
            mask &= chunk_valid;
        mask = member ? mask : 0ull;
        if (COUNT)
            pairs += (unsigned)__popcll(mask);
        /* The lowest set bit and its removal are computed on every lane (an
         * empty mask stays empty), so the mask needs no merge after the
         * body.  A 32-bit mask costs one find-first-bit and two operations
         * to clear per round, a 64-bit one about ten; walking the two
         * halves one after the other keeps the ascending order but takes
         * max(lo) + max(hi) rounds instead of max(lo + hi), so it is used
         * when that adds at most a quarter of the rounds (dense chunks). */
        const unsigned lo_m = (unsigned)mask, hi_m = (unsigned)(mask >> 32);
        const unsigned n_lo = (unsigned)__popc(lo_m), n_hi = (unsigned)__popc(hi_m);
        const unsigned r_lo = tay_wave_max_u7(n_lo), r_hi = tay_wave_max_u7(n_hi);
        const unsigned r_all = tay_wave_max_u7(n_lo + n_hi);
        if ((r_lo + r_hi) * 4u <= r_all * 5u) {
#pragma unroll 1
            for (unsigned h = 0; h < 2u; ++h) {
                unsigned m = h ? hi_m : lo_m;
                const unsigned base = 32u * h;
                const unsigned rounds = h ? r_hi : r_lo;
                for (unsigned round = 0; round < rounds; ++round) {
                    const bool has = m != 0u;
                    const unsigned c = base + ((unsigned)__builtin_ctz(m) & 31u);
                    m &= m - 1u;
                    if (has)
                        visit(c);
                }
            }
        } else {
            for (unsigned round = 0; round < r_all; ++round) {
                const bool has = mask != 0ull;
                const unsigned c = (unsigned)__builtin_ctzll(mask) & 63u;
                mask &= mask - 1ull;
                if (has)
                    visit(c);
            }
        }
        tay_wave_sync();
        fill = 0;
    };

    auto add_range = [&](unsigned beg, unsigned end) {
        while (beg < end) {
            const unsigned room = 64u - fill;
            const unsigned take = end - beg < room ? end - beg : room;
            if (lane - fill < take)
                cand = beg + (lane - fill);
            if (SKIP_SELF && i - beg < take)
                self_slot = fill + (i - beg);
            fill += take;
            beg += take;
            if (fill == 64u)
                process();
        }
    };

    if (SEEN_GRID) {
        /* A group of lanes whose bounding box would span more than `limit`
         * cells (about one cell plus the radius on every side) is split in
         * halves, recursively: aligned power-of-two lane ranges visited
         * depth first.  Seers far apart in space then stop scanning each
         * other's neighbourhoods; each seer still sees its cells in x->y->z
         * order. */
        const TayGridGeom *g = v.geom;
        /* Box agents are binned by their min corner: a seen agent whose box
         * reaches down to p - r has its min corner at most the group's
         * largest extent below that. */
        float ext[3] = {0.0f, 0.0f, 0.0f};
        if (SEEN_BOX) {
#pragma unroll
            for (int k = 0; k < DIMS; ++k)
                ext[k] = g->max_extent[k];
        }
        unsigned limit = 1;
#pragma unroll
        for (int k = 0; k < DIMS; ++k)
            limit *= (unsigned)ceilf((2.0f * r[k] + ext[k]) / g->cell_size[k]) + 3u;
        /* wave-uniform values held in scalar registers, so that the group
         * width and start and the split decision stay uniform too */
        limit = __builtin_amdgcn_readfirstlane(limit);
        const unsigned cx = (unsigned)g->count[0];
        const unsigned cxy = cx * (unsigned)g->count[1];
        const unsigned n_active =
            __builtin_amdgcn_readfirstlane(v.seer_end - i0 < seers ? v.seer_end - i0 : seers);
        unsigned beg = 0, width = width0;
        while (beg < n_active) {
            member = active && sl - beg < width;
            int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
            unsigned cells = 1;
#pragma unroll
            for (int k = 0; k < DIMS; ++k) {
                float tmin = tay_wave_min(member ? a_lo[k] : __builtin_inff());
                float tmax = tay_wave_max(member ? a_hi[k] : -__builtin_inff());
                /* margin (|p| + r + extent) * 2^-20 covers the rounding of
                 * p - r - extent, of the extents and of the box test, so no
                 * cell holding a passing pair is missed */
                float mlo = (fabsf(tmin) + r[k] + ext[k]) * 0x1p-20f;
                float mhi = (fabsf(tmax) + r[k]) * 0x1p-20f;
                lo[k] = tay_cell_axis(((tmin - r[k]) - ext[k]) - mlo, g->origin[k], g->cell_size[k], g->count[k]);
                hi[k] = tay_cell_axis((tmax + r[k]) + mhi, g->origin[k], g->cell_size[k], g->count[k]);
                cells *= (unsigned)(hi[k] - lo[k] + 1);
            }
            if (cells > limit && width > 1u && width > lane_walk) {
                width >>= 1;
                continue;
            }
            if (cells > limit && width > 1u) {
                /* A sparse group of at most lane_walk lanes: every member
                 * walks the cells of its own box, in x, y, z order and each
                 * cell in storage order, all members at once instead of one
                 * sub-group after another.  Memory latency is paid per batch.
                 * Plain tiles load the ranges of WALK consecutive z cells
                 * together, then each cell's candidates WALK at a time
                 * (positions loaded together, box-tested into a mask, the
                 * model called for the set bits in ascending order).  Batched
                 * tiles (ROW_WALK), mostly sparse instances with a few agents
                 * per cell, load the ranges of ZB z cells of a row together
                 * and take the row's candidates RW at a time across its
                 * cells, in the same order: fewer waits per seer, for about
                 * forty more registers, which the batched kernel has to spare
                 * and the plain kernels, compiled for seven or eight
                 * wavefronts per SIMD, would spill.  In a split pass the
                 * seer's lanes share its cell rows: lane `part` walks the
                 * rows (x, y) whose index in the seer's box is part mod
                 * SPLIT, into its partial. */
                static_assert(!ROW_WALK || SPLIT == 1u, "the row walk of batched tiles is not split");
                if (member) {
                    constexpr unsigned WALK = 2, ZB = 4, RW = 4;
                    int plo[3] = {0, 0, 0}, phi[3] = {0, 0, 0};
#pragma unroll
                    for (int k = 0; k < DIMS; ++k) {
                        /* computed here, not ahead of the lane-group loop
                         * (where the compiler would hoist the seer-only
                         * bounds and keep them in scratch for the whole tile) */
                        float al = a_lo[k], ah = a_hi[k];
                        asm volatile("" : "+v"(al), "+v"(ah));
                        const float mlo = (fabsf(al) + r[k] + ext[k]) * 0x1p-20f;
                        const float mhi = (fabsf(ah) + r[k]) * 0x1p-20f;
                        plo[k] = tay_cell_axis(((al - r[k]) - ext[k]) - mlo, g->origin[k], g->cell_size[k],
                                               g->count[k]);
                        phi[k] = tay_cell_axis((ah + r[k]) + mhi, g->origin[k], g->cell_size[k], g->count[k]);
                    }
                    if constexpr (ROW_WALK) {
                        for (int x = plo[0]; x <= phi[0]; ++x)
                            for (int y = plo[1]; y <= phi[1]; ++y)
                                for (int z0 = plo[2]; z0 <= phi[2]; z0 += (int)ZB) {
                                    /* ranges of up to ZB cells, then the row's
                                     * candidates as one sequence: position s lies
                                     * in the last cell q with first[q] <= s */
                                    unsigned beg[ZB], first[ZB], total = 0;
#pragma unroll
                                    for (unsigned q = 0; q < ZB; ++q) {
                                        const int z = z0 + (int)q <= phi[2] ? z0 + (int)q : phi[2];
                                        const TayCellRange cr =
                                            v.ranges[(unsigned)x + (unsigned)y * cx + (unsigned)z * cxy];
                                        beg[q] = cr.beg;
                                        first[q] = total;
                                        total += z0 + (int)q <= phi[2] ? cr.end - cr.beg : 0u;
                                    }
                                    for (unsigned s0 = 0; s0 < total; s0 += RW) {
                                        unsigned idx[RW];
                                        float b_lo[RW][3], b_hi[RW][3];
#pragma unroll
                                        for (unsigned t = 0; t < RW; ++t) {
                                            const unsigned s = s0 + t < total ? s0 + t : total - 1u;
                                            unsigned j = beg[0] + s;
#pragma unroll
                                            for (unsigned q = 1; q < ZB; ++q)
                                                j = first[q] <= s ? beg[q] + (s - first[q]) : j;
                                            idx[t] = j;
                                            const unsigned off = j * 4u;
#pragma unroll
                                            for (int k = 0; k < 3; ++k) {
                                                b_lo[t][k] = k < DIMS ? __uint_as_float(
                                                                            *(const unsigned *)(seen_col[2 + k] + off))
                                                                      : 0.0f;
                                                b_hi[t][k] = SEEN_BOX && k < DIMS
                                                                 ? __uint_as_float(
                                                                       *(const unsigned *)(seen_col[6 + k] + off))
                                                                 : b_lo[t][k];
                                            }
                                        }
                                        bool ok[RW];
#pragma unroll
                                        for (unsigned t = 0; t < RW; ++t) {
                                            ok[t] = s0 + t < total &&
                                                    tay_narrow<NARROW, DIMS>(a_lo, a_hi, b_lo[t], b_hi[t], r);
                                            if (SKIP_SELF)
                                                ok[t] = ok[t] && idx[t] != i;
                                        }
#pragma unroll
                                        for (unsigned t = 0; t < RW; ++t)
                                            if (ok[t]) {
                                                union {
                                                    typename T::seen agent;
                                                    unsigned w[SEEN_WORDS];
                                                } B;
                                                const unsigned off = idx[t] * 4u;
#pragma unroll
                                                for (unsigned k = 0; k < SEEN_WORDS; ++k)
                                                    B.w[k] = *(const unsigned *)(seen_col[k] + off);
                                                if (COUNT)
                                                    ++pairs;
                                                call(&A.agent, &B.agent, ctx);
                                            }
                                    }
                                }
                    } else {
                        const unsigned ny = (unsigned)(phi[1] - plo[1] + 1);
                        for (int x = plo[0]; x <= phi[0]; ++x)
                            for (int y = plo[1]; y <= phi[1]; ++y)
                                for (int z0 = plo[2]; z0 <= phi[2]; z0 += (int)WALK) {
                                    if (SPLIT > 1u &&
                                        ((unsigned)(x - plo[0]) * ny + (unsigned)(y - plo[1])) % SPLIT != part)
                                        break;
                                    TayCellRange crs[WALK];
#pragma unroll
                                    for (unsigned q = 0; q < WALK; ++q) {
                                        const int z = z0 + (int)q <= phi[2] ? z0 + (int)q : phi[2];
                                        crs[q] = v.ranges[(unsigned)x + (unsigned)y * cx + (unsigned)z * cxy];
                                    }
#pragma unroll
                                    for (unsigned q = 0; q < WALK; ++q) {
                                        if (z0 + (int)q > phi[2])
                                            break;
                                        const unsigned end = crs[q].end;
                                        for (unsigned j0 = crs[q].beg; j0 < end; j0 += WALK) {
                                            float b_lo[WALK][3], b_hi[WALK][3];
#pragma unroll
                                            for (unsigned t = 0; t < WALK; ++t) {
                                                const unsigned off = (j0 + t < end ? j0 + t : j0) * 4u;
#pragma unroll
                                                for (int k = 0; k < 3; ++k) {
                                                    b_lo[t][k] = k < DIMS ? __uint_as_float(
                                                                                *(const unsigned *)(seen_col[2 + k] + off))
                                                                          : 0.0f;
                                                    b_hi[t][k] = SEEN_BOX && k < DIMS
                                                                     ? __uint_as_float(
                                                                           *(const unsigned *)(seen_col[6 + k] + off))
                                                                     : b_lo[t][k];
                                                }
                                            }
                                            unsigned m = 0;
#pragma unroll
                                            for (unsigned t = 0; t < WALK; ++t) {
                                                const unsigned j = j0 + t;
                                                bool ok = j < end && tay_narrow<NARROW, DIMS>(a_lo, a_hi, b_lo[t], b_hi[t], r);
                                                if (SKIP_SELF)
                                                    ok = ok && j != i;
                                                m |= (unsigned)ok << t;
                                            }
                                            while (m) {
                                                const unsigned j = j0 + (unsigned)__builtin_ctz(m);
                                                m &= m - 1u;
                                                const unsigned off = j * 4u;
                                                union {
                                                    typename T::seen agent;
                                                    unsigned w[SEEN_WORDS];
                                                } B;
#pragma unroll
                                                for (unsigned k = 0; k < SEEN_WORDS; ++k)
                                                    B.w[k] = *(const unsigned *)(seen_col[k] + off);
                                                if (COUNT)
                                                    ++pairs;
                                                call(&A.agent, &B.agent, ctx);
                                            }
                                        }
                                    }
                                }
                    }
                }
            } else if (BALANCED) {
                /* The group's cells in x, y, z order (z fastest), 64 at a
                 * time, one cell per lane: the cell's first slot and the
                 * number of candidates before it (exclusive prefix sum) go
                 * to the wavefront's LDS slots; slot s of the segment's
                 * candidate sequence is then found by binary search. */
                const bool empty = hi[0] < lo[0] || hi[1] < lo[1] || hi[2] < lo[2];
                const unsigned nz = (unsigned)(hi[2] - lo[2] + 1), nyz = nz * (unsigned)(hi[1] - lo[1] + 1);
                const unsigned swept = empty ? 0u : nyz * (unsigned)(hi[0] - lo[0] + 1);
                /* the seer's own slot, excluded from its masks: its cell
                 * (the tag's Morton key, where the sort stored it) as an
                 * index into the group's cells, 0xFFFFFFFF outside them */
                unsigned own_cell = 0xFFFFFFFFu;
                if (SKIP_SELF && member && !empty) {
                    int oc[3];
                    tay_cell_of_tag(A.w[0], DIMS, oc);
                    if (oc[0] >= lo[0] && oc[0] <= hi[0] && oc[1] >= lo[1] && oc[1] <= hi[1] && oc[2] >= lo[2] &&
                        oc[2] <= hi[2])
                        own_cell = (unsigned)(oc[0] - lo[0]) * nyz + (unsigned)(oc[1] - lo[1]) * nz +
                                   (unsigned)(oc[2] - lo[2]);
                }
                for (unsigned c0 = 0; c0 < swept; c0 += 64u) {
                    const unsigned c = c0 + lane;
                    unsigned cbeg = 0, cnt = 0;
                    if (c < swept) {
                        const unsigned x = (unsigned)lo[0] + c / nyz, rem = c % nyz;
                        const unsigned y = (unsigned)lo[1] + rem / nz, z = (unsigned)lo[2] + rem % nz;
                        const TayCellRange cr = v.ranges[x + y * cx + z * cxy];
                        cbeg = cr.beg;
                        cnt = cr.end - cr.beg;
                    }
                    const unsigned incl = tay_wave_incl_sum(cnt);
                    const unsigned total = __builtin_amdgcn_readlane(incl, 63);
                    /* lanes past the segment's cells hold prefix `total`,
                     * which no slot reaches */
                    cell_ranges[wave][lane][0] = cbeg;
                    cell_ranges[wave][lane][1] = incl - cnt;
                    tay_wave_sync();
                    constexpr unsigned RUN = TAY_BALANCE_RUN, PARTS = 64u / TAY_BALANCE_RUN;
                    static_assert(RUN >= 1u && RUN <= 64u && (RUN & (RUN - 1u)) == 0u, "TAY_BALANCE_RUN");
                    const unsigned part_len = (total + PARTS - 1u) / PARTS;
                    const unsigned chunks = (part_len + RUN - 1u) / RUN;
                    /* the own seer's chunk and slot, as chunk * 64 + slot */
                    unsigned own = 0xFFFFFFFFu;
                    if (own_cell - c0 < 64u) {
                        const unsigned q = own_cell - c0;
                        const unsigned s_own = cell_ranges[wave][q][1] + (i - cell_ranges[wave][q][0]);
                        const unsigned o = s_own % part_len;
                        own = (o / RUN) * 64u + (s_own / part_len) * RUN + o % RUN;
                    }
                    {
                        /* each lane's first slot: the last cell whose prefix
                         * is at most its position (binary search); later
                         * chunks move the lane on through its part */
                        const unsigned s = (lane / RUN) * part_len + lane % RUN;
                        const unsigned s0 = s < total ? s : total - 1u;
                        unsigned q = 0;
#pragma unroll
                        for (unsigned step = 32; step; step >>= 1)
                            q = cell_ranges[wave][q + step][1] <= s0 ? q + step : q;
                        lane_cell[wave][lane] = q;
                    }
                    for (unsigned k = 0; k < chunks; ++k) {
                        const unsigned o = k * RUN + lane % RUN;
                        unsigned s = (lane / RUN) * part_len + o;
                        const bool valid = o < part_len && s < total;
                        chunk_valid = __ballot(valid);
                        fill = 64u - (unsigned)__clzll(chunk_valid);
                        if (valid) {
                            unsigned q = lane_cell[wave][lane];
                            while (q < 63u && cell_ranges[wave][q + 1u][1] <= s)
                                ++q;
                            lane_cell[wave][lane] = q;
                            cand = cell_ranges[wave][q][0] + (s - cell_ranges[wave][q][1]);
                        }
                        if (SKIP_SELF && own >> 6 == k)
                            self_slot = own & 63u;
                        process();
                    }
                }
            } else {
                /* The group's cells in x, y, z order (z fastest), their
                 * ranges loaded 64 at a time, one cell per lane, into the
                 * wavefront's LDS slots, so a sweep waits for one load per
                 * 64 cells instead of one per cell; the loop reads them back
                 * in order. */
                const bool empty = hi[0] < lo[0] || hi[1] < lo[1] || hi[2] < lo[2];
                const unsigned nz = (unsigned)(hi[2] - lo[2] + 1), nyz = nz * (unsigned)(hi[1] - lo[1] + 1);
                const unsigned swept = empty ? 0u : nyz * (unsigned)(hi[0] - lo[0] + 1);
                for (unsigned c0 = 0; c0 < swept; c0 += 64u) {
                    const unsigned c = c0 + lane;
                    if (c < swept) {
                        const unsigned x = (unsigned)lo[0] + c / nyz, rem = c % nyz;
                        const unsigned y = (unsigned)lo[1] + rem / nz, z = (unsigned)lo[2] + rem % nz;
                        const TayCellRange cr = v.ranges[x + y * cx + z * cxy];
                        cell_ranges[wave][lane][0] = cr.beg;
                        cell_ranges[wave][lane][1] = cr.end;
                    }
                    tay_wave_sync();
                    const unsigned n = swept - c0 < 64u ? swept - c0 : 64u;
                    for (unsigned k = 0; k < n; ++k)
                        add_range(__builtin_amdgcn_readfirstlane(cell_ranges[wave][k][0]),
                                  __builtin_amdgcn_readfirstlane(cell_ranges[wave][k][1]));
                    tay_wave_sync();
                }
                if (fill)
                    process();
            }
            beg += width;
            while (width < width0 && (beg & (2u * width - 1u)) == 0u)
                width <<= 1;
        }
    } else {
        add_range(v.seen_beg, v.seen_end);
        if (fill)
            process();
    }

    /* Words the model never writes fold to "unchanged" at compile time and
     * are not stored; every word it may write is stored, changed or not
     * (the storage is updated in place), so no copy of the loaded words
     * has to stay in registers through the sweep. */
    if constexpr (SPLIT > 1u) {
        /* merge the partials into part 0 in a fixed tree: (0 + 1), (2 + 3),
         * ..., then pairs of those */
#pragma unroll
        for (unsigned step = 1; step < SPLIT; step <<= 1) {
            union {
                typename T::seer agent;
                unsigned w[SEER_WORDS];
            } B;
#pragma unroll
            for (unsigned k = 0; k < SEER_WORDS; ++k)
                B.w[k] = __shfl_xor(A.w[k], (int)step);
            if ((part & (2u * step - 1u)) == 0u)
                combine(&A.agent, &B.agent);
        }
    }
    if (active && part == 0u) {
#pragma unroll
        for (unsigned k = 0; k < SEER_WORDS; ++k) {
            const bool unchanged = A.w[k] == orig[k];
            if (!(__builtin_constant_p(unchanged) && unchanged))
                v.seer_out[k * v.seer_cap + i] = A.w[k];
        }
    }
    return pairs;
}

__device__ __forceinline__ void tay_add_pairs(unsigned long long *counters, unsigned pairs) {
    const unsigned total = tay_wave_sum(pairs);
    if ((threadIdx.x & 63u) == 0u && total)
        atomicAdd(counters + blockIdx.x % TAY_PAIR_SLOTS, (unsigned long long)total);
}

/* Plain see kernel: wavefront k of the grid owns tile k. */
template <class T, unsigned SEER_WORDS, unsigned SEEN_WORDS, int NARROW, bool SKIP_SELF, bool SEEN_GRID, int DIMS,
          bool COUNT, unsigned CTX_WORDS, bool BALANCED, unsigned SPLIT, class Call, class Clear, class Combine>
__device__ __forceinline__ void tay_see_tiles(const TaySeeArgs &args, Call call, Clear clear, Combine combine) {
    constexpr unsigned WAVES = TAY_SEE_BLOCK / 64u;
    const unsigned tile = tay_xcd_block() * WAVES + (threadIdx.x >> 6);
    if (tile >= args.tiles)
        return;
    TaySeeView v;
    v.seer_cols = args.seer_cols;
    v.seer_out = args.seer_out;
    v.seer_cap = args.seer_cap;
    v.seer_end = args.seer_count;
    v.seen_cols = args.seen_cols;
    v.seen_cap = args.seen_cap;
    v.seen_beg = 0;
    v.seen_end = args.seen_count;
    v.ranges = args.ranges;
    v.geom = args.geom;
    v.context = args.context;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        v.radii[k] = args.radii[k];
    const unsigned pairs = tay_see_tile<T, SEER_WORDS, SEEN_WORDS, NARROW, SKIP_SELF, SEEN_GRID, DIMS, COUNT,
                                        CTX_WORDS, false, BALANCED, SPLIT>(v, tile * args.tile_seers,
                                                                         args.tile_seers, args.width,
                                                                         args.lane_walk, call, clear, combine);
    if (COUNT)
        tay_add_pairs(args.pair_count, pairs);
}

/* Batched see kernel: a persistent grid whose wavefronts take tiles in the
 * given order (heaviest first).  Each wavefront's first tile is fixed: the
 * w-th of the order for wavefront w, and the hardware places consecutive
 * workgroups on different compute units, so the heaviest tiles start one
 * per SIMD instead of wherever the first atomic requests happen to land;
 * after that, tiles come from a work queue (starting after the first
 * round) until it is empty.  Every wavefront reaches the exit test after
 * each tile, so the grid drains. */
template <class T, unsigned WORDS, bool SKIP_SELF, int DIMS, bool COUNT, unsigned CTX_WORDS, class Call>
__device__ __forceinline__ void tay_see_batch(const TayBatchSeeArgs &args, Call call) {
    unsigned pairs = 0;
    const unsigned waves = gridDim.x * (TAY_SEE_BLOCK / 64u);
    unsigned k = blockIdx.x * (TAY_SEE_BLOCK / 64u) + (threadIdx.x >> 6);
    for (;;) {
        if (k >= args.tiles)
            break;
        const unsigned tile = tay_uniform(args.order[k]);
        const unsigned b = tay_uniform(args.tile_inst[tile]);
        TayBatchInstance inst = args.inst[b];
        inst.off = tay_uniform(inst.off);
        inst.count = tay_uniform(inst.count);
        inst.ranges_off = ((unsigned long long)tay_uniform((unsigned)(inst.ranges_off >> 32)) << 32) |
                          tay_uniform((unsigned)inst.ranges_off);
        TaySeeView v;
        v.seer_cols = args.cols;
        v.seer_out = args.cols;
        v.seer_cap = args.cap;
        v.seer_end = inst.off + inst.count;
        v.seen_cols = args.cols;
        v.seen_cap = args.cap;
        v.seen_beg = inst.off;
        v.seen_end = inst.off + inst.count;
        v.ranges = args.ranges + inst.ranges_off;
        v.geom = args.geom + b;
        v.context = args.context + (unsigned long long)b * args.ctx_words;
#pragma unroll
        for (int d = 0; d < 4; ++d)
            v.radii[d] = tay_uniform(args.radii[4u * b + d]);
        pairs += tay_see_tile<T, WORDS, WORDS, TAY_NARROW_PP, SKIP_SELF, true, DIMS, COUNT, CTX_WORDS, true, false, 1u>(
            v, tile * 64u, 64u, args.width, args.lane_walk, call, [](typename T::seer *) {},
            [](typename T::seer *, typename T::seer *) {});
        unsigned q = 0;
        if ((threadIdx.x & 63u) == 0u)
            q = atomicAdd(args.queue, 1u);
        k = waves + __builtin_amdgcn_readfirstlane(__shfl(q, 0));
    }
    if (COUNT)
        tay_add_pairs(args.pair_count, pairs);
}

template <class T, unsigned WORDS, unsigned CTX_WORDS, class Call>
__device__ __forceinline__ void tay_act_agents(const TayActArgs &args, Call call) {
    static_assert(sizeof(typename T::agent) <= WORDS * 4u, "agent record larger than the group's agent size");
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= args.count)
        return;
    union {
        typename T::agent agent;
        unsigned w[WORDS];
    } A;
    unsigned orig[WORDS];
#pragma unroll
    for (unsigned k = 0; k < WORDS; ++k)
        orig[k] = A.w[k] = args.cols[k * args.cap + i];
    tay_ctx_holder<typename T::context, CTX_WORDS> ctx_holder(args.context);
    call(&A.agent, ctx_holder.get());
#pragma unroll
    for (unsigned k = 0; k < WORDS; ++k)
        if (A.w[k] != orig[k])
            args.cols[k * args.cap + i] = A.w[k];
}

/* Batched act kernel: one thread per slot of every instance; padding slots
 * do nothing. */
template <class T, unsigned WORDS, unsigned CTX_WORDS, class Call>
__device__ __forceinline__ void tay_act_batch(const TayBatchActArgs &args, Call call) {
    static_assert(sizeof(typename T::agent) <= WORDS * 4u, "agent record larger than the group's agent size");
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= args.slots)
        return;
    const unsigned b = args.tile_inst[i >> 6];
    const TayBatchInstance inst = args.inst[b];
    if (i - inst.off >= inst.count)
        return;
    union {
        typename T::agent agent;
        unsigned w[WORDS];
    } A;
    unsigned orig[WORDS];
#pragma unroll
    for (unsigned k = 0; k < WORDS; ++k)
        orig[k] = A.w[k] = args.cols[k * args.cap + i];
    tay_ctx_holder<typename T::context, CTX_WORDS> ctx_holder(args.context + (unsigned long long)b * args.ctx_words);
    call(&A.agent, ctx_holder.get());
#pragma unroll
    for (unsigned k = 0; k < WORDS; ++k)
        if (A.w[k] != orig[k])
            args.cols[k * args.cap + i] = A.w[k];
}
