/**
 *  @file szs_device.hpp
 *  @brief Layouts shared by the host runtime and the gfx950 kernels: engine parameters, segments, tasks, bins.
 *
 *  Data layout in HBM for one batch:
 *  - symbol tapes: the caller's byte tapes are used in place (or staged once); a UTF-8 segment additionally owns a
 *    rune tape `u32[bytes]` where the runes of string `s` start at the same index as its bytes;
 *  - `task_t[tasks]`: one 48-byte record per scored pair, rows = the shorter side;
 *  - `u32 order[tasks]`: task indices grouped by bin, heaviest bins first (device-side counting sort);
 *  - per-bin counters and prefix sums; a per-wave scratch row for pairs taller than one stripe.
 */
#pragma once
#include <stdint.h>

namespace szs_amd {

/** Recurrence family of one engine instance, derived from its kind and gap costs. */
enum family_t : int32_t {
    family_lev_linear_k = 0, /**< minimize, uniform costs, linear gaps */
    family_lev_affine_k = 1, /**< minimize, uniform costs, affine gaps */
    family_nw_linear_k = 2,  /**< maximize, class costs, global, linear */
    family_nw_affine_k = 3,  /**< maximize, class costs, global, affine */
    family_sw_linear_k = 4,  /**< maximize, class costs, local, linear */
    family_sw_affine_k = 5,  /**< maximize, class costs, local, affine */
    families_count_k = 6,
};

/** Symbol width of a segment: bytes, or runes decoded from UTF-8. */
enum symbols_t : int32_t { symbols_u8_k = 0, symbols_u32_k = 1 };
/**
 *  Class-cost families never score runes, so their variant code 1 marks byte pairs of an engine whose
 *  `cost - open` does not fit the int8 profile: those run a sweep that keeps scores unshifted.
 */
constexpr int32_t class_unshifted_k = 1;

constexpr int max_dense_classes_k = 32;

/** Device copy of one engine's parameters. */
struct engine_params_t {
    int32_t family;
    int32_t open, extend;     /**< gap costs (sign as given: positive for distances, negative for scores) */
    int32_t match, mismatch;  /**< uniform substitution costs */
    int32_t dense_count;      /**< number of distinct classes actually used by `byte_to_class` */
    /** Cost bias of the sweep's fast forms: `open + extend` for global families (their values live relative to
     *  the anti-diagonal), `open` for local ones (scores kept as H + open). The LDS profile holds
     *  `dense_costs - shift` as int8; 0 when some used cost would not fit (the plain recurrences then run). */
    int32_t shift;
    /** Most rows per lane (8, 6, 4, 2; 0 = none) for which the packed 16-bit sweep may hold one base for the whole
     *  wave instead of one per lane (runtime.hip, `packed_uniform_rows_limit`). */
    int32_t packed_uniform_rows;
    /** Bias of the packed 16-bit sweep's profile bytes `cost - packed_bias` (unsigned): `shift` for global families,
     *  min(open, smallest used cost) for local linear gaps, whose substitution term then subtracts
     *  `open - packed_bias` again. */
    int32_t packed_bias;
    uint8_t byte_to_dense[256];
    int8_t dense_costs[max_dense_classes_k * max_dense_classes_k]; /**< [dense(query)][dense(candidate)] */
};

/** One homogeneous slice of a batch, as the kernels see it. */
struct segment_view_t {
    engine_params_t const *params;
    void const *query_symbols;     /**< u8 tape data, or u32 rune tape for UTF-8 */
    void const *candidate_symbols; /**< same for candidates */
    int64_t *results;
    int32_t symbols;               /**< `symbols_t` */
    int32_t family;
};

/** One scored pair. `rows` is the shorter side. */
struct task_t {
    uint64_t rows_offset;
    uint64_t columns_offset;
    uint32_t rows_length;
    uint32_t columns_length;
    uint32_t segment;
    /** Bit 0: the rows are the candidate (transposed); bits 8-31: 1 + the pair's slot in the chain buffers when
     *  a chained shape scores it (0 otherwise). */
    uint32_t flags;
    uint64_t output;
    uint64_t mirror; /**< second output slot of a symmetric cross-product, or ~0 */
};
static_assert(sizeof(task_t) == 48, "task layout");

constexpr uint64_t no_mirror_k = ~0ull;
constexpr uint32_t task_transposed_k = 1u;

__host__ __device__ inline bool task_transposed(task_t const &task) { return (task.flags & task_transposed_k) != 0; }
/** Slot of a chained pair in the chain buffers. */
__host__ __device__ inline uint32_t task_chain_slot(task_t const &task) { return (task.flags >> 8) - 1u; }

/**
 *  A kernel shape: `group` lanes share one pair (64 / group pairs per wave), each lane holds `rows_per_lane`
 *  consecutive rows, so one stripe covers `group * rows_per_lane` rows. A 64-lane group sweeps taller pairs
 *  stripe after stripe (boundary row through HBM scratch). A `cooperative` shape gives one pair to all the
 *  waves of a workgroup, which pipeline its stripes in lock-step phases.
 */
struct shape_t {
    int32_t group;
    int32_t rows_per_lane;
    int32_t cooperative;
    int32_t bit_parallel; /**< unit-cost bytes: one 64-row Myers word per lane instead of R DP rows */
    int32_t packed;       /**< class costs: two 16-bit cells per register (two half-stripes per lane) */
    /** Cooperative shapes only: the pair's stripes are split into `chain_groups_k` consecutive work items, each
     *  a workgroup; a group reads the previous group's bottom row from HBM as it is published. */
    int32_t chained;
    /** int64 cells (one wave per pair, bottom rows in per-pair HBM rows): pairs whose values could leave int32. */
    int32_t wide = 0;
};

/** Waves per workgroup of the scoring kernels. */
constexpr int waves_per_group_k = 4;

/**
 *  The shapes compiled; index = shape id. Bit-parallel shapes exist for the unit-cost byte Levenshtein only, packed
 *  16-bit shapes for the class-cost families, for engines whose values provably fit (else their int32 twin runs).
 */
constexpr int shapes_count_k = 27;
constexpr shape_t shapes_k[shapes_count_k] = {
    {16, 1, 0, 0, 0, 0}, {16, 2, 0, 0, 0, 0}, {16, 4, 0, 0, 0, 0},  // up to 16 / 32 / 64 rows, four pairs per wave
    {64, 2, 0, 0, 0, 0}, {64, 3, 0, 0, 0, 0}, {64, 4, 0, 0, 0, 0},  // up to 128 / 192 / 256 rows
    {64, 6, 0, 0, 0, 0}, {64, 8, 0, 0, 0, 0}, {64, 16, 0, 0, 0, 0}, // up to 384 / 512 / 1024 rows per stripe
    {64, 8, 1, 0, 0, 0}, {64, 16, 1, 0, 0, 0},     // cooperative: 512- or 1024-row stripes, whole workgroup
    {4, 64, 0, 1, 0, 0}, {16, 64, 0, 1, 0, 0},     // Myers: up to 256 rows (16 pairs/wave), 1024 rows
    {64, 64, 0, 1, 0, 0}, {64, 64, 1, 1, 0, 0},    // Myers: 4096-row stripes on one wave, or cooperative
    {64, 4, 0, 0, 1, 0}, {64, 6, 0, 0, 1, 0}, {64, 8, 0, 0, 1, 0}, // packed 16-bit: 256 / 384 / 512-row stripes
    {64, 8, 1, 0, 1, 0},                           // packed 16-bit, cooperative 512-row stripes
    {64, 8, 1, 0, 1, 1}, {64, 8, 1, 0, 0, 1},      // chained cooperative 512-row stripes, packed or int32
    {16, 32, 0, 1, 0, 0}, {64, 32, 0, 1, 0, 0},    // Myers over 32-row words: up to 512 rows, 2048-row stripes
    {16, 8, 0, 0, 0, 0},                           // up to 128 rows, four pairs per wave
    {64, 8, 0, 0, 0, 0, 1},                        // int64 cells, 512-row stripes on one wave
    {4, 32, 0, 1, 0, 0},                           // Myers over 32-row words: up to 128 rows (16 pairs/wave)
    {16, 10, 0, 0, 0, 0},                          // up to 160 rows, four pairs per wave
};
/** The int64-cell shape, which pairs beyond the int32 range bound take (any family, bytes or runes). */
constexpr int wide_shape_k = 24;

/** The int32 shape with the same geometry as a packed one (itself for every other shape). */
__host__ __device__ constexpr int unpacked_twin(int shape) {
    if (!shapes_k[shape].packed) return shape;
    for (int s = 0; s < shapes_count_k; ++s)
        if (!shapes_k[s].packed && !shapes_k[s].bit_parallel && !shapes_k[s].wide &&
            shapes_k[s].group == shapes_k[shape].group &&
            shapes_k[s].rows_per_lane == shapes_k[shape].rows_per_lane &&
            shapes_k[s].cooperative == shapes_k[shape].cooperative && shapes_k[s].chained == shapes_k[shape].chained)
            return s;
    return shape;
}
static_assert(unpacked_twin(15) == 5 && unpacked_twin(16) == 6 && unpacked_twin(17) == 7 && unpacked_twin(18) == 9 &&
                  unpacked_twin(19) == 20,
              "every packed shape has an int32 twin");

/** The cooperative shape with the same geometry as a one-wave shape (-1 when there is none). */
__host__ __device__ constexpr int cooperative_twin(int shape) {
    if (shapes_k[shape].cooperative || shapes_k[shape].wide || shapes_k[shape].group != 64) return -1;
    for (int s = 0; s < shapes_count_k; ++s)
        if (shapes_k[s].cooperative && !shapes_k[s].chained && shapes_k[s].group == 64 &&
            shapes_k[s].rows_per_lane == shapes_k[shape].rows_per_lane &&
            shapes_k[s].bit_parallel == shapes_k[shape].bit_parallel && shapes_k[s].packed == shapes_k[shape].packed)
            return s;
    return -1;
}
static_assert(cooperative_twin(17) == 18 && cooperative_twin(13) == 14 && cooperative_twin(7) == 9 &&
                  cooperative_twin(8) == 10 && cooperative_twin(15) == -1 && cooperative_twin(22) == -1 &&
                  cooperative_twin(wide_shape_k) == -1 && unpacked_twin(wide_shape_k) == wide_shape_k,
              "one-wave shapes and their cooperative twins");

/**
 *  Work items (workgroups) a chained pair is split into: groups of at most 4 stripes, one per wave, so that a
 *  group's last stripe starts a few phases after its first and the next group can follow right behind; pairs of
 *  more than 4 x 64 stripes put more stripes in each group. Trailing groups of shorter pairs are empty.
 */
constexpr int chain_groups_k = 64;

/** Chain buffers of a batch: the bottom rows between the groups of each chained pair
 *  ([slot][group < chain_groups_k - 1][scores, gaps][pitch] int32), the columns each has published
 *  ([slot][group] u32, zeroed before every launch), and a word raised when a group gives up waiting. */
struct chain_view_t {
    int32_t *rows;
    uint32_t *progress;
    uint32_t pitch; /**< >= the longest chained pair's columns + 1 */
    uint32_t *timeout;
    /** Pairs of the int64 shape: per pair, its stripes' bottom row and vertical gaps ([slot][2][wide_pitch]). */
    int64_t *wide_rows;
    uint64_t wide_pitch; /**< >= the longest such pair's columns + 1 */
};

/** Shape-table slots of the unit-cost Levenshtein over bytes and over runes (after the family x symbols slots):
 *  the ones whose ladders may use bit-parallel shapes. */
constexpr int unit_bytes_slot_k = families_count_k * 2;
constexpr int unit_runes_slot_k = families_count_k * 2 + 1;
constexpr int shape_slots_k = families_count_k * 2 + 2;

/** Pairs one work item holds: one per wave slot of a non-cooperative shape, one pair for a cooperative one. */
/** Pairs of one work item: a cooperative item is one pair for the whole workgroup, any other item is the pairs of
 *  one wave (64 / lanes per pair), which each wave of a workgroup takes and sweeps on its own. */
__host__ __device__ inline uint32_t pairs_per_item(shape_t shape) {
    return shape.cooperative ? 1u : (uint32_t)(64 / shape.group);
}

/** Work items of `tasks` pairs of one shape: a chained pair is `chain_groups_k` items, one per group. */
__host__ __device__ inline uint32_t items_for(shape_t shape, uint32_t tasks) {
    if (shape.chained) return tasks * (uint32_t)chain_groups_k;
    uint32_t const per_item = pairs_per_item(shape);
    return (tasks + per_item - 1) / per_item;
}

/** Cost buckets per shape for binning: eighth-octaves of (stripes x (columns + group)), so that the pairs sharing
 *  a wave in a small-group shape (which runs until its longest pair ends) differ by at most 9% in work. */
constexpr int bucket_steps_per_octave_k = 8;
constexpr int column_buckets_k = 32 * bucket_steps_per_octave_k;
/** Items of this bucket and above (about 2.1 M lane-row-steps, e.g. a 65,536 x 65,536 pair) run at raised wave
 *  priority: their latency, not the batch's total work, would otherwise decide when the launch ends. */
constexpr int critical_bucket_k = 21 * bucket_steps_per_octave_k;
/** Variant = (family, symbols, shape); bins = variants x column buckets. */
constexpr int variants_count_k = families_count_k * 2 * shapes_count_k;
constexpr int bins_count_k = variants_count_k * column_buckets_k;

__host__ __device__ inline int variant_of(int family, int symbols, int shape) {
    return (family * 2 + symbols) * shapes_count_k + shape;
}

/** Per-batch counters and prefix sums the binning kernels fill and the persistent kernel reads. */
struct bins_view_t {
    uint32_t *histogram;     /**< [bins] tasks per bin */
    uint32_t *cursor;        /**< [bins] scatter cursor */
    uint32_t *task_start;    /**< [bins] first slot of the bin in `order`, in priority order */
    uint32_t *item_start;    /**< [bins + 1] first work item of the bin (prefix over priority order) */
    uint32_t *priority;      /**< [bins] bin ids sorted by estimated cost per item, heaviest first */
    uint32_t *order;         /**< [tasks] task ids grouped by bin */
    uint32_t *queue_head;    /**< persistent-kernel work counter */
    uint32_t *shape_of_task; /**< [tasks] chosen shape id (kept for the per-pair mode) */
    uint32_t *scan_blocks;   /**< [scan_blocks_k][tasks, items]: totals of the first scan pass's workgroups */
    /** Priority position of the first bin of one-wave items: the bins before it (chained, then cooperative) hold
     *  whole-workgroup items, so `item_start[wave_bins_from]` is where the waves start taking items alone. */
    uint32_t wave_bins_from;
};

/** Workgroups of 1,024 positions the bin scan uses at most. */
constexpr uint32_t scan_blocks_k = 128;

/** Shape choice per (family, symbols): single-pass shapes by ascending capacity, then the multi-pass shape. */
struct shape_table_t {
    int32_t limits[shape_slots_k][shapes_count_k]; /**< most rows a shape takes in this slot, -1 = unused */
    int32_t order[shape_slots_k][shapes_count_k];  /**< shapes to try, ascending capacity */
    int32_t used[shape_slots_k];                   /**< entries in `order` */
    int32_t multi_pass_shape[shape_slots_k];       /**< shape for rows beyond every listed limit */
};

/** Inputs of one segment for materialization. */
struct materialize_args_t {
    uint64_t const *query_offsets;
    uint64_t const *candidate_offsets;   /**< null for the symmetric cross-product */
    uint32_t const *query_rune_counts;   /**< UTF-8 only */
    uint32_t const *candidate_rune_counts;
    uint64_t queries_count;
    uint64_t candidates_count;
    uint64_t query_bytes;                /**< bytes of the query tape's data (`unknown_bytes_k`: its last offset) */
    uint64_t candidate_bytes;            /**< the same for the candidates; every string must end within them */
    uint64_t results_row_stride;
    uint64_t tasks_count;                /**< pairs, or cells of the (triangular) cross-product */
    uint32_t task_base;
    uint32_t segment;
    int32_t family;
    int32_t symbols;
    int32_t mode;                        /**< 0 pairs, 1 cross-product, 2 symmetric */
    int32_t open, extend;
    int32_t unit_costs;                  /**< match 0, mismatch 1, gap 1: bit-parallel shapes apply to bytes */
    int32_t runes;                       /**< lengths come from the rune counts (UTF-8) */
    int32_t cell_growth;                 /**< max |substitution| + 6 max |gap| of the engine, for the range bound */
    int32_t packed_rows;                 /**< most rows per lane of a packed 16-bit shape the engine admits (0: none) */
    int32_t latency_stripes;             /**< > 0: pairs of more stripes than this on one wave go cooperative */
};

/** Bytes of the per-wave LDS ring that stages lane 0's inputs of a 64-step chunk (64 entries of 16 bytes); it
 *  follows the wave's profile or Myers tables in dynamic LDS. */
constexpr uint32_t inject_ring_bytes_k = 64 * 16;
/** Myers LDS per wave: renumbering tables (one set for a 64-lane group, a 64-byte set per smaller group), then a
 *  [14 codes][64 lanes] u64 match table (split among the groups; code 0 = absent from the rows, so at most 13
 *  distinct symbols), then the ring: 9 KB, so that four workgroups of four waves fit a CU's 160 KB (with 16 codes
 *  the 10 KB per wave plus the kernel's static LDS left room for three). */
constexpr uint32_t myers_tables_bytes_k = 1024;
constexpr int32_t myers_eq_codes_k = 14;
constexpr uint32_t myers_lds_bytes_k = myers_tables_bytes_k + myers_eq_codes_k * 64 * 8 + inject_ring_bytes_k;

/** One UTF-8 tape to transcode: its strings are numbered from `string_base` across the batch. */
struct utf8_tape_t {
    uint8_t const *bytes;
    uint64_t const *offsets;
    uint32_t *runes;       /**< u32[capacity]: runes of string s start at index offsets[s] */
    uint32_t *rune_counts; /**< u32[count] */
    uint64_t string_base;
    uint64_t capacity;     /**< runes the buffer holds: the tape's byte count as the host knows it */
    uint32_t segment;
    uint32_t reserved;
};

/**
 *  Core recurrence work of one cell in the tier that scores it, in 32-bit VALU lane-ops per 64 cells: the ops of
 *  the cell update as each tier formulates it, without data movement (DPP hand-offs, LDS reads, addressing, loop
 *  control), so a launch issues at least this many. The roofline's algorithmic work is the cell count weighted by
 *  these figures:
 *  - bit-parallel (Hyyro's step on one word per lane, three-input terms as `v_bitop3`): 25 ops per 64-row word
 *    step on 32-bit halves, 15 per 32-row word; over runes the match word adds one `v_bfe` and one `v_bitop3` per
 *    half for each of at least 8 bit-planes (the smallest plane count compiled);
 *  - packed 16-bit NW: 6 packed ops per two cells (affine: perm, add, two max, three-way maximum, add), 3 for
 *    linear gaps (perm, add, three-way maximum); packed local 7 with linear gaps (perm, two adds, max against the
 *    zero floor, three-way maximum, add, max into the best) and 11 with affine ones (two adds and two maxima for
 *    the gap tracks);
 *  - int32 DP in the anti-diagonal domain (global families): linear 2 (add, max3), affine 5 (two max, add, max3,
 *    add); uniform costs add a compare and a select (Levenshtein 4 and 7); local families 4 (linear) and 8 (affine).
 */
__host__ __device__ constexpr uint32_t core_ops_per_64_cells(int family, int symbols, shape_t shape) {
    bool const affine = family == family_lev_affine_k || family == family_nw_affine_k || family == family_sw_affine_k;
    if (shape.bit_parallel) {
        bool const runes = symbols == symbols_u32_k;
        return shape.rows_per_lane == 64 ? (runes ? 25 + 3 * 8 : 25) : (runes ? 2 * (15 + 2 * 8) : 2 * 15);
    }
    if (shape.packed) return family == family_sw_linear_k ? 224 : family == family_sw_affine_k ? 352 : affine ? 192 : 96;
    if (shape.wide) return affine ? 20 * 64 : 8 * 64; // int64 adds and selects: two or three lane-ops each
    switch (family) {
    case family_lev_linear_k: return 4 * 64;
    case family_lev_affine_k: return 7 * 64;
    case family_nw_linear_k: return 2 * 64;
    case family_nw_affine_k: return 5 * 64;
    case family_sw_linear_k: return 4 * 64;
    default: return 8 * 64;
    }
}

/** What transcoding and materialization report back to the host: a few bytes instead of the task list. */
struct prepare_stats_t {
    unsigned long long cells;            /**< sum of n*m over non-empty tasks */
    unsigned long long core_ops_64;      /**< sum of n*m x `core_ops_per_64_cells` of each task's tier */
    uint32_t widest_multi_pass;          /**< most columns of a pair that needs several stripes */
    uint32_t overflow;                   /**< nonzero if some pair's scores could leave int32 */
    uint32_t bad_tape;                   /**< nonzero if some tape's offsets decrease or run past its byte count */
    uint32_t chained_pairs;              /**< pairs scored by a chained shape: slots of the chain buffers */
    uint32_t chained_columns;            /**< most columns of such a pair */
    uint32_t chain_timeout;              /**< nonzero if a chained group gave up waiting for its predecessor */
    uint32_t wide_pairs;                 /**< pairs scored by the int64 shape: slots of the wide rows */
    uint32_t wide_columns;               /**< most columns of such a pair */
    uint32_t decode_ticket;              /**< next ticket of the UTF-8 decode queue (8 strings each) */
};

/** A tape byte count the host did not read back from HBM: kernels that check offsets read `offsets[count]`. */
constexpr uint64_t unknown_bytes_k = ~0ull;

/** Up to 8 fills of u32 counters in one launch. */
struct fill_list_t {
    uint32_t *data[8];
    uint32_t count[8];
    uint32_t value[8];
    uint32_t entries;
};

/** Per-segment flags raised on the device (the batch fails, or only that segment's request when serving). */
enum segment_flag_t : uint32_t { segment_overflow_k = 1, segment_bad_tape_k = 2 };

} // namespace szs_amd
