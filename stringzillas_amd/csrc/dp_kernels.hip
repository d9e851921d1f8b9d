/**
 *  @file dp_kernels.hip
 *  @brief gfx950 kernels of the similarity engine: UTF-8 transcoding, pair materialization, length binning and
 *         the persistent dynamic-programming kernel.
 *
 *  Recurrences restated from include/stringzillas/similarities/serial.hpp:669-1079 (tile scorers) with the border
 *  conventions of its walkers (serial.hpp:1153-1158, 1310-1320, 1627-1633) and the empty-string rules of
 *  serial.hpp:142-151. The traversal is this file's own: a wave-synchronous systolic sweep.
 *
 *  Sweep: a group of G lanes scores one pair. The rows side (the shorter string) is split into stripes of G*R
 *  rows; lane k keeps rows [k*R, k*R+R) of the stripe in VGPRs, together with the score of each row in the last
 *  column it has seen. At step t lane k scores column j = t + 1 - k for all its R rows, then hands the score and
 *  vertical-gap value of its bottom row, plus the column symbol, to lane k+1 through one DPP shift each
 *  (`wave_shr:1` for G = 64, `row_shr:1` for G = 16). Lane 0 takes those values from the stripe above: the
 *  matrix border for the first stripe, the previous stripe's bottom row (spilled to per-wave HBM scratch)
 *  otherwise. Per step a lane does R cell updates and a fixed ~10 VALU ops of exchange and bookkeeping.
 *  Cells are int32; the host proves the reach bound of serial.hpp:117-140 fits before launching.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "szs_device.hpp"

namespace szs_amd {

#pragma region DPP helpers

/** DPP controls (GFX9 encoding). */
enum : int {
    dpp_row_shr1_k = 0x111,  /**< lane i <- lane i-1 within each row of 16; lane 0 of a row keeps `old` */
    dpp_row_ror15_k = 0x12F, /**< lane i <- lane (i+1) mod 16 within each row */
    dpp_wave_shl1_k = 0x130, /**< lane i <- lane i+1; lane 63 keeps `old` */
    dpp_wave_rol1_k = 0x134, /**< lane i <- lane (i+1) mod 64 */
    dpp_wave_shr1_k = 0x138, /**< lane i <- lane i-1; lane 0 keeps `old` */
};

template <int G>
__device__ __forceinline__ int32_t shift_down(int32_t old, int32_t src) {
    if constexpr (G == 64) return __builtin_amdgcn_update_dpp(old, src, dpp_wave_shr1_k, 0xF, 0xF, false);
    else return __builtin_amdgcn_update_dpp(old, src, dpp_row_shr1_k, 0xF, 0xF, false);
}

template <int G>
__device__ __forceinline__ int32_t rotate_up(int32_t src) {
    if constexpr (G == 64) return __builtin_amdgcn_update_dpp(src, src, dpp_wave_rol1_k, 0xF, 0xF, false);
    else return __builtin_amdgcn_update_dpp(src, src, dpp_row_ror15_k, 0xF, 0xF, false);
}

/** Shift register toward lane 0: lane 63 appends its own `fresh` value. Used for G = 64 only. */
__device__ __forceinline__ int32_t collect_from_last(int32_t fresh, int32_t stage) {
    return __builtin_amdgcn_update_dpp(fresh, stage, dpp_wave_shl1_k, 0xF, 0xF, false);
}

#pragma endregion

#pragma region Family traits

template <int family_>
struct family_traits {
    static constexpr bool minimize_k = family_ == family_lev_linear_k || family_ == family_lev_affine_k;
    static constexpr bool affine_k =
        family_ == family_lev_affine_k || family_ == family_nw_affine_k || family_ == family_sw_affine_k;
    static constexpr bool local_k = family_ == family_sw_linear_k || family_ == family_sw_affine_k;
    static constexpr bool uniform_k = minimize_k;
};

template <bool minimize_>
__device__ __forceinline__ int32_t opt(int32_t a, int32_t b) {
    if constexpr (minimize_) return a < b ? a : b;
    else return a > b ? a : b;
}

/** Score of the matrix border at distance `k` from the origin (serial.hpp:142-151, 716-727, 889-895). */
template <int family_>
__device__ __forceinline__ int32_t border_score(int32_t k, int32_t open, int32_t extend) {
    using traits = family_traits<family_>;
    if constexpr (traits::local_k) return 0;
    else if constexpr (traits::affine_k) return k ? open + extend * (k - 1) : 0;
    else return open * k;
}

#pragma endregion

#pragma region Scoring one work item

/** Everything the persistent kernel hands to one work item. */
struct item_args_t {
    task_t const *tasks;
    segment_view_t const *segments;
    uint32_t const *order; /**< task ids, `first .. first + count` belong to this item */
    uint32_t first;
    uint32_t count;
    int32_t *scratch;      /**< this wave's 2 x `scratch_capacity` int32: stripe-bottom scores and vertical gaps */
    uint32_t scratch_capacity;
    int8_t *profile;       /**< this wave's LDS substitution profile: [dense class][lane][R] int8 */
};

/**
 *  Scores up to 64 / G pairs with one wave. All lanes execute every step (the DPP exchange needs them); lanes
 *  whose column is outside [1, m] skip the cell updates under the exec mask.
 */
template <int family_, typename symbol_type_, int G, int R>
__device__ void score_item(item_args_t const &args) {
    using traits = family_traits<family_>;
    constexpr bool minimize_k = traits::minimize_k;
    constexpr bool affine_k = traits::affine_k;
    constexpr bool local_k = traits::local_k;
    constexpr bool uniform_k = traits::uniform_k;
    constexpr int RP = (R + 3) / 4 * 4; // profile bytes per lane and class, padded to whole dwords

    int const lane = (int)__lane_id();
    int const group = lane / G;
    int const k = lane % G;
    bool const active = (uint32_t)group < args.count;
    task_t const task = args.tasks[args.order[args.first + (active ? group : 0)]];
    segment_view_t const segment = args.segments[task.segment];
    engine_params_t const *params = segment.params;
    int32_t const n = active ? (int32_t)task.rows_length : 0;
    int32_t const m = active ? (int32_t)task.columns_length : 0;
    int32_t const open = params->open, extend = params->extend;
    int32_t const match = params->match, mismatch = params->mismatch;
    int32_t const dense_count = params->dense_count;
    bool const transposed = task.transposed != 0;
    symbol_type_ const *query = (symbol_type_ const *)segment.query_symbols;
    symbol_type_ const *candidate = (symbol_type_ const *)segment.candidate_symbols;
    symbol_type_ const *rows = (transposed ? candidate : query) + task.rows_offset;
    symbol_type_ const *columns = (transposed ? query : candidate) + task.columns_offset;

    // Wave-wide loop bounds: one pair per wave for G = 64, otherwise the longest of the wave's pairs.
    int32_t longest = m;
    if constexpr (G < 64) {
        for (int offset = G; offset < 64; offset <<= 1) longest = max(longest, __shfl_xor(longest, offset, 64));
    }
    int32_t const stripe_rows = G * R;
    int32_t const passes = G == 64 ? (n + stripe_rows - 1) / stripe_rows : 1;
    int32_t const steps = longest + G - 1;
    int32_t *scratch_scores = args.scratch;
    int32_t *scratch_gaps = args.scratch + args.scratch_capacity;
    int32_t running_best = 0; // local alignment: best cell over all passes (serial.hpp:781 starts it at 0)

    for (int32_t pass = 0; pass < passes; ++pass) {
        int32_t const base = pass * stripe_rows; // rows above this stripe
        int32_t const lane_top = base + k * R;   // rows above this lane's first row
        bool const from_scratch = pass > 0;

        // Per-row state: symbol (or dense class), score in the last column seen, horizontal gap track, best.
        uint32_t row_symbols[R];
        int32_t scores[R];
        int32_t horizontal[R];
        int32_t best[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            int32_t const i = lane_top + r + 1; // 1-based matrix row
            uint32_t symbol = 0xFFFFFFFFu;
            if (i <= n) symbol = (uint32_t)rows[i - 1];
            row_symbols[r] = symbol;
            scores[r] = border_score<family_>(i, open, extend);
            horizontal[r] = (open + extend) + scores[r]; // discard value of the gap track on column 0
            best[r] = 0;
        }
        if constexpr (!uniform_k) {
            // Profile: this lane's R row costs against every dense column class, packed 4 per dword.
            int8_t *mine = args.profile + lane * RP;
            for (int32_t c = 0; c < dense_count; ++c) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    uint32_t const row_class = row_symbols[r] == 0xFFFFFFFFu ? 0u : params->byte_to_dense[row_symbols[r] & 0xFF];
                    int8_t const cost = transposed ? params->dense_costs[c * max_dense_classes_k + row_class]
                                                   : params->dense_costs[row_class * max_dense_classes_k + c];
                    mine[c * 64 * RP + r] = cost;
                }
            }
        }

        int32_t diagonal_top = border_score<family_>(lane_top, open, extend); // H[top - 1][0]
        int32_t bottom_score = scores[R - 1];
        int32_t bottom_gap = 0;
        uint32_t column_symbol = 0;

        for (int32_t chunk = 0; chunk < steps; chunk += G) {
            // Next G column symbols for this group and the next G border values for its lane 0.
            int32_t const j_lane = chunk + k + 1; // 1-based column this lane loads
            uint32_t chunk_symbols = 0;
            if (j_lane <= m) {
                uint32_t const raw = (uint32_t)columns[j_lane - 1];
                chunk_symbols = uniform_k ? raw : (uint32_t)params->byte_to_dense[raw & 0xFF];
            }
            int32_t chunk_scores, chunk_gaps;
            if (from_scratch) {
                chunk_scores = j_lane <= m ? scratch_scores[j_lane] : 0;
                chunk_gaps = j_lane <= m ? scratch_gaps[j_lane] : 0;
            }
            else {
                chunk_scores = border_score<family_>(j_lane, open, extend);
                chunk_gaps = (open + extend) + chunk_scores; // discard value of the vertical track on row 0
            }
            int32_t stage_scores = 0, stage_gaps = 0;

#pragma unroll 2
            for (int32_t step = 0; step < G; ++step) {
                int32_t const t = chunk + step;
                int32_t const up_score = shift_down<G>(chunk_scores, bottom_score);
                int32_t up_gap = 0;
                if constexpr (affine_k) up_gap = shift_down<G>(chunk_gaps, bottom_gap);
                column_symbol = (uint32_t)shift_down<G>((int32_t)chunk_symbols, (int32_t)column_symbol);
                chunk_scores = rotate_up<G>(chunk_scores);
                if constexpr (affine_k) chunk_gaps = rotate_up<G>(chunk_gaps);
                chunk_symbols = (uint32_t)rotate_up<G>((int32_t)chunk_symbols);

                int32_t const j = t + 1 - k;
                if ((uint32_t)(j - 1) < (uint32_t)m) {
                    int32_t diagonal = diagonal_top;
                    int32_t up = up_score;
                    int32_t vertical = up_gap;
                    uint32_t costs_packed[RP / 4];
                    if constexpr (!uniform_k) {
                        uint32_t const *row_costs =
                            (uint32_t const *)(args.profile + (column_symbol * 64 + lane) * RP);
#pragma unroll
                        for (int w = 0; w < RP / 4; ++w) costs_packed[w] = row_costs[w];
                    }
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        int32_t cost;
                        if constexpr (uniform_k) cost = row_symbols[r] == column_symbol ? match : mismatch;
                        else cost = (int32_t)(int8_t)(costs_packed[r / 4] >> (8 * (r % 4)));
                        int32_t const left = scores[r];
                        int32_t const if_substitution = diagonal + cost;
                        int32_t cell;
                        if constexpr (affine_k) {
                            int32_t const h = opt<minimize_k>(left + open, horizontal[r] + extend);
                            int32_t const v = opt<minimize_k>(up + open, vertical + extend);
                            horizontal[r] = h;
                            vertical = v;
                            int32_t const gap = opt<minimize_k>(h, v);
                            if constexpr (local_k) cell = max(gap, max(if_substitution, 0));
                            else cell = opt<minimize_k>(gap, if_substitution);
                        }
                        else {
                            int32_t const gap = opt<minimize_k>(up, left) + open;
                            if constexpr (local_k) cell = max(gap, max(if_substitution, 0));
                            else cell = opt<minimize_k>(gap, if_substitution);
                        }
                        if constexpr (local_k) best[r] = max(best[r], cell);
                        diagonal = left;
                        up = cell;
                        scores[r] = cell;
                    }
                    bottom_score = up;
                    bottom_gap = vertical;
                }
                diagonal_top = up_score;
                if constexpr (G == 64) {
                    stage_scores = collect_from_last(bottom_score, stage_scores);
                    if constexpr (affine_k) stage_gaps = collect_from_last(bottom_gap, stage_gaps);
                }
            }
            // Lane 63's bottom row of the last G steps, oldest in lane 0: columns chunk - 62 .. chunk + 1.
            if constexpr (G == 64) {
                if (pass + 1 < passes) {
                    int32_t const j_out = chunk + lane - 62;
                    if (j_out >= 1 && j_out <= m) {
                        scratch_scores[j_out] = stage_scores;
                        if constexpr (affine_k) scratch_gaps[j_out] = stage_gaps;
                    }
                }
            }
        }
        if constexpr (G == 64) {
            if (pass + 1 < passes) __threadfence_block(); // lane 63's stores before lane 0's loads next pass
        }

        // Results.
        if constexpr (local_k) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (lane_top + r + 1 <= n) running_best = max(running_best, best[r]);
        }
        else if (pass + 1 == passes) {
            int32_t const last_row = n - 1 - base; // 0-based row within this stripe
            if (active && last_row / R == k) {
                int32_t value = scores[0];
#pragma unroll
                for (int r = 1; r < R; ++r)
                    if (last_row % R == r) value = scores[r];
                segment.results[task.output] = (int64_t)value;
                if (task.mirror != no_mirror_k) segment.results[task.mirror] = (int64_t)value;
            }
        }
    }
    if constexpr (local_k) {
        for (int offset = 1; offset < G; offset <<= 1) running_best = max(running_best, __shfl_xor(running_best, offset, 64));
        if (active && k == 0) {
            segment.results[task.output] = (int64_t)running_best;
            if (task.mirror != no_mirror_k) segment.results[task.mirror] = (int64_t)running_best;
        }
    }
}

#pragma endregion

#pragma region Dispatch

template <int family_, typename symbol_type_, int shape_>
__device__ __forceinline__ void run_shape(item_args_t const &args) {
    constexpr shape_t shape = shapes_k[shape_];
    score_item<family_, symbol_type_, shape.group, shape.rows_per_lane>(args);
}

template <int family_, typename symbol_type_>
__device__ void run_family(int shape, item_args_t const &args) {
    switch (shape) {
    case 0: run_shape<family_, symbol_type_, 0>(args); break;
    case 1: run_shape<family_, symbol_type_, 1>(args); break;
    case 2: run_shape<family_, symbol_type_, 2>(args); break;
    case 3: run_shape<family_, symbol_type_, 3>(args); break;
    case 4: run_shape<family_, symbol_type_, 4>(args); break;
    case 5: run_shape<family_, symbol_type_, 5>(args); break;
    case 6: run_shape<family_, symbol_type_, 6>(args); break;
    case 7: run_shape<family_, symbol_type_, 7>(args); break;
    case 8: run_shape<family_, symbol_type_, 8>(args); break;
    default: break;
    }
}
static_assert(shapes_count_k == 9, "dispatch switch lists every shape");

__device__ void run_variant(int variant, item_args_t const &args) {
    int const shape = variant % shapes_count_k;
    int const combo = variant / shapes_count_k; // family * 2 + symbols
    switch (combo) {
    case family_lev_linear_k * 2 + symbols_u8_k: run_family<family_lev_linear_k, uint8_t>(shape, args); break;
    case family_lev_linear_k * 2 + symbols_u32_k: run_family<family_lev_linear_k, uint32_t>(shape, args); break;
    case family_lev_affine_k * 2 + symbols_u8_k: run_family<family_lev_affine_k, uint8_t>(shape, args); break;
    case family_lev_affine_k * 2 + symbols_u32_k: run_family<family_lev_affine_k, uint32_t>(shape, args); break;
    case family_nw_linear_k * 2 + symbols_u8_k: run_family<family_nw_linear_k, uint8_t>(shape, args); break;
    case family_nw_affine_k * 2 + symbols_u8_k: run_family<family_nw_affine_k, uint8_t>(shape, args); break;
    case family_sw_linear_k * 2 + symbols_u8_k: run_family<family_sw_linear_k, uint8_t>(shape, args); break;
    case family_sw_affine_k * 2 + symbols_u8_k: run_family<family_sw_affine_k, uint8_t>(shape, args); break;
    default: break; // class-cost families never run over runes
    }
}

#pragma endregion

#pragma region Kernels

/** Arguments of the persistent scoring launch. */
struct persistent_args_t {
    task_t const *tasks;
    segment_view_t const *segments;
    bins_view_t bins;
    int32_t *scratch;
    uint32_t scratch_capacity;
};

/**
 *  Persistent scoring kernel: one 64-lane workgroup per resident wave slot; each wave pops work items (64/G pairs
 *  of one bin) from a device counter, heaviest bins first, until the queue is drained. The queue holds
 *  `item_start[bins]` items; every wave exits once the counter passes it.
 */
__global__ __launch_bounds__(64) void persistent_score_kernel(persistent_args_t args) {
    extern __shared__ int8_t profile_lds[];
    uint32_t const total_items = args.bins.item_start[bins_count_k];
    item_args_t item;
    item.tasks = args.tasks;
    item.segments = args.segments;
    item.order = args.bins.order;
    item.scratch = args.scratch + (size_t)blockIdx.x * 2 * args.scratch_capacity;
    item.scratch_capacity = args.scratch_capacity;
    item.profile = profile_lds;
    for (;;) {
        uint32_t taken = 0;
        if (__lane_id() == 0) taken = atomicAdd(args.bins.queue_head, 1u);
        taken = __shfl(taken, 0, 64);
        if (taken >= total_items) break;
        // Priority position p with item_start[p] <= taken < item_start[p + 1].
        uint32_t low = 0, high = bins_count_k;
        while (high - low > 1) {
            uint32_t const mid = (low + high) / 2;
            if (args.bins.item_start[mid] <= taken) low = mid;
            else high = mid;
        }
        uint32_t const bin = args.bins.priority[low];
        int const variant = (int)(bin / column_buckets_k);
        shape_t const shape = shapes_k[variant % shapes_count_k];
        uint32_t const per_item = 64u / (uint32_t)shape.group;
        uint32_t const item_in_bin = taken - args.bins.item_start[low];
        uint32_t const in_bin = args.bins.histogram[bin];
        item.first = args.bins.task_start[low] + item_in_bin * per_item;
        item.count = min(per_item, in_bin - item_in_bin * per_item);
        run_variant(variant, item);
    }
}

/** One-pair launch for the per-call baseline: the same scorer, one wave, no binning. */
__global__ __launch_bounds__(64) void single_task_kernel(task_t const *tasks, segment_view_t const *segments,
                                                         uint32_t const *identity, uint32_t task_index, int variant,
                                                         int32_t *scratch, uint32_t scratch_capacity) {
    extern __shared__ int8_t profile_lds[];
    item_args_t item;
    item.tasks = tasks;
    item.segments = segments;
    item.order = identity;
    item.first = task_index;
    item.count = 1;
    item.scratch = scratch;
    item.scratch_capacity = scratch_capacity;
    item.profile = profile_lds;
    run_variant(variant, item);
}

/** Fills `count` words with `value`; used instead of memset so the sequence can be graph-captured. */
__global__ void fill_u32_kernel(uint32_t *data, uint32_t count, uint32_t value) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) data[i] = value;
}

/**
 *  UTF-8 to runes, one thread per string, with `sz_rune_decode_unchecked` semantics
 *  (include/stringzilla/utf8_runes/serial.h:101-114); continuation bytes missing at the end of a string read as 0.
 *  Runes of string `s` start at index `offsets[s]` of `runes`; their count goes to `rune_counts[s]`.
 */
__global__ void utf8_decode_kernel(uint8_t const *bytes, uint64_t const *offsets, uint64_t count, uint32_t *runes,
                                   uint32_t *rune_counts) {
    for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < count; s += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t const begin = offsets[s], end = offsets[s + 1];
        uint64_t at = begin, written = 0;
        while (at < end) {
            uint32_t const lead = bytes[at];
            uint32_t const width = 1 + (lead >= 0xC0) + (lead >= 0xE0) + (lead >= 0xF0);
            uint32_t t0 = 0, t1 = 0, t2 = 0;
            if (width > 1 && at + 1 < end) t0 = bytes[at + 1] & 0x3Fu;
            if (width > 2 && at + 2 < end) t1 = bytes[at + 2] & 0x3Fu;
            if (width > 3 && at + 3 < end) t2 = bytes[at + 3] & 0x3Fu;
            uint32_t rune;
            if (width == 1) rune = lead;
            else if (width == 2) rune = (lead & 0x1Fu) << 6 | t0;
            else if (width == 3) rune = (lead & 0x0Fu) << 12 | t0 << 6 | t1;
            else rune = (lead & 0x07u) << 18 | t0 << 12 | t1 << 6 | t2;
            runes[begin + written++] = rune;
            at += width;
        }
        rune_counts[s] = (uint32_t)written;
    }
}

/** Widens u32 tape offsets to u64 so every later kernel reads one layout. */
__global__ void widen_offsets_kernel(uint32_t const *narrow, uint64_t *wide, uint64_t count) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
        wide[i] = narrow[i];
}

__device__ __forceinline__ uint32_t half_octave(uint64_t x) {
    if (x < 2) return 0;
    uint32_t const lg = 63 - __clzll(x);
    uint32_t const half = (x >> (lg - 1)) & 1u; // second-highest bit
    uint32_t const bucket = 2 * lg + half;
    return bucket < column_buckets_k ? bucket : column_buckets_k - 1;
}

/**
 *  Builds one task per scored cell and counts it into its bin. Degenerate cells (an empty side) are scored here
 *  with the closed forms of serial.hpp:142-151 and 2493-2496 and never reach the scoring kernel.
 */
__global__ void materialize_kernel(materialize_args_t a, shape_table_t const *table, task_t *tasks,
                                   uint32_t *bin_of_task, uint32_t *histogram, segment_view_t const *segments) {
    for (uint64_t index = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; index < a.tasks_count;
         index += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t q, c;
        if (a.mode == 0) q = c = index;
        else if (a.mode == 1) q = index / a.candidates_count, c = index % a.candidates_count;
        else {
            // Lower triangle including the diagonal, row-major: index = q (q + 1) / 2 + c with c <= q.
            uint64_t row = (uint64_t)((sqrt(8.0 * (double)index + 1.0) - 1.0) / 2.0);
            while (row * (row + 1) / 2 > index) --row;
            while ((row + 1) * (row + 2) / 2 <= index) ++row;
            q = row, c = index - row * (row + 1) / 2;
        }
        uint64_t const *c_offsets = a.candidate_offsets ? a.candidate_offsets : a.query_offsets;
        uint32_t const *c_runes = a.candidate_offsets ? a.candidate_rune_counts : a.query_rune_counts;
        uint64_t const q_begin = a.query_offsets[q], c_begin = c_offsets[c];
        uint64_t q_length = a.query_offsets[q + 1] - q_begin;
        uint64_t c_length = c_offsets[c + 1] - c_begin;
        if (a.symbols == symbols_u32_k) q_length = a.query_rune_counts[q], c_length = c_runes[c];

        task_t task;
        task.segment = a.segment;
        if (a.mode == 0) task.output = index, task.mirror = no_mirror_k;
        else {
            task.output = q * a.results_row_stride + c;
            task.mirror = (a.mode == 2 && q != c) ? c * a.results_row_stride + q : no_mirror_k;
        }
        bool const transposed = c_length < q_length;
        task.transposed = transposed;
        task.rows_offset = transposed ? c_begin : q_begin;
        task.columns_offset = transposed ? q_begin : c_begin;
        task.rows_length = (uint32_t)(transposed ? c_length : q_length);
        task.columns_length = (uint32_t)(transposed ? q_length : c_length);
        uint32_t bin = ~0u;
        if (q_length == 0 || c_length == 0) {
            bool const local = a.family == family_sw_linear_k || a.family == family_sw_affine_k;
            bool const affine = a.family == family_lev_affine_k || a.family == family_nw_affine_k ||
                                a.family == family_sw_affine_k;
            int64_t const longest = (int64_t)(q_length > c_length ? q_length : c_length);
            int64_t value = 0;
            if (!local) value = affine ? (longest ? a.open + (int64_t)a.extend * (longest - 1) : 0) : a.open * longest;
            int64_t *results = segments[a.segment].results;
            results[task.output] = value;
            if (task.mirror != no_mirror_k) results[task.mirror] = value;
        }
        else {
            uint32_t const rows = task.rows_length;
            int const family = a.family, symbols = a.symbols;
            int shape = table->multi_pass_shape[family][symbols];
            for (int s = 0; s < table->used[family][symbols]; ++s) {
                int const candidate_shape = table->order[family][symbols][s];
                if ((int64_t)rows <= table->limits[family][symbols][candidate_shape]) {
                    shape = candidate_shape;
                    break;
                }
            }
            shape_t const sh = shapes_k[shape];
            uint64_t const stripe = (uint64_t)sh.group * sh.rows_per_lane;
            uint64_t const passes = (rows + stripe - 1) / stripe;
            uint64_t const work = passes * ((uint64_t)task.columns_length + sh.group);
            bin = (uint32_t)variant_of(family, symbols, shape) * column_buckets_k + half_octave(work);
            atomicAdd(&histogram[bin], 1u);
        }
        tasks[a.task_base + index] = task;
        bin_of_task[a.task_base + index] = bin;
    }
}

/**
 *  Single-workgroup exclusive scans over the bins in priority order: first slot of each bin in `order`, and
 *  first work item of each bin (an item is 64 / G pairs of the bin).
 */
__global__ __launch_bounds__(1024) void scan_bins_kernel(bins_view_t bins) {
    __shared__ uint32_t task_carry, item_carry;
    __shared__ uint32_t partial_tasks[1024], partial_items[1024];
    if (threadIdx.x == 0) task_carry = 0, item_carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < bins_count_k; base += 1024) {
        uint32_t const p = base + threadIdx.x;
        uint32_t tasks_here = 0, items_here = 0;
        if (p < bins_count_k) {
            uint32_t const bin = bins.priority[p];
            shape_t const shape = shapes_k[(bin / column_buckets_k) % shapes_count_k];
            uint32_t const per_item = 64u / (uint32_t)shape.group;
            tasks_here = bins.histogram[bin];
            items_here = (tasks_here + per_item - 1) / per_item;
        }
        partial_tasks[threadIdx.x] = tasks_here;
        partial_items[threadIdx.x] = items_here;
        __syncthreads();
        for (uint32_t offset = 1; offset < 1024; offset <<= 1) { // Hillis-Steele inclusive scan
            uint32_t const add_tasks = threadIdx.x >= offset ? partial_tasks[threadIdx.x - offset] : 0;
            uint32_t const add_items = threadIdx.x >= offset ? partial_items[threadIdx.x - offset] : 0;
            __syncthreads();
            partial_tasks[threadIdx.x] += add_tasks;
            partial_items[threadIdx.x] += add_items;
            __syncthreads();
        }
        if (p < bins_count_k) {
            bins.task_start[p] = task_carry + partial_tasks[threadIdx.x] - tasks_here;
            bins.item_start[p] = item_carry + partial_items[threadIdx.x] - items_here;
        }
        __syncthreads();
        if (threadIdx.x == 1023) task_carry += partial_tasks[1023], item_carry += partial_items[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) bins.item_start[bins_count_k] = item_carry;
}

/** Scatters task ids into `order`, grouped by bin (order inside a bin is arbitrary and does not affect results). */
__global__ void scatter_bins_kernel(bins_view_t bins, uint32_t const *bin_of_task, uint32_t const *position_of_bin,
                                    uint32_t tasks_count) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < tasks_count; t += gridDim.x * blockDim.x) {
        uint32_t const bin = bin_of_task[t];
        if (bin == ~0u) continue;
        uint32_t const slot = bins.task_start[position_of_bin[bin]] + atomicAdd(&bins.cursor[bin], 1u);
        bins.order[slot] = t;
        bins.shape_of_task[t] = (bin / column_buckets_k) % shapes_count_k;
    }
}

#pragma endregion

} // namespace szs_amd

#pragma region Host-callable launchers

extern "C" {

hipError_t szs_amd_launch_fill(uint32_t *data, uint32_t count, uint32_t value, hipStream_t stream) {
    if (!count) return hipSuccess;
    uint32_t const blocks = (count + 255) / 256 < 1024 ? (count + 255) / 256 : 1024;
    hipLaunchKernelGGL(szs_amd::fill_u32_kernel, dim3(blocks), dim3(256), 0, stream, data, count, value);
    return hipGetLastError();
}

hipError_t szs_amd_launch_utf8_decode(uint8_t const *bytes, uint64_t const *offsets, uint64_t count, uint32_t *runes,
                                      uint32_t *rune_counts, hipStream_t stream) {
    if (!count) return hipSuccess;
    uint64_t const blocks64 = (count + 255) / 256;
    uint32_t const blocks = blocks64 < 65536 ? (uint32_t)blocks64 : 65536;
    hipLaunchKernelGGL(szs_amd::utf8_decode_kernel, dim3(blocks), dim3(256), 0, stream, bytes, offsets, count, runes,
                       rune_counts);
    return hipGetLastError();
}

hipError_t szs_amd_launch_widen_offsets(uint32_t const *narrow, uint64_t *wide, uint64_t count, hipStream_t stream) {
    if (!count) return hipSuccess;
    uint64_t const blocks64 = (count + 255) / 256;
    uint32_t const blocks = blocks64 < 65536 ? (uint32_t)blocks64 : 65536;
    hipLaunchKernelGGL(szs_amd::widen_offsets_kernel, dim3(blocks), dim3(256), 0, stream, narrow, wide, count);
    return hipGetLastError();
}

hipError_t szs_amd_launch_materialize(szs_amd::materialize_args_t const *args, szs_amd::shape_table_t const *table,
                                      szs_amd::task_t *tasks, uint32_t *bin_of_task, uint32_t *histogram,
                                      szs_amd::segment_view_t const *segments, hipStream_t stream) {
    if (!args->tasks_count) return hipSuccess;
    uint64_t const blocks64 = (args->tasks_count + 255) / 256;
    uint32_t const blocks = blocks64 < 65536 ? (uint32_t)blocks64 : 65536;
    hipLaunchKernelGGL(szs_amd::materialize_kernel, dim3(blocks), dim3(256), 0, stream, *args, table, tasks,
                       bin_of_task, histogram, segments);
    return hipGetLastError();
}

hipError_t szs_amd_launch_scan_bins(szs_amd::bins_view_t const *bins, hipStream_t stream) {
    hipLaunchKernelGGL(szs_amd::scan_bins_kernel, dim3(1), dim3(1024), 0, stream, *bins);
    return hipGetLastError();
}

hipError_t szs_amd_launch_scatter_bins(szs_amd::bins_view_t const *bins, uint32_t const *bin_of_task,
                                       uint32_t const *position_of_bin, uint32_t tasks_count, hipStream_t stream) {
    if (!tasks_count) return hipSuccess;
    uint32_t const blocks = (tasks_count + 255) / 256 < 65536 ? (tasks_count + 255) / 256 : 65536;
    hipLaunchKernelGGL(szs_amd::scatter_bins_kernel, dim3(blocks), dim3(256), 0, stream, *bins, bin_of_task,
                       position_of_bin, tasks_count);
    return hipGetLastError();
}

hipError_t szs_amd_launch_persistent(szs_amd::task_t const *tasks, szs_amd::segment_view_t const *segments,
                                     szs_amd::bins_view_t const *bins, int32_t *scratch, uint32_t scratch_capacity,
                                     uint32_t waves, uint32_t lds_bytes, hipStream_t stream) {
    szs_amd::persistent_args_t args;
    args.tasks = tasks;
    args.segments = segments;
    args.bins = *bins;
    args.scratch = scratch;
    args.scratch_capacity = scratch_capacity;
    hipLaunchKernelGGL(szs_amd::persistent_score_kernel, dim3(waves), dim3(64), lds_bytes, stream, args);
    return hipGetLastError();
}

hipError_t szs_amd_launch_single(szs_amd::task_t const *tasks, szs_amd::segment_view_t const *segments,
                                 uint32_t const *identity, uint32_t task_index, int variant, int32_t *scratch,
                                 uint32_t scratch_capacity, uint32_t lds_bytes, hipStream_t stream) {
    hipLaunchKernelGGL(szs_amd::single_task_kernel, dim3(1), dim3(64), lds_bytes, stream, tasks, segments, identity,
                       task_index, variant, scratch, scratch_capacity);
    return hipGetLastError();
}

hipError_t szs_amd_persistent_occupancy(int *blocks_per_cu, uint32_t lds_bytes) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, szs_amd::persistent_score_kernel, 64,
                                                        lds_bytes);
}

} // extern "C"

#pragma endregion
