/**
 *  @file dp_kernels.hip
 *  @brief gfx950 kernels of the similarity engine: the persistent scoring kernel (independent and cooperative
 *         work items over the sweep of dp_sweep.hpp), the one-pair baseline kernel, UTF-8 transcoding, pair
 *         materialization and the device-side length binning.
 *
 *  Workgroups have 4 waves. A work item of a non-cooperative shape is one wave's 64/G pairs of one bin, which a
 *  wave takes and sweeps on its own; a cooperative item gives the whole workgroup one tall pair, whose stripes the
 *  waves sweep in a lock-step
 *  pipeline: stripe s runs on wave s mod 4 during phases [s*D, s*D + C) (C = 64-column chunks per stripe,
 *  D = max(skew + 1, ceil(C / 4)), skew = 1, or 2 for the packed 16-bit sweep), reading stripe s-1's bottom row
 *  from an HBM ring of 4 rows that stripe s-1 wrote at least one phase (one workgroup barrier) earlier.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "dp_sweep.hpp"
#include "myers_sweep.hpp"
#include "packed_sweep.hpp"
#include "szs_device.hpp"
#include "wide_sweep.hpp"

namespace szs_amd {

#pragma region Work items

/** Chunks of G steps that sweep `columns` columns: lane G - 1 finishes column `columns` at step
 *  columns + skew * G - 2 (the packed sweep's lower half trails its upper half by one more column). */
template <typename sweep_type_>
__device__ __forceinline__ int32_t chunks_for(int32_t columns) {
    constexpr int32_t G = sweep_type_::group_k, skew = sweep_type_::skew_k;
    return (columns + skew * G - 1 + G - 1) / G;
}

/** A chunk of G steps in which every lane's columns stay within 1 .. columns - (skew - 1) on every step, so the
 *  per-step column check (and the packed sweep's edge fix-ups) can be skipped. */
template <typename sweep_type_>
__device__ __forceinline__ bool is_steady(int32_t chunk, int32_t columns) {
    constexpr int32_t G = sweep_type_::group_k, skew = sweep_type_::skew_k;
    return chunk >= skew && chunk * G + G + (skew - 1) <= columns;
}

/** Sweeps of several pairs per wave that take steady chunks, bounded by the wave's shortest pair. */
template <typename sweep_type_, typename = void>
struct steady_groups : std::false_type {};
template <typename sweep_type_>
struct steady_groups<sweep_type_, std::void_t<decltype(sweep_type_::steady_groups_k)>>
    : std::integral_constant<bool, sweep_type_::steady_groups_k> {};

/**
 *  Shorter sweeps for the last stripe of a tall pair whose remaining rows fit one: the same lanes, skew and
 *  boundary-row format (int32 rows of the anti-diagonal form, or Myers delta codes), fewer rows per lane, so the
 *  stripe costs fewer instructions per step. `void` where there is none.
 */
template <typename sweep_type_>
struct tail_sweeps_t {
    using tiny = void;
    using shorter = void;
    using middle = void;
};
template <int family_, bool uniform_base_, bool biased_>
struct tail_sweeps_t<packed_sweep_t<family_, 8, uniform_base_, biased_>> { // 512-row stripes: 128/256/384-row tails
    using tiny = packed_sweep_t<family_, 2, uniform_base_, biased_>;
    using shorter = packed_sweep_t<family_, 4, uniform_base_, biased_>;
    using middle = packed_sweep_t<family_, 6, uniform_base_, biased_>;
};
template <bool runes_>
struct tail_sweeps_t<myers_sweep_t<64, runes_, 64>> { // 4096-row stripes: 2048-row tails of 32-row words
    using tiny = void;
    using shorter = myers_sweep_t<64, runes_, 32>;
    using middle = void;
};

template <typename tail_type_>
__device__ constexpr int32_t tail_rows() {
    if constexpr (std::is_void<tail_type_>::value) return 0;
    else return 64 * tail_type_::rows_per_lane;
}

/**
 *  Sweeps the last stripe (rows from `base`) of one pair with `tail_type_`, reading the row above from `in_row`
 *  and `gaps_row` (the matrix border when `in_row` is null). Returns the stripe's share of a summed result.
 */
template <typename tail_type_>
__device__ __forceinline__ int32_t run_tail(item_args_t const &a, uint32_t task_index, uint32_t wave, int32_t base,
                                            int32_t const *in_row, int32_t const *gaps_row) {
    static_assert(tail_type_::group_k == 64, "tails are whole-wave stripes");
    tail_type_ tail;
    tail.load(a, task_index, true, wave);
    int32_t const columns = __builtin_amdgcn_readfirstlane(tail.m);
    int32_t const chunks = chunks_for<tail_type_>(columns);
    tail.begin_stripe(base);
    for (int32_t chunk = 0; chunk < chunks; ++chunk) {
        if (is_steady<tail_type_>(chunk, columns))
            tail.template run_chunk<true, false>(chunk, in_row, gaps_row, nullptr, nullptr);
        else tail.template run_chunk<false, false>(chunk, in_row, gaps_row, nullptr, nullptr);
    }
    tail.end_stripe(base, true);
    return tail.partial();
}

/**
 *  Each wave scores 64/G pairs of the item on its own, stripe after stripe for tall pairs (G = 64); a tall pair's
 *  last stripe runs with a shorter sweep when one covers its rows.
 */
template <typename sweep_type_>
__device__ void run_independent(item_args_t const &a) {
    constexpr int G = sweep_type_::group_k;
    constexpr int R = sweep_type_::rows_per_lane;
    constexpr uint32_t per_wave = 64 / G;
    uint32_t const wave = threadIdx.x / 64;
    uint32_t const first_slot = a.wave_item ? 0u : wave * per_wave; // this wave's pairs within the item
    if (first_slot >= a.count) return;                              // wave-uniform: no pairs left for this wave
    uint32_t const slot = first_slot + (uint32_t)__lane_id() / G;
    bool const active = slot < a.count;
    sweep_type_ sweep;
    sweep.load(a, a.order[a.first + (active ? slot : first_slot)], active, wave);
    int32_t longest = sweep.m, shortest = sweep.m;
    if constexpr (G < 64) {
        for (int offset = G; offset < 64; offset <<= 1) longest = max(longest, __shfl_xor(longest, offset, 64));
        if constexpr (steady_groups<sweep_type_>::value) {
            for (int offset = G; offset < 64; offset <<= 1) shortest = min(shortest, __shfl_xor(shortest, offset, 64));
        }
    }
    int32_t const stripe_rows = G * R;
    int32_t const passes = G == 64 ? (sweep.n + stripe_rows - 1) / stripe_rows : 1;
    int32_t const chunks = chunks_for<sweep_type_>(longest);
    int32_t *scores_row = a.scratch + (size_t)wave * 2 * a.scratch_capacity;
    int32_t *gaps_row = scores_row + a.scratch_capacity;
    // Chunks whose columns are valid for every lane skip the per-step column check: with one pair per wave, or with
    // several whose sweep supports it (every group's columns valid: the wave's shortest pair bounds them).
    constexpr bool steady_chunks = G == 64 || steady_groups<sweep_type_>::value;
    int32_t const columns = G == 64 ? __builtin_amdgcn_readfirstlane(sweep.m)
                                    : steady_chunks ? __builtin_amdgcn_readfirstlane(shortest) : 0;
    // The last stripe's rows pick its sweep (wave-uniform: one pair per wave): 1 = shorter, 2 = middle, 3 = tiny tail.
    using tiny_t = typename tail_sweeps_t<sweep_type_>::tiny;
    using shorter_t = typename tail_sweeps_t<sweep_type_>::shorter;
    using middle_t = typename tail_sweeps_t<sweep_type_>::middle;
    int32_t const tail_base = (passes - 1) * stripe_rows;
    int32_t tail = 0;
    if constexpr (tail_rows<shorter_t>() > 0) {
        int32_t const last_rows = __builtin_amdgcn_readfirstlane(sweep.n) - tail_base;
        if (passes > 0 && last_rows <= tail_rows<tiny_t>()) tail = 3;
        else if (passes > 0 && last_rows <= tail_rows<shorter_t>()) tail = 1;
        else if (passes > 0 && last_rows <= tail_rows<middle_t>()) tail = 2;
    }
    int32_t const full_passes = passes - (tail ? 1 : 0);
    for (int32_t pass = 0; pass < full_passes; ++pass) {
        int32_t const base = pass * stripe_rows;
        bool const more = pass + 1 < passes;
        int32_t const *in_row = pass ? scores_row : nullptr;
        sweep.begin_stripe(base);
        for (int32_t chunk = 0; chunk < chunks; ++chunk) {
            bool const steady = steady_chunks && is_steady<sweep_type_>(chunk, columns);
            if (more) {
                if (steady) sweep.template run_chunk<true, true>(chunk, in_row, gaps_row, scores_row, gaps_row);
                else sweep.template run_chunk<false, true>(chunk, in_row, gaps_row, scores_row, gaps_row);
            }
            else if (steady) sweep.template run_chunk<true, false>(chunk, in_row, gaps_row, nullptr, nullptr);
            else sweep.template run_chunk<false, false>(chunk, in_row, gaps_row, nullptr, nullptr);
        }
        if (more) __threadfence_block(); // lane 63's stores before lane 0's loads in the next pass
        sweep.end_stripe(base, !more);
    }
    int32_t tail_sum = 0;
    if constexpr (tail_rows<shorter_t>() > 0) {
        int32_t const *in_row = full_passes ? scores_row : nullptr;
        uint32_t const task_index = a.order[a.first + first_slot];
        if (tail == 1) tail_sum = run_tail<shorter_t>(a, task_index, wave, tail_base, in_row, gaps_row);
        if constexpr (tail_rows<tiny_t>() > 0) {
            if (tail == 3) tail_sum = run_tail<tiny_t>(a, task_index, wave, tail_base, in_row, gaps_row);
        }
        if constexpr (tail_rows<middle_t>() > 0) {
            if (tail == 2) tail_sum = run_tail<middle_t>(a, task_index, wave, tail_base, in_row, gaps_row);
        }
    }
    // A tail stripe's share: its best (local alignment, at least 0 like every best) or its sum.
    if constexpr (sweep_type_::reduce_k == reduce_max_k) sweep.write_result(max(sweep.partial(), tail_sum));
    else if constexpr (sweep_type_::reduce_k != reduce_none_k) sweep.write_result(sweep.partial() + tail_sum);
}

/** Polls of a predecessor's progress (each a device-scope load plus a ~128-cycle sleep) a group may make on top of
 *  the work it waits for: seconds, far beyond any real wait, so a kernel always ends even if a predecessor never
 *  published. */
constexpr uint64_t chain_poll_floor_k = 1ull << 22;
/** Polls allowed per 64-column chunk of one stripe of predecessor work: ~14 us of sleeping, several times what a
 *  chunk takes on a fully loaded CU. */
constexpr uint64_t chain_polls_per_chunk_k = 256;

/** Waits until the group before this one has published column `needed` of its bottom row. Returns false, and
 *  raises the batch's timeout word, if the wait was given up (after `limit` polls): the caller then waits no more
 *  for this pair. */
__device__ __forceinline__ bool wait_for_columns(uint32_t const *progress, uint32_t needed, uint64_t limit,
                                                 uint32_t *timeout) {
    for (uint64_t polls = 0;; ++polls) {
        uint32_t const published = __hip_atomic_load(progress, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (published >= needed) return true;
        if (polls == limit) {
            if (__lane_id() == 0) atomicOr(timeout, 1u);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

/** Makes this wave's bottom-row stores visible device-wide, then publishes how many columns are complete. */
__device__ __forceinline__ void publish_columns(uint32_t *progress, uint32_t columns) {
    __threadfence();
    if (__lane_id() == 0) __hip_atomic_store(progress, columns, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

/**
 *  All waves of the workgroup pipeline the stripes of one pair (see the file comment for the schedule). With
 *  `chained_`, the pair's stripes are split into `chain_groups_k` groups of consecutive stripes, one work item
 *  (one workgroup) each: a group's first stripe reads the previous group's bottom row from the chain buffer as
 *  that group publishes it, and its last stripe publishes its own. Items are taken in queue order and a group
 *  only waits for earlier groups, which are already running, so every wait ends.
 */
template <typename sweep_type_, bool chained_ = false>
__device__ void run_cooperative(item_args_t const &a) {
    constexpr int32_t W = waves_per_group_k;
    static_assert(sweep_type_::group_k == 64, "cooperative items sweep 64-lane stripes");
    static_assert(!chained_ || sweep_type_::reduce_k == reduce_none_k, "chained groups combine no partial results");
    int32_t const wave = (int32_t)(threadIdx.x / 64);
    sweep_type_ sweep;
    sweep.load(a, a.order[a.first], true, (uint32_t)wave);
    int32_t const stripe_rows = 64 * sweep_type_::rows_per_lane;
    int32_t const stripes = (sweep.n + stripe_rows - 1) / stripe_rows;
    int32_t const chunks = chunks_for<sweep_type_>(sweep.m);
    int32_t const columns = __builtin_amdgcn_readfirstlane(sweep.m);
    // Stripe s - 1 has written the bottom-row columns that stripe s's chunk x reads by the end of its own chunk
    // x + skew: a delay of skew + 1 phases keeps the reader behind.
    // A wave that sweeps several stripes of the item (more than W) runs them one after another: the stagger then
    // spreads them over the waves, ceil(chunks / W) phases apart.
    int32_t const stagger = max(sweep_type_::skew_k + 1, (chunks + W - 1) / W);
    size_t const row_pitch = 2 * (size_t)a.scratch_capacity;
    if constexpr (sweep_type_::reduce_k != reduce_none_k) {
        if (threadIdx.x == 0) *a.shared_result = 0;
    }
    int32_t groups = 1, per_group = stripes, first_group = 0, last_group = 0;
    uint32_t slot = 0;
    bool waiting = chained_; // cleared if a wait is given up
    if constexpr (chained_) {
        groups = chain_groups_k;
        per_group = max(W, (stripes + groups - 1) / groups);
        first_group = a.group >= 0 ? a.group : 0;
        last_group = a.group >= 0 ? a.group : groups - 1;
        slot = task_chain_slot(a.tasks[a.order[a.first]]);
    }
    for (int32_t group = first_group; group <= last_group; ++group) {
        int32_t const first_stripe = group * per_group;
        int32_t const end_stripe = min(stripes, first_stripe + per_group);
        if (first_stripe >= end_stripe) break; // trailing groups of a shorter pair hold no stripes
        int32_t const local_stripes = end_stripe - first_stripe;
        // At most one stripe per wave: the minimal delay that keeps each stripe behind the one above.
        int32_t const delay = local_stripes <= W ? sweep_type_::skew_k + 1 : stagger;
        // Chain buffers: the bottom row entering this group and the one leaving it, with their column counts.
        int32_t *chain_in = nullptr, *chain_out = nullptr;
        uint32_t *progress_in = nullptr, *progress_out = nullptr;
        if constexpr (chained_) {
            size_t const pitch = 2 * (size_t)a.chain_pitch;
            int32_t *base = a.chain_rows + (size_t)slot * (chain_groups_k - 1) * pitch;
            if (group > 0) chain_in = base + (size_t)(group - 1) * pitch, progress_in = a.chain_progress + slot * chain_groups_k + group - 1;
            if (end_stripe < stripes) chain_out = base + (size_t)group * pitch, progress_out = a.chain_progress + slot * chain_groups_k + group;
        }
        int32_t const phases = (local_stripes - 1) * delay + chunks;
        // A group may wait, at worst, for all the stripes of the groups before it: a pair of more than 4 x 64
        // stripes puts more stripes in each group, which runs them W at a time, so the bound grows with them.
        uint64_t const poll_limit = chain_poll_floor_k + chain_polls_per_chunk_k * (uint64_t)group *
                                                             (uint64_t)per_group * (uint64_t)chunks;
        for (int32_t phase = 0; phase < phases; ++phase) {
            int32_t local = -1, chunk = 0;
            if (phase >= wave * delay) {
                int32_t const since = phase - wave * delay;
                if (local_stripes <= W) { // wave w sweeps stripe w alone, `delay` phases behind wave w - 1
                    if (wave < local_stripes && since < chunks) local = wave, chunk = since;
                }
                else { // wave w sweeps stripes w, w + W, ... one after another (W * delay >= chunks)
                    int32_t const candidate = wave + (since / (W * delay)) * W;
                    int32_t const at = since % (W * delay);
                    if (candidate < local_stripes && at < chunks) local = candidate, chunk = at;
                }
            }
            if (local >= 0) { // wave-uniform
                int32_t const stripe = first_stripe + local;
                int32_t const base = stripe * stripe_rows;
                if (chunk == 0) sweep.begin_stripe(base);
                int32_t *in_row = local ? a.scratch + ((local - 1) % W) * row_pitch : chain_in;
                int32_t *out_row = local + 1 < local_stripes ? a.scratch + (local % W) * row_pitch : chain_out;
                size_t const in_gap_offset = local ? a.scratch_capacity : a.chain_pitch;
                size_t const out_gap_offset = local + 1 < local_stripes ? a.scratch_capacity : a.chain_pitch;
                int32_t const *in_gaps = in_row ? in_row + in_gap_offset : nullptr;
                if (chained_ && !local && chain_in && waiting)
                    waiting = wait_for_columns(progress_in, (uint32_t)min(chunk * 64 + 64, columns), poll_limit,
                                               a.chain_timeout);
                bool const steady = is_steady<sweep_type_>(chunk, columns);
                if (out_row) {
                    int32_t *out_gaps = out_row + out_gap_offset;
                    if (steady) sweep.template run_chunk<true, true>(chunk, in_row, in_gaps, out_row, out_gaps);
                    else sweep.template run_chunk<false, true>(chunk, in_row, in_gaps, out_row, out_gaps);
                    if (chained_ && out_row == chain_out) {
                        // Columns up to lane 63's last lower-half column of this chunk are final; all of them after
                        // the last chunk.
                        int32_t const done = chunk == chunks - 1 ? columns
                                                                 : chunk * 64 + 1 - 64 * (sweep_type_::skew_k - 1);
                        publish_columns(progress_out, (uint32_t)max(0, min(done, columns)));
                    }
                }
                else if (steady) sweep.template run_chunk<true, false>(chunk, in_row, in_gaps, nullptr, nullptr);
                else sweep.template run_chunk<false, false>(chunk, in_row, in_gaps, nullptr, nullptr);
                if (chunk == chunks - 1) sweep.end_stripe(base, stripe == stripes - 1);
            }
            __syncthreads(); // publishes this phase's boundary rows to the next stripe's wave
        }
    }
    if constexpr (sweep_type_::reduce_k != reduce_none_k) {
        int32_t const mine = sweep.partial();
        if (__lane_id() == 0) {
            if constexpr (sweep_type_::reduce_k == reduce_max_k) atomicMax(a.shared_result, mine);
            else atomicAdd(a.shared_result, mine);
        }
        __syncthreads();
        if (wave == 0) sweep.write_result(*a.shared_result);
    }
}

#pragma endregion

#pragma region Dispatch

/** One packed shape with the sweep's compile-time options. */
template <int family_, int shape_, bool uniform_base_, bool biased_>
__device__ __forceinline__ void run_packed(item_args_t const &a) {
    constexpr shape_t shape = shapes_k[shape_];
    using sweep = packed_sweep_t<family_, shape.rows_per_lane, uniform_base_, biased_>;
    if constexpr (shape.chained) run_cooperative<sweep, true>(a);
    else if constexpr (shape.cooperative) run_cooperative<sweep>(a);
    else run_independent<sweep>(a);
}

template <int family_, typename symbol_type_, int shape_, bool shifted_>
__device__ __forceinline__ void run_shape(item_args_t const &a) {
    constexpr shape_t shape = shapes_k[shape_];
    constexpr bool uniform = family_traits<family_>::uniform_k;
    if constexpr (shape.wide) run_wide<family_, symbol_type_, shape.rows_per_lane>(a);
    else if constexpr (shape.bit_parallel) {
        // Myers shapes exist for the unit-cost Levenshtein over bytes and runes only (the binning never picks them
        // otherwise).
        if constexpr (family_ == family_lev_linear_k) {
            constexpr bool runes = sizeof(symbol_type_) == 4;
            if constexpr (shape.cooperative) run_cooperative<myers_sweep_t<64, runes, shape.rows_per_lane>>(a);
            else run_independent<myers_sweep_t<shape.group, runes, shape.rows_per_lane>>(a);
        }
    }
    else if constexpr (shape.packed) {
        // Packed 16-bit shapes exist for the global class-cost families in their anti-diagonal form and for the
        // local ones in their `H + o` form (the binning never picks them otherwise); chained groups for the global
        // ones only (a local score needs every stripe's best).
        constexpr bool global = family_ == family_nw_linear_k || family_ == family_nw_affine_k;
        constexpr bool local = family_ == family_sw_linear_k || family_ == family_sw_affine_k;
        if constexpr ((global || (local && !shape.chained)) && sizeof(symbol_type_) == 1 && shifted_) {
            // One base for the whole wave when the pair's engine admits it at these rows (no per-step moves of
            // values between lane bases), else one base per lane. Every pair of a packed item has one wave to
            // itself or the whole workgroup, so the item's first pair decides.
            // Local alignment: a profile biased below o costs one more add per two cells (`biased_`).
            engine_params_t const *params = a.segments[a.tasks[a.order[a.first]].segment].params;
            bool const uniform_base = params->packed_uniform_rows >= shape.rows_per_lane;
            if constexpr (local) {
                if (params->packed_bias != params->open) {
                    if (uniform_base) run_packed<family_, shape_, true, true>(a);
                    else run_packed<family_, shape_, false, true>(a);
                    return;
                }
            }
            if (uniform_base) run_packed<family_, shape_, true, false>(a);
            else run_packed<family_, shape_, false, false>(a);
        }
    }
    // Class-cost families keep their LDS profile at <= 12 rows per lane; larger shapes are never binned for them.
    else if constexpr (!uniform && shape.rows_per_lane > 12) return;
    else if constexpr (shape.chained) {
        // Chained groups pass only bottom rows along: global families, whose result is the last cell.
        if constexpr (!family_traits<family_>::local_k)
            run_cooperative<sweep_t<family_, symbol_type_, 64, shape.rows_per_lane, shifted_>, true>(a);
    }
    else if constexpr (shape.cooperative)
        run_cooperative<sweep_t<family_, symbol_type_, 64, shape.rows_per_lane, shifted_>>(a);
    else run_independent<sweep_t<family_, symbol_type_, shape.group, shape.rows_per_lane, shifted_>>(a);
}

template <int family_, typename symbol_type_, bool shifted_ = true>
__device__ void run_family(int shape, item_args_t const &a) {
    switch (shape) {
    case 0: run_shape<family_, symbol_type_, 0, shifted_>(a); break;
    case 1: run_shape<family_, symbol_type_, 1, shifted_>(a); break;
    case 2: run_shape<family_, symbol_type_, 2, shifted_>(a); break;
    case 3: run_shape<family_, symbol_type_, 3, shifted_>(a); break;
    case 4: run_shape<family_, symbol_type_, 4, shifted_>(a); break;
    case 5: run_shape<family_, symbol_type_, 5, shifted_>(a); break;
    case 6: run_shape<family_, symbol_type_, 6, shifted_>(a); break;
    case 7: run_shape<family_, symbol_type_, 7, shifted_>(a); break;
    case 8: run_shape<family_, symbol_type_, 8, shifted_>(a); break;
    case 9: run_shape<family_, symbol_type_, 9, shifted_>(a); break;
    case 10: run_shape<family_, symbol_type_, 10, shifted_>(a); break;
    case 11: run_shape<family_, symbol_type_, 11, shifted_>(a); break;
    case 12: run_shape<family_, symbol_type_, 12, shifted_>(a); break;
    case 13: run_shape<family_, symbol_type_, 13, shifted_>(a); break;
    case 14: run_shape<family_, symbol_type_, 14, shifted_>(a); break;
    case 15: run_shape<family_, symbol_type_, 15, shifted_>(a); break;
    case 16: run_shape<family_, symbol_type_, 16, shifted_>(a); break;
    case 17: run_shape<family_, symbol_type_, 17, shifted_>(a); break;
    case 18: run_shape<family_, symbol_type_, 18, shifted_>(a); break;
    case 19: run_shape<family_, symbol_type_, 19, shifted_>(a); break;
    case 20: run_shape<family_, symbol_type_, 20, shifted_>(a); break;
    case 21: run_shape<family_, symbol_type_, 21, shifted_>(a); break;
    case 22: run_shape<family_, symbol_type_, 22, shifted_>(a); break;
    case 23: run_shape<family_, symbol_type_, 23, shifted_>(a); break;
    case 24: run_shape<family_, symbol_type_, 24, shifted_>(a); break;
    case 25: run_shape<family_, symbol_type_, 25, shifted_>(a); break;
    case 26: run_shape<family_, symbol_type_, 26, shifted_>(a); break;
    default: break;
    }
}
static_assert(shapes_count_k == 27, "dispatch switch lists every shape");

__device__ void run_variant(int variant, item_args_t const &a) {
    int const shape = variant % shapes_count_k;
    int const combo = variant / shapes_count_k; // family * 2 + symbols
    switch (combo) {
    case family_lev_linear_k * 2 + symbols_u8_k: run_family<family_lev_linear_k, uint8_t>(shape, a); break;
    case family_lev_linear_k * 2 + symbols_u32_k: run_family<family_lev_linear_k, uint32_t>(shape, a); break;
    case family_lev_affine_k * 2 + symbols_u8_k: run_family<family_lev_affine_k, uint8_t>(shape, a); break;
    case family_lev_affine_k * 2 + symbols_u32_k: run_family<family_lev_affine_k, uint32_t>(shape, a); break;
    case family_nw_linear_k * 2 + symbols_u8_k: run_family<family_nw_linear_k, uint8_t>(shape, a); break;
    case family_nw_affine_k * 2 + symbols_u8_k: run_family<family_nw_affine_k, uint8_t>(shape, a); break;
    case family_sw_linear_k * 2 + symbols_u8_k: run_family<family_sw_linear_k, uint8_t>(shape, a); break;
    case family_sw_affine_k * 2 + symbols_u8_k: run_family<family_sw_affine_k, uint8_t>(shape, a); break;
    // Class-cost families never run over runes; their second code is the unshifted-profile sweep.
    case family_nw_linear_k * 2 + class_unshifted_k: run_family<family_nw_linear_k, uint8_t, false>(shape, a); break;
    case family_nw_affine_k * 2 + class_unshifted_k: run_family<family_nw_affine_k, uint8_t, false>(shape, a); break;
    case family_sw_linear_k * 2 + class_unshifted_k: run_family<family_sw_linear_k, uint8_t, false>(shape, a); break;
    case family_sw_affine_k * 2 + class_unshifted_k: run_family<family_sw_affine_k, uint8_t, false>(shape, a); break;
    default: break;
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
    uint32_t wave_lds_bytes;
    chain_view_t chain;
};

/**
 *  Priority position p of work item `taken`: item_start[p] <= taken < item_start[p + 1], searched upward from
 *  `from`, a position at or before it (a wave's items only move forward through the queue). One load when the item
 *  is in the same bin as the wave's last one; otherwise the whole wave searches: each round its 64 lanes probe
 *  evenly spaced positions and a ballot keeps the span between the last probe at or below `taken` and the next, so
 *  the 70,656 bins take 3 rounds of loads instead of a chain of 17-34 dependent ones (a wave's first item of a
 *  small batch waited tens of microseconds on that chain). Called by whole waves with a wave-uniform `taken`.
 */
__device__ __forceinline__ uint32_t position_of_item(bins_view_t const &bins, uint32_t taken, uint32_t from = 0) {
    if (bins.item_start[from + 1] > taken) return from; // from + 1 <= bins_count_k: item_start has one more entry
    uint32_t low = from + 1, high = bins_count_k; // item_start[low] <= taken < item_start[high]
    uint32_t const lane = __lane_id();
    while (high - low > 1) {
        uint32_t const stride = (high - low + 63) / 64;
        uint32_t const probe = low + lane * stride; // lane 0 probes `low`, which qualifies
        bool const below = probe < high && bins.item_start[probe] <= taken;
        uint64_t const qualifying = __ballot(below); // lanes 0 .. L - 1: item_start is nondecreasing
        uint32_t const last = 63 - (uint32_t)__builtin_clzll(qualifying);
        low += last * stride;
        high = min(high, low + stride);
    }
    return low;
}

/** Items far longer than the rest (the batch's tallest pairs, and the groups of chained ones) set its critical
 *  path: their waves take precedence in the SIMD arbiter, and the other waves on those SIMDs fill the cycles they
 *  leave. */
__device__ __forceinline__ void set_item_priority(uint32_t bin) {
    if ((int)(bin % column_buckets_k) >= critical_bucket_k ||
        shapes_k[(bin / column_buckets_k) % shapes_count_k].chained)
        __builtin_amdgcn_s_setprio(3);
    else __builtin_amdgcn_s_setprio(0);
}

/**
 *  One-wave items a wave draws with one atomic on the queue counter. Same-address device atomics serialize (about
 *  10 ns each), so a counter serving thousands of waves one item at a time caps the launch when items are short
 *  (16 pairs of 100 x 100 cells take ~17 us): items whose bucket is below `draw_bucket_k` are drawn 2, 4, 8 or 16
 *  at a time, about 100-200 us of work per draw (the bucket counts eighth-octaves of ~30 ns work units). `limit`
 *  keeps a small batch at one item per draw, so that its few items still spread over every wave.
 */
constexpr int draw_bucket_k = 12 * bucket_steps_per_octave_k;
__device__ __forceinline__ uint32_t items_per_draw(uint32_t bin, uint32_t limit) {
    int const family = (int)(bin / column_buckets_k) / shapes_count_k / 2;
    bool const affine = family == family_lev_affine_k || family == family_nw_affine_k || family == family_sw_affine_k;
    // Affine cells cost ~1.8x linear ones (7 eighth-octaves), as the queue's priority order also weighs them.
    int const below = draw_bucket_k - (int)(bin % column_buckets_k) - (affine ? 7 : 0);
    uint32_t const count = below <= 0 ? 1u : 1u << min(below / bucket_steps_per_octave_k, 4);
    return min(count, limit);
}

/**
 *  Persistent scoring kernel: work items are popped from one device counter in priority order until the queue
 *  (`item_start[bins]` items) is drained. The queue holds whole-workgroup items first (chained groups, then
 *  cooperative pairs, heaviest first), taken by a workgroup at a time, then one-wave items (heaviest first): once
 *  a workgroup draws past the whole-workgroup part, its waves each take items on their own, with no barrier, so
 *  a wave never idles behind a sibling's longer item.
 */
__global__ __launch_bounds__(256, 4) void persistent_score_kernel(persistent_args_t args) {
    __shared__ uint32_t taken_shared;
    __shared__ int32_t result_shared;
    uint32_t const total_items = args.bins.item_start[bins_count_k];
    uint32_t const wave_items_from = args.bins.item_start[args.bins.wave_bins_from];
    item_args_t item;
    item.tasks = args.tasks;
    item.segments = args.segments;
    item.order = args.bins.order;
    item.scratch = args.scratch + (size_t)blockIdx.x * waves_per_group_k * 2 * args.scratch_capacity;
    item.scratch_capacity = args.scratch_capacity;
    item.wave_lds_bytes = args.wave_lds_bytes;
    item.shared_result = &result_shared;
    item.chain_rows = args.chain.rows;
    item.chain_progress = args.chain.progress;
    item.chain_pitch = args.chain.pitch;
    item.chain_timeout = args.chain.timeout;
    item.wide_rows = args.chain.wide_rows;
    item.wide_pitch = args.chain.wide_pitch;
    uint32_t taken;
    for (;;) { // whole-workgroup items
        if (threadIdx.x == 0) taken_shared = atomicAdd(args.bins.queue_head, 1u);
        __syncthreads();
        taken = taken_shared;
        __syncthreads();
        if (taken >= total_items) return;
        if (taken >= wave_items_from) break;
        uint32_t const low = position_of_item(args.bins, taken);
        uint32_t const bin = args.bins.priority[low];
        int const variant = (int)(bin / column_buckets_k);
        set_item_priority(bin);
        uint32_t const item_in_bin = taken - args.bins.item_start[low];
        if (shapes_k[variant % shapes_count_k].chained) { // consecutive items are the groups of one pair, in order
            item.first = args.bins.task_start[low] + item_in_bin / chain_groups_k;
            item.group = (int32_t)(item_in_bin % chain_groups_k);
        }
        else {
            item.first = args.bins.task_start[low] + item_in_bin;
            item.group = 0;
        }
        item.count = 1;
        item.wave_item = false;
        item.same_variant = false;
        run_variant(variant, item);
    }
    // One-wave items: wave 0 sweeps the item the workgroup drew, every wave then draws its own.
    uint32_t const wave = threadIdx.x / 64;
    item.wave_item = true;
    item.group = 0;
    if (wave != 0) {
        uint32_t drawn = 0;
        if (__lane_id() == 0) drawn = atomicAdd(args.bins.queue_head, 1u);
        taken = (uint32_t)__builtin_amdgcn_readfirstlane((int)drawn);
    }
    uint32_t low = 0, next = 0, end = 0; // items drawn and not yet swept: next .. end - 1
    uint32_t const draw_limit = max(1u, total_items / (32u * gridDim.x));
    int last_variant = -1;
    while (taken < total_items) {
        low = position_of_item(args.bins, taken, low);
        uint32_t const bin = args.bins.priority[low];
        int const variant = (int)(bin / column_buckets_k);
        item.same_variant = variant == last_variant;
        last_variant = variant;
        set_item_priority(bin);
        uint32_t const per_item = pairs_per_item(shapes_k[variant % shapes_count_k]);
        uint32_t const item_in_bin = taken - args.bins.item_start[low];
        uint32_t const in_bin = args.bins.histogram[bin];
        item.first = args.bins.task_start[low] + item_in_bin * per_item;
        item.count = min(per_item, in_bin - item_in_bin * per_item);
        run_variant(variant, item);
        if (next < end) { // the rest of the last draw (past the queue's end, the loop stops)
            taken = next++;
            continue;
        }
        // Drawn only now: a wave holding the next item while it sweeps a long one would leave that item waiting
        // for it while other waves run dry at the end of the queue. Later items are no heavier than this one.
        uint32_t const count = items_per_draw(bin, draw_limit);
        uint32_t drawn = 0;
        if (__lane_id() == 0) drawn = atomicAdd(args.bins.queue_head, count);
        taken = (uint32_t)__builtin_amdgcn_readfirstlane((int)drawn);
        next = taken + 1, end = taken + count;
    }
}

/** One-pair launch for the per-call baseline: the same scorers, one workgroup, no binning. */
__global__ __launch_bounds__(256, 4) void single_task_kernel(task_t const *tasks, segment_view_t const *segments,
                                                          uint32_t const *identity, uint32_t task_index, int variant,
                                                          int32_t *scratch, uint32_t scratch_capacity,
                                                          uint32_t wave_lds_bytes, chain_view_t chain) {
    __shared__ int32_t result_shared;
    item_args_t item;
    item.tasks = tasks;
    item.segments = segments;
    item.order = identity;
    item.first = task_index;
    item.count = 1;
    item.wave_item = false;
    item.same_variant = false;
    item.scratch = scratch;
    item.scratch_capacity = scratch_capacity;
    item.wave_lds_bytes = wave_lds_bytes;
    item.shared_result = &result_shared;
    item.group = -1; // a chained pair runs all its groups in turn here
    item.chain_rows = chain.rows;
    item.chain_progress = chain.progress;
    item.chain_pitch = chain.pitch;
    item.chain_timeout = chain.timeout;
    item.wide_rows = chain.wide_rows;
    item.wide_pitch = chain.wide_pitch;
    run_variant(variant, item);
}

/** Fills `count` words with `value`; used instead of memset so the sequence can be graph-captured. */
__global__ void fill_u32_kernel(uint32_t *data, uint32_t count, uint32_t value) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) data[i] = value;
}

/** Several fills in one launch (the counters a batch clears before each run). */
__global__ void fill_many_kernel(fill_list_t list) {
    for (uint32_t e = 0; e < list.entries; ++e)
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < list.count[e]; i += gridDim.x * blockDim.x)
            list.data[e][i] = list.value[e];
}

/** Byte `k` (0..15) of a 16-byte block held in registers. */
__device__ __forceinline__ uint32_t block_byte(uint4 block, uint32_t k) {
    uint32_t const word = (k & 8) ? ((k & 4) ? block.w : block.z) : ((k & 4) ? block.y : block.x);
    return (word >> ((k & 3) * 8)) & 0xFFu;
}

/** The rune of a lead byte of `width` bytes and the continuation bytes that exist (missing ones are 0). */
__device__ __forceinline__ uint32_t rune_of(uint32_t lead, uint32_t width, uint32_t t0, uint32_t t1, uint32_t t2) {
    if (width == 1) return lead;
    if (width == 2) return (lead & 0x1Fu) << 6 | (t0 & 0x3Fu);
    if (width == 3) return (lead & 0x0Fu) << 12 | (t0 & 0x3Fu) << 6 | (t1 & 0x3Fu);
    return (lead & 0x07u) << 18 | (t0 & 0x3Fu) << 12 | (t1 & 0x3Fu) << 6 | (t2 & 0x3Fu);
}

__device__ __forceinline__ uint32_t utf8_width(uint32_t lead) {
    return 1 + (lead >= 0xC0) + (lead >= 0xE0) + (lead >= 0xF0);
}

/**
 *  One lane walks one string, the way `sz_rune_decode_unchecked` is applied in a loop, reading it through two
 *  16-byte aligned blocks held in registers. A block is loaded only if it holds a byte of the string, so it never
 *  leaves the pages of the tape. At most `room` runes are written; returns the count, or `room + 1` if the
 *  string needed more.
 */
__device__ uint64_t decode_walk(uint8_t const *first, uint8_t const *last, uint32_t *out, uint64_t room) {
    uint64_t written = 0;
    if (first >= last) return 0;
    uint8_t const *block_start = (uint8_t const *)((uintptr_t)first & ~(uintptr_t)15);
    uint4 current = *(uint4 const *)block_start, next = make_uint4(0, 0, 0, 0);
    if (block_start + 16 < last) next = *(uint4 const *)(block_start + 16);
    uint8_t const *at = first;
    while (at < last) {
        if (written == room) return room + 1;
        while (at >= block_start + 16) { // slide the pair of blocks forward
            block_start += 16;
            current = next;
            next = make_uint4(0, 0, 0, 0);
            if (block_start + 16 < last) next = *(uint4 const *)(block_start + 16);
        }
        uint32_t const k = (uint32_t)(at - block_start); // 0..15; at + 3 stays within the two blocks
        auto byte_at = [&](uint32_t d) {
            uint32_t const j = k + d;
            return j < 16 ? block_byte(current, j) : block_byte(next, j - 16);
        };
        uint32_t const lead = byte_at(0);
        uint32_t const width = utf8_width(lead);
        uint32_t const t0 = width > 1 && at + 1 < last ? byte_at(1) : 0u;
        uint32_t const t1 = width > 2 && at + 2 < last ? byte_at(2) : 0u;
        uint32_t const t2 = width > 3 && at + 3 < last ? byte_at(3) : 0u;
        out[written++] = rune_of(lead, width, t0, t1, t2);
        at += width;
    }
    return written;
}

/** Bytes a decode pass stages in LDS: 64 lanes x 16. */
constexpr int32_t decode_pass_bytes_k = 1024;
/** Strings a wave draws per ticket of the decode queue (their offsets load in one instruction). */
constexpr uint32_t decode_strings_per_ticket_k = 8;

typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) u32x4_t const global_block_t;
typedef __attribute__((address_space(1))) uint32_t global_rune_t;
typedef __attribute__((address_space(1))) uint8_t global_byte_t;

/** The 16-byte aligned block at `block` when it holds a byte before `end` (so no load leaves the tape's pages), else
 *  0x80 bytes; `past_end` replaces its bytes from `end` on when they are staged, not here (a load waited for at once
 *  would stall the wave for a whole memory round trip). Global-address-space loads, so the kernel's LDS waits never
 *  wait for them. */
__device__ __forceinline__ uint4 tape_block(uintptr_t block, uintptr_t end) {
    constexpr uint32_t past_k = 0x80808080u;
    uint4 data = make_uint4(past_k, past_k, past_k, past_k);
    if (block < end) {
        u32x4_t const loaded = *(global_block_t *)block;
        data = make_uint4(loaded.x, loaded.y, loaded.z, loaded.w);
    }
    return data;
}

/** `data`, the block at `block`, with every byte from `end` on read as 0x80: a continuation byte with no payload
 *  bits, so a sequence cut by the end of the string decodes as if its missing bytes were 0 and still shows its
 *  continuation pattern. */
__device__ __forceinline__ uint4 past_end(uint4 data, uintptr_t block, uintptr_t end) {
    constexpr uint32_t past_k = 0x80808080u;
    if (block < end && end - block < 16) {
        uint32_t const kept = (uint32_t)(end - block); // 1..15 bytes
        auto keep = [&](uint32_t word, uint32_t value) { // bytes of dword `word` from `kept` on become 0x80
            uint32_t const bytes = kept > 4 * word ? min(kept - 4 * word, 4u) : 0u;
            uint32_t const mask = bytes == 4 ? ~0u : (1u << (8 * bytes)) - 1u;
            return (value & mask) | (past_k & ~mask);
        };
        data = make_uint4(keep(0, data.x), keep(1, data.y), keep(2, data.z), keep(3, data.w));
    }
    return data;
}

/** This lane's block of the first pass of `decode_wave` over the string at `first`. */
__device__ __forceinline__ uint4 decode_head(uint8_t const *first, int64_t length) {
    uintptr_t const start = (uintptr_t)first & ~(uintptr_t)15;
    return tape_block(start + 16 * __lane_id(), (uintptr_t)first + (uintptr_t)length);
}

/**
 *  The whole wave decodes one string (the same arguments in every lane). When the string is well formed (every lead
 *  byte followed by exactly its continuation bytes, no stray continuation byte; a sequence cut by the end of the
 *  string is allowed), the walk above visits exactly the bytes that are not continuation bytes, so a rune's index
 *  is the count of such bytes before it.
 *
 *  The string is read in passes of 1 KB from the 16-byte boundary at or before its first byte, one aligned block
 *  per lane, through an LDS ring of two slots (the next pass's loads are in flight while this one decodes), and
 *  read back as rows of 64 byte positions, one per lane: a lane takes the 4 bytes from its position (a lead and the
 *  continuation bytes it may need), the wave's ballot of continuation bytes gives the leads, and the lead lanes of
 *  a row store their runes to consecutive indices, one coalesced store.
 *
 *  Well-formedness, checked on the way: (a) each lead's next `more` bytes (1-3 by its top bits) are continuation
 *  bytes, or past the string's end; (b) the continuation bytes the leads claim this way, counted up to the end,
 *  number exactly the string's continuation bytes (its length minus its leads). Claims cannot overlap (a claimed
 *  byte is a continuation, never a lead), so (b) means no continuation byte is left unclaimed, and with (a) the
 *  walk steps from lead to lead. Leads 0xF8-0xFF (a 3-bit payload where 0xF0-0xF7 have 3 as well, but the field
 *  below would take 4) also send the string to the walk. Returns the rune count, or -1 if the string is not well
 *  formed (the caller then walks it in one lane). The output must have room for one rune per byte. `head` is this
 *  lane's block of the first pass (`decode_head`), which the caller loads ahead of time.
 */
__device__ int64_t decode_wave(uint8_t const *first, int64_t length, uint32_t *out, uint4 head) {
    constexpr int32_t slot_words = decode_pass_bytes_k / 4;
    // Per wave of the 256-thread workgroup: two pass slots, then a copy of slot 0's first 16 bytes, so that the
    // 4 bytes read from any position of slot 1 continue into slot 0 without wrapping.
    __shared__ __align__(16) uint32_t ring_lds[4][2 * slot_words + 4];
    uint32_t *ring = ring_lds[threadIdx.x / 64];
    uint32_t const lane = __lane_id();
    uintptr_t const start = (uintptr_t)first & ~(uintptr_t)15, end = (uintptr_t)first + (uintptr_t)length;
    int64_t const skew = (int64_t)((uintptr_t)first - start), span = skew + length; // positions from `start`
    auto block_of = [&](int64_t pass) { return start + (uintptr_t)(pass * decode_pass_bytes_k) + 16 * lane; };
    auto stage = [&](int32_t slot, uint4 block, int64_t pass) {
        block = past_end(block, block_of(pass), end);
        *(uint4 *)(ring + slot * slot_words + 4 * lane) = block;
        if (slot == 0 && lane == 0) *(uint4 *)(ring + 2 * slot_words) = block;
    };
    stage(0, head, 0);
    uint4 next = tape_block(block_of(1), end);
    global_rune_t *runes = (global_rune_t *)out;
    int64_t written = 0;
    uint32_t violations = 0; // nonzero if a lead of this lane broke rule (a)
    uint64_t high_leads = 0; // lanes that saw a lead 0xF8-0xFF
    int32_t claimed = 0;     // this lane's share of the continuation bytes the leads claim, rule (b)
    for (int64_t base = 0, pass = 0; base < span; base += decode_pass_bytes_k, ++pass) {
        int32_t const slot = (int32_t)(pass & 1);
        __builtin_amdgcn_wave_barrier(); // the previous pass's reads of the other slot come first
        stage(slot ^ 1, next, pass + 1); // this pass's last runes may need the next pass's first 3 bytes
        next = tape_block(block_of(pass + 2), end);
        __builtin_amdgcn_wave_barrier(); // the wave's LDS operations complete in order: stores, then these loads
        global_rune_t *pass_out = runes + written; // a pass writes at most 1 KB of runes: 32-bit offsets from here
        // Positions [from, to) of this pass hold the string's bytes.
        int32_t const from = (int32_t)max((int64_t)0, skew - base);
        int32_t const to = (int32_t)min((int64_t)decode_pass_bytes_k, span - base);
        uint32_t const *words = ring + slot * slot_words + lane / 4;
        auto row_bytes = [&](int32_t row) { // the 4 bytes from this lane's position of `row`
            return __builtin_amdgcn_alignbyte(words[row * 16 + 1], words[row * 16], lane & 3u);
        };
        // The string's bytes after position 0 of this pass (more than a lead can claim once the string goes on).
        int32_t const after_row0 = (int32_t)min(span - 1 - base, (int64_t)decode_pass_bytes_k + 3);
        uint32_t pass_written = 0;
        auto decode_row = [&](int32_t row, uint32_t bytes, uint64_t in_mask) __attribute__((always_inline)) {
            uint32_t const lead = bytes & 0xFFu;
            uint64_t const leads = ~__builtin_amdgcn_ballot_w64((bytes & 0xC0u) == 0x80u) & in_mask;
            high_leads |= __builtin_amdgcn_ballot_w64(lead >= 0xF8u) & in_mask;
            if (__builtin_amdgcn_inverse_ballot_w64(leads)) {
                // Continuation bytes the lead announces: 0 below 0xC0, 1 for 0xC_ and 0xD_, 2 for 0xE_, 3 for 0xF_.
                uint32_t const more = __builtin_amdgcn_ubfe(0xE5000000u, __builtin_amdgcn_ubfe(bytes, 4, 4) * 2, 2);
                // (a): bytes 1 .. more are continuation bytes (past the end they read as 0x80).
                uint32_t const pattern = (bytes ^ 0x80808080u) & 0xC0C0C0C0u; // zero bytes: continuation bytes
                violations |= __builtin_amdgcn_ubfe(pattern, 8, 8 * more);
                // (b): the claim, cut at the string's end.
                claimed += min((int32_t)more, after_row0 - row * 64 - (int32_t)lane);
                // The lead byte over the continuation bytes' 6 bits each, as if the sequence were 4 bytes long: the
                // rune is its field from bit 6 * (3 - more), 5 * more + 7 bits wide (the lead's marker bits above
                // the payload are dropped; the one bit past the payload is zero in every lead below 0xF8).
                uint32_t const full = lead << 18 | __builtin_amdgcn_ubfe(bytes, 8, 6) << 12 |
                                      __builtin_amdgcn_ubfe(bytes, 16, 6) << 6 | __builtin_amdgcn_ubfe(bytes, 24, 6);
                uint32_t const rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(leads >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)leads, 0u));
                *(global_rune_t *)((global_byte_t *)pass_out + ((pass_written + rank) << 2)) =
                    __builtin_amdgcn_ubfe(full, 18 - 6 * more, 5 * more + 7);
            }
            pass_written += (uint32_t)__builtin_popcountll(leads);
        };
        auto partial_row = [&](int32_t row) { // lanes of positions [from, to) only
            int32_t const low = max(from - row * 64, 0), high = min(to - row * 64, 64);
            uint64_t const upto = high >= 64 ? ~0ull : (1ull << high) - 1u;
            decode_row(row, row_bytes(row), upto & ~((1ull << low) - 1u));
        };
        // Rows in order: a partial first row, whole rows (four at a time, their LDS reads issued together), a
        // partial last row.
        int32_t const whole_begin = (from + 63) / 64, whole_end = to / 64;
        if (whole_begin > whole_end) partial_row(from / 64); // the pass's bytes lie inside one row
        else {
            if (from % 64) partial_row(from / 64);
            int32_t row = whole_begin;
            for (; row + 4 <= whole_end; row += 4) {
                uint32_t const a = row_bytes(row), b = row_bytes(row + 1), c = row_bytes(row + 2);
                uint32_t const d = row_bytes(row + 3);
                decode_row(row, a, ~0ull);
                decode_row(row + 1, b, ~0ull);
                decode_row(row + 2, c, ~0ull);
                decode_row(row + 3, d, ~0ull);
            }
            for (; row < whole_end; ++row) decode_row(row, row_bytes(row), ~0ull);
            if (to % 64) partial_row(whole_end);
        }
        written += pass_written;
        if (__builtin_amdgcn_ballot_w64(violations != 0u) | high_leads) return -1;
    }
    int64_t claims = claimed;
    for (int offset = 32; offset; offset >>= 1) claims += __shfl_xor(claims, offset, 64);
    return claims == length - written ? written : -1;
}

/**
 *  UTF-8 to runes for every tape of a batch in one launch, with `sz_rune_decode_unchecked` semantics
 *  (include/stringzilla/utf8_runes/serial.h:101-114): the lead byte fixes the width, continuation bytes are not
 *  validated, and continuation bytes missing at the end of a string read as 0. Runes of string `s` start at
 *  index `offsets[s]` of the tape's `runes`; their count goes to `rune_counts[s]`.
 *
 *  Waves draw tickets of 8 strings, T apart (T tickets in all), from a queue (`prepare_stats_t::decode_ticket`,
 *  zeroed before the launch), so a wave that drew long strings never holds the launch back while others idle; a
 *  ticket's offsets load at once, then the whole wave decodes its strings one after another (`decode_wave`, with the next
 *  string's first block and the next ticket requested ahead; a malformed string, or one the buffer has too little
 *  room for, is walked by one lane). Runes never go past the buffer's
 *  capacity: a tape whose offsets run further than the byte count the host sized it for keeps its strings
 *  truncated there and raises its segment's flag.
 */
__global__ __launch_bounds__(256) void utf8_decode_kernel(utf8_tape_t const *tapes, uint32_t tapes_count,
                                                          uint64_t strings_total, prepare_stats_t *stats,
                                                          uint32_t *segment_flags) {
    uint32_t const lane = __lane_id();
    auto uniform = [](uint64_t value, uint32_t from) { // lane `from`'s value, in every lane (a scalar)
        uint32_t const lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)value, (int)from);
        uint32_t const hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(value >> 32), (int)from);
        return (uint64_t)hi << 32 | lo;
    };
    auto draw = [&]() { // the first string of a new ticket
        uint32_t ticket = 0;
        if (lane == 0) ticket = atomicAdd(&stats->decode_ticket, 1u);
        return (uint64_t)__builtin_amdgcn_readfirstlane(ticket) * decode_strings_per_ticket_k;
    };
    uint64_t const tickets = (strings_total + decode_strings_per_ticket_k - 1) / decode_strings_per_ticket_k;
    uint64_t first_string = draw();
    while (first_string < strings_total) {
        uint64_t const following = draw(); // drawn now, so its latency hides behind this ticket's strings
        // Ticket t holds strings t, t + T, t + 2T, ... (T tickets in all): runs of long strings, as tapes often
        // hold, are dealt to different waves instead of one.
        uint64_t const global = first_string / decode_strings_per_ticket_k + lane * tickets;
        bool const mine = lane < decode_strings_per_ticket_k && global < strings_total;
        uint32_t low = 0, high = tapes_count - 1;
        while (mine && low < high) {
            uint32_t const middle = (low + high + 1) / 2;
            if (tapes[middle].string_base <= global) low = middle;
            else high = middle - 1;
        }
        utf8_tape_t const &tape = tapes[low];
        uint64_t const s = mine ? global - tape.string_base : 0;
        uint64_t const begin_read = mine ? tape.offsets[s] : 0, end_read = mine ? tape.offsets[s + 1] : 0;
        // A string whose offsets decrease or run past the tape's bytes is decoded as empty (nothing outside the
        // tape is read) and fails its segment.
        bool const out_of_tape = end_read < begin_read || end_read > tape.capacity;
        uint64_t const begin = out_of_tape ? 0 : begin_read, end = out_of_tape ? 0 : end_read;
        uint64_t const room = begin < tape.capacity ? tape.capacity - begin : 0;
        uint32_t const strings = (uint32_t)__builtin_popcountll(__ballot(mine)); // lanes 0 .. strings - 1
        // String i's bytes, rune output, length, and whether the whole wave decodes it (room for one rune per byte).
        auto first_of = [&](uint32_t i) { return (uint8_t const *)uniform((uint64_t)(tape.bytes + begin), i); };
        auto length_of = [&](uint32_t i) { return (int64_t)uniform(end - begin, i); };
        auto by_wave = [&](uint32_t i) {
            int64_t const length = length_of(i);
            return length > 0 && uniform(room, i) >= (uint64_t)length;
        };
        // The first pass's block of the next string is loaded while this one decodes.
        uint4 head = by_wave(0) ? decode_head(first_of(0), length_of(0)) : make_uint4(0, 0, 0, 0);
        for (uint32_t i = 0; i < strings; ++i) {
            uint8_t const *first = first_of(i);
            uint32_t *out = (uint32_t *)uniform((uint64_t)(tape.runes + begin), i);
            int64_t const length = length_of(i);
            uint64_t const space = uniform(room, i);
            bool const wave_path = by_wave(i);
            uint4 const current = head;
            if (i + 1 < strings && by_wave(i + 1)) head = decode_head(first_of(i + 1), length_of(i + 1));
            uint64_t written = 0;
            if (length > 0) {
                int64_t const decoded = wave_path ? decode_wave(first, length, out, current) : -1;
                if (decoded >= 0) written = (uint64_t)decoded;
                else {
                    uint64_t walked = 0;
                    if (lane == 0) walked = decode_walk(first, first + length, out, space);
                    written = uniform(walked, 0);
                }
            }
            if (lane == i) { // this lane loaded string i's offsets
                if (written > room || out_of_tape) { // the offsets disagree with the byte count the buffer was sized for
                    written = out_of_tape ? 0 : room;
                    atomicOr(&stats->bad_tape, 1u);
                    atomicOr(&segment_flags[tape.segment], (uint32_t)segment_bad_tape_k);
                }
                tape.rune_counts[s] = (uint32_t)written;
            }
        }
        first_string = following;
    }
}

/** Widens u32 tape offsets to u64 so every later kernel reads one layout. */
__global__ void widen_offsets_kernel(uint32_t const *narrow, uint64_t *wide, uint64_t count) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
        wide[i] = narrow[i];
}

/**
 *  Adds each valid lane's `count` to `counters[key]` with one atomic per distinct key of the wave, and returns the
 *  counter's value before this lane's share (lanes of one key in lane order). Every lane of the wave must call it.
 */
__device__ __forceinline__ uint32_t wave_counted_add_n(uint32_t *counters, uint32_t key, uint32_t count, bool valid) {
    uint32_t const lane = __lane_id();
    uint64_t pending = __ballot(valid);
    uint32_t mine = 0;
    while (pending) { // one pass per distinct key
        int const leader = __builtin_ctzll(pending);
        uint32_t const leader_key = (uint32_t)__shfl((int)key, leader, 64);
        bool const same = ((pending >> lane) & 1ull) && key == leader_key;
        uint32_t const share = same ? count : 0u;
        uint32_t inclusive = share;
        for (int d = 1; d < 64; d <<= 1) {
            uint32_t const other = (uint32_t)__shfl_up((int)inclusive, d, 64);
            if ((int)lane >= d) inclusive += other;
        }
        uint32_t const total = (uint32_t)__shfl((int)inclusive, 63, 64);
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(&counters[leader_key], total);
        base = (uint32_t)__shfl((int)base, leader, 64);
        if (same) mine = base + inclusive - share;
        pending &= ~__ballot(same);
    }
    return mine;
}

template <typename value_type_>
__device__ __forceinline__ value_type_ wave_sum(value_type_ x) {
    for (int offset = 32; offset; offset >>= 1) x += __shfl_xor(x, offset, 64);
    return x;
}

template <typename value_type_>
__device__ __forceinline__ value_type_ wave_max(value_type_ x) {
    for (int offset = 32; offset; offset >>= 1) x = max(x, __shfl_xor(x, offset, 64));
    return x;
}

/** The cost bucket of `x`: its octave and the bits after the leading one (`bucket_steps_per_octave_k` steps). */
__device__ __forceinline__ uint32_t cost_bucket(uint64_t x) {
    constexpr uint32_t steps = bucket_steps_per_octave_k;
    constexpr uint32_t bits = steps == 8 ? 3 : steps == 4 ? 2 : steps == 2 ? 1 : 0;
    static_assert((1u << bits) == steps, "a power of two steps per octave");
    if (x < steps) return (uint32_t)x;
    uint32_t const lg = 63 - __clzll(x);
    uint32_t const fraction = (uint32_t)(x >> (lg - bits)) & (steps - 1u); // the bits after the leading one
    uint32_t const bucket = steps * lg + fraction;
    return bucket < column_buckets_k ? bucket : column_buckets_k - 1;
}

/**
 *  Builds one task per scored cell of every segment (one launch for the whole batch) and counts it into its bin.
 *  Degenerate cells (an empty side) are scored here with the closed forms of serial.hpp:142-151 and 2493-2496 and
 *  never reach the scoring kernel. A pair whose scores could leave int32 (the reach bound of serial.hpp:117-140)
 *  is not scored: its segment is flagged instead. Cells and the widest multi-stripe pair are reduced per block
 *  into `stats`, so the host reads back a few bytes rather than the task list.
 */
__global__ void materialize_kernel(materialize_args_t const *all, uint32_t segments_count, uint64_t tasks_total,
                                   shape_table_t const *table, task_t *tasks, uint32_t *bin_of_task,
                                   uint32_t *histogram, segment_view_t const *segments, prepare_stats_t *stats,
                                   uint32_t *segment_overflow) {
    __shared__ unsigned long long block_cells, block_core_ops;
    __shared__ uint32_t block_widest, block_overflow;
    if (threadIdx.x == 0) block_cells = 0, block_core_ops = 0, block_widest = 0, block_overflow = 0;
    __syncthreads();
    unsigned long long cells = 0, core_ops = 0;
    uint32_t widest = 0, overflow = 0;
    // This lane's run of tasks in one bin (consecutive iterations), counted into the histogram when the bin
    // changes and, wave-aggregated, at the end: a batch of millions of pairs in one bin (a cross-product tile)
    // would otherwise queue one atomic per wave on one counter.
    uint32_t run_bin = ~0u, run_tasks = 0;
    for (uint64_t global = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; global < tasks_total;
         global += (uint64_t)gridDim.x * blockDim.x) {
        // The segment holding this task: the last one whose first task is at or before it.
        uint32_t low = 0, high = segments_count - 1;
        while (low < high) {
            uint32_t const middle = (low + high + 1) / 2;
            if (all[middle].task_base <= global) low = middle;
            else high = middle - 1;
        }
        materialize_args_t const &a = all[low];
        uint64_t const index = global - a.task_base;
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
        uint64_t const q_end = a.query_offsets[q + 1], c_end = c_offsets[c + 1];
        // A byte count the host did not read back is the tape's last offset, read here.
        uint64_t const q_bytes = a.query_bytes != unknown_bytes_k ? a.query_bytes : a.query_offsets[a.queries_count];
        uint64_t const c_bytes = !a.candidate_offsets ? q_bytes
                                 : a.candidate_bytes != unknown_bytes_k ? a.candidate_bytes
                                                                        : a.candidate_offsets[a.candidates_count];
        // Offsets that decrease or run past the tape's data would make the sweeps read outside it: such a pair
        // is never scored and its segment fails.
        bool const bad_tape = q_end < q_begin || q_end > q_bytes || c_end < c_begin || c_end > c_bytes;
        uint64_t q_length = bad_tape ? 0 : q_end - q_begin;
        uint64_t c_length = bad_tape ? 0 : c_end - c_begin;
        if (a.runes && !bad_tape) q_length = a.query_rune_counts[q], c_length = c_runes[c];

        task_t task;
        task.segment = a.segment;
        if (a.mode == 0) task.output = index, task.mirror = no_mirror_k;
        else {
            task.output = q * a.results_row_stride + c;
            task.mirror = (a.mode == 2 && q != c) ? c * a.results_row_stride + q : no_mirror_k;
        }
        bool const transposed = c_length < q_length;
        task.flags = transposed ? task_transposed_k : 0u;
        task.rows_offset = transposed ? c_begin : q_begin;
        task.columns_offset = transposed ? q_begin : c_begin;
        task.rows_length = (uint32_t)(transposed ? c_length : q_length);
        task.columns_length = (uint32_t)(transposed ? q_length : c_length);
        uint32_t bin = ~0u;
        bool const local = a.family == family_sw_linear_k || a.family == family_sw_affine_k;
        bool const affine = a.family == family_lev_affine_k || a.family == family_nw_affine_k ||
                            a.family == family_sw_affine_k;
        uint64_t const longest = q_length > c_length ? q_length : c_length;
        uint64_t const shorter = q_length < c_length ? q_length : c_length;
        // Range of the int32 cells (the reference picks its cell width by the reach bound of serial.hpp:117-140
        // instead). Every value the int32 sweeps keep for cell (i, j) is a path score over at most min(i, j)
        // substitutions and 2 min(i, j) + 2 gap runs, taken relative to its anti-diagonal (global families in the
        // shifted form: H - e(i+j)), or for local families with non-positive gaps an absolute score in
        // [0, s_max min(i, j)]: either way below (shorter + 1,100) x (max|s| + 6 max|gap|), the rows past n of a last
        // stripe (at most 1,024) included. The plain recurrences (engines whose biased costs leave int8) and local
        // families with positive gaps keep absolute scores, which grow with n + m. A pair whose bound reaches 2^31
        // runs the int64 shape instead; only the final e(n+m) of a global score is added in 64 bits.
        bool const unshifted = a.family >= family_nw_linear_k && a.symbols == class_unshifted_k;
        bool const relative = local ? a.open <= 0 && a.extend <= 0 : !unshifted;
        uint64_t const reach = (relative ? shorter : q_length + c_length) + 1100;
        bool const wide = reach * (uint64_t)a.cell_growth >= 0x7FFFFFFFull;
        bool const overflows = longest > 0xFFFFFFFFull; // task lengths are 32-bit
        if (bad_tape) {
            atomicOr(&stats->bad_tape, 1u);
            atomicOr(&segment_overflow[a.segment], (uint32_t)segment_bad_tape_k);
        }
        else if (q_length == 0 || c_length == 0) {
            int64_t value = 0;
            if (!local)
                value = affine ? (longest ? a.open + (int64_t)a.extend * ((int64_t)longest - 1) : 0)
                               : a.open * (int64_t)longest;
            int64_t *results = segments[a.segment].results;
            results[task.output] = value;
            if (task.mirror != no_mirror_k) results[task.mirror] = value;
        }
        else if (overflows) {
            overflow = 1;
            atomicOr(&segment_overflow[a.segment], (uint32_t)segment_overflow_k);
        }
        else {
            uint32_t const rows = task.rows_length;
            int const family = a.family, symbols = a.symbols;
            int const slot = (a.unit_costs && family == family_lev_linear_k)
                                 ? (symbols == symbols_u8_k ? unit_bytes_slot_k : unit_runes_slot_k)
                                 : family * 2 + symbols;
            int shape = table->multi_pass_shape[slot];
            for (int s = 0; s < table->used[slot]; ++s) {
                int const candidate_shape = table->order[slot][s];
                if ((int64_t)rows <= table->limits[slot][candidate_shape]) {
                    shape = candidate_shape;
                    break;
                }
            }
            // Packed 16-bit shapes only for engines whose score range provably fits them; else the int32 twin.
            if (shapes_k[shape].packed && shapes_k[shape].rows_per_lane > a.packed_rows) shape = unpacked_twin(shape);
            // Latency-oriented batches give a pair that one wave would sweep in many stripes the whole workgroup.
            if (a.latency_stripes > 0 && cooperative_twin(shape) >= 0) {
                uint64_t const stripe_rows = (uint64_t)shapes_k[shape].group * shapes_k[shape].rows_per_lane;
                if ((rows + stripe_rows - 1) / stripe_rows > (uint64_t)a.latency_stripes) shape = cooperative_twin(shape);
            }
            if (wide) shape = wide_shape_k;
            shape_t const sh = shapes_k[shape];
            // Latency of the item in lane-row-steps, the key for heaviest-first ordering.
            uint64_t const stripe = (uint64_t)sh.group * sh.rows_per_lane;
            uint64_t passes = (rows + stripe - 1) / stripe;
            if (sh.cooperative) {
                if (sh.chained) passes = max<uint64_t>(waves_per_group_k, (passes + chain_groups_k - 1) / chain_groups_k);
                passes = (passes + waves_per_group_k - 1) / waves_per_group_k + 1;
            }
            // ~DP rows of VALU work per step; a packed step costs about 3/4 of an int32 one and its stripes are
            // swept in columns + 2 * group steps.
            uint64_t const per_step = sh.bit_parallel ? (symbols == symbols_u8_k ? 8u : 20u) * sh.rows_per_lane / 64 // ~17 rune planes
                                      : sh.packed     ? (uint64_t)(sh.rows_per_lane * 3 + 3) / 4
                                      : sh.wide       ? 4ull * sh.rows_per_lane // int64 cells
                                                      : (uint64_t)sh.rows_per_lane;
            uint64_t const ramp = (uint64_t)sh.group * (sh.packed ? 2u : 1u);
            // A tall pair's last stripe may run a shorter sweep (`tail_sweeps_t`): half the Myers word, or 2, 4 or
            // 6 packed rows per lane.
            uint64_t last_step = per_step;
            if (!sh.cooperative && sh.group == 64 && passes > 0) {
                uint64_t const last_rows = rows - (passes - 1) * stripe;
                if (sh.bit_parallel && sh.rows_per_lane == 64 && last_rows <= 64 * 32) last_step = per_step / 2;
                else if (sh.packed && sh.rows_per_lane == 8 && last_rows <= 64 * 6)
                    last_step = last_rows <= 64 * 2 ? 2 : last_rows <= 64 * 4 ? 3 : 5;
            }
            uint64_t const work = (passes ? (passes - 1) * per_step + last_step : 0) *
                                  ((uint64_t)task.columns_length + ramp);
            if (sh.chained) { // a slot in the chain buffers, sized by the host from these stats
                uint32_t const chain_slot = atomicAdd(&stats->chained_pairs, 1u);
                atomicMax(&stats->chained_columns, task.columns_length);
                task.flags |= (chain_slot + 1u) << 8;
            }
            if (sh.wide) { // a slot in the int64 rows, likewise
                uint32_t const wide_slot = atomicAdd(&stats->wide_pairs, 1u);
                atomicMax(&stats->wide_columns, task.columns_length);
                task.flags |= (wide_slot + 1u) << 8;
            }
            bin = (uint32_t)variant_of(family, symbols, shape) * column_buckets_k + cost_bucket(work);
            cells += (unsigned long long)rows * task.columns_length;
            core_ops += (unsigned long long)rows * task.columns_length * core_ops_per_64_cells(family, symbols, sh);
            if (rows > stripe && !sh.wide && task.columns_length > widest) widest = task.columns_length;
        }
        if (bin != run_bin) {
            if (run_bin != ~0u) atomicAdd(&histogram[run_bin], run_tasks);
            run_bin = bin, run_tasks = 0;
        }
        ++run_tasks;
        tasks[global] = task;
        bin_of_task[global] = bin;
    }
    (void)wave_counted_add_n(histogram, run_bin, run_tasks, run_bin != ~0u);
    // Every lane is active again here: one LDS atomic per wave.
    cells = wave_sum(cells), core_ops = wave_sum(core_ops);
    widest = wave_max(widest), overflow = wave_max(overflow);
    if (__lane_id() == 0) {
        if (cells) atomicAdd(&block_cells, cells);
        if (core_ops) atomicAdd(&block_core_ops, core_ops);
        if (widest) atomicMax(&block_widest, widest);
        if (overflow) atomicOr(&block_overflow, overflow);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (block_cells) atomicAdd(&stats->cells, block_cells);
        if (block_core_ops) atomicAdd(&stats->core_ops_64, block_core_ops);
        if (block_widest) atomicMax(&stats->widest_multi_pass, block_widest);
        if (block_overflow) atomicOr(&stats->overflow, block_overflow);
    }
}

/**
 *  Single-workgroup exclusive scans over the bins in priority order: first slot of each bin in `order`, and
 *  first work item of each bin (an item is 64 / G pairs of the bin).
 */
__global__ __launch_bounds__(1024) void scan_bins_kernel(bins_view_t bins) {
    // Pass 1 of 2: workgroup b scans priority positions [1024 b, 1024 b + 1024) on its own and leaves its totals in
    // `scan_blocks`; pass 2 adds the totals of the workgroups before it. Every position has its own thread, so
    // the gathers of the histogram run in parallel across the chip.
    __shared__ uint32_t wave_tasks[16], wave_items[16];
    uint32_t const p = blockIdx.x * 1024 + threadIdx.x;
    uint32_t tasks = 0, items = 0;
    if (p < bins_count_k) {
        uint32_t const bin = bins.priority[p];
        tasks = bins.histogram[bin];
        items = items_for(shapes_k[(bin / column_buckets_k) % shapes_count_k], tasks);
    }
    uint32_t const lane = __lane_id(), wave = threadIdx.x / 64;
    uint32_t tasks_inclusive = tasks, items_inclusive = items;
    for (int offset = 1; offset < 64; offset <<= 1) {
        uint32_t const t = (uint32_t)__shfl_up((int)tasks_inclusive, offset, 64);
        uint32_t const n = (uint32_t)__shfl_up((int)items_inclusive, offset, 64);
        if ((int)lane >= offset) tasks_inclusive += t, items_inclusive += n;
    }
    if (lane == 63) wave_tasks[wave] = tasks_inclusive, wave_items[wave] = items_inclusive;
    __syncthreads();
    uint32_t tasks_before = tasks_inclusive - tasks, items_before = items_inclusive - items;
    for (uint32_t w = 0; w < wave; ++w) tasks_before += wave_tasks[w], items_before += wave_items[w];
    if (p < bins_count_k) bins.task_start[p] = tasks_before, bins.item_start[p] = items_before;
    if (threadIdx.x == 1023) {
        bins.scan_blocks[2 * blockIdx.x] = tasks_before + tasks;
        bins.scan_blocks[2 * blockIdx.x + 1] = items_before + items;
    }
}

/** Pass 2 of the bin scan: each position adds the totals of the workgroups of pass 1 before its own; the last
 *  workgroup also writes the total item count. */
__global__ __launch_bounds__(1024) void scan_bins_carry_kernel(bins_view_t bins) {
    __shared__ uint32_t carry_tasks, carry_items;
    if (threadIdx.x < 64) {
        uint32_t t = 0, n = 0;
        for (uint32_t b = threadIdx.x; b < blockIdx.x; b += 64) t += bins.scan_blocks[2 * b], n += bins.scan_blocks[2 * b + 1];
        t = wave_sum(t), n = wave_sum(n);
        if (threadIdx.x == 0) carry_tasks = t, carry_items = n;
    }
    __syncthreads();
    uint32_t const p = blockIdx.x * 1024 + threadIdx.x;
    if (p < bins_count_k) bins.task_start[p] += carry_tasks, bins.item_start[p] += carry_items;
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
        bins.item_start[bins_count_k] = carry_items + bins.scan_blocks[2 * blockIdx.x + 1];
}

/**
 *  Scatters task ids into `order`, grouped by bin (order inside a bin is arbitrary and does not affect results).
 *  Each lane reserves the slots of its run of tasks in one bin (consecutive iterations) at once, when the bin changes
 *  and, wave-aggregated, at the end, then writes their ids: a bin of millions of tasks takes a few thousand atomics.
 */
__global__ void scatter_bins_kernel(bins_view_t bins, uint32_t const *bin_of_task, uint32_t const *position_of_bin,
                                    uint32_t tasks_count) {
    // Each wave places one block of consecutive tasks (lane l takes every 64th from the block's l-th), so a bin's
    // order keeps runs of nearby tasks: the pairs of an all-pairs item share their query string, whose tables a
    // sweep can keep from one item to the next. With one task per lane the blocks are 64 tasks.
    uint32_t const waves = gridDim.x * (blockDim.x / 64);
    uint32_t const per_wave = (uint32_t)(((uint64_t)tasks_count + waves - 1) / waves + 63) & ~63u;
    uint64_t const block = (uint64_t)(blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64) * per_wave;
    uint32_t const block_end = (uint32_t)min(block + per_wave, (uint64_t)tasks_count);
    uint32_t run_bin = ~0u, run_tasks = 0, run_first = 0;
    auto place = [&](uint32_t rank) { // the run's ids into its slots from `rank` in its bin
        uint32_t const at = bins.task_start[position_of_bin[run_bin]] + rank;
        for (uint32_t i = 0; i < run_tasks; ++i) bins.order[at + i] = run_first + i * 64;
    };
    for (uint64_t t = block + __lane_id(); t < block_end; t += 64) {
        uint32_t const bin = bin_of_task[t];
        if (bin != run_bin) {
            if (run_bin != ~0u) place(atomicAdd(&bins.cursor[run_bin], run_tasks));
            run_bin = bin, run_tasks = 0, run_first = t;
        }
        ++run_tasks;
        if (bin != ~0u) bins.shape_of_task[t] = (bin / column_buckets_k) % shapes_count_k;
    }
    uint32_t const rank = wave_counted_add_n(bins.cursor, run_bin, run_tasks, run_bin != ~0u);
    if (run_bin != ~0u) place(rank);
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

hipError_t szs_amd_launch_fill_many(szs_amd::fill_list_t const *list, hipStream_t stream) {
    uint32_t most = 0;
    for (uint32_t e = 0; e < list->entries; ++e) most = list->count[e] > most ? list->count[e] : most;
    if (!most) return hipSuccess;
    uint32_t const blocks = (most + 255) / 256 < 1024 ? (most + 255) / 256 : 1024;
    hipLaunchKernelGGL(szs_amd::fill_many_kernel, dim3(blocks), dim3(256), 0, stream, *list);
    return hipGetLastError();
}

hipError_t szs_amd_launch_utf8_decode(szs_amd::utf8_tape_t const *tapes, uint32_t tapes_count, uint64_t strings_total,
                                      szs_amd::prepare_stats_t *stats, uint32_t *segment_flags, hipStream_t stream) {
    if (!strings_total || !tapes_count) return hipSuccess;
    // Enough waves to fill every SIMD several times over; they draw tickets of 8 strings until none are left.
    uint64_t const tickets = (strings_total + szs_amd::decode_strings_per_ticket_k - 1) / szs_amd::decode_strings_per_ticket_k;
    uint64_t const blocks64 = (tickets + 3) / 4;
    uint32_t const blocks = blocks64 < 2048 ? (uint32_t)blocks64 : 2048;
    hipLaunchKernelGGL(szs_amd::utf8_decode_kernel, dim3(blocks), dim3(256), 0, stream, tapes, tapes_count,
                       strings_total, stats, segment_flags);
    return hipGetLastError();
}

hipError_t szs_amd_launch_widen_offsets(uint32_t const *narrow, uint64_t *wide, uint64_t count, hipStream_t stream) {
    if (!count) return hipSuccess;
    uint64_t const blocks64 = (count + 255) / 256;
    uint32_t const blocks = blocks64 < 65536 ? (uint32_t)blocks64 : 65536;
    hipLaunchKernelGGL(szs_amd::widen_offsets_kernel, dim3(blocks), dim3(256), 0, stream, narrow, wide, count);
    return hipGetLastError();
}

hipError_t szs_amd_launch_materialize(szs_amd::materialize_args_t const *args, uint32_t segments_count,
                                      uint64_t tasks_total, szs_amd::shape_table_t const *table, szs_amd::task_t *tasks,
                                      uint32_t *bin_of_task, uint32_t *histogram,
                                      szs_amd::segment_view_t const *segments, szs_amd::prepare_stats_t *stats,
                                      uint32_t *segment_overflow, hipStream_t stream) {
    if (!tasks_total || !segments_count) return hipSuccess;
    // One task per lane up to 256 K tasks, more beyond: each lane's run of tasks in one bin is counted with one
    // atomic, and a small batch keeps one short dependent chain of loads per lane.
    uint64_t const blocks64 = (tasks_total + 255) / 256;
    uint32_t const blocks = blocks64 < 1024 ? (uint32_t)blocks64 : 1024;
    hipLaunchKernelGGL(szs_amd::materialize_kernel, dim3(blocks), dim3(256), 0, stream, args, segments_count,
                       tasks_total, table, tasks, bin_of_task, histogram, segments, stats, segment_overflow);
    return hipGetLastError();
}

hipError_t szs_amd_launch_scan_bins(szs_amd::bins_view_t const *bins, hipStream_t stream) {
    constexpr uint32_t blocks = (szs_amd::bins_count_k + 1023) / 1024;
    static_assert(blocks <= szs_amd::scan_blocks_k, "the scan's block totals fit their buffer");
    hipLaunchKernelGGL(szs_amd::scan_bins_kernel, dim3(blocks), dim3(1024), 0, stream, *bins);
    hipLaunchKernelGGL(szs_amd::scan_bins_carry_kernel, dim3(blocks), dim3(1024), 0, stream, *bins);
    return hipGetLastError();
}

hipError_t szs_amd_launch_scatter_bins(szs_amd::bins_view_t const *bins, uint32_t const *bin_of_task,
                                       uint32_t const *position_of_bin, uint32_t tasks_count, hipStream_t stream) {
    if (!tasks_count) return hipSuccess;
    // One task per lane up to 256 K tasks, more beyond, as for materialization.
    uint32_t const blocks64 = (tasks_count + 255) / 256;
    uint32_t const blocks = blocks64 < 1024 ? blocks64 : 1024;
    hipLaunchKernelGGL(szs_amd::scatter_bins_kernel, dim3(blocks), dim3(256), 0, stream, *bins, bin_of_task,
                       position_of_bin, tasks_count);
    return hipGetLastError();
}

hipError_t szs_amd_launch_persistent(szs_amd::task_t const *tasks, szs_amd::segment_view_t const *segments,
                                     szs_amd::bins_view_t const *bins, int32_t *scratch, uint32_t scratch_capacity,
                                     uint32_t workgroups, uint32_t profile_bytes, szs_amd::chain_view_t const *chain,
                                     hipStream_t stream) {
    szs_amd::persistent_args_t args;
    args.tasks = tasks;
    args.segments = segments;
    args.bins = *bins;
    args.scratch = scratch;
    args.scratch_capacity = scratch_capacity;
    args.wave_lds_bytes = profile_bytes;
    args.chain = *chain;
    hipLaunchKernelGGL(szs_amd::persistent_score_kernel, dim3(workgroups), dim3(64 * szs_amd::waves_per_group_k),
                       profile_bytes * szs_amd::waves_per_group_k, stream, args);
    return hipGetLastError();
}

hipError_t szs_amd_launch_single(szs_amd::task_t const *tasks, szs_amd::segment_view_t const *segments,
                                 uint32_t const *identity, uint32_t task_index, int variant, int32_t *scratch,
                                 uint32_t scratch_capacity, uint32_t profile_bytes, szs_amd::chain_view_t const *chain,
                                 hipStream_t stream) {
    hipLaunchKernelGGL(szs_amd::single_task_kernel, dim3(1), dim3(64 * szs_amd::waves_per_group_k),
                       profile_bytes * szs_amd::waves_per_group_k, stream, tasks, segments, identity, task_index,
                       variant, scratch, scratch_capacity, profile_bytes, *chain);
    return hipGetLastError();
}

hipError_t szs_amd_persistent_occupancy(int *workgroups_per_cu, uint32_t profile_bytes) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(workgroups_per_cu, szs_amd::persistent_score_kernel,
                                                        64 * szs_amd::waves_per_group_k,
                                                        profile_bytes * szs_amd::waves_per_group_k);
}

} // extern "C"

#pragma endregion
