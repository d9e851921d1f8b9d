/**
 *  @file wide_sweep.hpp
 *  @brief The 64-bit sweep: one wave per pair, int64 cells, for pairs whose values could leave the int32 cells
 *         of the other sweeps (the reference moves to 8-byte cells there, serial.hpp:293-356).
 *
 *  The recurrences are the plain ones of dp_sweep.hpp (serial.hpp:669-1079, borders and discard values of its
 *  walkers), with no domain shift: every stored value is the cell's own H, E or F. Layout as the int32 sweep: lane
 *  k holds R consecutive rows of a 64R-row stripe and updates them at column t + 1 - k on step t; the row above
 *  and its vertical gap arrive from lane k - 1 through DPP shifts of both 32-bit halves, lane 0 taking them (and
 *  the column symbol) from registers the wave loaded for the current 64 columns, one broadcast per step. Lane 63
 *  stores the stripe's bottom row and gaps into the pair's own int64 rows in HBM, which the next stripe reads in
 *  place (the columns a chunk reads are loaded before the stores of that chunk can reach them).
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dp_sweep.hpp"
#include "szs_device.hpp"

namespace szs_amd {

template <bool minimize_>
__device__ __forceinline__ int64_t opt64(int64_t a, int64_t b) {
    if constexpr (minimize_) return a < b ? a : b;
    else return a > b ? a : b;
}

/** H of the matrix border at distance `k` from the origin, in 64 bits (serial.hpp:142-151). */
template <int family_>
__device__ __forceinline__ int64_t border_score64(int64_t k, int64_t open, int64_t extend) {
    using traits = family_traits<family_>;
    if constexpr (traits::local_k) return 0;
    else if constexpr (traits::affine_k) return k ? open + extend * (k - 1) : 0;
    else return open * k;
}

/** Lane i takes lane i - 1's 64-bit value; lane 0 takes `old`. */
__device__ __forceinline__ int64_t shift_down64(int64_t old, int64_t src) {
    uint32_t const low = (uint32_t)__builtin_amdgcn_update_dpp((int32_t)(uint32_t)old, (int32_t)(uint32_t)src,
                                                              dpp_wave_shr1_k, 0xF, 0xF, false);
    uint32_t const high = (uint32_t)__builtin_amdgcn_update_dpp((int32_t)(uint32_t)((uint64_t)old >> 32),
                                                               (int32_t)(uint32_t)((uint64_t)src >> 32),
                                                               dpp_wave_shr1_k, 0xF, 0xF, false);
    return (int64_t)((uint64_t)high << 32 | low);
}

/** Lane `source`'s 64-bit value in every lane (`source` wave-uniform). */
__device__ __forceinline__ int64_t broadcast64(int64_t value, int32_t source) {
    uint32_t const low = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)value, source);
    uint32_t const high = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)((uint64_t)value >> 32), source);
    return (int64_t)((uint64_t)high << 32 | low);
}

/**
 *  Scores one pair of the item with int64 cells: the wave's own pair for a one-wave item, else pair `wave` of the
 *  workgroup's item. `rows_per_lane_` rows per lane, 64 lanes.
 */
template <int family_, typename symbol_type_, int rows_per_lane_>
__device__ void run_wide(item_args_t const &a) {
    using traits = family_traits<family_>;
    constexpr int R = rows_per_lane_;
    constexpr bool minimize = traits::minimize_k, affine = traits::affine_k, local = traits::local_k;
    constexpr bool uniform = traits::uniform_k;
    uint32_t const wave = threadIdx.x / 64;
    int32_t const lane = (int32_t)__lane_id();
    uint32_t const slot = a.wave_item ? 0u : wave;
    if (slot >= a.count) return; // wave-uniform
    task_t const task = a.tasks[a.order[a.first + slot]];
    segment_view_t const segment = a.segments[task.segment];
    engine_params_t const *params = segment.params;
    int64_t const n = task.rows_length, m = task.columns_length;
    int64_t const open = params->open, extend = params->extend;
    int64_t const match = params->match, mismatch = params->mismatch;
    int32_t const dense_count = params->dense_count;
    bool const transposed = task_transposed(task);
    symbol_type_ const *query = (symbol_type_ const *)segment.query_symbols;
    symbol_type_ const *candidate = (symbol_type_ const *)segment.candidate_symbols;
    symbol_type_ const *rows = (transposed ? candidate : query) + task.rows_offset;
    symbol_type_ const *columns = (transposed ? query : candidate) + task.columns_offset;
    int64_t *above = a.wide_rows + (size_t)task_chain_slot(task) * 2 * a.wide_pitch; // bottom row, then its gaps
    int64_t *above_gaps = above + a.wide_pitch;

    // Class costs in this wave's LDS as [row class][column class] (transposed pairs swap the operands here, so
    // the cost is always C[class(query)][class(candidate)]).
    int8_t *costs = wave_lds(wave, a.wave_lds_bytes);
    if constexpr (!uniform) {
        for (int32_t e = lane; e < dense_count * dense_count; e += 64) {
            int32_t const r = e / dense_count, c = e % dense_count;
            costs[e] = transposed ? params->dense_costs[c * max_dense_classes_k + r]
                                  : params->dense_costs[r * max_dense_classes_k + c];
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    auto symbol_of = [&](symbol_type_ raw) -> uint32_t {
        if constexpr (uniform) return (uint32_t)raw;
        else return (uint32_t)params->byte_to_dense[(uint32_t)raw & 0xFFu];
    };

    int64_t const stripe_rows = 64 * R;
    int64_t const stripes = (n + stripe_rows - 1) / stripe_rows;
    int64_t const chunks = (m + 63 + 63) / 64; // lane 63 finishes column m on step m + 62
    int64_t best = 0;                          // local families: serial.hpp:781 starts the best at 0
    int64_t result = 0;
    for (int64_t stripe = 0; stripe < stripes; ++stripe) {
        int64_t const base = stripe * stripe_rows;
        int64_t const lane_top = base + (int64_t)lane * R; // the row above this lane's first row
        bool const more = stripe + 1 < stripes;
        uint32_t row_symbol[R];
        int64_t scores[R], horizontal[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            int64_t const i = lane_top + r + 1;
            row_symbol[r] = i <= n ? symbol_of(rows[i - 1]) : 0xFFFFFFFFu;
            int64_t const border = border_score64<family_>(i, open, extend);
            scores[r] = border;
            horizontal[r] = (open + extend) + border; // the discard value of the track on column 0
        }
        int64_t saved_up = border_score64<family_>(lane_top, open, extend); // H(lane_top, 0)
        int64_t bottom_score = 0, bottom_gap = 0;
        uint32_t symbol_below = 0;
        for (int64_t chunk = 0; chunk < chunks; ++chunk) {
            // Lane l loads lane 0's inputs for column chunk * 64 + l + 1: the row above the stripe and its gap.
            int64_t const j_load = chunk * 64 + lane + 1;
            int64_t in_score = 0, in_gap = 0;
            uint32_t in_symbol = 0;
            if (stripe == 0) {
                in_score = border_score64<family_>(j_load, open, extend);
                in_gap = (open + extend) + in_score; // discard value of the vertical track on row 0
            }
            else if (j_load <= m) {
                in_score = above[j_load];
                in_gap = above_gaps[j_load];
            }
            if (j_load <= m) in_symbol = symbol_of(columns[j_load - 1]);
            for (int32_t step = 0; step < 64; ++step) {
                int64_t const up_in = shift_down64(broadcast64(in_score, step), bottom_score);
                int64_t const gap_in = shift_down64(broadcast64(in_gap, step), bottom_gap);
                uint32_t const symbol = (uint32_t)__builtin_amdgcn_update_dpp(
                    __builtin_amdgcn_readlane((int32_t)in_symbol, step), (int32_t)symbol_below, dpp_wave_shr1_k,
                    0xF, 0xF, false);
                symbol_below = symbol;
                int64_t const j = chunk * 64 + step + 1 - lane;
                if (j >= 1 && j <= m) {
                    int64_t diagonal = saved_up;
                    saved_up = up_in;
                    int64_t up = up_in, vertical = gap_in;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        int64_t cost;
                        if constexpr (uniform) cost = row_symbol[r] == symbol ? match : mismatch;
                        else cost = row_symbol[r] == 0xFFFFFFFFu ? 0 : (int64_t)costs[row_symbol[r] * dense_count + symbol];
                        int64_t const left = scores[r];
                        int64_t const if_substitution = diagonal + cost;
                        int64_t cell;
                        if constexpr (affine) {
                            horizontal[r] = opt64<minimize>(left + open, horizontal[r] + extend);
                            vertical = opt64<minimize>(up + open, vertical + extend);
                            cell = opt64<minimize>(opt64<minimize>(horizontal[r], vertical), if_substitution);
                        }
                        else cell = opt64<minimize>(opt64<minimize>(up, left) + open, if_substitution);
                        if constexpr (local) {
                            cell = cell > 0 ? cell : 0;
                            if (lane_top + r + 1 <= n) best = cell > best ? cell : best;
                        }
                        diagonal = left;
                        up = cell;
                        scores[r] = cell;
                    }
                    bottom_score = up;
                    bottom_gap = vertical;
                    if (more && lane == 63) above[j] = up, above_gaps[j] = vertical;
                }
            }
        }
        if (more) __threadfence_block(); // lane 63's stores before the next stripe's loads
        if constexpr (!local) {
            if (!more) { // the lane holding row n keeps H(n, m) in that row's register
                int64_t const last_row = n - 1 - base;
                if (last_row / R == lane) {
                    int32_t const wanted = (int32_t)(last_row % R);
#pragma unroll
                    for (int r = 0; r < R; ++r) result += scores[r] & -(int64_t)(wanted == r);
                }
            }
        }
    }
    if constexpr (local) {
        for (int offset = 1; offset < 64; offset <<= 1) {
            int64_t const other = (int64_t)__shfl_xor((long long)best, offset, 64);
            best = other > best ? other : best;
        }
        if (lane == 0) {
            segment.results[task.output] = best;
            if (task.mirror != no_mirror_k) segment.results[task.mirror] = best;
        }
    }
    else {
        int64_t const last_row = n - 1 - (stripes - 1) * stripe_rows;
        if (last_row / R == lane) {
            segment.results[task.output] = result;
            if (task.mirror != no_mirror_k) segment.results[task.mirror] = result;
        }
    }
}

} // namespace szs_amd
