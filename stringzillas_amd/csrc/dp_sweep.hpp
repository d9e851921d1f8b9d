/**
 *  @file dp_sweep.hpp
 *  @brief The wave-systolic dynamic-programming sweep: one stripe of rows of one pair, column after column.
 *
 *  Recurrences restated from include/stringzillas/similarities/serial.hpp:669-1079 (the four tile scorers) with
 *  the border conventions of its walkers (serial.hpp:1153-1158, 1310-1320, 1627-1633):
 *  - linear global:  H = opt(opt(up, left) + g, diag + s)
 *  - linear local:   H = max(max(up, left) + g, max(diag + s, 0)), best over all cells, starting at 0
 *  - affine global:  E = opt(H_left + o, E_left + e), V = opt(H_up + o, V_up + e), H = opt(opt(E, V), diag + s)
 *  - affine local:   H = max(max(E, V), max(diag + s, 0))
 *  Border scores: H[i][0] = g*i (linear) or o + e*(i-1) (affine, 0 at i = 0), 0 for local; the gap tracks start
 *  from the discard value (o + e) + H_border.
 *
 *  Layout: a group of G lanes holds one stripe of G*R rows of the shorter string, R consecutive rows per lane,
 *  in VGPRs. At step t lane k updates its R rows at column j = t + 1 - k, then passes its bottom score, bottom
 *  vertical-gap value and the column symbol to lane k+1 with one DPP shift each. Lane 0 receives the row above
 *  the stripe: the matrix border, or the previous stripe's bottom row read back from HBM scratch. Lane G-1's
 *  bottom row is stored step by step into the consumed entries of the LDS ring and written to scratch 64 columns
 *  at a time (G = 64) for the next stripe.
 *
 *  Global families keep every value relative to its anti-diagonal: Ĉ = H - e*(i+j), ĥ = E - e*(i+j) + (e - o),
 *  v̂ = F - e*(i+j) + (e - o) (e = extend, o = open; for linear gaps e = o = g). In that domain the recurrences lose
 *  their gap constants: ĥ = opt(Ĉ_left, ĥ_left), v̂ = opt(Ĉ_up, v̂_up), Ĉ = opt(ĥ, v̂, Ĉ_diag + s - o - e) + (o - e),
 *  and linear Ĉ = opt(Ĉ_up, Ĉ_left, Ĉ_diag + s - 2g): 5 and 2 VALU per cell. Each value lives in its own cell's
 *  domain, so lanes, stripes and the HBM boundary rows exchange them unchanged; the border values enter in the
 *  same domain and the result leaves it with + e*(n+m). Local families cannot (their zero floor is not constant
 *  along anti-diagonals) and keep scores shifted by S = open instead: the gap terms read the stored value, the
 *  substitution term reads `cost - S`, one add per cell stores `cell + S`. The LDS profile holds `cost - bias`
 *  with the bias of the form used (o + e, 2g or o). With G = 64 lane 0's inputs for a
 *  64-step chunk (row above the stripe, its vertical gaps, column symbols) are staged in a 1 KB LDS ring, one
 *  16-byte entry per step read back as a broadcast, so each step injects them with the DPP shift alone. Chunks in
 *  which every lane has a valid column skip the per-step column check.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "szs_device.hpp"

namespace szs_amd {

#pragma region DPP helpers

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

/**
 *  `from` of the lane below plus `step` in lanes 1..63, `lane0` in lane 0: one `v_add_u32_dpp` (wave_shr:1; lane 0
 *  has no source lane, so it keeps the value its destination register already held). Written out because the
 *  compiler fuses a DPP move into an add only with a zero for lane 0. The `s_nop 1` gives the two wait states a DPP
 *  read needs after a vector write of its source, which the compiler cannot see inside the asm.
 */
__device__ __forceinline__ uint32_t lane_above_plus(uint32_t from, uint32_t step, uint32_t lane0) {
    uint32_t result = lane0;
    asm("s_nop 1\n\tv_add_u32_dpp %0, %1, %2 wave_shr:1 row_mask:0xf bank_mask:0xf"
        : "+v"(result)
        : "v"(from), "v"(step));
    return result;
}

/** Shift register toward lane 0 in which lane 63 appends its own `fresh` value (G = 64 only). */
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

/** `a + b` modulo 2^32. Border scores and anti-diagonal offsets grow with i + j and may wrap; their sums, the
 *  stored values, are small (the range bound guarantees it), so modular arithmetic gives them exactly. */
__device__ __forceinline__ int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
__device__ __forceinline__ int32_t wrap_mul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }

/** Score of the matrix border at distance `k` from the origin (serial.hpp:142-151, 716-727, 889-895), modulo
 *  2^32 (see `wrap_add`). */
template <int family_>
__device__ __forceinline__ int32_t border_score(int32_t k, int32_t open, int32_t extend) {
    using traits = family_traits<family_>;
    if constexpr (traits::local_k) return 0;
    else if constexpr (traits::affine_k) return k ? wrap_add(open, wrap_mul(extend, k - 1)) : 0;
    else return wrap_mul(open, k);
}

#pragma endregion

struct myers_tables_t;

/** Everything a workgroup needs to score one work item. */
struct item_args_t {
    task_t const *tasks;
    segment_view_t const *segments;
    uint32_t const *order; /**< task ids; `first .. first + count` belong to this item */
    uint32_t first;
    uint32_t count;
    bool wave_item;        /**< the item is this wave's alone (else the workgroup's: wave w takes its w-th share) */
    /** The wave's previous item ran the same variant (one-wave items only), so its LDS still holds what that
     *  sweep left there: a sweep may reuse tables built for the same rows string. */
    bool same_variant;
    int32_t *scratch;      /**< this workgroup's 4 x 2 x `scratch_capacity` int32 rows */
    uint32_t scratch_capacity;
    uint32_t wave_lds_bytes; /**< dynamic LDS per wave: profile or Myers tables, then the injection ring */
    int32_t *shared_result; /**< LDS word for cooperative reductions */
    // Chained shapes: group `group` of the pair (-1: every group in turn, in one workgroup), the bottom rows
    // between groups ([slot][group][scores, gaps][chain_pitch] int32) and their published column counts
    // ([slot][group] u32), and the word that records a wait given up.
    int32_t group;
    int32_t *chain_rows;
    uint32_t *chain_progress;
    uint32_t chain_pitch;
    uint32_t *chain_timeout;
    int64_t *wide_rows;  /**< the int64 shape's per-pair bottom rows ([slot][scores, gaps][wide_pitch]) */
    uint64_t wide_pitch;
};

/**
 *  This wave's slice of the dynamic LDS. It is addressed through the `extern __shared__` array in the function
 *  that uses it, so the compiler emits `ds_*` instructions; a pointer handed over through memory would lose its
 *  address space and turn every profile read into a flat load.
 */
__device__ __forceinline__ int8_t *wave_lds(uint32_t wave, uint32_t bytes) {
    extern __shared__ __align__(16) int8_t szs_dynamic_lds[];
    return szs_dynamic_lds + wave * bytes;
}

/** How a sweep's per-lane partial results combine into the pair's score. */
enum reduce_kind_t { reduce_none_k = 0, reduce_max_k = 1, reduce_sum_k = 2 };

/**
 *  Per-lane state of one group sweeping one stripe of one pair. `shifted_` selects the anti-diagonal domain
 *  (global families) or the `H + open` shift (local families), whose profiles hold `cost - bias`; without it the
 *  plain recurrences run with explicit gap adds, for engines whose biased costs would not fit int8.
 */
template <int family_, typename symbol_type_, int G, int R, bool shifted_ = true>
struct sweep_t {
    using traits = family_traits<family_>;
    static constexpr bool minimize_k = traits::minimize_k;
    static constexpr bool affine_k = traits::affine_k;
    static constexpr bool local_k = traits::local_k;
    static constexpr bool uniform_k = traits::uniform_k;
    static constexpr int RP = (R + 3) / 4 * 4; // profile bytes per lane and class, padded to whole dwords
    static constexpr int group_k = G;
    static constexpr int rows_per_lane = R;
    static constexpr int skew_k = 1; // lane k + 1 runs one column behind lane k
    static constexpr reduce_kind_t reduce_k = local_k ? reduce_max_k : reduce_none_k;
    static constexpr bool diagonal_k = shifted_ && !local_k; /**< values relative to their anti-diagonal */

    // The pair.
    int lane, k;
    bool active;
    int32_t n, m;
    int32_t open, extend, match, mismatch, dense_count;
    int32_t shift;                  /**< cost bias of the profile and of `match`/`mismatch` below */
    int32_t match_shifted, mismatch_shifted;
    int32_t open_minus_extend;
    bool transposed;
    engine_params_t const *params;
    symbol_type_ const *rows;
    symbol_type_ const *columns;
    int64_t *results;
    uint64_t output, mirror;
    int8_t *profile; /**< this wave's LDS profile: [dense class][lane][RP] int8 of `cost - S` */
    int8_t *ring;    /**< this wave's injection ring (G = 64) */

    // The stripe.
    int32_t lane_top;
    uint32_t row_symbols[R];
    int32_t scores[R];
    int32_t horizontal[R];
    int32_t best[R];
    int32_t diagonal_top, bottom_score, bottom_gap;
    uint32_t column_symbol;
    int32_t running_best;

    __device__ __forceinline__ void load(item_args_t const &a, uint32_t task_index, bool is_active, uint32_t wave) {
        int8_t *wave_profile = wave_lds(wave, a.wave_lds_bytes);
        lane = (int)__lane_id();
        k = lane % G;
        active = is_active;
        task_t const task = a.tasks[task_index];
        segment_view_t const segment = a.segments[task.segment];
        params = segment.params;
        n = active ? (int32_t)task.rows_length : 0;
        m = active ? (int32_t)task.columns_length : 0;
        open = params->open, extend = params->extend;
        match = params->match, mismatch = params->mismatch;
        dense_count = params->dense_count;
        shift = shifted_ ? params->shift : 0;
        match_shifted = match - shift, mismatch_shifted = mismatch - shift;
        open_minus_extend = open - extend;
        transposed = task_transposed(task);
        symbol_type_ const *query = (symbol_type_ const *)segment.query_symbols;
        symbol_type_ const *candidate = (symbol_type_ const *)segment.candidate_symbols;
        rows = (transposed ? candidate : query) + task.rows_offset;
        columns = (transposed ? query : candidate) + task.columns_offset;
        results = segment.results;
        output = task.output;
        mirror = task.mirror;
        profile = wave_profile;
        ring = wave_profile + (a.wave_lds_bytes - inject_ring_bytes_k);
        running_best = 0; // serial.hpp:781 starts the local best at 0
    }

    /** Offset between a stored score and H at cell (i, j): -e*(i+j) (modulo 2^32), S = open (local families) or
     *  0. */
    __device__ __forceinline__ int32_t domain(int32_t i, int32_t j) const {
        if constexpr (diagonal_k) return wrap_mul(-extend, wrap_add(i, j));
        else return shift;
    }

    /** The same offset in 64 bits, for the pair's score. */
    __device__ __forceinline__ int64_t domain64(int64_t i, int64_t j) const {
        if constexpr (diagonal_k) return -(int64_t)extend * (i + j);
        else return shift;
    }

    /** Offset between a stored gap value and its track at cell (i, j). */
    __device__ __forceinline__ int32_t gap_domain(int32_t i, int32_t j) const {
        if constexpr (diagonal_k) return wrap_add(wrap_mul(-extend, wrap_add(i, j)), extend - open);
        else return 0;
    }

    /** Loads the rows `base + k*R .. + R` of this lane and resets the stripe state to the column-0 border. */
    __device__ __forceinline__ void begin_stripe(int32_t base) {
        lane_top = base + k * R;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            int32_t const i = lane_top + r + 1; // 1-based matrix row
            uint32_t symbol = 0xFFFFFFFFu;
            if (i <= n) symbol = (uint32_t)rows[i - 1];
            row_symbols[r] = symbol;
            int32_t const border = border_score<family_>(i, open, extend);
            scores[r] = wrap_add(border, domain(i, 0));
            // The horizontal track starts from its discard value on column 0.
            horizontal[r] = wrap_add(wrap_add(open + extend, border), gap_domain(i, 0));
            best[r] = 0;
        }
        if constexpr (!uniform_k) {
            int8_t *mine = profile + lane * RP;
            for (int32_t c = 0; c < dense_count; ++c) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    uint32_t const row_class =
                        row_symbols[r] == 0xFFFFFFFFu ? 0u : params->byte_to_dense[row_symbols[r] & 0xFF];
                    int8_t const cost = transposed ? params->dense_costs[c * max_dense_classes_k + row_class]
                                                   : params->dense_costs[row_class * max_dense_classes_k + c];
                    mine[c * 64 * RP + r] = (int8_t)(cost - shift); // fits: checked when the engine was made
                }
            }
        }
        diagonal_top = wrap_add(border_score<family_>(lane_top, open, extend), domain(lane_top, 0)); // above-left
        bottom_score = scores[R - 1];
        bottom_gap = 0;
        column_symbol = 0;
    }

    /** Lane-0 inputs of column `j` (1-based): the row above the stripe and its vertical gap (stored form), and
     *  the column symbol. */
    __device__ __forceinline__ void column_inputs(int32_t j, int32_t const *in_scores, int32_t const *in_gaps,
                                                  int32_t &score, int32_t &gap, uint32_t &symbol) const {
        symbol = 0;
        if (j <= m) {
            uint32_t const raw = (uint32_t)columns[j - 1];
            symbol = uniform_k ? raw : (uint32_t)params->byte_to_dense[raw & 0xFF];
        }
        if (in_scores) {
            score = j <= m ? in_scores[j] : 0;
            gap = (affine_k && j <= m) ? in_gaps[j] : 0;
        }
        else {
            int32_t const border = border_score<family_>(j, open, extend);
            score = wrap_add(border, domain(0, j));
            gap = wrap_add(wrap_add(open + extend, border), gap_domain(0, j)); // the vertical track's discard value
        }
    }

    /** One column of this lane's R rows: `up_score` is H + S of the cell above, `up_gap` its vertical gap. */
    __device__ __forceinline__ void update_rows(int32_t up_score, int32_t up_gap) {
        int32_t diagonal = diagonal_top;
        int32_t up = up_score;
        int32_t vertical = up_gap;
        uint32_t costs_packed[RP / 4];
        if constexpr (!uniform_k) {
            uint32_t const *row_costs = (uint32_t const *)(profile + (column_symbol * 64 + lane) * RP);
#pragma unroll
            for (int w = 0; w < RP / 4; ++w) costs_packed[w] = row_costs[w];
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            int32_t cost; // cost - bias
            if constexpr (uniform_k) cost = row_symbols[r] == column_symbol ? match_shifted : mismatch_shifted;
            else cost = (int32_t)(int8_t)(costs_packed[r / 4] >> (8 * (r % 4)));
            int32_t const left = scores[r];
            int32_t const if_substitution = diagonal + cost;
            int32_t stored;
            if constexpr (diagonal_k) {
                if constexpr (affine_k) {
                    int32_t const h = opt<minimize_k>(left, horizontal[r]);
                    int32_t const v = opt<minimize_k>(up, vertical);
                    horizontal[r] = h;
                    vertical = v;
                    stored = opt<minimize_k>(opt<minimize_k>(h, v), if_substitution) + open_minus_extend;
                }
                else stored = opt<minimize_k>(opt<minimize_k>(up, left), if_substitution);
            }
            else {
                int32_t cell; // H[i][j]; `left`, `up` and `diagonal` hold H + S, `if_substitution` is unshifted
                if constexpr (affine_k) {
                    int32_t const open_left = shifted_ ? left : left + open; // H[i][j-1] + open
                    int32_t const open_up = shifted_ ? up : up + open;
                    int32_t const h = opt<minimize_k>(open_left, horizontal[r] + extend);
                    int32_t const v = opt<minimize_k>(open_up, vertical + extend);
                    horizontal[r] = h;
                    vertical = v;
                    if constexpr (local_k) cell = max(max(h, v), max(if_substitution, 0));
                    else cell = opt<minimize_k>(opt<minimize_k>(h, v), if_substitution);
                }
                else {
                    int32_t gap = opt<minimize_k>(up, left); // + g, folded into the stored scores when shifted
                    if constexpr (!shifted_) gap += open;
                    if constexpr (local_k) cell = max(gap, max(if_substitution, 0));
                    else cell = opt<minimize_k>(gap, if_substitution);
                }
                if constexpr (local_k) best[r] = max(best[r], cell);
                stored = cell + shift;
            }
            diagonal = left;
            up = stored;
            scores[r] = stored;
        }
        bottom_score = up;
        bottom_gap = vertical;
    }

    /**
     *  Steps `chunk * G .. + G`. The row above the stripe comes from `in_scores`/`in_gaps` (indexed by 1-based
     *  column, holding H + S) or, when null, from the matrix border. With `collect_` (G = 64 only), lane 63's
     *  bottom row is stored to `out_scores`/`out_gaps` for the next stripe. `steady_` promises that every lane's
     *  column is within 1..m for all steps of the chunk.
     */
    template <bool steady_, bool collect_>
    __device__ __forceinline__ void run_chunk(int32_t chunk, int32_t const *in_scores, int32_t const *in_gaps,
                                              int32_t *out_scores, int32_t *out_gaps) {
        int32_t const first_step = chunk * G;
        int32_t const j_lane = first_step + k + 1; // 1-based column this lane stages for its group
        int32_t chunk_scores, chunk_gaps;
        uint32_t chunk_symbols;
        column_inputs(j_lane, in_scores, in_gaps, chunk_scores, chunk_gaps, chunk_symbols);
        if constexpr (G == 64) {
            *(int4 *)(ring + lane * 16) = make_int4(chunk_scores, chunk_gaps, (int32_t)chunk_symbols, 0);
            int4 inject = *(int4 const *)ring;
#pragma unroll 8
            for (int32_t step = 0; step < 64; ++step) {
                int4 const next = *(int4 const *)(ring + ((step + 1) & 63) * 16); // prefetch the next entry
                int32_t const up_score = shift_down<64>(inject.x, bottom_score);
                int32_t up_gap = 0;
                if constexpr (affine_k) up_gap = shift_down<64>(inject.y, bottom_gap);
                column_symbol = (uint32_t)shift_down<64>(inject.z, (int32_t)column_symbol);
                inject = next;
                int32_t const j = first_step + step + 1 - k;
                if (steady_ || (uint32_t)(j - 1) < (uint32_t)m) update_rows(up_score, up_gap);
                diagonal_top = up_score;
                // Lane 63's bottom row goes to ring entry `step`, read during the previous step: an LDS store
                // under a one-lane mask instead of shift registers through every lane. Ring words are accessed as
                // 32-bit integers throughout, so the compiler sees these stores alias the reads below.
                if (collect_ && lane == 63) {
                    int32_t *entry = (int32_t *)(ring + step * 16);
                    entry[0] = bottom_score;
                    if constexpr (affine_k) entry[1] = bottom_gap;
                }
            }
            // Entry l holds step l's bottom row: column first_step + l - 62.
            if constexpr (collect_) {
                int32_t const j_out = first_step + lane - 62;
                int32_t const *entry = (int32_t const *)(ring + lane * 16);
                if (j_out >= 1 && j_out <= m) {
                    out_scores[j_out] = entry[0];
                    if constexpr (affine_k) out_gaps[j_out] = entry[1];
                }
            }
        }
        else {
#pragma unroll 8
            for (int32_t step = 0; step < G; ++step) {
                int32_t const up_score = shift_down<G>(chunk_scores, bottom_score);
                int32_t up_gap = 0;
                if constexpr (affine_k) up_gap = shift_down<G>(chunk_gaps, bottom_gap);
                column_symbol = (uint32_t)shift_down<G>((int32_t)chunk_symbols, (int32_t)column_symbol);
                chunk_scores = rotate_up<G>(chunk_scores);
                if constexpr (affine_k) chunk_gaps = rotate_up<G>(chunk_gaps);
                chunk_symbols = (uint32_t)rotate_up<G>((int32_t)chunk_symbols);
                int32_t const j = first_step + step + 1 - k;
                if ((uint32_t)(j - 1) < (uint32_t)m) update_rows(up_score, up_gap);
                diagonal_top = up_score;
            }
        }
    }

    /** Folds the stripe into the result: local best of its valid rows, or the final score of a global pair. */
    __device__ __forceinline__ void end_stripe(int32_t base, bool last_stripe) {
        if constexpr (local_k) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (lane_top + r + 1 <= n) running_best = max(running_best, best[r]);
        }
        else if (last_stripe) {
            int32_t const last_row = n - 1 - base; // 0-based row within this stripe
            if (active && last_row / R == k) {
                // A masked sum rather than `if (row == r) value = scores[r]`: the compiler folds that chain into
                // a dynamic index, which demotes the whole row state to scratch memory.
                int32_t const wanted = last_row % R;
                int32_t value = 0;
#pragma unroll
                for (int r = 0; r < R; ++r) value += scores[r] & -(int32_t)(wanted == r);
                int64_t const score = (int64_t)value - domain64(n, m); // may leave int32: H = stored + e(n+m)
                results[output] = score;
                if (mirror != no_mirror_k) results[mirror] = score;
            }
        }
    }

    /** Local alignment: maximum of `running_best` over the group's lanes. */
    __device__ __forceinline__ int32_t partial() const {
        int32_t value = running_best;
        for (int offset = 1; offset < G; offset <<= 1) value = max(value, __shfl_xor(value, offset, 64));
        return value;
    }

    __device__ __forceinline__ void write_result(int32_t value) const {
        if (active && k == 0) {
            results[output] = (int64_t)value;
            if (mirror != no_mirror_k) results[mirror] = (int64_t)value;
        }
    }
};

} // namespace szs_amd
