/**
 *  @file packed_sweep.hpp
 *  @brief Needleman-Wunsch and Smith-Waterman sweeps with two 16-bit cells per 32-bit register (packed integer
 *  VALU).
 *
 *  The global recurrences are those of include/stringzillas/similarities/serial.hpp:669-749 (linear) and 854-964
 *  (affine, Gotoh), in the anti-diagonal domain of dp_sweep.hpp: Ĉ = H - e*(i+j), ĥ and v̂ likewise plus (e - o),
 *  so that
 *      ĥ = max(Ĉ_left, ĥ_left),  v̂ = max(Ĉ_up, v̂_up),  Ĉ = max(ĥ, v̂, Ĉ_diag + s - o - e) + (o - e)
 *      linear: Ĉ = max(Ĉ_up, Ĉ_left, Ĉ_diag + s - 2g).
 *  Local alignment (serial.hpp:762-841 linear, 977-1079 affine; best over all cells from 0) keeps the int32
 *  sweep's `H + o` form instead, whose zero floor is one constant per lane and chunk: X = H + o and
 *      linear: X = max(X_up, X_left, max(X_diag + s - o, 0)) + o,
 *      affine: E = max(X_left, E_left + e), F = max(X_up, F_up + e), X = max(E, F, max(X_diag + s - o, 0)) + o,
 *  the gap terms read the stored values as they are. Per two cells: the cost `v_perm_b32`, one `v_pk_add_u16`
 *  (a second one when the profile bytes hold `cost - min(o, s_min)` below o, so that they read as unsigned), one
 *  `v_pk_max_u16` (substitution against the floor; all off the row-to-row chain), the three-way maximum, the `+ o`
 *  add and half a `v_pk_max3` into the lane's running best (two registers per maximum), which is folded into an
 *  int32 best at the end of every 64-step chunk (the base changes between chunks); affine gaps add an add and a
 *  maximum per track.
 *
 *  Packing. The rows of one lane form a serial chain (each row needs the row above at the same column), so two
 *  cells of one register must come from independent chains: lane k holds R rows, and its register p carries row
 *  p of the lane at column c in the low half and row p + R/2 at column c - 1 in the high half. The lower half
 *  runs one column behind the upper one, so its inputs are the upper half's last row from the previous step; a
 *  lane's bottom row (high half of the last register) reaches lane k + 1 one column late, hence lane k + 1 runs two
 *  columns behind lane k (lane k is at column t + 1 - 2k at step t) and a stripe needs m + 127 steps.
 *  Per two cells the affine step costs one `v_perm_b32` (both halves' costs from the int8 LDS profile: this
 *  column's byte for the upper row, the previous column's for the lower row), one `v_pk_add_u16` (diagonal +
 *  cost), two `v_pk_max_u16` (ĥ, v̂), one `v_pk_maximum3_f16` (the three-way maximum, see below) and one
 *  `v_pk_add_u16` (+ o - e); linear two fewer. Per step add three DPP moves, two `v_alignbit_b32` that assemble
 *  the packed top inputs (the lane above's bottom row in the low half, this lane's upper-half bottom row in the
 *  high half) and two `v_pk_add_u16` that move the lane above's values into this lane's base.
 *
 *  Value range. Cells are stored as u16 = value - base + 0x3E00, where `base` is an int32 value of each lane,
 *  chosen at the start of every 64-step chunk (the lane's own first row at its current column) and every register
 *  is moved to the new base with one packed add. Values then lie in [0, 0x7BFF]: the bit patterns of the finite,
 *  non-negative halves, whose float order is their integer order, so the float `maximum3` (gfx950 has no 3-input
 *  integer packed maximum) is exact on them; kernels run with half denormals preserved. The engine admits this form
 *  only when every value a lane holds during a chunk provably stays within 15871 of its base: adjacent cells differ
 *  by at most delta = max(|o|, max(0, s_max - min(o, e)) + max(0, o - e)) + |e| in this domain, the cells a lane
 *  holds during a chunk (its R rows and the row above, 66 columns) lie within a Manhattan distance of R + 70 of its
 *  base cell, and gap tracks and intermediate sums stay within a few more deltas of their cell (runtime.hip,
 *  `packed_rows_limit`). Otherwise pairs run the int32 sweep. A lane whose upper half has passed column m holds no
 *  cell that is kept and keeps its base; the row above the stripe repeats column m's values beyond it; cells beyond
 *  column m or row n only feed cells that are discarded, and the two halves of a register never mix. Profile
 *  bytes hold `cost - (o + e)` (global) or `cost - min(o, s_min)` (local), which the engine requires in [0, 127],
 *  so they read as unsigned bytes.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dp_sweep.hpp"
#include "szs_device.hpp"

namespace szs_amd {

typedef unsigned short u16x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pk_max_u16(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t,
                              __builtin_elementwise_max(__builtin_bit_cast(u16x2_t, a), __builtin_bit_cast(u16x2_t, b)));
}

__device__ __forceinline__ uint32_t pk_add_u16(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2_t, a) + __builtin_bit_cast(u16x2_t, b));
}

/** The u16 that stands for int32 `value - base`: centred in [0, 0x7BFF], the finite non-negative halves. */
constexpr int32_t packed_centre_k = 0x3E00;

typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));

/** Packed maximum of three, exact on u16 values in [0, 0x7BFF] read as halves (one `v_pk_maximum3_f16`). */
__device__ __forceinline__ uint32_t pk_max3_ordered(uint32_t a, uint32_t b, uint32_t c) {
    f16x2_t const x = __builtin_bit_cast(f16x2_t, a), y = __builtin_bit_cast(f16x2_t, b);
    f16x2_t const z = __builtin_bit_cast(f16x2_t, c);
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_maximum(__builtin_elementwise_maximum(x, y), z));
}

/** `biased_` (local families only): the profile bias lies below o, so each substitution term takes o - bias off. */
template <int family_, int R, bool uniform_base_ = false, bool biased_ = false>
struct packed_sweep_t {
    static_assert(family_ == family_nw_linear_k || family_ == family_nw_affine_k || family_ == family_sw_linear_k ||
                      family_ == family_sw_affine_k,
                  "class-cost families only");
    static_assert(R % 2 == 0 && R >= 2 && R <= 8, "two half-stripes of R/2 rows; profiles hold at most 8 rows");
    static constexpr int P = R / 2;               // packed registers per lane
    static constexpr int B = (P + 1) / 2;         // local best registers per lane: one per two score registers
    static constexpr int RP = (R + 3) / 4 * 4;    // profile bytes per lane and class (the int32 sweep's layout)
    static constexpr bool affine_k = family_ == family_nw_affine_k || family_ == family_sw_affine_k;
    static constexpr bool local_k = family_ == family_sw_linear_k || family_ == family_sw_affine_k;
    static constexpr int group_k = 64;
    static constexpr int rows_per_lane = R;
    static constexpr int skew_k = 2;              // lane k + 1 runs two columns behind lane k
    static constexpr reduce_kind_t reduce_k = local_k ? reduce_max_k : reduce_none_k;

    // The pair.
    int lane;
    int32_t n, m;
    int32_t open, extend, dense_count, shift; // shift: the profile bias (`packed_bias`)
    bool transposed;
    engine_params_t const *params;
    uint8_t const *rows;
    uint8_t const *columns;
    int64_t *results;
    uint64_t output, mirror;
    int8_t *profile; /**< this wave's LDS profile: [dense class][lane][RP] bytes of `cost - (o + e)` */
    int8_t *ring;    /**< this wave's injection ring: 64 entries of 16 bytes */

    // The stripe.
    int32_t lane_top;
    int32_t base;                     /**< this lane's: the int32 value that the u16 `packed_centre_k` stands for */
    uint32_t scores[P], horizontal[P];
    uint32_t diagonal;                /**< the packed top input of the previous step: this step's diagonal */
    uint32_t vertical;                /**< vertical gaps below the last register after the previous step */
    uint32_t symbol_now, symbol_low;  /**< dense classes of this step's upper and lower columns */
    uint32_t costs_now[RP / 4], costs_low[RP / 4];
    bool result_upper;                /**< the pair's last row is in this lane's upper half */
    int32_t result_register;
    int32_t captured;                 /**< the last row's score at column m, when it sits in the upper half */
    // Local alignment.
    uint32_t best[B];                 /**< per half of two registers: the chunk's largest cell (u16 of H in its base) */
    int32_t running_best;             /**< the lane's best H over the chunks folded so far (0 from the start) */

    __device__ __forceinline__ void load(item_args_t const &a, uint32_t task_index, bool, uint32_t wave) {
        int8_t *wave_profile = wave_lds(wave, a.wave_lds_bytes);
        lane = (int)__lane_id();
        task_t const task = a.tasks[task_index];
        segment_view_t const segment = a.segments[task.segment];
        params = segment.params;
        n = (int32_t)task.rows_length;
        m = (int32_t)task.columns_length;
        open = params->open, extend = params->extend;
        dense_count = params->dense_count;
        shift = params->packed_bias;
        transposed = task_transposed(task);
        uint8_t const *query = (uint8_t const *)segment.query_symbols;
        uint8_t const *candidate = (uint8_t const *)segment.candidate_symbols;
        rows = (transposed ? candidate : query) + task.rows_offset;
        columns = (transposed ? query : candidate) + task.columns_offset;
        results = segment.results;
        output = task.output;
        mirror = task.mirror;
        profile = wave_profile;
        ring = wave_profile + (a.wave_lds_bytes - inject_ring_bytes_k);
        captured = 0;
        running_best = 0; // serial.hpp:781 starts the local best at 0
    }

    /** The stored int32 form minus H at cell (i, j): -e*(i+j) modulo 2^32 (linear: e = g; see `wrap_add`), or the
     *  constant o of the local form. */
    __device__ __forceinline__ int32_t domain(int32_t i, int32_t j) const {
        if constexpr (local_k) return open;
        else return wrap_mul(-extend, wrap_add(i, j));
    }
    /** The same for the gap tracks: the local form keeps them as they are. */
    __device__ __forceinline__ int32_t gap_domain(int32_t i, int32_t j) const {
        if constexpr (local_k) return 0;
        else return wrap_add(domain(i, j), extend - open);
    }

    __device__ __forceinline__ uint32_t to_u16(int32_t value) const {
        return (uint32_t)(value - base + packed_centre_k) & 0xFFFFu;
    }
    __device__ __forceinline__ int32_t from_u16(uint32_t half) const {
        return (int32_t)(half & 0xFFFFu) - packed_centre_k + base;
    }

    /** Column-0 values of any stripe row (constant in the anti-diagonal domain), as stored int32 values. */
    __device__ __forceinline__ int32_t border_cell() const {
        return wrap_add(border_score<family_>(1, open, extend), domain(1, 0));
    }
    __device__ __forceinline__ int32_t border_gap() const {
        return wrap_add(wrap_add(open + extend, border_score<family_>(1, open, extend)), gap_domain(1, 0));
    }

    __device__ __forceinline__ void read_costs(uint32_t symbol, uint32_t (&out)[RP / 4]) const {
        uint32_t const *row_costs = (uint32_t const *)(profile + (symbol * 64 + (uint32_t)lane) * RP);
#pragma unroll
        for (int w = 0; w < RP / 4; ++w) out[w] = row_costs[w];
    }

    /** Loads rows `base_row + lane*R .. + R`, builds the profile and resets the stripe to the column-0 border. */
    __device__ __forceinline__ void begin_stripe(int32_t base_row) {
        lane_top = base_row + lane * R;
        uint32_t row_classes[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            int32_t const i = lane_top + r + 1; // 1-based matrix row
            row_classes[r] = i <= n ? params->byte_to_dense[rows[i - 1]] : 0u;
        }
        int8_t *mine = profile + lane * RP;
        for (int32_t c = 0; c < dense_count; ++c) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                int8_t const cost = transposed ? params->dense_costs[c * max_dense_classes_k + row_classes[r]]
                                               : params->dense_costs[row_classes[r] * max_dense_classes_k + c];
                // In [0, 127]: checked when the engine was made. Local alignment gives rows past n the lowest cost,
                // min(o, s_min) <= 0: every value such a row takes is then the floor or a cell of the pair plus
                // steps <= 0 (gap costs are <= 0 too), never above the pair's best, so the best needs no row mask.
                bool const past = local_k && lane_top + r + 1 > n;
                mine[c * 64 * RP + r] = past ? (int8_t)0 : (int8_t)(cost - shift);
            }
        }
        base = border_cell();
        uint32_t const cell = to_u16(border_cell()) * 0x10001u;
        uint32_t const gap = to_u16(border_gap()) * 0x10001u;
#pragma unroll
        for (int p = 0; p < P; ++p) scores[p] = cell, horizontal[p] = gap;
        // Lane 0's first diagonal is the cell above the stripe at column 0 (the origin for the first stripe).
        int32_t const above = wrap_add(border_score<family_>(lane_top, open, extend), domain(lane_top, 0));
        diagonal = to_u16(above) | (cell & 0xFFFF0000u);
        vertical = gap;
        symbol_now = symbol_low = 0;
#pragma unroll
        for (int w = 0; w < RP / 4; ++w) costs_now[w] = costs_low[w] = 0;
        int32_t const last_row = n - 1 - base_row; // beyond this stripe unless it is the last one
        result_upper = last_row >= 0 && last_row / R == lane && last_row % R < P;
        result_register = (last_row % R) % P;
        if constexpr (local_k) {
#pragma unroll
            for (int b = 0; b < B; ++b) best[b] = 0;
        }
    }

    /** Lane-0 inputs of column `j`: the row above the stripe (stored int32 form), its vertical gap, its symbol.
     *  Boundary rows hold, per column, the u16 score (low) and vertical gap (high) in the base of the chunk that
     *  wrote them (`in_scores`, u32 per column), and that chunk's base (`in_bases`, int32 per chunk): column j comes
     *  from chunk (j + 126) / 64, whose lane 63 held it. */
    __device__ __forceinline__ void column_inputs(int32_t j, int32_t const *in_scores, int32_t const *in_bases,
                                                  int32_t &score, int32_t &gap, uint32_t &symbol) const {
        symbol = j <= m ? (uint32_t)params->byte_to_dense[columns[j - 1]] : 0u;
        // Beyond column m the row above repeats column m's values: near every lane's base, and they only feed
        // cells that are discarded.
        int32_t const at = min(j, m);
        if (in_scores) {
            uint32_t const both = (uint32_t)in_scores[at];
            int32_t const written_base = in_bases[(at + 126) >> 6] - packed_centre_k;
            score = (int32_t)(both & 0xFFFFu) + written_base;
            gap = affine_k ? (int32_t)(both >> 16) + written_base : 0;
        }
        else {
            int32_t const border = border_score<family_>(at, open, extend);
            score = wrap_add(border, domain(0, at));
            gap = wrap_add(wrap_add(open + extend, border), gap_domain(0, at)); // the vertical track's discard value
        }
    }

    /** Moves every held value to a new base (exact modulo 2^16; the range bound keeps the result in range). */
    __device__ __forceinline__ void rebase(int32_t new_base) {
        uint32_t const delta = ((uint32_t)(base - new_base) & 0xFFFFu) * 0x10001u;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            scores[p] = pk_add_u16(scores[p], delta);
            if constexpr (affine_k) horizontal[p] = pk_add_u16(horizontal[p], delta);
        }
        diagonal = pk_add_u16(diagonal, delta);
        if constexpr (affine_k) vertical = pk_add_u16(vertical, delta);
        base = new_base;
    }

    /** One step of both half-stripes: `up` and `vertical_in` are the packed top inputs. Local alignment: `zero_floor`
     *  is the packed H = 0, `keep` the halves whose columns lie within the pair (their cells count for the best). */
    __device__ __forceinline__ void update_rows(uint32_t up, uint32_t vertical_in, uint32_t kappa, uint32_t zero_floor,
                                                uint32_t unbias, uint32_t extension, uint32_t keep) {
        uint32_t diag = diagonal;
        uint32_t vert = vertical_in;
        uint32_t held = 0; // local alignment: the previous register's cell, paired with this one's into the best
#pragma unroll
        for (int p = 0; p < P; ++p) {
            // Upper row p reads this column's profile byte p, lower row P + p the previous column's byte P + p;
            // selector bytes 0x0C give zeros, so both bytes widen to u16 unsigned.
            uint32_t const upper_byte = p % 4, lower_byte = (P + p) % 4; // constants once unrolled
            uint32_t const selector = upper_byte | 0x0C00u | ((4u + lower_byte) << 16) | 0x0C000000u;
            uint32_t const cost = __builtin_amdgcn_perm(costs_low[(P + p) / 4], costs_now[p / 4], selector);
            uint32_t const left = scores[p];
            uint32_t const if_substitution = pk_add_u16(diag, cost);
            // The chain from one row to the next is v̂, the three-way maximum and + (o - e) (affine) or the maximum
            // alone (linear), so the scheduler interleaves the rows instead of padding back-to-back dependent
            // packed ops with wait states.
            uint32_t stored;
            if constexpr (local_k) {
                // - (o - packed_bias) when the profile is biased below o
                uint32_t const substitution = biased_ ? pk_add_u16(if_substitution, unbias) : if_substitution;
                uint32_t const floored = pk_max_u16(substitution, zero_floor);
                uint32_t cell;
                if constexpr (affine_k) { // E = max(X_left, E_left + e), F = max(X_up, F_up + e)
                    uint32_t const h = pk_max_u16(left, pk_add_u16(horizontal[p], extension));
                    uint32_t const v = pk_max_u16(up, pk_add_u16(vert, extension));
                    horizontal[p] = h;
                    vert = v;
                    cell = pk_max3_ordered(h, v, floored);
                }
                else cell = pk_max3_ordered(up, left, floored);
                uint32_t const counted = cell & keep; // all ones on steady steps
                if (p % 2 == 1) best[p / 2] = pk_max3_ordered(best[p / 2], held, counted);
                else if (p == P - 1) best[p / 2] = pk_max_u16(best[p / 2], counted);
                held = counted;
                stored = pk_add_u16(cell, kappa);
            }
            else if constexpr (affine_k) {
                uint32_t const h = pk_max_u16(left, horizontal[p]);
                uint32_t const v = pk_max_u16(up, vert);
                horizontal[p] = h;
                vert = v;
                stored = pk_add_u16(pk_max3_ordered(h, if_substitution, v), kappa);
            }
            else stored = pk_max3_ordered(up, left, if_substitution);
            diag = left;
            up = stored;
            scores[p] = stored;
        }
        if constexpr (affine_k) vertical = vert;
    }

    /** The masked sum reads register `index` without a dynamic index (which would demote the array to scratch). */
    __device__ __forceinline__ uint32_t register_at(int32_t index) const {
        uint32_t value = 0;
#pragma unroll
        for (int p = 0; p < P; ++p) value += scores[p] & -(uint32_t)(index == p);
        return value;
    }

    /**
     *  Steps `chunk * 64 .. + 64`. The row above the stripe comes from `in_scores`/`in_gaps` (1-based columns, the
     *  packed form of `column_inputs`) or the matrix border; with `collect_`, lane 63's bottom row goes to
     *  `out_scores` (u16 score and gap per column) and its base to `out_gaps[chunk]`. `steady_` promises both halves
     *  of every lane a column within 1..m - 1 on every step.
     */
    template <bool steady_, bool collect_>
    __device__ __forceinline__ void run_chunk(int32_t chunk, int32_t const *in_scores, int32_t const *in_gaps,
                                              int32_t *out_scores, int32_t *out_gaps) {
        int32_t const first_step = chunk * 64;
        uint32_t into_mine = 0;
        if constexpr (uniform_base_) {
            // One base for the wave: lane 32's first row at its current column, the middle of what the wave holds.
            rebase(__builtin_amdgcn_readlane(from_u16(scores[0]), 32));
        }
        else {
            // Each lane's base for the chunk: its first row at its current column, a cell every value the lane
            // holds during the chunk is close to. A lane past column m keeps its base (it holds no kept cell).
            int32_t const current = first_step - 2 * lane;
            rebase(current <= m ? from_u16(scores[0]) : base);
            // The lane above's values arrive in its base: one packed add of (base above - base, 0) moves them into
            // this lane's (lane 0 reads the ring, staged in its own base).
            int32_t const base_above = shift_down<64>(base, base);
            into_mine = (uint32_t)(base_above - base) & 0xFFFFu;
        }
        int32_t const ring_base = __builtin_amdgcn_readfirstlane(base);

        int32_t const j_lane = first_step + lane + 1;
        int32_t chunk_score, chunk_gap;
        uint32_t chunk_symbol;
        column_inputs(j_lane, in_scores, in_gaps, chunk_score, chunk_gap, chunk_symbol);
        *(int4 *)(ring + lane * 16) =
            make_int4((int32_t)(((uint32_t)(chunk_score - ring_base + packed_centre_k) & 0xFFFFu) << 16),
                      (int32_t)(((uint32_t)(chunk_gap - ring_base + packed_centre_k) & 0xFFFFu) << 16),
                      (int32_t)chunk_symbol, 0);
        int4 inject = *(int4 const *)ring;
        // Lane 0's column symbol comes from the ring; the other lanes fetched theirs during the previous step.
        if (lane == 0) symbol_now = (uint32_t)inject.z;
        read_costs(symbol_now, costs_now);
        // The add after the maximum: o - e (affine global), o (local: stores H + o).
        uint32_t const kappa = ((uint32_t)(local_k ? open : open - extend) & 0xFFFFu) * 0x10001u;
        uint32_t const extension = ((uint32_t)extend & 0xFFFFu) * 0x10001u; // local affine gap tracks
        // The packed H = 0 of the local form in this chunk's base; below every value the lane holds (and so clamped
        // to 0) once the base is far above zero.
        uint32_t zero_floor = 0;
        if constexpr (local_k) zero_floor = (uint32_t)max(0, packed_centre_k - base) * 0x10001u;
        uint32_t const unbias = ((uint32_t)(shift - open) & 0xFFFFu) * 0x10001u;
        uint32_t const border_high = to_u16(border_cell()) << 16, border_gap_high = to_u16(border_gap()) << 16;

#pragma unroll 8
        for (int32_t step = 0; step < 64; ++step) {
            int4 const next = *(int4 const *)(ring + ((step + 1) & 63) * 16);
            // Packed top inputs: the lane above's bottom row (its high half) low, this lane's upper-half bottom row
            // from the previous step high.
            uint32_t const from_above = (uint32_t)shift_down<64>(inject.x, (int32_t)scores[P - 1]);
            uint32_t up = __builtin_amdgcn_alignbit(scores[P - 1], from_above, 16);
            if constexpr (!uniform_base_) up = pk_add_u16(up, into_mine);
            uint32_t vertical_in = 0;
            if constexpr (affine_k) {
                uint32_t const gaps_above = (uint32_t)shift_down<64>(inject.y, (int32_t)vertical);
                vertical_in = __builtin_amdgcn_alignbit(vertical, gaps_above, 16);
                if constexpr (!uniform_base_) vertical_in = pk_add_u16(vertical_in, into_mine);
            }
            // Next step's upper column is the lane above's current lower column: fetch its class and costs now.
            uint32_t const symbol_next = (uint32_t)shift_down<64>(next.z, (int32_t)symbol_low);
            uint32_t costs_next[RP / 4];
            read_costs(symbol_next, costs_next);
            int32_t const c = first_step + step + 1 - 2 * lane; // upper column; the lower half is at c - 1
            if (steady_ || (uint32_t)(c - 1) <= (uint32_t)m) {
                // Upper half at column c, lower half at c - 1: both within 1..m on steady steps.
                uint32_t const keep = steady_ ? ~0u : (c <= m ? 0xFFFFu : 0u) | (c >= 2 ? 0xFFFF0000u : 0u);
                update_rows(up, vertical_in, kappa, zero_floor, unbias, extension, keep);
                if (!steady_) {
                    if (c == 1) { // the lower half stays on the column-0 border
#pragma unroll
                        for (int p = 0; p < P; ++p) {
                            scores[p] = (scores[p] & 0xFFFFu) | border_high;
                            if constexpr (affine_k) horizontal[p] = (horizontal[p] & 0xFFFFu) | border_gap_high;
                        }
                    }
                    // The upper half moves on to column m + 1 next step: keep the last row's score at column m.
                    if (!local_k && result_upper && c == m) captured = from_u16(register_at(result_register));
                }
            }
            diagonal = up;
            symbol_low = symbol_now;
            symbol_now = symbol_next;
#pragma unroll
            for (int w = 0; w < RP / 4; ++w) costs_low[w] = costs_now[w], costs_now[w] = costs_next[w];
            inject = next;
            // Lane 63's bottom row (high halves: score low, vertical gap high) goes to the spare word of ring entry
            // `step`, which was read during the previous step: one LDS store under a one-lane mask instead of shift
            // registers through every lane. Every ring access is a 32-bit integer one, so they all alias.
            if (collect_ && lane == 63)
                *(uint32_t *)(ring + step * 16 + 12) =
                    affine_k ? __builtin_amdgcn_perm(vertical, scores[P - 1], 0x07060302u) : scores[P - 1] >> 16;
        }
        // Entry l holds step l's bottom row (lane 63's, in its base): column first_step + l - 126. It goes out as it
        // is, with that base once per chunk (`column_inputs` reads both).
        if constexpr (collect_) {
            int32_t const j_out = first_step + lane - 126;
            uint32_t const bottom = *(uint32_t const *)(ring + lane * 16 + 12);
            if (j_out >= 1 && j_out <= m) out_scores[j_out] = (int32_t)bottom;
            if (lane == 0) out_gaps[chunk] = __builtin_amdgcn_readlane(base, 63);
        }
        if constexpr (local_k) fold_best();
    }

    /** Local alignment: moves the chunk's best cells (u16 of H in this chunk's base) into the int32 best, and clears
     *  them for the next chunk. A half that took no cell in the chunk holds 0. */
    __device__ __forceinline__ void fold_best() {
#pragma unroll
        for (int b = 0; b < B; ++b) {
            uint32_t const low = best[b] & 0xFFFFu, high = best[b] >> 16;
            if (low) running_best = max(running_best, (int32_t)low - packed_centre_k + base);
            if (high) running_best = max(running_best, (int32_t)high - packed_centre_k + base);
            best[b] = 0;
        }
    }

    /** Writes the pair's score from the lane holding its last row, on the last stripe (global families; the local
     *  best is folded chunk by chunk and reduced by `partial`). */
    __device__ __forceinline__ void end_stripe(int32_t base_row, bool last_stripe) {
        if (local_k || !last_stripe) return;
        int32_t const last_row = n - 1 - base_row;
        if (last_row / R != lane) return;
        int32_t const value = result_upper ? captured : from_u16(register_at(result_register) >> 16);
        int64_t const score = (int64_t)value + (int64_t)extend * ((int64_t)n + m); // may leave int32
        results[output] = score;
        if (mirror != no_mirror_k) results[mirror] = score;
    }

    /** Local alignment: the largest `running_best` of the wave's lanes. */
    __device__ __forceinline__ int32_t partial() const {
        if constexpr (!local_k) return 0;
        int32_t value = running_best;
        for (int offset = 1; offset < 64; offset <<= 1) value = max(value, __shfl_xor(value, offset, 64));
        return value;
    }

    __device__ __forceinline__ void write_result(int32_t value) const {
        if constexpr (local_k) {
            if (lane == 0) {
                results[output] = (int64_t)value;
                if (mirror != no_mirror_k) results[mirror] = (int64_t)value;
            }
        }
    }
};

} // namespace szs_amd
