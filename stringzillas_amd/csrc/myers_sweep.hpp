/**
 *  @file myers_sweep.hpp
 *  @brief Bit-parallel unit-cost Levenshtein (match 0, mismatch 1, gap 1) over bytes: Myers' algorithm in
 *         Hyyrö's block form, laid out on the same wave-systolic sweep as the DP kernels.
 *
 *  The reference routes unit costs to Myers as well (include/stringzillas/similarities/serial.hpp:1657-1865,
 *  2126-2140); results equal the unit-cost DP, which is what the tests pin against the oracle.
 *
 *  Layout: lane k of a group owns one 64-row word of the stripe (rows 64k .. 64k+63, bit 0 = top row; 32-row words
 *  for the half-height shapes) as the vertical-delta vectors Pv/Mv. At step t it processes column j = t + 1 - k and
 *  passes the horizontal delta of its bottom row (+1, 0 or -1: the top bits of its Ph and Mh words) to lane k+1 by DPP;
 *  lane 0 receives +1 from the matrix border or the previous stripe's bottom deltas, coded in 2 bits, from HBM scratch
 *  (one byte per column). The match vector of a column symbol is built from bit-planes: plane b holds bit b of every
 *  row's symbol code, and Eq = AND_b (plane_b XNOR bit b of the column code), one `v_bitop3` per plane and half-word.
 *  For 64-lane groups the symbols are first renumbered to a dense code (0 = absent from the rows) so that only
 *  ceil(log2(distinct + 1)) planes are needed, e.g. 3 for DNA; every group renumbers its own pair (4- and 16-lane
 *  groups compute codes from a presence set and its prefix counts). When every group of the wave has at most 13
 *  distinct symbols, each stripe instead writes the match word of every code to a 7 KB LDS table ([14 codes][lanes]
 *  u64 per group), and a step reads its word back (for 64-lane groups prefetched one step ahead) in place of the plane
 *  arithmetic; 4- and 16-lane groups keep that table from one item to the next while their rows strings stay the same
 *  (the queries of an all-pairs product). For 64-lane groups lane 0's inputs of a 64-step chunk (horizontal deltas
 *  from the stripe above, column codes) are staged in the LDS ring and read back as one broadcast entry per step. The
 *  distance is m + the sum of the vertical deltas of the last column over all rows.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "dp_sweep.hpp"
#include "szs_device.hpp"

namespace szs_amd {

enum : int {
    dpp_quad_shr1_k = 0x90,  /**< quad_perm [0,0,1,2]: lane i <- lane i-1 within each quad (lane 0 <- itself) */
    dpp_quad_rol1_k = 0x39,  /**< quad_perm [1,2,3,0]: lane i <- lane (i+1) mod 4 within each quad */
};

/** Shift toward higher lanes within groups of G; the group's first lane takes `old`. */
template <int G>
__device__ __forceinline__ uint32_t group_shift_down(uint32_t old, uint32_t src, bool first_in_group) {
    if constexpr (G == 4) {
        // Every lane has a source lane in its quad; the first lane's (itself) is replaced by the select, so the move
        // can overwrite `src` in place instead of a copy of `old`.
        uint32_t const shifted = (uint32_t)__builtin_amdgcn_update_dpp((int)src, (int)src, dpp_quad_shr1_k, 0xF, 0xF, false);
        return first_in_group ? old : shifted;
    }
    else return (uint32_t)shift_down<G>((int32_t)old, (int32_t)src);
}

template <int G>
__device__ __forceinline__ uint32_t group_rotate_up(uint32_t src) {
    if constexpr (G == 4) return (uint32_t)__builtin_amdgcn_update_dpp((int)src, (int)src, dpp_quad_rol1_k, 0xF, 0xF, false);
    else return (uint32_t)rotate_up<G>((int32_t)src);
}

/** LDS a smaller group uses to renumber its pair's symbols: codes are computed from these on demand. Entry w covers
 *  byte values 32w .. 32w + 31: x is their presence set in the rows string, y one more than the number of values
 *  present before 32w (so a value's code is y plus the present values below it in x, read with one 64-bit load). */
struct myers_group_tables_t {
    uint2 entries[8];
};

static_assert(sizeof(myers_group_tables_t) * 16 <= myers_tables_bytes_k, "16 four-lane groups per wave");

/** What a smaller group's LDS tables were last built for (kept where 64-lane groups stage their chunk inputs): the
 *  rows string and its distinct byte values, or `myers_eq_codes_k` and above when no match-word table was built. */
struct myers_group_record_t {
    uint64_t rows;
    int32_t n;
    int32_t distinct;
};
static_assert(sizeof(myers_group_record_t) * 16 <= inject_ring_bytes_k, "16 four-lane groups per wave");

/** Global-address-space views of the tapes: loads through them never count against the waits for LDS operations. */
typedef __attribute__((address_space(1))) uint32_t const global_dword_t;
typedef __attribute__((address_space(1))) uint8_t const global_u8_t;

/** LDS a 64-lane Myers group uses to renumber its pair's symbols. */
struct myers_tables_t {
    uint32_t presence[8]; /**< 256-bit set of the byte values in the rows string */
    uint8_t codes[256];   /**< byte -> dense code, 0 for bytes absent from the rows */
};
static_assert(sizeof(myers_tables_t) <= myers_tables_bytes_k, "the match table follows the renumbering tables");

/**
 *  `runes_` scores UTF-8 runes (decoded to u32): each rune is its own code, so nothing is renumbered; the planes
 *  cover the bits of the largest rune in the wave's rows (at most 21), and a column rune with a higher bit set
 *  matches no row.
 */
template <int G, bool runes_ = false, int word_bits_ = 64>
struct myers_sweep_t {
    static_assert(word_bits_ == 64 || word_bits_ == 32, "a lane's word holds 64 or 32 rows");
    using symbol_t = typename std::conditional<runes_, uint32_t, uint8_t>::type;
    /** One lane's rows as bits: 64-row words, or 32-row words (half-height stripes for pairs that would leave most
     *  of a 64-row stripe empty; a step then costs about half). */
    using word_t = typename std::conditional<word_bits_ == 64, uint64_t, uint32_t>::type;
    static constexpr int group_k = G;
    static constexpr int rows_per_lane = word_bits_;
    static constexpr int skew_k = 1; // lane k + 1 runs one column behind lane k
    static constexpr bool local_k = false;
    static constexpr bool renumber_k = !runes_ && G == 64; /**< one pair per wave: a full byte -> code table */
    static constexpr int planes_k = runes_ ? 21 : 8;        /**< bits of a code: a rune has at most 21 */
    static constexpr reduce_kind_t reduce_k = reduce_sum_k;
    /** Groups of several pairs per wave over bytes: a lane's word of row bytes loads once as dwords. A small group's
     *  pair fits one stripe (the ladder gives these shapes no taller pairs), so its rows are read once. */
    static constexpr bool small_bytes_k = !runes_ && G < 64;
    static constexpr int row_dwords_k = word_bits_ / 4;
    /** Groups of several pairs per wave skip the per-step column check in chunks where every group's columns are
     *  all valid (the wave's shortest pair bounds them). */
    static constexpr bool steady_groups_k = G < 64;

    int lane, k;
    bool active;
    int32_t n, m;
    symbol_t const *rows, *columns;
    int64_t *results;
    uint64_t output, mirror;
    myers_tables_t *tables;
    myers_group_tables_t *group_tables; /**< this lane's group (G < 64) */
    int32_t distinct;    /**< distinct byte values in the rows */
    word_t *eq_table;    /**< [14 codes][G lanes] match words of this group's current stripe (table mode) */
    int8_t *ring;        /**< lane-0 inputs of a 64-step chunk (G = 64) */
    int32_t planes_used;
    bool table_mode;     /**< wave-uniform: at most 13 distinct symbols, match words come from `eq_table` */
    bool reused;         /**< wave-uniform, smaller groups over bytes: the LDS tables of the wave's previous item
                              were built for the same rows strings, and stand */

    int32_t lane_top;
    word_t planes[planes_k];
    word_t pv, mv;
    uint32_t column_code;
    uint32_t top_p, top_n; /**< Ph's and Mh's top 32 bits before the shift (bit 31: the bottom row's delta) */
    uint32_t column_at;    /**< 64-lane groups in table mode: LDS address of this lane's match word of `column_code` */
    int32_t vertical_sum;
    uint32_t row_dwords[small_bytes_k ? row_dwords_k : 1]; /**< bytes of this lane's rows, 4 per dword, top row first */
    symbol_t pending; /**< groups of several pairs: this lane's column symbol of the next chunk, loaded a chunk ahead */

    __device__ __forceinline__ uint32_t code_of(uint32_t byte) const {
        if constexpr (runes_) return byte;
        else if constexpr (renumber_k) return tables->codes[byte];
        else {
            if (distinct == 256) return byte; // no absent code left: bytes stay as they are
            uint2 const entry = group_tables->entries[byte >> 5];
            uint32_t const bit = byte & 31u;
            uint32_t const code = __builtin_popcount(entry.x & ((1u << bit) - 1u)) + entry.y;
            return (entry.x >> bit) & 1u ? code : 0u;
        }
    }

    /** Byte `r` of this lane's rows (`row_dwords`). */
    __device__ __forceinline__ uint32_t row_byte(int r) const {
        return __builtin_amdgcn_ubfe(row_dwords[r >> 2], 8 * (r & 3), 8);
    }

    /**
     *  Loads the bytes of rows k·W .. k·W + `valid` - 1 into `row_dwords` with aligned dword loads issued together
     *  (a dword holding one of those bytes lies inside the tape's pages); bytes past `valid` are left unspecified.
     */
    __device__ __forceinline__ void load_row_bytes(int32_t valid) {
        uintptr_t const at = (uintptr_t)(rows + k * word_bits_);
        uint32_t const shift = (uint32_t)at & 3u;
        global_dword_t *aligned = (global_dword_t *)(at - shift);
        uint32_t raw[row_dwords_k + 1];
#pragma unroll
        for (int j = 0; j <= row_dwords_k; ++j) raw[j] = valid > 0 && 4 * j < valid + (int32_t)shift ? aligned[j] : 0u;
#pragma unroll
        for (int i = 0; i < row_dwords_k; ++i) row_dwords[i] = __builtin_amdgcn_alignbyte(raw[i + 1], raw[i], shift);
    }

    /** Symbol `i` of a tape string, loaded through the global address space. */
    __device__ __forceinline__ static symbol_t column_symbol_at(symbol_t const *string, int32_t i) {
        if constexpr (runes_) return *(global_dword_t *)(uintptr_t)(string + i);
        else return *(global_u8_t *)(uintptr_t)(string + i);
    }

    /** This lane's column symbol `j` (0-based), 0 past the columns' end. */
    __device__ __forceinline__ symbol_t column_symbol(int32_t j) const {
        return j < m ? column_symbol_at(columns, j) : (symbol_t)0;
    }

    __device__ __forceinline__ void load(item_args_t const &a, uint32_t task_index, bool is_active, uint32_t wave) {
        myers_tables_t *wave_tables = reinterpret_cast<myers_tables_t *>(wave_lds(wave, a.wave_lds_bytes));
        lane = (int)__lane_id();
        k = lane % G;
        active = is_active;
        task_t const task = a.tasks[task_index];
        segment_view_t const segment = a.segments[task.segment];
        n = active ? (int32_t)task.rows_length : 0;
        m = active ? (int32_t)task.columns_length : 0;
        symbol_t const *query = (symbol_t const *)segment.query_symbols;
        symbol_t const *candidate = (symbol_t const *)segment.candidate_symbols;
        // Uniform costs are symmetric in the operands, so the transposition needs no compensation here.
        rows = (task_transposed(task) ? candidate : query) + task.rows_offset;
        columns = (task_transposed(task) ? query : candidate) + task.columns_offset;
        results = segment.results;
        output = task.output;
        mirror = task.mirror;
        tables = wave_tables;
        uint32_t const group = (uint32_t)lane / G;
        group_tables = reinterpret_cast<myers_group_tables_t *>(wave_lds(wave, a.wave_lds_bytes)) + group;
        eq_table = reinterpret_cast<word_t *>(wave_lds(wave, a.wave_lds_bytes) + myers_tables_bytes_k) +
                   group * (myers_eq_codes_k * G);
        ring = wave_lds(wave, a.wave_lds_bytes) + (a.wave_lds_bytes - inject_ring_bytes_k);
        vertical_sum = 0;
        planes_used = 8;
        table_mode = false;
        reused = false;
        if constexpr (runes_) {
            // Runes are their own codes: plane b holds bit b, up to the highest bit of the wave's largest row rune.
            // Loads go 16 at a time, so a pair's rows cost one memory round trip per 16 G rows.
            uint32_t largest = 0;
            for (int32_t first = k; first < n; first += 16 * G) {
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    int32_t const i = first + q * G;
                    largest = max(largest, i < n ? (uint32_t)*(global_dword_t *)(uintptr_t)(rows + i) : 0u);
                }
            }
            for (int offset = 1; offset < 64; offset <<= 1)
                largest = max(largest, (uint32_t)__shfl_xor((int)largest, offset, 64));
            planes_used = __builtin_amdgcn_readfirstlane(largest ? 32 - __builtin_clz(largest) : 1);
            distinct = 0;
        }
        else if constexpr (renumber_k) {
            // One pair per wave: collect the byte values of the rows, then number them 1..D in byte order.
            if (lane < 8) tables->presence[lane] = 0;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            for (int32_t i = lane; i < n; i += 64) {
                uint32_t const byte = rows[i];
                atomicOr(&tables->presence[byte >> 5], 1u << (byte & 31));
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            uint32_t words[8];
            int32_t distinct = 0;
#pragma unroll
            for (int w = 0; w < 8; ++w) words[w] = tables->presence[w], distinct += __builtin_popcount(words[w]);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                uint32_t const value = (uint32_t)lane * 4 + q;
                uint32_t const word = value >> 5, bit = value & 31;
                int32_t below = 0;
#pragma unroll
                for (int w = 0; w < 8; ++w)
                    if ((uint32_t)w < word) below += __builtin_popcount(words[w]);
                below += __builtin_popcount(words[word] & ((1u << bit) - 1u));
                bool const present = (words[word] >> bit) & 1u;
                // With all 256 values present there is no absent code left: keep bytes as they are.
                uint32_t const code = distinct == 256 ? value : present ? (uint32_t)below + 1 : 0u;
                tables->codes[value] = (uint8_t)code;
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            int32_t const largest = distinct == 256 ? 255 : distinct;
            planes_used = largest ? 32 - __builtin_clz((uint32_t)largest) : 1;
            this->distinct = distinct;
            table_mode = distinct < myers_eq_codes_k && a.wave_lds_bytes >= myers_lds_bytes_k;
        }
        else {
            // Several pairs per wave: each group collects its rows' byte values and keeps the per-word prefix
            // counts; codes 1..D follow byte order, as for the 64-lane groups. Lane k's word holds rows
            // k·W .. k·W + W - 1, and the group's words cover its pair. When the wave's previous item ran this
            // sweep on the same rows strings in table mode (the pairs of an all-pairs product share their query),
            // the tables it left in LDS stand and nothing is rebuilt.
            myers_group_record_t *record = reinterpret_cast<myers_group_record_t *>(ring) + group;
            if (a.same_variant) {
                myers_group_record_t const kept = *record;
                bool const same = kept.rows == (uint64_t)(uintptr_t)rows && kept.n == n &&
                                  kept.distinct < myers_eq_codes_k;
                reused = __ballot(!same) == 0 && a.wave_lds_bytes >= myers_lds_bytes_k;
                if (reused) {
                    distinct = kept.distinct;
                    planes_used = distinct ? 32 - __builtin_clz((uint32_t)distinct) : 1;
                    table_mode = true;
                    return;
                }
            }
            int32_t const valid = min(max(n - k * word_bits_, 0), word_bits_);
            load_row_bytes(valid);
            for (int w = k; w < 8; w += G) group_tables->entries[w] = make_uint2(0u, 0u);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            if (valid > 0) {
                uint32_t const fill = row_byte(0); // rows past the pair's end repeat this lane's first row
#pragma unroll
                for (int r = 0; r < word_bits_; ++r) {
                    uint32_t const byte = r < valid ? row_byte(r) : fill;
                    atomicOr(&group_tables->entries[byte >> 5].x, 1u << (byte & 31u));
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            uint32_t count = 0;
#pragma unroll
            for (int w = 0; w < 8; ++w) {
                uint32_t const word = group_tables->entries[w].x;
                if ((w % G) == k) group_tables->entries[w].y = count + 1u;
                count += __builtin_popcount(word);
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            distinct = (int32_t)count;
            int32_t const largest = distinct == 256 ? 255 : distinct;
            planes_used = largest ? 32 - __builtin_clz((uint32_t)largest) : 1;
            bool const wide = distinct >= myers_eq_codes_k;
            table_mode = __ballot(wide) == 0 && a.wave_lds_bytes >= myers_lds_bytes_k;
            // The tables now stand for these rows (the match words only in table mode).
            if (k == 0 && a.wave_lds_bytes >= myers_lds_bytes_k)
                *record = myers_group_record_t{(uint64_t)(uintptr_t)rows, n, table_mode ? distinct : myers_eq_codes_k};
        }
    }

    __device__ __forceinline__ void begin_stripe(int32_t base) {
        lane_top = base + k * word_bits_;
        pv = ~(word_t)0; // column 0: D[i][0] = i, every vertical delta is +1
        mv = 0;
        top_p = 1u << 31, top_n = 0;
        column_at = eq_address(0);
        column_code = 0;
        if constexpr (G < 64) pending = column_symbol(k); // the first chunk's, loaded while the words are built
        if constexpr (small_bytes_k) {
            // One stripe from row 0 (`base` is 0), rows in `row_dwords`. Rows past the pair's end (bytes left unspecified) only
            // reach words below the pair's last row, which no result reads. The dwords pass through an empty asm so
            // that no byte's lookup address is kept live from the presence loop (32-64 of them would spill).
#pragma unroll
            for (int i = 0; i < row_dwords_k; ++i) asm volatile("" : "+v"(row_dwords[i]));
            if (table_mode && !reused) { // each row sets its bit in the match word of its code
#pragma unroll
                for (int code = 0; code < myers_eq_codes_k; ++code) eq_table[code * G + k] = 0;
#pragma unroll
                for (int r = 0; r < word_bits_; ++r) {
                    word_t *entry = &eq_table[code_of(row_byte(r)) * G + k];
                    if constexpr (word_bits_ == 64) atomicOr((unsigned long long *)entry, 1ull << r);
                    else atomicOr((unsigned int *)entry, 1u << r);
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
            else if (!table_mode) {
#pragma unroll
                for (int b = 0; b < planes_k; ++b) planes[b] = 0;
#pragma unroll
                for (int r = 0; r < word_bits_; ++r) {
                    uint32_t const code = code_of(row_byte(r)); // below 2^planes_used: higher planes stay zero
#pragma unroll
                    for (int b = 0; b < planes_k; ++b) planes[b] |= (word_t)((code >> b) & 1u) << r;
                }
            }
            return;
        }
#pragma unroll
        for (int b = 0; b < planes_k; ++b) planes[b] = 0;
        if constexpr (G == 64) { // one pair per wave: a stripe's setup is spread over all its columns
            for (int32_t r = 0; r < word_bits_; ++r) {
                int32_t const i = lane_top + r; // 0-based row
                uint32_t const code = i < n ? code_of(rows[i]) : 0u;
#pragma unroll
                for (int b = 0; b < planes_k; ++b)
                    if (b < planes_used) planes[b] |= (word_t)((code >> b) & 1u) << r;
            }
        }
        // Groups of several pairs (runes) sweep few columns: 16 rows at a time, so that their loads issue together
        // (one memory round trip per 16 rows), then each plane takes their bits (`planes_used` is wave-uniform).
        else for (int32_t first = 0; first < word_bits_; first += 16) {
            uint32_t codes[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                int32_t const i = lane_top + first + q; // 0-based row
                uint32_t symbol = 0;
                if (i < n) symbol = column_symbol_at(rows, i);
                codes[q] = i < n ? code_of(symbol) : 0u;
            }
#pragma unroll
            for (int b = 0; b < planes_k; ++b) {
                if (b < planes_used) {
                    uint32_t bits = 0;
#pragma unroll
                    for (int q = 0; q < 16; ++q) bits |= ((codes[q] >> b) & 1u) << q;
                    planes[b] |= (word_t)bits << first;
                }
            }
        }
        if (table_mode) { // the match word of every code for this lane's 64 rows
            for (uint32_t code = 0; code < (uint32_t)myers_eq_codes_k; ++code)
                eq_table[code * G + k] = planes_eq<4>(code); // codes below 16 use planes 0..3
        }
    }

    /**
     *  Eq of `code` over this lane's 64 rows: AND over planes 0 .. PB - 1 of (plane XNOR code bit). Any PB of at
     *  least `planes_used` is exact, because the rows' bits from `planes_used` up are zero: codes of bytes never
     *  reach them, and a rune with a bit at PB or above matches no row.
     */
    template <int PB>
    __device__ __forceinline__ word_t planes_eq(uint32_t code) const {
        static_assert(PB <= planes_k, "planes beyond the array");
        // One 32-bit half at a time; the all-ones or all-zeros mask is formed with a subtraction: a sign-extended
        // 1-bit field widened to 64 bits was compiled as zero-extended (the upper half of Eq came out wrong).
        uint32_t eq_low = ~0u, eq_high = ~0u;
        if constexpr (runes_ && PB < 32) { // a rune wider than every plane matches no row
            uint32_t const fits = 0u - (uint32_t)((code >> PB) == 0u);
            eq_low = fits, eq_high = fits;
        }
#pragma unroll
        for (int b = 0; b < PB; ++b) {
            uint32_t const bit = 0u - ((code >> b) & 1u); // all ones where the code has bit b
            // eq & ~(plane ^ bit), one v_bitop3 per half (truth table 0x90 over the operands' bits 4a + 2b + c)
            eq_low = __builtin_amdgcn_bitop3_b32(eq_low, (uint32_t)planes[b], bit, 0x90);
            if constexpr (word_bits_ == 64)
                eq_high = __builtin_amdgcn_bitop3_b32(eq_high, (uint32_t)((uint64_t)planes[b] >> 32), bit, 0x90);
        }
        if constexpr (word_bits_ == 64) return (word_t)((uint64_t)eq_high << 32 | eq_low);
        else return (word_t)eq_low;
    }

    /** A three-input bitwise function per 32-bit half: bit i of the result is bit 4a + 2b + c of `table_`. */
    template <uint32_t table_>
    __device__ __forceinline__ static word_t bitop3(word_t a, word_t b, word_t c) {
        uint32_t const low = __builtin_amdgcn_bitop3_b32((uint32_t)a, (uint32_t)b, (uint32_t)c, table_);
        if constexpr (word_bits_ == 64) {
            uint32_t const high = __builtin_amdgcn_bitop3_b32((uint32_t)((uint64_t)a >> 32), (uint32_t)((uint64_t)b >> 32),
                                                              (uint32_t)((uint64_t)c >> 32), table_);
            return (word_t)((uint64_t)high << 32 | low);
        }
        else return (word_t)low;
    }

    /** The LDS address of this lane's match word of `code` in the table (the low 32 bits of a generic address in
     *  the LDS aperture are the LDS address). */
    __device__ __forceinline__ uint32_t eq_address(uint32_t code) const {
        return (uint32_t)(uintptr_t)(eq_table + code * G + k);
    }

    /** `(x << 1) | (from >> 31)`: a word shifted up one row, its top bit from bit 31 of `from`. */
    __device__ __forceinline__ static word_t shift_in(word_t x, uint32_t from) {
        uint32_t const low = __builtin_amdgcn_alignbit((uint32_t)x, from, 31);
        if constexpr (word_bits_ == 64) {
            uint32_t const high = __builtin_amdgcn_alignbit((uint32_t)((uint64_t)x >> 32), (uint32_t)x, 31);
            return (word_t)((uint64_t)high << 32 | low);
        }
        else return (word_t)low;
    }

    /**
     *  One column of this lane's word (Hyyro's step), with the horizontal delta entering the top row given as bit 31
     *  of two words (`hp`: +1, `hn`: -1; the lane above's `top_p`/`top_n`, or the chunk's inputs or the border for
     *  a group's first lane), and the one leaving the bottom row left the same way in `top_p`/`top_n`: the words
     *  move between lanes as they are and shift straight into Ph/Mh, with no 2-bit code to pack and unpack on every
     *  step. The three-input terms are single `v_bitop3` per half: (x ^ y) | z is table 0xBE, x | ~(y | z) is 0xF1.
     */
    __device__ __forceinline__ void step_word_bits(word_t eq, uint32_t hp, uint32_t hn) {
        word_t const xv = eq | mv;
        eq |= (word_t)(hn >> 31);
        word_t const xh = bitop3<0xBE>((eq & pv) + pv, pv, eq); // (((eq & pv) + pv) ^ pv) | eq
        word_t const ph = bitop3<0xF1>(mv, xh, pv);             // mv | ~(xh | pv)
        word_t const mh = pv & xh;
        top_p = (uint32_t)((uint64_t)ph >> (word_bits_ - 32));
        top_n = (uint32_t)((uint64_t)mh >> (word_bits_ - 32));
        word_t const ph_in = shift_in(ph, hp), mh_in = shift_in(mh, hn);
        pv = bitop3<0xF1>(mh_in, xv, ph_in); // mh | ~(xv | ph)
        mv = ph_in & xv;
    }

    /**
     *  Horizontal deltas are coded as bit 0 = +1, bit 1 = -1, one byte per column in the boundary rows (`in_codes`
     *  and `out_codes` point at byte arrays indexed by column). `steady_` promises every lane a valid column on
     *  every step of the chunk; `collect_` (G = 64) stores lane 63's bottom codes for the next stripe.
     */
    template <bool steady_, bool collect_>
    __device__ __forceinline__ void run_chunk(int32_t chunk, int32_t const *in_codes, int32_t const *,
                                              int32_t *out_codes, int32_t *) {
        // The plane count is a compile-time constant of the step loop (a guard per plane would cost selects):
        // the smallest compiled count covering the wave's `planes_used`.
        if constexpr (runes_) {
            if (planes_used <= 8) chunk_of<steady_, collect_, false, 8>(chunk, in_codes, out_codes);
            else if (planes_used <= 12) chunk_of<steady_, collect_, false, 12>(chunk, in_codes, out_codes);
            else if (planes_used <= 17) chunk_of<steady_, collect_, false, 17>(chunk, in_codes, out_codes);
            else chunk_of<steady_, collect_, false, 21>(chunk, in_codes, out_codes);
        }
        else if (table_mode) chunk_of<steady_, collect_, true, 4>(chunk, in_codes, out_codes);
        else if (planes_used <= 4) chunk_of<steady_, collect_, false, 4>(chunk, in_codes, out_codes);
        else chunk_of<steady_, collect_, false, 8>(chunk, in_codes, out_codes);
    }

    template <bool steady_, bool collect_, bool table_, int PB>
    __device__ __forceinline__ void chunk_of(int32_t chunk, int32_t const *in_codes, int32_t *out_codes) {
        if constexpr (G == 64) chunk_64<steady_, collect_, table_, PB>(chunk, in_codes, out_codes);
        else chunk_small<steady_, table_, PB>(chunk);
    }

    template <bool steady_, bool collect_, bool table_, int PB>
    __device__ __forceinline__ void chunk_64(int32_t chunk, int32_t const *in_codes, int32_t *out_codes) {
        int32_t const first_step = chunk * 64;
        int32_t const j_lane = first_step + lane + 1;
        uint32_t const symbol = j_lane <= m ? code_of(columns[j_lane - 1]) : 0u;
        uint32_t const entering = in_codes ? (j_lane <= m ? (uint32_t)((uint8_t const *)in_codes)[j_lane] : 0u) : 1u;
        // Ring entry l: lane 0's entering delta at step l as the words of `step_word_bits`, and its column code (in
        // table mode the LDS address of lane 0's match word of it: lane k's is lane k - 1's plus one word).
        constexpr uint32_t word_bytes = sizeof(word_t);
        *(uint4 *)(ring + lane * 16) = make_uint4(entering << 31, (entering >> 1) << 31,
                                                  table_ ? eq_address(symbol) - word_bytes * lane : symbol, 0u);
        // Column codes do not depend on the scores, so step s + 1's code and match word are fetched during step s.
        uint4 now = *(uint4 const *)ring, after = *(uint4 const *)(ring + 16);
        typedef __attribute__((address_space(3))) word_t const lds_word_t;
        uint32_t code = 0, at = 0;
        word_t eq_next = 0;
        if constexpr (table_) {
            at = lane_above_plus(column_at, word_bytes, now.z);
            eq_next = *(lds_word_t *)(uintptr_t)at;
        }
        else code = (uint32_t)shift_down<64>((int32_t)now.z, (int32_t)column_code);
#pragma unroll 8
        for (int32_t step = 0; step < 64; ++step) {
            uint4 const later = *(uint4 const *)(ring + ((step + 2) & 63) * 16);
            uint32_t const hp = (uint32_t)shift_down<64>((int32_t)now.x, (int32_t)top_p);
            uint32_t const hn = (uint32_t)shift_down<64>((int32_t)now.y, (int32_t)top_n);
            uint32_t code_after = 0, at_after = 0;
            word_t const eq_now = eq_next;
            if constexpr (table_) {
                at_after = lane_above_plus(at, word_bytes, after.z);
                eq_next = *(lds_word_t *)(uintptr_t)at_after;
                column_at = at;
            }
            else {
                code_after = (uint32_t)shift_down<64>((int32_t)after.z, (int32_t)code);
                column_code = code;
            }
            int32_t const j = first_step + step + 1 - lane;
            if (steady_ || (uint32_t)(j - 1) < (uint32_t)m) step_word_bits(table_ ? eq_now : planes_eq<PB>(code), hp, hn);
            // Lane 63's bottom words go to ring entry `step`, read two steps earlier: LDS stores under a one-lane
            // mask instead of a shift register through every lane.
            if (collect_ && lane == 63) *(uint2 *)(ring + step * 16) = make_uint2(top_p, top_n);
            code = code_after, at = at_after;
            now = after;
            after = later;
        }
        if constexpr (collect_) { // entry l holds step l's bottom words: column first_step + l - 62
            int32_t const j_out = first_step + lane - 62;
            uint2 const bottom = *(uint2 const *)(ring + lane * 16);
            if (j_out >= 1 && j_out <= m) ((uint8_t *)out_codes)[j_out] = (uint8_t)(bottom.x >> 31 | (bottom.y >> 31) << 1);
        }
    }

    /**
     *  A chunk of G steps of a group of several pairs per wave. Such a group sweeps its pair in one stripe, so lane 0
     *  always enters from the matrix border (+1). Lane k's symbol of the chunk (column first + k + 1) rotates toward
     *  lane 0, which takes one per step; the next chunk's is loaded now, a chunk before it is needed.
     */
    template <bool steady_, bool table_, int PB>
    __device__ __forceinline__ void chunk_small(int32_t chunk) {
        int32_t const first_step = chunk * G;
        // Past the columns' end the symbol is 0; those steps are masked below (and never occur in steady chunks).
        uint32_t chunk_symbols = code_of(pending);
        pending = column_symbol(first_step + G + k);
        bool const first_in_group = k == 0;
#pragma unroll 4
        for (int32_t step = 0; step < G; ++step) {
            // The delta entering the top row: the lane above's bottom-row Ph/Mh bits, or +1 from the border.
            uint32_t const hp = group_shift_down<G>(1u << 31, top_p, first_in_group);
            uint32_t const hn = group_shift_down<G>(0u, top_n, first_in_group);
            column_code = group_shift_down<G>(chunk_symbols, column_code, first_in_group);
            chunk_symbols = group_rotate_up<G>(chunk_symbols);
            int32_t const j = first_step + step + 1 - k;
            if (steady_ || (uint32_t)(j - 1) < (uint32_t)m)
                step_word_bits(table_ ? eq_table[column_code * G + k] : planes_eq<PB>(column_code), hp, hn);
        }
    }

    __device__ __forceinline__ void end_stripe(int32_t, bool) {
        int32_t const rows_here = min(max(n - lane_top, 0), word_bits_);
        word_t const mask = rows_here == word_bits_ ? ~(word_t)0 : (word_t)(((word_t)1 << rows_here) - 1u);
        vertical_sum += __builtin_popcountll(pv & mask) - __builtin_popcountll(mv & mask);
    }

    __device__ __forceinline__ int32_t partial() const {
        int32_t value = vertical_sum;
        for (int offset = 1; offset < G; offset <<= 1) value += __shfl_xor(value, offset, 64);
        return value;
    }

    __device__ __forceinline__ void write_result(int32_t sum) const {
        if (active && k == 0) {
            int64_t const distance = (int64_t)m + sum;
            results[output] = distance;
            if (mirror != no_mirror_k) results[mirror] = distance;
        }
    }
};

} // namespace szs_amd
