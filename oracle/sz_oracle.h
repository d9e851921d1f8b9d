/**
 *  @file oracle/sz_oracle.h
 *  @brief TEST INFRASTRUCTURE ONLY: the CPU oracle for the similarity scores and the rolling fingerprints.
 *
 *  This is a plain-C restatement of the reference's scoring semantics, used by `tests/`, by the package smoke
 *  check and by the CPU-baseline leg of `bench.py` as the checker. The shipped library never links, loads or
 *  calls it. Every function cites the reference file:line whose behaviour it restates.
 *
 *  Pinning (what keeps a plausible mistake here from going unnoticed):
 *  - the reference's own known-answer tests (test/similarities.py:134-143, 160-168, 188-197, 219-227),
 *    committed as fixtures under tests/golden/;
 *  - differential checks against the reference's serial and Ice Lake scorers compiled into oracle/_ref/
 *    (tests/test_oracle.py), on random inputs that cross every cell-width tier of serial.hpp:293-356;
 *  - brute-force enumeration of all alignments for strings of length <= 5 (tests/test_oracle.py);
 *  - closed forms: Lev(a,a)=0, empty-string costs, NW(0/-1/-1) = -Lev, SW >= 0, ASCII UTF-8 == bytes.
 */
#ifndef SZ_ORACLE_H_
#define SZ_ORACLE_H_
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/** Which score a call computes; mirrors the four reference engines. */
typedef enum {
    szo_levenshtein_k = 0,      /**< byte Levenshtein, minimize, uniform costs (serial.hpp:2049-2181) */
    szo_levenshtein_utf8_k = 1, /**< rune Levenshtein, minimize, uniform costs (serial.hpp:2191-2341) */
    szo_needleman_wunsch_k = 2, /**< global, maximize, class costs (serial.hpp:2352-2433) */
    szo_smith_waterman_k = 3,   /**< local, maximize, class costs (serial.hpp:2444-2532) */
} szo_kind_t;

typedef struct szo_params_t {
    int32_t kind;
    int8_t match, mismatch; /**< uniform substitution costs (Levenshtein kinds only) */
    int8_t open, extend;    /**< gap costs; `open == extend` selects the linear recurrence */
    uint8_t byte_to_class[256];
    int8_t class_costs[32 * 32]; /**< row-major, `class_costs[class(query) * 32 + class(candidate)]` */
} szo_params_t;

/**
 *  Decodes UTF-8 the way `sz_rune_decode_unchecked` does (include/stringzilla/utf8_runes/serial.h:101-114):
 *  the lead byte alone sets the length `1 + [b>=0xC0] + [b>=0xE0] + [b>=0xF0]`, nothing is validated.
 *  A sequence truncated by the end of the string reads its missing continuation bytes as 0x00 (the reference
 *  over-reads there); this is the one behaviour defined here rather than restated.
 *  @return number of runes written to @p runes (capacity >= @p length).
 */
size_t szo_utf8_decode(uint8_t const *text, size_t length, uint32_t *runes);

/** Score of one pair, query @p a (rows) against candidate @p b (columns), in int64 cells. */
int64_t szo_score(szo_params_t const *params, uint8_t const *a, size_t a_length, uint8_t const *b, size_t b_length);

/** Same over already-decoded symbols: uniform costs compare symbols, class costs map them through byte_to_class. */
int64_t szo_score_symbols(szo_params_t const *params, uint32_t const *a, size_t a_length, uint32_t const *b,
                          size_t b_length);

/** Scores `count` independent pairs from two u64-offset tapes (Arrow layout, count+1 offsets) on @p threads. */
int szo_pairs(szo_params_t const *params, uint8_t const *a_data, uint64_t const *a_offsets, uint8_t const *b_data,
              uint64_t const *b_offsets, uint64_t count, int64_t *results, int threads);

/**
 *  Cross-product of two tapes into `results[q * stride + c]`; `c_data == NULL` requests the symmetric mode of
 *  serial.hpp:2545-2580: only `c <= q` is scored as `score(queries[q], queries[c])` and mirrored to `[c][q]`.
 */
int szo_cross(szo_params_t const *params, uint8_t const *q_data, uint64_t const *q_offsets, uint64_t q_count,
              uint8_t const *c_data, uint64_t const *c_offsets, uint64_t c_count, int64_t *results, uint64_t stride,
              int threads);

/*
 *  Rolling MinHash / Count-Min fingerprints (sz_fingerprints.c), restating
 *  include/stringzillas/fingerprints/serial.hpp and c/stringzillas/fingerprints.cuh.
 *
 *  Pinning: published splitmix64 outputs; a closed form (each window's hash is the Rabin-Karp polynomial
 *  sum of (byte + 1) * multiplier^(window - 1 - k) mod modulo, evaluated with exact integers and a brute-force
 *  minimum and count) in tests/test_oracle.py; differential checks against the reference's serial sliced,
 *  serial per-dimension and AVX-512 sliced engines compiled into oracle/_ref/.
 */

/** One dimension of the reference's f64 rolling hasher (serial.hpp:364-431). */
typedef struct szo_hasher_t {
    double multiplier, modulo, inverse_modulo;
    double negative_discarding_multiplier; /**< -(multiplier^(window-1) mod modulo) */
    double discarding_multiplier;          /**< (-multiplier^window) mod modulo, in [0, modulo) */
    uint64_t window_width;
} szo_hasher_t;

uint64_t szo_splitmix64(uint64_t state);

/** Fills `hashers[dimensions]` in output order; returns 1 for the sliced layout, 0 for the per-dimension one. */
int szo_fingerprint_hashers(uint64_t dimensions, uint64_t const *widths, uint64_t widths_count, uint64_t seed,
                            szo_hasher_t *hashers);

/** Min-hashes and their counts of one text over all dimensions. */
void szo_fingerprint(szo_hasher_t const *hashers, uint64_t dimensions, uint8_t const *text, size_t length,
                     uint32_t *min_hashes, uint32_t *min_counts);

/** Fingerprints of a u64-offset tape into strided rows (strides in bytes) on @p threads. */
int szo_fingerprints(szo_hasher_t const *hashers, uint64_t dimensions, uint8_t const *data, uint64_t const *offsets,
                     uint64_t count, uint32_t *min_hashes, uint64_t hashes_stride, uint32_t *min_counts,
                     uint64_t counts_stride, int threads);

#ifdef __cplusplus
}
#endif
#endif
