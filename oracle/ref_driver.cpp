/**
 *  @file oracle/ref_driver.cpp
 *  @brief Test infrastructure only: a thin C shim over the reference's own CPU scorers, compiled from the
 *         reference headers where they lie (never copied) into `oracle/_ref/`. It is used to pin the oracle
 *         restatement (`oracle/sz_oracle.c`) and as the reported CPU baseline. Nothing in the product links it.
 *
 *  Scorers called (reference files):
 *  - `levenshtein_distance<char, gaps, caps>`       include/stringzillas/similarities/serial.hpp:2049-2181
 *  - `levenshtein_distance_utf8<gaps, caps>`        serial.hpp:2191-2341
 *  - `needleman_wunsch_score<char, 32x32, gaps>`    serial.hpp:2352-2433
 *  - `smith_waterman_score<char, 32x32, gaps>`      serial.hpp:2444-2532
 *  - Ice Lake specialisations                       include/stringzillas/similarities/icelake.hpp
 *  Gap model selection follows c/stringzillas/levenshtein.cuh:104 (`open == extend` selects the linear engine).
 */
#include <stringzillas/similarities.hpp>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace szs = ashvardanian::stringzillas;
namespace sz = ashvardanian::stringzilla;

#ifndef SZREF_CAPS
#define SZREF_CAPS sz_cap_serial_k
#endif
static constexpr sz_capability_t caps_k = (sz_capability_t)(SZREF_CAPS);

struct szref_params_t {
    int8_t match, mismatch, open, extend;
    uint8_t byte_to_class[256];
    int8_t class_costs[32 * 32];
};

enum { szref_lev_k = 0, szref_lev_utf8_k = 1, szref_nw_k = 2, szref_sw_k = 3 };

template <typename scorer_type_, typename result_type_>
static int run_one_(scorer_type_ const &scorer, char const *a, size_t n, char const *b, size_t m,
                    std::vector<std::byte> &scratch, result_type_ &out) {
    sz::cpu_specs_t specs;
    szs::span<char const> first {a, n}, second {b, m};
    size_t needed = scorer.scratch_space_needed(first, second, specs);
    if (scratch.size() < needed + 64) scratch.resize(needed + 64);
    szs::dummy_executor_t executor;
    auto status = scorer(first, second, out, szs::scratch_space_t {scratch.data(), scratch.size()}, executor, specs);
    return (int)status;
}

static szs::error_costs_32x32_t make_matrix_(szref_params_t const *p) {
    szs::error_costs_32x32_t costs;
    std::memcpy(costs.byte_to_class, p->byte_to_class, 256);
    std::memcpy(costs.class_substitution_costs, p->class_costs, 1024);
    return costs;
}

static int score_pair_(int kind, szref_params_t const *p, char const *a, size_t n, char const *b, size_t m,
                       std::vector<std::byte> &scratch, int64_t *out) {
    bool const linear = p->open == p->extend;
    szs::uniform_substitution_costs_t uniform {p->match, p->mismatch};
    int status = 0;
    if (kind == szref_lev_k || kind == szref_lev_utf8_k) {
        size_t result = 0;
        if (kind == szref_lev_k) {
            if (linear)
                status = run_one_(szs::levenshtein_distance<char, szs::linear_gap_costs_t, caps_k> {
                                      uniform, szs::linear_gap_costs_t {p->open}},
                                  a, n, b, m, scratch, result);
            else
                status = run_one_(szs::levenshtein_distance<char, szs::affine_gap_costs_t, caps_k> {
                                      uniform, szs::affine_gap_costs_t {p->open, p->extend}},
                                  a, n, b, m, scratch, result);
        }
        else {
            if (linear)
                status = run_one_(szs::levenshtein_distance_utf8<szs::linear_gap_costs_t, caps_k> {
                                      uniform, szs::linear_gap_costs_t {p->open}},
                                  a, n, b, m, scratch, result);
            else
                status = run_one_(szs::levenshtein_distance_utf8<szs::affine_gap_costs_t, sz_cap_serial_k> {
                                      uniform, szs::affine_gap_costs_t {p->open, p->extend}},
                                  a, n, b, m, scratch, result);
        }
        *out = (int64_t)result;
        return status;
    }
    szs::error_costs_32x32_t const costs = make_matrix_(p);
    ptrdiff_t result = 0;
    if (kind == szref_nw_k) {
        if (linear)
            status = run_one_(szs::needleman_wunsch_score<char, szs::error_costs_32x32_t, szs::linear_gap_costs_t,
                                                          caps_k> {costs, szs::linear_gap_costs_t {p->open}},
                              a, n, b, m, scratch, result);
        else
            status = run_one_(szs::needleman_wunsch_score<char, szs::error_costs_32x32_t, szs::affine_gap_costs_t,
                                                          caps_k> {costs, szs::affine_gap_costs_t {p->open, p->extend}},
                              a, n, b, m, scratch, result);
    }
    else {
        if (linear)
            status = run_one_(szs::smith_waterman_score<char, szs::error_costs_32x32_t, szs::linear_gap_costs_t,
                                                        caps_k> {costs, szs::linear_gap_costs_t {p->open}},
                              a, n, b, m, scratch, result);
        else
            status = run_one_(szs::smith_waterman_score<char, szs::error_costs_32x32_t, szs::affine_gap_costs_t,
                                                        caps_k> {costs, szs::affine_gap_costs_t {p->open, p->extend}},
                              a, n, b, m, scratch, result);
    }
    *out = (int64_t)result;
    return status;
}

extern "C" {

/** @brief Scores `count` independent pairs `(a[i], b[i])` from two u64-offset tapes on `threads` OpenMP threads. */
int szref_pairs(int kind, szref_params_t const *params, char const *a_data, uint64_t const *a_offsets,
                char const *b_data, uint64_t const *b_offsets, uint64_t count, int64_t *results, int threads) {
    int failed = 0;
    if (threads < 1) threads = 1;
#pragma omp parallel num_threads(threads)
    {
        std::vector<std::byte> scratch;
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < (int64_t)count; ++i) {
            int status = score_pair_(kind, params, a_data + a_offsets[i], a_offsets[i + 1] - a_offsets[i],
                                     b_data + b_offsets[i], b_offsets[i + 1] - b_offsets[i], scratch, results + i);
            if (status != 0) {
#pragma omp atomic write
                failed = status;
            }
        }
    }
    return failed;
}

/** @brief Capability mask this shim was compiled for (1 = serial, 129 = serial + Ice Lake). */
int szref_capabilities(void) { return (int)caps_k; }
}
