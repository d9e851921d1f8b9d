/**
 *  @file oracle/ref_fingerprints.cpp
 *  @brief Test infrastructure only: a thin C shim over the reference's own CPU fingerprint engines, compiled from
 *         the reference headers where they lie (never copied) into `oracle/_ref/`. It pins the oracle restatement
 *         (`oracle/sz_fingerprints.c`) and serves as the reported CPU baseline of the fingerprints benchmark.
 *         Nothing in the product links it.
 *
 *  The reference's C entry point (c/stringzillas/fingerprints.cuh:27-140) cannot be linked here: its device
 *  scopes call the Fork Union C library, whose sources are not part of the reference tree. This shim therefore
 *  builds the same engines that entry point builds, with the same dimension layout (fingerprints.cuh:39-47,
 *  117-140), and drives them with the reference's single-threaded `dummy_executor_t`:
 *  - `floating_rolling_hashers<caps, 64>` slices, seeded by `try_seed(width, alphabet, 64 * i, seed)`
 *    (serial.hpp:876-1054; skylake.hpp / haswell.hpp for the AVX-512 / AVX2 builds), when every width gets
 *    the same multiple of 64 dimensions;
 *  - otherwise `basic_rolling_hashers<floating_rolling_hasher<f64_t>, u32_t>` grown one dimension at a time
 *    by `try_extend(widths[d % count], 1, alphabet, seed)` (serial.hpp:497-870).
 *  Texts are sharded over OpenMP threads, one text per call of the engine's batch operator.
 */
#include <stringzillas/fingerprints.hpp>

#include <cstdint>
#include <string_view>
#include <vector>
#include <omp.h>

namespace szs = ashvardanian::stringzillas;
namespace sz = ashvardanian::stringzilla;

#ifndef SZREF_FP_CAPS
#define SZREF_FP_CAPS sz_cap_serial_k
#endif
static constexpr sz_capability_t caps_k = (sz_capability_t)(SZREF_FP_CAPS);
static constexpr size_t slice_k = 64;

/** One text as the engines' batch operator reads it. */
struct one_text_t {
    std::string_view text;
    size_t size() const noexcept { return 1; }
    std::string_view operator[](size_t) const noexcept { return text; }
};

/** One output row of `extent_` u32 values, as the engines' batch operator writes it. */
template <size_t extent_>
struct one_row_t {
    uint32_t *row;
    size_t length;
    size_t size() const noexcept { return 1; }
    szs::span<uint32_t, extent_> operator[](size_t) const noexcept { return {row, length}; }
};

extern "C" {

/**
 *  Fingerprints `count` texts of a u64-offset tape into rows of `dimensions` u32 (row strides in elements).
 *  `layout` 0 picks what `szs_fingerprints_init` picks; 1 forces the per-dimension engine.
 *  @return 0 on success, otherwise the reference's status code.
 */
int szref_fingerprints(uint64_t dimensions, uint64_t alphabet_size, uint64_t const *widths, uint64_t widths_count,
                       uint64_t seed, char const *data, uint64_t const *offsets, uint64_t count, uint32_t *hashes,
                       uint32_t *counts, uint64_t row_stride, int layout, int threads) {
    static uint64_t const default_widths[] = {3, 4, 5, 7, 9, 11, 15, 31};
    if (!widths || widths_count == 0) widths = default_widths, widths_count = 8;
    uint64_t const per_width_min = dimensions / widths_count;
    uint64_t const per_width_max = (dimensions + widths_count - 1) / widths_count;
    bool const sliced = layout == 0 && per_width_min == per_width_max && per_width_min % slice_k == 0;
    int failed = 0;
    if (threads < 1) threads = 1;
    if (sliced) {
        using hasher_t = szs::floating_rolling_hashers<caps_k, slice_k>;
        std::vector<hasher_t> slices(dimensions / slice_k);
        for (size_t i = 0; i < slices.size(); ++i)
            if (slices[i].try_seed(widths[i % widths_count], alphabet_size, i * slice_k, seed) !=
                sz::status_t::success_k)
                return 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
        for (int64_t t = 0; t < (int64_t)count; ++t) {
            one_text_t text {{data + offsets[t], (size_t)(offsets[t + 1] - offsets[t])}};
            for (size_t i = 0; i < slices.size(); ++i) {
                one_row_t<slice_k> h {hashes + t * row_stride + i * slice_k, slice_k};
                one_row_t<slice_k> c {counts + t * row_stride + i * slice_k, slice_k};
                auto status = slices[i](text, h, c);
                if (status != sz::status_t::success_k) {
#pragma omp atomic write
                    failed = (int)status;
                }
            }
        }
        return failed;
    }
    using engine_t = szs::basic_rolling_hashers<szs::floating_rolling_hasher<sz::f64_t>, sz::u32_t>;
    engine_t engine;
    for (size_t d = 0; d < dimensions; ++d)
        if (engine.try_extend(widths[d % widths_count], 1, alphabet_size, seed) != sz::status_t::success_k) return 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int64_t t = 0; t < (int64_t)count; ++t) {
        one_text_t text {{data + offsets[t], (size_t)(offsets[t + 1] - offsets[t])}};
        one_row_t<SZ_SIZE_MAX> h {hashes + t * row_stride, dimensions};
        one_row_t<SZ_SIZE_MAX> c {counts + t * row_stride, dimensions};
        auto status = engine(text, h, c);
        if (status != sz::status_t::success_k) {
#pragma omp atomic write
            failed = (int)status;
        }
    }
    return failed;
}

/** Capability mask this shim was compiled for. */
int szref_fingerprints_capabilities(void) { return (int)caps_k; }
}
