/**
 *  @file oracle/sz_oracle.c
 *  @brief TEST INFRASTRUCTURE ONLY: CPU oracle restating the reference's similarity recurrences.
 *
 *  Only the test-suite, the smoke check and the CPU-baseline leg of the benchmark use this file, as the
 *  checker. Cells are int64 and the matrix is walked row by row; the reference walks anti-diagonals in the
 *  narrowest safe cell type (serial.hpp:293-356, 2126-2180), which yields the same integers whenever no cell
 *  overflows — the reach bound there guarantees that.
 */
#include "sz_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/** `sz_rune_decode_unchecked`, include/stringzilla/utf8_runes/serial.h:101-114 (missing tail bytes read as 0). */
size_t szo_utf8_decode(uint8_t const *text, size_t length, uint32_t *runes) {
    size_t count = 0;
    size_t at = 0;
    while (at < length) {
        uint8_t lead = text[at];
        size_t width = 1 + (lead >= 0xC0) + (lead >= 0xE0) + (lead >= 0xF0);
        uint32_t tail[3] = {0, 0, 0};
        for (size_t k = 1; k < width; ++k) tail[k - 1] = at + k < length ? (uint32_t)(text[at + k] & 0x3F) : 0u;
        uint32_t rune;
        switch (width) {
        case 1: rune = lead; break;
        case 2: rune = (uint32_t)(lead & 0x1F) << 6 | tail[0]; break;
        case 3: rune = (uint32_t)(lead & 0x0F) << 12 | tail[0] << 6 | tail[1]; break;
        default: rune = (uint32_t)(lead & 0x07) << 18 | tail[0] << 12 | tail[1] << 6 | tail[2]; break;
        }
        runes[count++] = rune;
        at += width;
    }
    return count;
}

static inline int64_t opt2(int maximize, int64_t x, int64_t y) {
    return maximize ? (x > y ? x : y) : (x < y ? x : y);
}

/**
 *  `s(i, j)`: uniform costs (serial.hpp:91-103) compare symbols; class costs (serial.hpp:167-190) look up
 *  `C[class(a)][class(b)]`, always with the query symbol as the row operand (serial.hpp:693-698, 721-722).
 */
static inline int64_t substitution(szo_params_t const *p, uint32_t a, uint32_t b) {
    if (p->kind == szo_levenshtein_k || p->kind == szo_levenshtein_utf8_k) return a == b ? p->match : p->mismatch;
    return p->class_costs[p->byte_to_class[a & 0xFF] * 32 + p->byte_to_class[b & 0xFF]];
}

/**
 *  The four recurrences of serial.hpp:669-1079 with the boundary conventions of the walkers
 *  (serial.hpp:1153-1158, 1310-1320, 1482-1488, 1627-1633) and the empty-string rules of serial.hpp:142-151.
 *
 *  - linear global (tile_scorer 669-749):  H = opt(opt(up, left) + g, diag + s), H[i][0] = g*i, H[0][j] = g*j
 *  - linear local  (762-841):              H = max(max(up, left) + g, max(diag + s, 0)), borders 0, best >= 0
 *  - affine global (854-964):  E = opt(M_left + o, E_left + e), V = opt(M_up + o, V_up + e),
 *                              M = opt(opt(V, E), M_diag + s); M[i][0] = o + e*(i-1), M[0][0] = 0;
 *                              the gap tracks at the border hold the discard value (o + e) + M_border (892-895)
 *  - affine local  (977-1079): M = max(max(V, E), max(M_diag + s, 0)), M borders 0, gap borders o + e
 */
int64_t szo_score_symbols(szo_params_t const *p, uint32_t const *a, size_t n, uint32_t const *b, size_t m) {
    int const maximize = p->kind == szo_needleman_wunsch_k || p->kind == szo_smith_waterman_k;
    int const local = p->kind == szo_smith_waterman_k;
    int const affine = p->open != p->extend; /* c/stringzillas/levenshtein.cuh:104, needleman_wunsch.cuh:95 */
    int64_t const o = p->open, e = p->extend;

    /* Border of the score matrix at distance `k` from the origin. */
#define SZO_BORDER(k) (local ? 0 : affine ? ((k) ? o + e * (int64_t)((k)-1) : 0) : o * (int64_t)(k))
    int64_t *row = (int64_t *)malloc((m + 1) * sizeof(int64_t));
    int64_t *vertical = (int64_t *)malloc((m + 1) * sizeof(int64_t));
    if (!row || !vertical) {
        free(row), free(vertical);
        return INT64_MIN;
    }
    for (size_t j = 0; j <= m; ++j) {
        row[j] = SZO_BORDER(j);
        vertical[j] = (o + e) + row[j]; /* discard value of the vertical track on row 0 */
    }
    int64_t best = 0;
    for (size_t i = 1; i <= n; ++i) {
        int64_t diag = row[0];
        row[0] = SZO_BORDER(i);
        int64_t horizontal = (o + e) + row[0]; /* discard value of the horizontal track on column 0 */
        for (size_t j = 1; j <= m; ++j) {
            int64_t const up = row[j], left = row[j - 1];
            int64_t const if_substitution = diag + substitution(p, a[i - 1], b[j - 1]);
            int64_t cell;
            if (affine) {
                horizontal = opt2(maximize, left + o, horizontal + e);
                vertical[j] = opt2(maximize, up + o, vertical[j] + e);
                int64_t const gap = opt2(maximize, vertical[j], horizontal);
                cell = local ? opt2(1, gap, opt2(1, if_substitution, 0)) : opt2(maximize, gap, if_substitution);
            }
            else {
                int64_t const gap = opt2(maximize, up, left) + o;
                cell = local ? opt2(1, gap, opt2(1, if_substitution, 0)) : opt2(maximize, gap, if_substitution);
            }
            if (local && cell > best) best = cell;
            diag = up;
            row[j] = cell;
        }
    }
#undef SZO_BORDER
    int64_t const result = local ? best : row[m];
    free(row), free(vertical);
    return result;
}

int64_t szo_score(szo_params_t const *p, uint8_t const *a, size_t n, uint8_t const *b, size_t m) {
    uint32_t *as = (uint32_t *)malloc((n + 1) * sizeof(uint32_t));
    uint32_t *bs = (uint32_t *)malloc((m + 1) * sizeof(uint32_t));
    if (!as || !bs) {
        free(as), free(bs);
        return INT64_MIN;
    }
    size_t an = n, bn = m;
    if (p->kind == szo_levenshtein_utf8_k) {
        /* serial.hpp:2271-2273 takes the byte path when both sides are ASCII; decoding ASCII is the identity, so
           decoding unconditionally yields the same runes. */
        an = szo_utf8_decode(a, n, as);
        bn = szo_utf8_decode(b, m, bs);
    }
    else {
        for (size_t i = 0; i < n; ++i) as[i] = a[i];
        for (size_t j = 0; j < m; ++j) bs[j] = b[j];
    }
    int64_t const result = szo_score_symbols(p, as, an, bs, bn);
    free(as), free(bs);
    return result;
}

typedef struct {
    szo_params_t const *params;
    uint8_t const *q_data, *c_data;
    uint64_t const *q_offsets, *c_offsets;
    uint64_t q_count, c_count, stride;
    int64_t *results;
    int mode; /* 0 pairs, 1 cross, 2 symmetric */
    uint64_t begin, end, step;
} szo_job_t;

static void *szo_worker(void *arg) {
    szo_job_t *job = (szo_job_t *)arg;
    for (uint64_t cell = job->begin; cell < job->end; cell += job->step) {
        uint64_t q, c;
        if (job->mode == 0) q = c = cell;
        else {
            q = cell / job->c_count;
            c = cell % job->c_count;
            if (job->mode == 2 && c > q) continue;
        }
        uint8_t const *qs = job->q_data + job->q_offsets[q];
        uint8_t const *cs = job->c_data + job->c_offsets[c];
        int64_t const score = szo_score(job->params, qs, job->q_offsets[q + 1] - job->q_offsets[q], cs,
                                        job->c_offsets[c + 1] - job->c_offsets[c]);
        if (job->mode == 0) job->results[q] = score;
        else {
            job->results[q * job->stride + c] = score;
            if (job->mode == 2 && c != q) job->results[c * job->stride + q] = score;
        }
    }
    return NULL;
}

static int szo_run(szo_job_t const *base, uint64_t total, int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t handles[256];
    szo_job_t jobs[256];
    for (int t = 0; t < threads; ++t) {
        jobs[t] = *base;
        jobs[t].begin = (uint64_t)t;
        jobs[t].end = total;
        jobs[t].step = (uint64_t)threads;
    }
    if (threads == 1) {
        szo_worker(&jobs[0]);
        return 0;
    }
    for (int t = 0; t < threads; ++t) pthread_create(&handles[t], NULL, szo_worker, &jobs[t]);
    for (int t = 0; t < threads; ++t) pthread_join(handles[t], NULL);
    return 0;
}

int szo_pairs(szo_params_t const *params, uint8_t const *a_data, uint64_t const *a_offsets, uint8_t const *b_data,
              uint64_t const *b_offsets, uint64_t count, int64_t *results, int threads) {
    szo_job_t job;
    memset(&job, 0, sizeof(job));
    job.params = params, job.q_data = a_data, job.q_offsets = a_offsets, job.c_data = b_data;
    job.c_offsets = b_offsets, job.q_count = job.c_count = count, job.results = results, job.mode = 0;
    return szo_run(&job, count, threads);
}

int szo_cross(szo_params_t const *params, uint8_t const *q_data, uint64_t const *q_offsets, uint64_t q_count,
              uint8_t const *c_data, uint64_t const *c_offsets, uint64_t c_count, int64_t *results, uint64_t stride,
              int threads) {
    szo_job_t job;
    memset(&job, 0, sizeof(job));
    int const symmetric = c_data == NULL;
    job.params = params, job.q_data = q_data, job.q_offsets = q_offsets, job.q_count = q_count;
    job.c_data = symmetric ? q_data : c_data, job.c_offsets = symmetric ? q_offsets : c_offsets;
    job.c_count = symmetric ? q_count : c_count, job.results = results, job.stride = stride;
    job.mode = symmetric ? 2 : 1;
    if (q_count == 0 || job.c_count == 0) return 0;
    return szo_run(&job, q_count * job.c_count, threads);
}
