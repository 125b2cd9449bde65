/*
 * esth_layout.h - sizes of the per-step random-number buffer.
 *
 * The reference draws all random numbers of a step from one MTGP32 buffer
 * (pf_generic.c:267-282):
 *   Nn = N * V normals, followed by uniforms:
 *     multinomial / systematic : Nu = (m + 1) * G   (selection + gate)
 *     alias                    : Nu = (2m + 1) * G  (draw pairs + gate)
 *     alias, adaptive          : Nu = 2m * G
 *   total = roundup(Nn + Nu, 768 * K), stream j writes [j*size, (j+1)*size),
 *   size = total / K, K = 96 (MTGP_WORKGROUPS, pf_generic.c:59).
 * ESTH_RNG_MTGP_PER_FILTER keeps the rule but picks K from the filter's own
 * need (about four 768-word rounds per stream, 1 <= K <= 96), so a filter's
 * numbers depend on nothing but its own size and seed.
 */
#ifndef ESTH_LAYOUT_H
#define ESTH_LAYOUT_H

#include <stdint.h>
#include "esthera.h"

#define ESTH_MTGP_N          351
#define ESTH_MTGP_THREADS    256
#define ESTH_MTGP_LARGE      768
#define ESTH_MTGP_SETS       200
#define ESTH_MTGP_PARAM_WORDS 52
#define ESTH_MTGP_COMPAT_STREAMS 96

static inline uint64_t esth_rng_uniform_count(int resampler, int adaptive, uint64_t m, uint64_t G) {
    if (resampler == ESTH_RESAMPLE_ALIAS) return adaptive ? 2 * m * G : (2 * m + 1) * G;
    return (m + 1) * G;
}

static inline uint64_t esth_rng_needed(int V, int resampler, int adaptive, uint64_t m, uint64_t G) {
    return (uint64_t)V * m * G + esth_rng_uniform_count(resampler, adaptive, m, G);
}

static inline int esth_rng_streams(int rng_mode, uint64_t needed) {
    if (rng_mode == ESTH_RNG_MTGP_COMPAT) return ESTH_MTGP_COMPAT_STREAMS;
    uint64_t k = (needed + 4 * ESTH_MTGP_LARGE - 1) / (4 * ESTH_MTGP_LARGE);
    if (k < 1) k = 1;
    if (k > ESTH_MTGP_COMPAT_STREAMS) k = ESTH_MTGP_COMPAT_STREAMS;
    return (int)k;
}

/* words per stream per step (multiple of 768) */
static inline uint64_t esth_rng_stream_size(int K, uint64_t needed) {
    uint64_t unit = (uint64_t)ESTH_MTGP_LARGE * (uint64_t)K;
    uint64_t total = ((needed + unit - 1) / unit) * unit;
    return total / (uint64_t)K;
}

/* one filter of a batch, as the batched kernels see it (device pointers;
 * R and frame are offsets into the per-parity buffers of the launch) */
struct esth_bfilter {
    float* state;
    float* tmp;
    float* w;
    float* part;
    float* bt_w;
    float* bt_s;
    void* winner;      /* float2[G] */
    void* result;      /* float2 */
    double* est;
    uint32_t* top;
    uint64_t r_off;    /* words into R[parity] */
    uint64_t f_off;    /* floats into frames[parity] */
    uint32_t N, G, m, t;
    int32_t dimx, dimy;
    uint32_t uoff;     /* N * V: first uniform of the step's words */
    uint32_t flags;    /* bit 0: carried weights */
};

/* one MTGP32 stream of a launch */
struct esth_mtgp_stream {
    uint32_t set;       /* parameter set */
    uint32_t size;      /* words this launch, multiple of 768 */
    uint64_t out_off;   /* first output word */
    uint64_t status;    /* index of its 351-word status in status_in / status_out */
};

/* reference rule for the 2-D torus (cuda/pf_generic.cu:682-683):
 * dimx = 2^floor(log2(sqrt(G))), dimy = G / dimx */
static inline void esth_torus_dims(uint32_t G, int* dimx, int* dimy) {
    uint32_t x = 1;
    while ((uint64_t)(2 * x) * (uint64_t)(2 * x) <= G) x *= 2;
    *dimx = (int)x;
    *dimy = (int)(G / x);
}

#endif /* ESTH_LAYOUT_H */
