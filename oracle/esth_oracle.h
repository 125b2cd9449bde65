/*
 * esth_oracle.h - CPU oracle of the filter step.  TEST INFRASTRUCTURE ONLY:
 * used by tests/, the smoke check and the CPU baseline leg of bench.py as the
 * checker, never by the engine.
 */
#ifndef ESTH_ORACLE_H
#define ESTH_ORACLE_H

#include <stdint.h>
#include <stddef.h>
#include "esthera.h"

#ifdef __cplusplus
extern "C" {
#endif

/* MTGP32 */
void or_mtgp_seed(const uint32_t* param52, uint32_t seed, uint32_t* st351);
void or_mtgp_generate(const uint32_t* param52, uint32_t mask, uint32_t* st351, int n,
                      uint32_t* out, int float_mode);
void or_mtgp_fill(const uint32_t* params, int nstreams, int first_set, uint32_t* status,
                  int size, uint32_t* out, int float_mode);

/* stages */
void or_boxmuller(float* d, int64_t n);
int  or_build_frame(int model, const void* params, const void* ctrl, const void* meas, float dt,
                    float* frame);
int  or_sample_weight(int model, float* state, uint32_t N, const float* normals, const float* frame,
                      int weights, float* w);
int  or_sample_weight_f64(int model, const float* state_in, uint32_t N, const float* normals,
                          const void* params, const void* ctrl, const void* meas, float dt,
                          double* state_out, double* lw_out);
void or_block_sort(float* w, const float* state_in, float* state_out, int S, uint32_t m,
                   uint32_t G, float* winner, int* perm);
void or_argmax(const float* winner, uint32_t G, int weights, float* result);
int  or_exchange(float* state, float* w, int S, uint32_t m, uint32_t G, uint32_t t, int topo,
                 int dimx, int dimy);
int  or_resample(int resampler, float* dst, const float* src, int S, const float* w,
                 const float* rnd, uint32_t m, uint32_t G, float thr, int adaptive, int weights,
                 int* sel);
void or_alias_table(const float* w, uint32_t m, uint32_t G, int weights, float* prob,
                    uint32_t* alias);

/* math, exported for error-bound tests */
float or_expf(float x);
float or_logf(float x);
void  or_sincosf(float x, float* s, float* c);
float or_atan2f(float y, float x);
void  or_polar_abs(float u, float v, float* r, float* a);
float or_rcpf(float x);

/* whole filter (FP32 restatement, bit-exact target) with an FP64 posterior */
typedef struct or_filter or_filter;
or_filter* or_filter_create(const esth_filter_desc* d, const void* params, size_t params_size,
                            const uint32_t* mtgp_params);
void or_filter_destroy(or_filter* f);
void or_filter_init_uniform(or_filter* f, const float* state);
void or_filter_set_state(or_filter* f, const float* state_soa);
int  or_filter_step(or_filter* f, const void* ctrl, const void* meas, float dt);
/* state (S planes of N), sorted weights (N) */
void or_filter_read(or_filter* f, float* state, float* w);
/* FP64 posterior mean of the last step computed from the FP32 step's own
 * sampled particles (mean32) and from an FP64 propagation of the same prior
 * with the same normals (mean64); best = (w, g) of the argmax. */
void or_filter_estimate(or_filter* f, double* mean_f32path, double* mean_f64path, float* best);
/* state of the best particle (sorted block winner_best[1], position 0) */
void or_filter_best_particle(or_filter* f, float* out);
uint64_t or_filter_rng_words(or_filter* f);
void or_filter_set_f64(or_filter* f, int on);
void or_set_threads(int n);

#ifdef __cplusplus
}
#endif

#endif
