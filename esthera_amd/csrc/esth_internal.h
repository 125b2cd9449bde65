/*
 * esth_internal.h - host-side internals shared by the engine's translation units.
 */
#ifndef ESTH_INTERNAL_H
#define ESTH_INTERNAL_H

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "esthera.h"
#include "esth_layout.h"
#include "esth_models.h"

int esth_fail(int code, const char* fmt, ...);

#define ESTH_HIP(x)                                                                        \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return esth_fail(ESTH_ERR_DEVICE, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), \
                             __FILE__, __LINE__);                                          \
    } while (0)

#define ESTH_TRY(x)               \
    do {                          \
        int rc_ = (x);            \
        if (rc_ != ESTH_OK) return rc_; \
    } while (0)

struct esth_engine {
    int device = 0;
    size_t budget = 0;
    size_t used = 0;
    std::mutex mu;
    uint32_t* d_params = nullptr;   /* 200 x 52 MTGP parameter words */
    std::atomic<int> batch_fusion[11];   /* per log2(m): ESTH_FUSION_* */
    std::atomic<int> batch_lanes[11];    /* per log2(m): lanes per sub-filter of the resampling, 0 auto */
    esth_engine() {
        for (auto& f : batch_fusion) f.store(ESTH_FUSION_AUTO, std::memory_order_relaxed);
        for (auto& l : batch_lanes) l.store(0, std::memory_order_relaxed);
    }
};

int esth_engine_alloc(esth_engine* e, size_t bytes, void** p);
/* charge (bytes > 0) or return (bytes < 0) budget without allocating */
int esth_engine_account(esth_engine* e, int64_t bytes);
void esth_engine_release(esth_engine* e, void* p, size_t bytes);
int esth_check_desc(const esth_filter_desc* d, int* S, int* V);
int esth_params_size(int model, size_t* sz);
const uint32_t* esth_param_table();

/* launchers shared by filters, batches and stage entry points */
int esth_launch_mtgp(const uint32_t* d_params, const uint32_t* d_status_in, uint32_t* d_status_out,
                     const esth_mtgp_stream* d_streams, int nstreams, void* d_out, int float_mode, hipStream_t s,
                     int nsteps = 1, uint64_t out_step = 0, uint64_t status_step = 0);
/* frame: the step's frame in device memory; host_frame (non-null): the
 * frame passed by value as a kernel argument instead (frame unused) */
int esth_launch_sample_weight(int model, float* state, uint32_t N, const float* R, const float* frame,
                              int log_w, float* w, int carry, int layout, hipStream_t s,
                              const float* host_frame = nullptr, const esth_mtgp_inline* mtgp = nullptr);
/* one launch across the step boundary (k_step_x): step k's resampling, argmax
 * and posterior from its sorted buffers (*_k), then step k + 1's
 * sample+weight and sort into the other buffers (*_k1); the unicycle at
 * m = 512 on a ring or torus with t > 0, SOA state, no carried weights.
 * host_frame (non-null): the frame as a kernel argument; otherwise the
 * kernel reads it from `frame` in device memory (graph replay) */
int esth_launch_step_x(int model, uint32_t m, float* tmp, float* w, const float* rnd_k, float thr, int systematic,
                       const float* bt_w_k, const float* bt_s_k, const float2* winner_k, const float* part_k,
                       float2* result, double* est, float* state, uint32_t N, uint32_t G, uint32_t t, int topo,
                       int dimx, int dimy, const float* R_k1, const float* frame, const float* host_frame, int log_w,
                       float2* winner_k1,
                       float* bt_w_k1, float* bt_s_k1, float* part_k1, const float* gate_k1, const esth_mtgp_inline* mtgp,
                       hipStream_t s);
int esth_launch_sort(int S, uint32_t m, float* w, const float* st_in, float* st_out, uint32_t N,
                     uint32_t G, float2* winner, int* perm, uint32_t t, float* bt_w, float* bt_s,
                     float* part, int log_w, int layout, hipStream_t s);
/* sample+weight and block sort in one launch (one work-group per sub-filter,
 * the unsorted state kept in LDS); for 128 <= m <= 1024.  The propagated
 * state is stored to `state` only for sub-filters g with keep_all or
 * gate[g] - 1 > keep_thr (those the resampling leaves untouched). */
int esth_waves_per_simd(const void* kernel, int threads);   /* resident waves per SIMD (registers, LDS) */
/* table-driven (batch, server) sample+weight + sort for size class lm: taken
 * where kb_sample_sort keeps kb_sample_weight's occupancy (cached answer) */
/* whether batches and servers of engine e sample and sort size class lm
 * (m = 2^lm) in one launch: the engine's setting (esth_engine_set_batch_fusion),
 * by default the fused kernel where it keeps as many waves resident as the
 * separate sampling */
bool esth_kb_fuse_pays(const esth_engine* e, int model, int S, int lm);
/* the same decision from a mode read once (esth_kb_fuse_mode), so that the
 * tables and launches of one server wave agree */
int esth_kb_fuse_mode(const esth_engine* e, int lm);
/* lanes per sub-filter of the batched resampling of class lm (m = 2^lm) on
 * engine e: 32 (m = 64: two sub-filters per wave), 64 (one wave) or 128 (two
 * waves, m >= 256), as set by esth_engine_set_batch_resample_lanes, 64 by
 * default */
int esth_kb_resample_lanes(const esth_engine* e, int lm);
/* the batched resampling of n entries of class lm at that tiling; work-groups
 * below argmax_blocks run the argmax of filters 0 .. argmax_blocks - 1 */
int esth_launch_kb_resample(int S, int lm, int lanes, const esth_bfilter* F, const uint2* sub, uint32_t n,
                            const float* R, float thr, int systematic, int log_w, int topo, uint32_t argmax_blocks,
                            hipStream_t st);
/* mean device time of n whole steps of f (events on its stream around them) */
int esth_filter_time_steps(esth_filter* f, int n, const void* control, const void* measurement, float dt,
                           float* us_per_step);
bool esth_kb_fuse_with(int mode, int model, int S, int lm);
/* the m = 64 fused class works on units of two consecutive sub-filters of one
 * filter (kb_sample_sort64): from a class list ordered by filter and
 * sub-filter, writes the units to `units` (room for n) and returns their count */
size_t esth_pair_units(const uint2* cls, size_t n, uint2* units);
int esth_launch_kb_sample_sort(int model, int S, int lm, const esth_bfilter* F, const uint2* sub, uint32_t n,
                               const float* R, const float* frames, int log_w, float thr, int keep_all,
                               hipStream_t st);
bool esth_sample_sort_fits(uint32_t m);
bool esth_sample_sort_pays(int model, int layout, uint32_t m);   /* fits, and costs the sampling no occupancy */
int esth_launch_sample_sort(int model, uint32_t m, float* state, uint32_t N, uint32_t G, const float* R,
                            const float* frame, int log_w, float* w, int carry, int layout, float* st_out,
                            float2* winner, uint32_t t, float* bt_w, float* bt_s, float* part, const float* gate,
                            float keep_thr, int keep_all, hipStream_t s,
                            const float* host_frame = nullptr, const esth_mtgp_inline* mtgp = nullptr);
/* whether esth_launch_sample_sort can carry MTGP32 work-groups at this m */
bool esth_sample_sort_inline_rng(uint32_t m);
int esth_launch_argmax(int S, const float2* winner, uint32_t G, int log_w, const float* part,
                       const float* sorted, uint32_t N, uint32_t m, float2* result, double* est,
                       hipStream_t s);
int esth_launch_exchange(float* st, float* w, int S, uint32_t N, uint32_t m, uint32_t G, uint32_t t,
                         int topo, int dimx, int dimy, const float* bt_w, const float* bt_s,
                         uint32_t* top, hipStream_t s);
int esth_launch_resample(int resampler, int S, float* st_main, const float* st_tmp, const float* w,
                         const float* rnd, uint32_t N, uint32_t m, uint32_t G, float thr,
                         int adaptive, int log_w, int* sel, int carry, int layout, hipStream_t s);
/* scan resampling with the step's argmax + posterior and the exchange (its
 * receiving side, plus the all-to-all top t) in the same launch; the best
 * particle is read from best_src[c * plane_stride + g * g_stride] */
int esth_launch_resample_argmax(int resampler, int S, float* st_main, const float* st_tmp, float* w,
                                const float* rnd, uint32_t N, uint32_t m, uint32_t G, float thr, int log_w,
                                const float2* winner, const float* part, const float* best_src, size_t plane_stride,
                                uint32_t g_stride, float2* result, double* est, uint32_t t, int topo, int dimx,
                                int dimy, const float* bt_w, const float* bt_s, uint32_t* top, int carry,
                                int layout, hipStream_t s);
/* the same with the one-wave Vose alias resampling (m >= 64) */
int esth_launch_alias_argmax(int S, float* st_main, const float* st_tmp, float* w, const float* rnd, uint32_t N,
                             uint32_t m, uint32_t G, float thr, int adaptive, int log_w, const float2* winner,
                             const float* part, const float* best_src, size_t plane_stride, uint32_t g_stride,
                             float2* result, double* est, uint32_t t, int topo, int dimx, int dimy, const float* bt_w,
                             const float* bt_s, uint32_t* top, int carry, int layout, hipStream_t s);
/* the global top t of the compact best-t buffer (all-to-all exchange) */
int esth_launch_topt(const float* bt_w, uint32_t n, uint32_t t, uint32_t* top, hipStream_t s);

#endif
