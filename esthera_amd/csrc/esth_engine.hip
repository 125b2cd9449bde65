/*
 * esth_engine.hip - host side of the C ABI: engine, kernel launchers, stage
 * entry points, single filters, the reference-compatible driver, trace I/O
 * and synthetic scenarios.
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <cstring>

#include "esth_internal.h"
#include "esth_kernels.h"
#include "esth_dispatch.h"

/* ================================================================ errors */
static thread_local std::string g_last_error;

int esth_fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

extern "C" const char* esth_last_error(void) { return g_last_error.c_str(); }
extern "C" int esth_version(void) { return ESTH_VERSION; }

/* ================================================== MTGP parameter data
 * The 200 MTGP32-11213 parameter sets (BSD-licensed data of the MTGP
 * authors, esthera_amd/data/mtgp32dc_11213.txt), compiled in by the build. */
static const uint32_t g_mtgp_table[ESTH_MTGP_SETS * ESTH_MTGP_PARAM_WORDS] = {
#include "mtgp32dc_11213.inc"
};

const uint32_t* esth_param_table() { return g_mtgp_table; }
extern "C" const uint32_t* esth_mtgp_params(void) { return g_mtgp_table; }

static int esth_noalloc_scratch(float** out);

/* mtgp32_init_state, mtgp32-fast.c:119-137 */
extern "C" int esth_mtgp_seed_state(int set, uint32_t seed, uint32_t* st) {
    if (set < 0 || set >= ESTH_MTGP_SETS || !st) return esth_fail(ESTH_ERR_INVALID, "bad MTGP set %d", set);
    const uint32_t* tbl = g_mtgp_table + (size_t)set * ESTH_MTGP_PARAM_WORDS + 4;
    uint32_t hidden = tbl[4] ^ (tbl[8] << 16);
    uint32_t tmp = hidden;
    tmp += tmp >> 16;
    tmp += tmp >> 8;
    uint32_t fill = (tmp & 0xffu) * 0x01010101u;
    for (int i = 0; i < ESTH_MTGP_N; ++i) st[i] = fill;
    st[0] = seed;
    st[1] = hidden;
    for (int i = 1; i < ESTH_MTGP_N; ++i) st[i] ^= 1812433253u * (st[i - 1] ^ (st[i - 1] >> 30)) + (uint32_t)i;
    return ESTH_OK;
}

/* ================================================================ models */
extern "C" int esth_model_lookup(const char* name) {
    if (!name) return -1;
    if (!strcmp(name, "robotarm")) return ESTH_MODEL_ROBOTARM5;
#define ESTH_X(ID, NAME, PFX, S_, V_, PSZ, CSZ, MSZ) \
    if (!strcmp(name, NAME)) return ID;
    ESTH_MODEL_LIST(ESTH_X)
#undef ESTH_X
    return -1;
}

int esth_params_size(int model, size_t* sz) {
    switch (model) {
#define ESTH_X(ID, NAME, PFX, S_, V_, PSZ, CSZ, MSZ) \
    case ID: *sz = PSZ; return ESTH_OK;
    ESTH_MODEL_LIST(ESTH_X)
#undef ESTH_X
    default: return esth_fail(ESTH_ERR_INVALID, "unknown model %d", model);
    }
}

extern "C" int esth_model_info(int model, int* S, int* V, size_t* ps, size_t* cs, size_t* ms) {
    switch (model) {
#define ESTH_X(ID, NAME, PFX, S_, V_, PSZ, CSZ, MSZ) \
    case ID:                                        \
        if (S) *S = S_;                             \
        if (V) *V = V_;                             \
        if (ps) *ps = PSZ;                          \
        if (cs) *cs = CSZ;                          \
        if (ms) *ms = MSZ;                          \
        return ESTH_OK;
    ESTH_MODEL_LIST(ESTH_X)
#undef ESTH_X
    default: return esth_fail(ESTH_ERR_INVALID, "unknown model %d", model);
    }
}

/* particle_filter.cu:60-74 */
extern "C" void esth_unicycle10_params_default(esth_unicycle10_params* p) {
    if (!p) return;
    static const float sx[10] = {12, 12, 13, 14, 16, 16, 18, 20, 20, 20};
    static const float sy[10] = {12, 37, 25, 18, 12, 37, 30, 12, 25, 37};
    for (int j = 0; j < 10; ++j) {
        p->sx[j] = sx[j];
        p->sy[j] = sy[j];
        p->sz[j] = 3.0f;
    }
    p->noise_v = 0.5f;
    p->noise_omega = 0.1f;
    p->noise_gamma = 0.01f;
}

/* pf_model_soa_packed_code.cuh:33-43 */
extern "C" void esth_camera_params_default(esth_camera_params* p) {
    if (!p) return;
    p->fc[0] = 408.65831f;
    p->fc[1] = 409.69059f;
    p->cc[0] = 166.32611f;
    p->cc[1] = 108.93539f;
    p->kc[0] = -0.21561f;
    p->kc[1] = 1.60704f;
    p->kc[2] = -0.01291f;
    p->kc[3] = 0.00007f;
}

extern "C" int esth_model_build_frame(int model, const void* params, const void* control,
                                      const void* measurement, float dt, float* frame) {
    if (!frame) return esth_fail(ESTH_ERR_INVALID, "null frame");
    if (esth_build_frame_impl(model, params, control, measurement, dt, frame))
        return esth_fail(ESTH_ERR_INVALID, "model %d: bad params/control/measurement", model);
    return ESTH_OK;
}

static bool is_pow2(uint32_t x) { return x && !(x & (x - 1)); }

int esth_check_desc(const esth_filter_desc* d, int* S, int* V) {
    if (!d) return esth_fail(ESTH_ERR_INVALID, "null descriptor");
    if (esth_model_dims(d->model, S, V)) return esth_fail(ESTH_ERR_INVALID, "unknown model %d", d->model);
    const uint32_t m = d->num_particles, G = d->num_blocks, t = d->num_transfer;
    if (!is_pow2(m) || m < 2 || m > ESTH_MAX_NUM_PARTICLES)
        return esth_fail(ESTH_ERR_INVALID, "num_particles %u: need a power of two in [2, %d]", m, ESTH_MAX_NUM_PARTICLES);
    if (G < 1 || G > ESTH_MAX_NUM_BLOCKS)
        return esth_fail(ESTH_ERR_INVALID, "num_blocks %u exceeds maximum %d", G, ESTH_MAX_NUM_BLOCKS);
    if (t > ESTH_MAX_NUM_TRANSFER) return esth_fail(ESTH_ERR_INVALID, "num_transfer %u exceeds maximum", t);
    if (d->topology < ESTH_TOPO_NONE || d->topology > ESTH_TOPO_ALL2ALL)
        return esth_fail(ESTH_ERR_INVALID, "bad topology %d", d->topology);
    if (t > 0 && d->topology != ESTH_TOPO_NONE) {
        if (d->topology == ESTH_TOPO_TORUS) {
            if (4ull * t > m) return esth_fail(ESTH_ERR_INVALID, "torus exchange needs 4t <= m (t=%u, m=%u)", t, m);
            int dx = d->torus_dimx, dy = d->torus_dimy;
            if (dx <= 0 || dy <= 0) esth_torus_dims(G, &dx, &dy);
            if ((uint64_t)dx * (uint64_t)dy != G)
                return esth_fail(ESTH_ERR_INVALID, "torus %dx%d does not cover %u sub-filters", dx, dy, G);
        } else if (2ull * t > m) {
            return esth_fail(ESTH_ERR_INVALID, "exchange needs 2t <= m (t=%u, m=%u)", t, m);
        }
        if (d->topology == ESTH_TOPO_ALL2ALL && t > 64)
            return esth_fail(ESTH_ERR_UNSUPPORTED, "all-to-all exchange supports t <= 64");
    }
    if (d->resampler < ESTH_RESAMPLE_MULTINOMIAL || d->resampler > ESTH_RESAMPLE_ALIAS)
        return esth_fail(ESTH_ERR_INVALID, "bad resampler %d", d->resampler);
    if (d->adaptive && d->resampler != ESTH_RESAMPLE_ALIAS)
        return esth_fail(ESTH_ERR_UNSUPPORTED, "adaptive resampling is defined for the alias resampler");
    if (!(d->resampling_threshold >= 0.0f) || d->resampling_threshold > 1.0f)
        return esth_fail(ESTH_ERR_INVALID, "resampling threshold %f outside [0, 1]", (double)d->resampling_threshold);
    if (d->weights != ESTH_WEIGHTS_LINEAR && d->weights != ESTH_WEIGHTS_LOG)
        return esth_fail(ESTH_ERR_INVALID, "bad weights mode %d", d->weights);
    if (d->rng != ESTH_RNG_MTGP_COMPAT && !esth_rng_is_per_filter(d->rng))
        return esth_fail(ESTH_ERR_INVALID, "bad rng mode %d", d->rng);
    if (d->state_layout < ESTH_LAYOUT_SOA || d->state_layout > ESTH_LAYOUT_PACKED)
        return esth_fail(ESTH_ERR_INVALID, "bad state layout %d", d->state_layout);
    return ESTH_OK;
}

extern "C" int esth_state_convert(int S, size_t n, int from, const float* src, int to, float* dst) {
    if (S < 1 || S > ESTH_MAX_STATE_DIM) return esth_fail(ESTH_ERR_INVALID, "state dimension %d", S);
    if (from < ESTH_LAYOUT_SOA || from > ESTH_LAYOUT_PACKED || to < ESTH_LAYOUT_SOA || to > ESTH_LAYOUT_PACKED)
        return esth_fail(ESTH_ERR_INVALID, "bad layout %d -> %d", from, to);
    if (!src || !dst) return esth_fail(ESTH_ERR_INVALID, "null buffer");
    const size_t total = (size_t)S * n;
    if (src < dst + total && dst < src + total) return esth_fail(ESTH_ERR_INVALID, "overlapping buffers");
    for (int c = 0; c < S; ++c)
        for (size_t i = 0; i < n; ++i) dst[esth_state_index(to, S, n, c, i)] = src[esth_state_index(from, S, n, c, i)];
    return ESTH_OK;
}

extern "C" void esth_filter_desc_default(esth_filter_desc* d, int model) {
    if (!d) return;
    memset(d, 0, sizeof(*d));
    d->model = model;
    d->num_particles = 256;
    d->num_blocks = 16;
    d->num_transfer = 1;
    d->topology = ESTH_TOPO_RING;
    d->resampler = ESTH_RESAMPLE_MULTINOMIAL;
    d->resampling_threshold = 1.0f;
    d->weights = ESTH_WEIGHTS_LINEAR;
    d->rng = ESTH_RNG_MTGP_COMPAT;
    d->seed = 1234;
}

/* ================================================================ engine */
extern "C" int esth_set_device(int device) {
    ESTH_HIP(hipSetDevice(device));
    return ESTH_OK;
}

extern "C" int esth_engine_create(int device, size_t budget, esth_engine** out) {
    if (!out) return esth_fail(ESTH_ERR_INVALID, "null out");
    int n = 0;
    ESTH_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return esth_fail(ESTH_ERR_INVALID, "device %d of %d", device, n);
    ESTH_HIP(hipSetDevice(device));
    size_t fr = 0, tot = 0;
    ESTH_HIP(hipMemGetInfo(&fr, &tot));
    esth_engine* e = new esth_engine();
    e->device = device;
    e->budget = budget ? budget : (size_t)(0.9 * (double)fr);
    hipError_t err = hipMalloc((void**)&e->d_params, sizeof(g_mtgp_table));
    if (err != hipSuccess) { delete e; return esth_fail(ESTH_ERR_DEVICE, "hipMalloc params: %s", hipGetErrorString(err)); }
    err = hipMemcpy(e->d_params, g_mtgp_table, sizeof(g_mtgp_table), hipMemcpyHostToDevice);
    if (err != hipSuccess) { hipFree(e->d_params); delete e; return esth_fail(ESTH_ERR_DEVICE, "upload params: %s", hipGetErrorString(err)); }
    float* sc = nullptr;
    int rc = esth_noalloc_scratch(&sc);   /* this device's allocation-free scratch address */
    if (rc != ESTH_OK) { hipFree(e->d_params); delete e; return rc; }
    *out = e;
    return ESTH_OK;
}

extern "C" int esth_engine_destroy(esth_engine* e) {
    if (!e) return ESTH_OK;
    hipSetDevice(e->device);
    hipFree(e->d_params);
    delete e;
    return ESTH_OK;
}

extern "C" int esth_engine_memory(const esth_engine* e, size_t* used, size_t* budget) {
    if (!e) return esth_fail(ESTH_ERR_INVALID, "null engine");
    if (used) *used = e->used;
    if (budget) *budget = e->budget;
    return ESTH_OK;
}

int esth_engine_alloc(esth_engine* e, size_t bytes, void** p) {
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->used + bytes > e->budget)
        return esth_fail(ESTH_ERR_BUDGET, "HBM budget: %zu used + %zu requested > %zu", e->used, bytes, e->budget);
    hipError_t err = hipMalloc(p, bytes ? bytes : 4);
    if (err != hipSuccess) return esth_fail(ESTH_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(err));
    e->used += bytes;
    return ESTH_OK;
}

int esth_engine_account(esth_engine* e, int64_t bytes) {
    std::lock_guard<std::mutex> lk(e->mu);
    if (bytes > 0 && e->used + (size_t)bytes > e->budget)
        return esth_fail(ESTH_ERR_BUDGET, "HBM budget: %zu used + %lld requested > %zu", e->used, (long long)bytes,
                         e->budget);
    if (bytes >= 0)
        e->used += (size_t)bytes;
    else
        e->used -= (size_t)(-bytes);
    return ESTH_OK;
}

void esth_engine_release(esth_engine* e, void* p, size_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(e->mu);
    hipFree(p);
    e->used -= bytes;
}

/* ============================================================ launchers */
static int grid_for(uint64_t threads, int block) { return (int)((threads + block - 1) / block); }

int esth_launch_mtgp(const uint32_t* d_params, const uint32_t* d_status_in, uint32_t* d_status_out,
                     const esth_mtgp_stream* d_streams, int nstreams, void* d_out, int float_mode, hipStream_t s,
                     int nsteps, uint64_t out_step, uint64_t status_step) {
    if (nstreams <= 0 || nsteps <= 0) return ESTH_OK;
    if (nsteps > 1 && d_status_in == d_status_out)
        return esth_fail(ESTH_ERR_INVALID, "multi-step MTGP32 launch needs separate status buffers");
    const uint32_t mask = g_mtgp_table[3];
    /* at most 8 resident work-groups per CU; more streams loop in the grid */
    static std::atomic<int> cus[64];
    int dev = 0;
    ESTH_HIP(hipGetDevice(&dev));
    int ncu = (dev >= 0 && dev < 64) ? cus[dev].load(std::memory_order_relaxed) : 0;
    if (ncu <= 0) {
        ESTH_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        if (dev >= 0 && dev < 64) cus[dev].store(ncu, std::memory_order_relaxed);
    }
    const int grid = std::min(nstreams, 8 * std::max(ncu, 1));
    hipLaunchKernelGGL(k_mtgp, dim3(grid), dim3(ESTH_MTGP_THREADS), 0, s, d_params, mask, d_status_in, d_status_out,
                       d_streams, (uint32_t)nstreams, (uint32_t*)d_out, float_mode, nsteps, out_step, status_step);
    ESTH_HIP(hipGetLastError());
    return ESTH_OK;
}

template <int MODEL, int L>
static void launch_sample_weight_l(int grid, float* state, uint32_t N, const float* R, const float* frame,
                                   const esth_frame_arg* fa, int log_w, float* w, int carry, const esth_mtgp_inline& mi,
                                   hipStream_t s) {
    if (fa)
        hipLaunchKernelGGL((k_sample_weight<MODEL, L, true>), dim3(grid + mi.n), dim3(256), 0, s, state, N, R, nullptr,
                           *fa, log_w, w, carry, mi);
    else
        hipLaunchKernelGGL((k_sample_weight<MODEL, L, false>), dim3(grid + mi.n), dim3(256), 0, s, state, N, R, frame,
                           esth_frame_arg{}, log_w, w, carry, mi);
}

template <int MODEL>
static void launch_sample_weight_t(int layout, int grid, float* state, uint32_t N, const float* R, const float* frame,
                                   const esth_frame_arg* fa, int log_w, float* w, int carry, const esth_mtgp_inline& mi,
                                   hipStream_t s) {
    if (layout == ESTH_LAYOUT_AOS)
        launch_sample_weight_l<MODEL, ESTH_LAYOUT_AOS>(grid, state, N, R, frame, fa, log_w, w, carry, mi, s);
    else if (layout == ESTH_LAYOUT_PACKED)
        launch_sample_weight_l<MODEL, ESTH_LAYOUT_PACKED>(grid, state, N, R, frame, fa, log_w, w, carry, mi, s);
    else
        launch_sample_weight_l<MODEL, ESTH_LAYOUT_SOA>(grid, state, N, R, frame, fa, log_w, w, carry, mi, s);
}

int esth_launch_sample_weight(int model, float* state, uint32_t N, const float* R, const float* frame, int log_w,
                              float* w, int carry, int layout, hipStream_t s, const float* host_frame,
                              const esth_mtgp_inline* mtgp) {
    const esth_mtgp_inline mi = mtgp ? *mtgp : esth_mtgp_inline{};
    esth_frame_arg fa_v;
    const esth_frame_arg* fa = nullptr;
    if (host_frame) {
        memcpy(fa_v.v, host_frame, sizeof(fa_v.v));
        fa = &fa_v;
    }
    const int grid = grid_for(N / 2, 256);
    ESTH_TRY(esth_with_model(model, [&](auto id) {
        launch_sample_weight_t<decltype(id)::value>(layout, grid, state, N, R, frame, fa, log_w, w, carry, mi, s);
        return ESTH_OK;
    }));
    ESTH_HIP(hipGetLastError());
    return ESTH_OK;
}

static int log2u(uint32_t m) { int k = 0; while ((1u << k) < m) ++k; return k; }

template <int LOGM, int S>
static void launch_sort_t(float* w, const float* st_in, float* st_out, uint32_t N, uint32_t G, float2* winner,
                          int* perm, uint32_t t, float* bt_w, float* bt_s, float* part, int log_w, int layout,
                          hipStream_t s) {
    using TL = esth_tile<LOGM>;
    const uint64_t waves = (G + (ESTH_WAVE / TL::LPS) - 1) / (ESTH_WAVE / TL::LPS);
    hipLaunchKernelGGL((k_sort<LOGM, S>), dim3(grid_for(waves * ESTH_WAVE, 256)), dim3(256), 0, s, w, st_in,
                       st_out, N, G, winner, perm, t, bt_w, bt_s, part, log_w, layout);
}

template <int LOGM, int S>
static void launch_resample_t(float* st_main, const float* st_tmp, const float* w, const float* rnd, uint32_t N,
                              uint32_t G, float thr, int systematic, int log_w, int* sel, int layout, hipStream_t s) {
    using TL = esth_tile<LOGM>;
    const uint64_t waves = (G + (ESTH_WAVE / TL::LPS) - 1) / (ESTH_WAVE / TL::LPS);
    hipLaunchKernelGGL((k_resample<LOGM, S>), dim3(grid_for(waves * ESTH_WAVE, 256)), dim3(256), 0, s, st_main,
                       st_tmp, w, rnd, N, G, thr, systematic, log_w, sel, layout);
}

#define ESTH_DISPATCH_LOGM(FN, S, ...)                 \
    switch (lm) {                                      \
    case 1: FN<1, S>(__VA_ARGS__); break;              \
    case 2: FN<2, S>(__VA_ARGS__); break;              \
    case 3: FN<3, S>(__VA_ARGS__); break;              \
    case 4: FN<4, S>(__VA_ARGS__); break;              \
    case 5: FN<5, S>(__VA_ARGS__); break;              \
    case 6: FN<6, S>(__VA_ARGS__); break;              \
    case 7: FN<7, S>(__VA_ARGS__); break;              \
    case 8: FN<8, S>(__VA_ARGS__); break;              \
    case 9: FN<9, S>(__VA_ARGS__); break;              \
    case 10: FN<10, S>(__VA_ARGS__); break;            \
    default: return esth_fail(ESTH_ERR_INVALID, "m"); \
    }

int esth_launch_sort(int S, uint32_t m, float* w, const float* st_in, float* st_out, uint32_t N, uint32_t G,
                     float2* winner, int* perm, uint32_t t, float* bt_w, float* bt_s, float* part, int log_w,
                     int layout, hipStream_t s) {
    const int lm = log2u(m);
#define ESTH_SORT_ARGS w, st_in, st_out, N, G, winner, perm, t, bt_w, bt_s, part, log_w, layout, s
    ESTH_TRY(esth_with_state_dim(S, [&](auto sv) {
        constexpr int SV = decltype(sv)::value;
        ESTH_DISPATCH_LOGM(launch_sort_t, SV, ESTH_SORT_ARGS);
        return ESTH_OK;
    }));
#undef ESTH_SORT_ARGS
    ESTH_HIP(hipGetLastError());
    return ESTH_OK;
}

template <int MODEL, int L, int LOGM>
static void launch_sample_sort_m(float* state, uint32_t N, uint32_t G, const float* R, const float* frame,
                                 const esth_frame_arg* fa, int log_w, float* w, int carry, float* st_out,
                                 float2* winner, uint32_t t, float* bt_w, float* bt_s, float* part, const float* gate,
                                 float keep_thr, int keep_all, const esth_mtgp_inline& mi, hipStream_t s) {
    if (fa)
        hipLaunchKernelGGL((k_sample_sort<MODEL, L, LOGM, true>), dim3(G + mi.n),
                           dim3(ESTH_WAVE * esth_ss_waves<LOGM>()), 0, s, state, N, G, R, nullptr, *fa, log_w, w, carry,
                           st_out, winner, t, bt_w, bt_s, part, gate, keep_thr, keep_all, mi);
    else
        hipLaunchKernelGGL((k_sample_sort<MODEL, L, LOGM, false>), dim3(G + mi.n),
                           dim3(ESTH_WAVE * esth_ss_waves<LOGM>()), 0, s, state, N, G, R, frame, esth_frame_arg{},
                           log_w, w, carry, st_out, winner, t, bt_w, bt_s, part, gate, keep_thr, keep_all, mi);
}

template <int MODEL, int L>
static int launch_sample_sort_l(int lm, float* state, uint32_t N, uint32_t G, const float* R,
                                const float* frame, const esth_frame_arg* fa, int log_w, float* w, int carry,
                                float* st_out, float2* winner, uint32_t t, float* bt_w, float* bt_s, float* part,
                                const float* gate, float keep_thr, int keep_all, const esth_mtgp_inline& mi,
                                hipStream_t s) {
#define ESTH_SS_ARGS state, N, G, R, frame, fa, log_w, w, carry, st_out, winner, t, bt_w, bt_s, part, gate, \
                     keep_thr, keep_all, mi, s
    switch (lm) {
    case 7: launch_sample_sort_m<MODEL, L, 7>(ESTH_SS_ARGS); break;
    case 8: launch_sample_sort_m<MODEL, L, 8>(ESTH_SS_ARGS); break;
    case 9: launch_sample_sort_m<MODEL, L, 9>(ESTH_SS_ARGS); break;
    case 10: launch_sample_sort_m<MODEL, L, 10>(ESTH_SS_ARGS); break;
    default: return esth_fail(ESTH_ERR_UNSUPPORTED, "fused sample+sort needs 128 <= m <= 1024");
    }
#undef ESTH_SS_ARGS
    return ESTH_OK;
}

template <int MODEL>
static int launch_sample_sort_t(int layout, int lm, float* state, uint32_t N, uint32_t G, const float* R,
                                const float* frame, const esth_frame_arg* fa, int log_w, float* w, int carry,
                                float* st_out, float2* winner, uint32_t t, float* bt_w, float* bt_s, float* part,
                                const float* gate, float keep_thr, int keep_all, const esth_mtgp_inline& mi,
                                hipStream_t s) {
#define ESTH_SS_ARGS lm, state, N, G, R, frame, fa, log_w, w, carry, st_out, winner, t, bt_w, bt_s, part, gate, \
                     keep_thr, keep_all, mi, s
    if (layout == ESTH_LAYOUT_AOS) return launch_sample_sort_l<MODEL, ESTH_LAYOUT_AOS>(ESTH_SS_ARGS);
    if (layout == ESTH_LAYOUT_PACKED) return launch_sample_sort_l<MODEL, ESTH_LAYOUT_PACKED>(ESTH_SS_ARGS);
    return launch_sample_sort_l<MODEL, ESTH_LAYOUT_SOA>(ESTH_SS_ARGS);
#undef ESTH_SS_ARGS
}

bool esth_sample_sort_fits(uint32_t m) { return m >= 128 && m <= 1024 && (m & (m - 1)) == 0; }
/* 256-thread work-groups (4 waves) with LDS for an MTGP32 ring: m >= 512 */
bool esth_sample_sort_inline_rng(uint32_t m) { return esth_sample_sort_fits(m) && m >= 512; }

/* resident waves per SIMD of a kernel launched with `threads`-thread
 * work-groups: 512 VGPRs per lane slot (granule 8), at most 8 waves, and
 * 160 KiB of LDS per CU shared by its work-groups; 0 for a kernel that spills
 * to scratch memory */
int esth_waves_per_simd(const void* fn, int threads) {
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, fn) != hipSuccess) return 0;
    if (a.localSizeBytes > 0) return 0;
    const int regs = a.numRegs > 0 ? (a.numRegs + 7) / 8 * 8 : 512;
    int occ = 512 / regs;
    if (occ > 8) occ = 8;
    if (a.sharedSizeBytes > 0) {
        const int by_lds = (int)((160u * 1024u) / a.sharedSizeBytes) * ((threads + ESTH_WAVE - 1) / ESTH_WAVE) / 4;
        if (by_lds < occ) occ = by_lds;
    }
    return occ;
}

template <int MODEL, int L>
static bool sample_sort_pays_l(int lm) {
    const void* fn = nullptr;
    switch (lm) {
    case 7: fn = (const void*)&k_sample_sort<MODEL, L, 7, true>; break;
    case 8: fn = (const void*)&k_sample_sort<MODEL, L, 8, true>; break;
    case 9: fn = (const void*)&k_sample_sort<MODEL, L, 9, true>; break;
    case 10: fn = (const void*)&k_sample_sort<MODEL, L, 10, true>; break;
    default: return false;
    }
    const int ss_waves[11] = {0, 0, 0, 0, 0, 0, 0, esth_ss_waves<7>(), esth_ss_waves<8>(), esth_ss_waves<9>(),
                              esth_ss_waves<10>()};
    const int fused = esth_waves_per_simd(fn, ESTH_WAVE * ss_waves[lm]);
    return fused > 0 &&
           (fused >= 3 || fused >= esth_waves_per_simd((const void*)&k_sample_weight<MODEL, L, true>, 256));
}

template <int MODEL>
static bool sample_sort_pays_t(int layout, int lm) {
    if (layout == ESTH_LAYOUT_AOS) return sample_sort_pays_l<MODEL, ESTH_LAYOUT_AOS>(lm);
    if (layout == ESTH_LAYOUT_PACKED) return sample_sort_pays_l<MODEL, ESTH_LAYOUT_PACKED>(lm);
    return sample_sort_pays_l<MODEL, ESTH_LAYOUT_SOA>(lm);
}

/* the fused launch is taken where it keeps 3 waves per SIMD resident, or as
 * many as the sampling alone.  Measured (whole steps, fused against two
 * launches): with 3-4 waves it wins even where the sampling alone keeps 5-8
 * (dh6 at m = 512 -11 %, robotarm5 and camera at 512 -8 %, unicycle at 1024
 * -9 %, randomwalk3 at 1024 -18 %: the state's round trip through HBM
 * between the two launches outweighs the occupancy), with 2 it loses or
 * ties (dh6 at 1024 +13 %, robotarm5 at 1024 even): the sort's registers
 * (E = m/64 keys, indices and gathered planes per lane) then leave too few
 * waves to hide the sampling's memory and arithmetic latency */
bool esth_sample_sort_pays(int model, int layout, uint32_t m) {
    if (!esth_sample_sort_fits(m)) return false;
    const int lm = log2u(m);
    return esth_with_model_or(model, false,
                              [&](auto id) { return sample_sort_pays_t<decltype(id)::value>(layout, lm); });
}

int esth_launch_sample_sort(int model, uint32_t m, float* state, uint32_t N, uint32_t G, const float* R,
                            const float* frame, int log_w, float* w, int carry, int layout, float* st_out,
                            float2* winner, uint32_t t, float* bt_w, float* bt_s, float* part, const float* gate,
                            float keep_thr, int keep_all, hipStream_t s, const float* host_frame,
                            const esth_mtgp_inline* mtgp) {
    if (!esth_sample_sort_fits(m) || (uint64_t)m * G != N)
        return esth_fail(ESTH_ERR_UNSUPPORTED, "fused sample+sort needs 128 <= m <= 1024 and N = m G");
    const int lm = log2u(m);
    const esth_mtgp_inline mi = mtgp ? *mtgp : esth_mtgp_inline{};
    if (mi.n && !esth_sample_sort_inline_rng(m))
        return esth_fail(ESTH_ERR_UNSUPPORTED, "MTGP32 work-groups ride only in 256-thread sample+sort launches");
    esth_frame_arg fa_v;
    const esth_frame_arg* fa = nullptr;
    if (host_frame) {
        memcpy(fa_v.v, host_frame, sizeof(fa_v.v));
        fa = &fa_v;
    }
#define ESTH_SS_ARGS layout, lm, state, N, G, R, frame, fa, log_w, w, carry, st_out, winner, t, bt_w, bt_s, part, \
                     gate, keep_thr, keep_all, mi, s
    const int rc = esth_with_model(model, [&](auto id) { return launch_sample_sort_t<decltype(id)::value>(ESTH_SS_ARGS); });
#undef ESTH_SS_ARGS
    if (rc != ESTH_OK) return rc;
    ESTH_HIP(hipGetLastError());
    return ESTH_OK;
}

int esth_launch_step_x(int model, uint32_t m, float* tmp, float* w, const float* rnd_k, float thr, int systematic,
                       const float* bt_w_k, const float* bt_s_k, const float2* winner_k, const float* part_k,
                       float2* result, double* est, float* state, uint32_t N, uint32_t G, uint32_t t, int topo,
                       int dimx, int dimy, const float* R_k1, const float* frame, const float* host_frame, int log_w,
                       float2* winner_k1, float* bt_w_k1, float* bt_s_k1, float* part_k1, const float* gate_k1,
                       const esth_mtgp_inline* mtgp, hipStream_t s) {
    if (model != ESTH_MODEL_UNICYCLE || m != 512 || (uint64_t)m * G != N || (!host_frame && !frame) || t == 0 ||
        (topo != ESTH_TOPO_RING && topo != ESTH_TOPO_TORUS))
        return esth_fail(ESTH_ERR_UNSUPPORTED, "cross-step launch: unicycle, m = 512, ring or torus with t > 0");
    const esth_mtgp_inline mi = mtgp ? *mtgp : esth_mtgp_inline{};
    esth_frame_arg fa;
    if (host_frame) memcpy(fa.v, host_frame, sizeof(fa.v));
    else memset(fa.v, 0, sizeof(fa.v));
    const esth_xch xk{bt_w_k, bt_s_k, nullptr, w, G, t, topo == ESTH_TOPO_TORUS ? 4u : 1u, topo, dimx, dimy};
    if (host_frame)
        hipLaunchKernelGGL((k_step_x<ESTH_MODEL_UNICYCLE, 9, true>), dim3(mi.n + 1 + G), dim3(256), 0, s, tmp, w, rnd_k,
                           thr, systematic, xk, winner_k, part_k, bt_s_k, result, est, state, N, G, R_k1, nullptr, fa,
                           log_w, winner_k1, bt_w_k1, bt_s_k1, part_k1, gate_k1, thr, mi);
    else
        hipLaunchKernelGGL((k_step_x<ESTH_MODEL_UNICYCLE, 9, false>), dim3(mi.n + 1 + G), dim3(256), 0, s, tmp, w, rnd_k,
                           thr, systematic, xk, winner_k, part_k, bt_s_k, result, est, state, N, G, R_k1, frame, fa,
                           log_w, winner_k1, bt_w_k1, bt_s_k1, part_k1, gate_k1, thr, mi);
    ESTH_HIP(hipGetLastError());
    return ESTH_OK;
}

int esth_launch_argmax(int S, const float2* winner, uint32_t G, int log_w, const float* part, const float* sorted,
                       uint32_t N, uint32_t m, float2* result, double* est, hipStream_t s) {
    ESTH_TRY(esth_with_state_dim(S, [&](auto sv) {
        hipLaunchKernelGGL(k_argmax<decltype(sv)::value>, dim3(1), dim3(1024), 0, s, winner, G, log_w, part, sorted, N, m,
                           result, est);
        return ESTH_OK;
    }));
    ESTH_HIP(hipGetLastError());
    return ESTH_OK;
}

int esth_launch_topt(const float* bt_w, uint32_t n, uint32_t t, uint32_t* top, hipStream_t s) {
    hipLaunchKernelGGL(k_topt, dim3(1), dim3(1024), 0, s, bt_w, n, t, top);
    ESTH_HIP(hipGetLastError());
    return ESTH_OK;
}

int esth_launch_exchange(float* st, float* w, int S, uint32_t N, uint32_t m, uint32_t G, uint32_t t, int topo,
                         int dimx, int dimy, const float* bt_w, const float* bt_s, uint32_t* top, hipStream_t s) {
    if (t == 0 || topo == ESTH_TOPO_NONE) return ESTH_OK;
    if (topo == ESTH_TOPO_ALL2ALL) {
        hipLaunchKernelGGL(k_topt, dim3(1), dim3(1024), 0, s, bt_w, G * t, t, top);
        ESTH_HIP(hipGetLastError());
    }
    const uint64_t total = (uint64_t)G * (topo == ESTH_TOPO_TORUS ? 4u : 1u) * t;
    hipLaunchKernelGGL(k_exchange, dim3(grid_for(total, 256)), dim3(256), 0, s, st, w, S, N, m, G, t, topo, dimx,
                       dimy, bt_w, bt_s, top);
    ESTH_HIP(hipGetLastError());
    return ESTH_OK;
}

template <int LOGM, int S>
static void launch_alias_w_t(float* st_main, const float* st_tmp, const float* w, const float* rnd, uint32_t N,
                             uint32_t G, float thr, int adaptive, int log_w, int* sel, float* prob_out,
                             uint32_t* alias_out, int carry, int layout, hipStream_t s) {
    if constexpr (LOGM >= 6) {
        constexpr int WPB = esth_alias_wpb<LOGM>::value;
        hipLaunchKernelGGL((k_alias_w<LOGM, S>), dim3((G + WPB - 1) / WPB), dim3(WPB * ESTH_WAVE), 0, s, st_main, st_tmp,
                           w, rnd, N, G, thr, adaptive, log_w, sel, prob_out, alias_out, carry, layout);
    }
}

/* Vose alias: one wave per sub-filter for m >= 64, one work-group of m below */
static int launch_alias(int S, float* st_main, const float* st_tmp, const float* w, const float* rnd, uint32_t N,
                        uint32_t m, uint32_t G, float thr, int adaptive, int log_w, int* sel, float* prob_out,
                        uint32_t* alias_out, int carry, int layout, hipStream_t s) {
    if (m >= ESTH_WAVE) {
        const int lm = log2u(m);
#define ESTH_AW_ARGS st_main, st_tmp, w, rnd, N, G, thr, adaptive, log_w, sel, prob_out, alias_out, carry, layout, s
        ESTH_TRY(esth_with_state_dim(S, [&](auto sv) {
            constexpr int SV = decltype(sv)::value;
            ESTH_DISPATCH_LOGM(launch_alias_w_t, SV, ESTH_AW_ARGS);
            return ESTH_OK;
        }));
#undef ESTH_AW_ARGS
        ESTH_HIP(hipGetLastError());
        return ESTH_OK;
    }
    ESTH_TRY(esth_with_state_dim(S, [&](auto sv) {
        constexpr int SV = decltype(sv)::value;
        hipLaunchKernelGGL(k_alias<SV>, dim3(G), dim3(m), 0, s, st_main, st_tmp, w, rnd, N, m, thr, adaptive, log_w, sel, prob_out, alias_out, carry, layout);
        return ESTH_OK;
    }));
    ESTH_HIP(hipGetLastError());
    return ESTH_OK;
}

int esth_launch_resample(int resampler, int S, float* st_main, const float* st_tmp, const float* w, const float* rnd,
                         uint32_t N, uint32_t m, uint32_t G, float thr, int adaptive, int log_w, int* sel, int carry,
                         int layout, hipStream_t s) {
    if (resampler == ESTH_RESAMPLE_ALIAS)
        return launch_alias(S, st_main, st_tmp, w, rnd, N, m, G, thr, adaptive, log_w, sel, nullptr, nullptr, carry,
                            layout, s);
    const int lm = log2u(m);
    const int sys = resampler == ESTH_RESAMPLE_SYSTEMATIC;
    ESTH_TRY(esth_with_state_dim(S, [&](auto sv) {
        constexpr int SV = decltype(sv)::value;
        ESTH_DISPATCH_LOGM(launch_resample_t, SV, st_main, st_tmp, w, rnd, N, G, thr, sys, log_w, sel, layout, s);
        return ESTH_OK;
    }));
    ESTH_HIP(hipGetLastError());
    return ESTH_OK;
}

template <int LOGM, int S>
static void launch_resample_argmax_t(float* st_main, const float* st_tmp, const float* w, const float* rnd, uint32_t N,
                                     uint32_t G, float thr, int systematic, int log_w, const float2* winner,
                                     const float* part, const float* best_src, size_t plane_stride, uint32_t g_stride,
                                     float2* result, double* est, const esth_xch& xch, int carry, int layout,
                                     hipStream_t s) {
    /* two waves per sub-filter for m >= 512 where the work-group's two
     * sub-filters' CDFs and staged planes fit in 48 KB of LDS (at m = 256 in
     * the batched classes the halves were too short to pay for the barriers) */
    if constexpr (LOGM >= 9 && (1 + S) * (1 << LOGM) <= 6144) {
        hipLaunchKernelGGL((k_resample2_argmax<LOGM, S>), dim3(1 + (G + 1) / 2), dim3(256), 0, s, st_main, st_tmp, w,
                           rnd, N, G, thr, systematic, log_w, winner, part, best_src, plane_stride, g_stride, result,
                           est, xch, carry, layout);
    } else {
        using TL = esth_tile<LOGM>;
        const uint64_t waves = (G + (ESTH_WAVE / TL::LPS) - 1) / (ESTH_WAVE / TL::LPS);
        hipLaunchKernelGGL((k_resample_argmax<LOGM, S>), dim3(1 + grid_for(waves * ESTH_WAVE, 256)), dim3(256), 0, s,
                           st_main, st_tmp, w, rnd, N, G, thr, systematic, log_w, winner, part, best_src,
                           plane_stride, g_stride, result, est, xch, carry, layout);
    }
}

int esth_launch_resample_argmax(int resampler, int S, float* st_main, const float* st_tmp, float* w,
                                const float* rnd, uint32_t N, uint32_t m, uint32_t G, float thr, int log_w,
                                const float2* winner, const float* part, const float* best_src, size_t plane_stride,
                                uint32_t g_stride, float2* result, double* est, uint32_t t, int topo, int dimx,
                                int dimy, const float* bt_w, const float* bt_s, uint32_t* top, int carry,
                                int layout, hipStream_t s) {
    if (resampler == ESTH_RESAMPLE_ALIAS) return esth_fail(ESTH_ERR_UNSUPPORTED, "fused argmax: scan resamplers only");
    const int lm = log2u(m);
    const int sys = resampler == ESTH_RESAMPLE_SYSTEMATIC;
    esth_xch xch{};
    if (t > 0 && topo != ESTH_TOPO_NONE) {
        if (topo == ESTH_TOPO_ALL2ALL) {   /* the global top t the receivers read */
            hipLaunchKernelGGL(k_topt, dim3(1), dim3(1024), 0, s, bt_w, G * t, t, top);
            ESTH_HIP(hipGetLastError());
        }
        xch = {bt_w, bt_s, top, w, G, t, topo == ESTH_TOPO_TORUS ? 4u : 1u, topo, dimx, dimy};
    }
#define ESTH_RA_ARGS st_main, st_tmp, w, rnd, N, G, thr, sys, log_w, winner, part, best_src, plane_stride, g_stride, result, est, xch, carry, layout, s
    ESTH_TRY(esth_with_model_state_dim(S, [&](auto sv) {
        constexpr int SV = decltype(sv)::value;
        ESTH_DISPATCH_LOGM(launch_resample_argmax_t, SV, ESTH_RA_ARGS);
        return ESTH_OK;
    }));
#undef ESTH_RA_ARGS
    ESTH_HIP(hipGetLastError());
    return ESTH_OK;
}

template <int LOGM, int S>
static void launch_alias_argmax_t(float* st_main, const float* st_tmp, const float* w, const float* rnd, uint32_t N,
                                  uint32_t G, float thr, int adaptive, int log_w, const float2* winner,
                                  const float* part, const float* best_src, size_t plane_stride, uint32_t g_stride,
                                  float2* result, double* est, const esth_xch& xch, int carry, int layout,
                                  hipStream_t s) {
    if constexpr (LOGM >= 6) {
        constexpr int WPB = esth_alias_wpb<LOGM>::value;
        hipLaunchKernelGGL((k_alias_argmax_w<LOGM, S>), dim3(1 + (G + WPB - 1) / WPB), dim3(WPB * ESTH_WAVE), 0, s,
                           st_main, st_tmp, w, rnd, N, G, thr, adaptive, log_w, winner, part, best_src, plane_stride,
                           g_stride, result, est, xch, carry, layout);
    }
}

int esth_launch_alias_argmax(int S, float* st_main, const float* st_tmp, float* w, const float* rnd, uint32_t N,
                             uint32_t m, uint32_t G, float thr, int adaptive, int log_w, const float2* winner,
                             const float* part, const float* best_src, size_t plane_stride, uint32_t g_stride,
                             float2* result, double* est, uint32_t t, int topo, int dimx, int dimy, const float* bt_w,
                             const float* bt_s, uint32_t* top, int carry, int layout, hipStream_t s) {
    if (m < ESTH_WAVE) return esth_fail(ESTH_ERR_UNSUPPORTED, "fused alias: m >= 64");
    const int lm = log2u(m);
    esth_xch xch{};
    if (t > 0 && topo != ESTH_TOPO_NONE) {
        if (topo == ESTH_TOPO_ALL2ALL) {   /* the global top t the receivers read */
            hipLaunchKernelGGL(k_topt, dim3(1), dim3(1024), 0, s, bt_w, G * t, t, top);
            ESTH_HIP(hipGetLastError());
        }
        xch = {bt_w, bt_s, top, w, G, t, topo == ESTH_TOPO_TORUS ? 4u : 1u, topo, dimx, dimy};
    }
#define ESTH_AA_ARGS st_main, st_tmp, w, rnd, N, G, thr, adaptive, log_w, winner, part, best_src, plane_stride, g_stride, result, est, xch, carry, layout, s
    ESTH_TRY(esth_with_model_state_dim(S, [&](auto sv) {
        constexpr int SV = decltype(sv)::value;
        ESTH_DISPATCH_LOGM(launch_alias_argmax_t, SV, ESTH_AA_ARGS);
        return ESTH_OK;
    }));
#undef ESTH_AA_ARGS
    ESTH_HIP(hipGetLastError());
    return ESTH_OK;
}

/* ======================================================= device helpers */
extern "C" int esth_dev_alloc(size_t bytes, void** p) {
    if (!p) return esth_fail(ESTH_ERR_INVALID, "null out");
    ESTH_HIP(hipMalloc(p, bytes ? bytes : 4));
    return ESTH_OK;
}
extern "C" int esth_dev_free(void* p) { ESTH_HIP(hipFree(p)); return ESTH_OK; }
extern "C" int esth_memcpy_htod(void* d, const void* s, size_t n) { ESTH_HIP(hipMemcpy(d, s, n, hipMemcpyHostToDevice)); return ESTH_OK; }
extern "C" int esth_memcpy_dtoh(void* d, const void* s, size_t n) { ESTH_HIP(hipMemcpy(d, s, n, hipMemcpyDeviceToHost)); return ESTH_OK; }
extern "C" int esth_memset_d(void* d, int v, size_t n) { ESTH_HIP(hipMemset(d, v, n)); return ESTH_OK; }
extern "C" int esth_dev_sync(void) { ESTH_HIP(hipDeviceSynchronize()); return ESTH_OK; }

/* ================================================================ stages */
extern "C" int esth_stage_mtgp(uint32_t* d_status, int nstreams, int first_set, int size, void* d_out, int float_mode) {
    if (nstreams <= 0 || first_set < 0 || first_set + nstreams > ESTH_MTGP_SETS)
        return esth_fail(ESTH_ERR_INVALID, "streams %d from set %d", nstreams, first_set);
    if (size <= 0 || size % ESTH_MTGP_LARGE) return esth_fail(ESTH_ERR_INVALID, "size %d not a multiple of 768", size);
    std::vector<esth_mtgp_stream> hs(nstreams);
    for (int j = 0; j < nstreams; ++j) hs[j] = {(uint32_t)(first_set + j), (uint32_t)size, (uint64_t)j * (uint64_t)size, (uint64_t)j};
    esth_mtgp_stream* ds = nullptr;
    uint32_t* dp = nullptr;
    ESTH_HIP(hipMalloc((void**)&ds, sizeof(esth_mtgp_stream) * nstreams));
    ESTH_HIP(hipMalloc((void**)&dp, sizeof(g_mtgp_table)));
    ESTH_HIP(hipMemcpy(ds, hs.data(), sizeof(esth_mtgp_stream) * nstreams, hipMemcpyHostToDevice));
    ESTH_HIP(hipMemcpy(dp, g_mtgp_table, sizeof(g_mtgp_table), hipMemcpyHostToDevice));
    int rc = esth_launch_mtgp(dp, d_status, d_status, ds, nstreams, d_out, float_mode, 0);
    ESTH_HIP(hipDeviceSynchronize());
    hipFree(ds);
    hipFree(dp);
    return rc;
}

extern "C" int esth_stage_mtgp_steps(const uint32_t* d_status_in, uint32_t* d_status_out, int nstreams, int first_set,
                                     int size, int nsteps, void* d_out, int float_mode) {
    if (nstreams <= 0 || first_set < 0 || first_set + nstreams > ESTH_MTGP_SETS)
        return esth_fail(ESTH_ERR_INVALID, "streams %d from set %d", nstreams, first_set);
    if (size <= 0 || size % ESTH_MTGP_LARGE) return esth_fail(ESTH_ERR_INVALID, "size %d not a multiple of 768", size);
    if (nsteps <= 0 || !d_status_in || !d_status_out || !d_out) return esth_fail(ESTH_ERR_INVALID, "steps/buffers");
    std::vector<esth_mtgp_stream> hs(nstreams);
    for (int j = 0; j < nstreams; ++j) hs[j] = {(uint32_t)(first_set + j), (uint32_t)size, (uint64_t)j * (uint64_t)size, (uint64_t)j};
    esth_mtgp_stream* ds = nullptr;
    uint32_t* dp = nullptr;
    ESTH_HIP(hipMalloc((void**)&ds, sizeof(esth_mtgp_stream) * nstreams));
    ESTH_HIP(hipMalloc((void**)&dp, sizeof(g_mtgp_table)));
    ESTH_HIP(hipMemcpy(ds, hs.data(), sizeof(esth_mtgp_stream) * nstreams, hipMemcpyHostToDevice));
    ESTH_HIP(hipMemcpy(dp, g_mtgp_table, sizeof(g_mtgp_table), hipMemcpyHostToDevice));
    int rc = esth_launch_mtgp(dp, d_status_in, d_status_out, ds, nstreams, d_out, float_mode, 0, nsteps,
                              (uint64_t)nstreams * (uint64_t)size, (uint64_t)nstreams * ESTH_MTGP_N);
    ESTH_HIP(hipDeviceSynchronize());
    hipFree(ds);
    hipFree(dp);
    return rc;
}

extern "C" int esth_stage_boxmuller(float* d, int n) {
    if (n <= 0) return ESTH_OK;
    hipLaunchKernelGGL(k_boxmuller, dim3(grid_for((uint64_t)(n + 1) / 2, 256)), dim3(256), 0, 0, d, n);
    ESTH_HIP(hipGetLastError());
    ESTH_HIP(hipDeviceSynchronize());
    return ESTH_OK;
}

extern "C" int esth_stage_sample_weight(int model, float* d_state, uint32_t N, const float* d_random, const float* frame,
                                        int weights, float* d_w) {
    if (N == 0 || (N & 1u)) return esth_fail(ESTH_ERR_INVALID, "N=%u must be even and positive", N);
    if (!frame) return esth_fail(ESTH_ERR_INVALID, "null frame");
    /* the frame travels as a kernel argument */
    int rc = esth_launch_sample_weight(model, d_state, N, d_random, nullptr, weights == ESTH_WEIGHTS_LOG, d_w, 0,
                                       ESTH_LAYOUT_SOA, 0, frame);
    ESTH_HIP(hipDeviceSynchronize());
    return rc;
}

extern "C" int esth_stage_sort(float* d_w, const float* d_in, float* d_out, int S, uint32_t m, uint32_t G, float* d_winner,
                               int* d_perm) {
    if (!is_pow2(m) || m < 2 || m > ESTH_MAX_NUM_PARTICLES) return esth_fail(ESTH_ERR_INVALID, "m=%u", m);
    if (!esth_state_dim_known(S))
        return esth_fail(ESTH_ERR_UNSUPPORTED, "state dimension %d", S);
    const uint32_t N = m * G;
    float* part = nullptr;
    float2* win = nullptr;
    ESTH_HIP(hipMalloc((void**)&part, sizeof(float) * (size_t)G * (2 + S)));
    if (!d_winner) ESTH_HIP(hipMalloc((void**)&win, sizeof(float2) * G));
    int rc = esth_launch_sort(S, m, d_w, d_in, d_out, N, G, d_winner ? (float2*)d_winner : win, d_perm, 0, nullptr,
                              nullptr, part, 0, ESTH_LAYOUT_SOA, 0);
    ESTH_HIP(hipDeviceSynchronize());
    hipFree(part);
    if (win) hipFree(win);
    return rc;
}

/* ----------------------------------------- allocation-free block sort
 * esth_stage_sort_noalloc runs esth_stage_sort's kernel with its scratch (the
 * per-sub-filter partial sums, G (2 + S) floats, and the winners, 2 G floats,
 * when the caller passes none) in this file-scope device array instead of
 * hipMalloc'd buffers.  Each GPU holds its own copy of the array, at its own
 * address, for the life of the process; host code reaches it only through
 * the address hipGetSymbolAddress returns for the current device, kept per
 * device in g_scratch_dev (the array's own name in host code is not a device
 * address).  The sort writes every scratch word it later reads, so nothing
 * needs clearing between calls.  Calls that overlap on one GPU share the
 * array and race. */
#define ESTH_NOALLOC_SCRATCH_FLOATS (1u << 19)   /* 2 MiB per GPU */
#define ESTH_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) float scratch[ESTH_NOALLOC_SCRATCH_FLOATS];
static float* g_scratch_dev[ESTH_NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

/* the current device's scratch address; looked up once per device (at
 * engine creation, or at the first allocation-free call on that device) */
static int esth_noalloc_scratch(float** out) {
    int dev = -1;
    ESTH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= ESTH_NOALLOC_MAX_DEVICES) return esth_fail(ESTH_ERR_UNSUPPORTED, "device %d", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !p)
            return esth_fail(ESTH_ERR_DEVICE, "scratch address on device %d: %s", dev, hipGetErrorString(err));
        g_scratch_dev[dev] = (float*)p;
    }
    *out = g_scratch_dev[dev];
    return ESTH_OK;
}

extern "C" int esth_stage_sort_noalloc(float* d_w, const float* d_in, float* d_out, int S, uint32_t m, uint32_t G,
                                       float* d_winner, int* d_perm, void* stream) {
    if (!is_pow2(m) || m < 2 || m > ESTH_MAX_NUM_PARTICLES) return esth_fail(ESTH_ERR_INVALID, "m=%u", m);
    if (G == 0 || G > ESTH_MAX_NUM_BLOCKS) return esth_fail(ESTH_ERR_INVALID, "G=%u", G);
    if (!esth_state_dim_known(S))
        return esth_fail(ESTH_ERR_UNSUPPORTED, "state dimension %d", S);
    if (!d_w || !d_in || !d_out) return esth_fail(ESTH_ERR_INVALID, "null buffer");
    const uint64_t need = (uint64_t)G * (2 + S) + (d_winner ? 0 : 2ull * G);
    if (need > ESTH_NOALLOC_SCRATCH_FLOATS)
        return esth_fail(ESTH_ERR_BUDGET, "allocation-free sort: G=%u, S=%d needs %llu scratch floats, at most %u",
                         G, S, (unsigned long long)need, ESTH_NOALLOC_SCRATCH_FLOATS);
    float* sc = nullptr;
    ESTH_TRY(esth_noalloc_scratch(&sc));
    float* part = sc;
    float2* win = d_winner ? (float2*)d_winner : (float2*)(sc + (size_t)G * (2 + S));
    return esth_launch_sort(S, m, d_w, d_in, d_out, m * G, G, win, d_perm, 0, nullptr, nullptr, part, 0,
                            ESTH_LAYOUT_SOA, (hipStream_t)stream);
}

extern "C" int esth_stage_argmax(const float* d_winner, uint32_t G, int weights, float* d_result) {
    if (G == 0) return esth_fail(ESTH_ERR_INVALID, "G=0");
    float* part = nullptr;
    double* est = nullptr;
    ESTH_HIP(hipMalloc((void**)&part, sizeof(float) * (size_t)G * 3));
    ESTH_HIP(hipMemset(part, 0, sizeof(float) * (size_t)G * 3));
    ESTH_HIP(hipMalloc((void**)&est, sizeof(double) * 8));
    int rc = esth_launch_argmax(1, (const float2*)d_winner, G, weights == ESTH_WEIGHTS_LOG, part, nullptr, 0, 1,
                                (float2*)d_result, est, 0);
    ESTH_HIP(hipDeviceSynchronize());
    hipFree(part);
    hipFree(est);
    return rc;
}

extern "C" int esth_stage_exchange(float* d_state, float* d_w, int S, uint32_t m, uint32_t G, uint32_t t, int topo,
                                   int dimx, int dimy) {
    if (t == 0 || topo == ESTH_TOPO_NONE) return ESTH_OK;
    esth_filter_desc d;
    esth_filter_desc_default(&d, ESTH_MODEL_ROBOTARM5);
    d.num_particles = m; d.num_blocks = G; d.num_transfer = t; d.topology = topo; d.torus_dimx = dimx; d.torus_dimy = dimy;
    int s_, v_;
    ESTH_TRY(esth_check_desc(&d, &s_, &v_));
    if (topo == ESTH_TOPO_TORUS && (dimx <= 0 || dimy <= 0)) esth_torus_dims(G, &dimx, &dimy);
    const uint32_t N = m * G;
    float *bw = nullptr, *bs = nullptr;
    uint32_t* top = nullptr;
    ESTH_HIP(hipMalloc((void**)&bw, sizeof(float) * (size_t)G * t));
    ESTH_HIP(hipMalloc((void**)&bs, sizeof(float) * (size_t)G * t * S));
    ESTH_HIP(hipMalloc((void**)&top, sizeof(uint32_t) * t));
    /* compact the best t of every (already sorted) block */
    for (uint32_t k = 0; k < t; ++k) {
        ESTH_HIP(hipMemcpy2D(bw + k, sizeof(float) * t, d_w + k, sizeof(float) * m, sizeof(float), G, hipMemcpyDeviceToDevice));
        for (int c = 0; c < S; ++c)
            ESTH_HIP(hipMemcpy2D(bs + (size_t)c * G * t + k, sizeof(float) * t, d_state + (size_t)c * N + k, sizeof(float) * m,
                                 sizeof(float), G, hipMemcpyDeviceToDevice));
    }
    int rc = esth_launch_exchange(d_state, d_w, S, N, m, G, t, topo, dimx, dimy, bw, bs, top, 0);
    ESTH_HIP(hipDeviceSynchronize());
    hipFree(bw);
    hipFree(bs);
    hipFree(top);
    return rc;
}

extern "C" int esth_stage_resample(int resampler, float* dst, const float* src, int S, const float* d_w, const float* d_random,
                                   uint32_t m, uint32_t G, float thr, int adaptive, int weights, int* d_sel) {
    esth_filter_desc d;
    esth_filter_desc_default(&d, ESTH_MODEL_ROBOTARM5);
    d.num_particles = m; d.num_blocks = G; d.num_transfer = 0; d.resampler = resampler;
    d.resampling_threshold = thr; d.adaptive = adaptive; d.weights = weights;
    int s_, v_;
    ESTH_TRY(esth_check_desc(&d, &s_, &v_));
    int rc = esth_launch_resample(resampler, S, dst, src, d_w, d_random, m * G, m, G, thr, adaptive,
                                  weights == ESTH_WEIGHTS_LOG, d_sel, 0, ESTH_LAYOUT_SOA, 0);
    ESTH_HIP(hipDeviceSynchronize());
    return rc;
}

extern "C" int esth_stage_alias_table(const float* d_w, uint32_t m, uint32_t G, int weights, float* d_prob, uint32_t* d_alias) {
    if (!is_pow2(m) || m < 2 || m > ESTH_MAX_NUM_PARTICLES) return esth_fail(ESTH_ERR_INVALID, "m=%u", m);
    if (!d_prob || !d_alias) return esth_fail(ESTH_ERR_INVALID, "null output");
    ESTH_TRY(launch_alias(1, nullptr, nullptr, d_w, nullptr, m * G, m, G, 1.0f, 0, weights == ESTH_WEIGHTS_LOG, nullptr,
                          d_prob, d_alias, 0, ESTH_LAYOUT_SOA, 0));
    ESTH_HIP(hipDeviceSynchronize());
    return ESTH_OK;
}

extern "C" int esth_stage_copy(void* dst, const void* src, size_t bytes) {
    if (bytes % 16) return esth_fail(ESTH_ERR_INVALID, "copy size %zu not a multiple of 16", bytes);
    const size_t n = bytes / 16;
    int grid = (int)std::min<size_t>((n + 255) / 256, 256 * 64);
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, (float4*)dst, (const float4*)src, n);
    ESTH_HIP(hipGetLastError());
    ESTH_HIP(hipDeviceSynchronize());
    return ESTH_OK;
}

extern "C" int esth_stage_math(int fn, const float* x, const float* x2, float* out, int n) {
    if (n <= 0) return ESTH_OK;
    if (fn < 0 || fn > 7) return esth_fail(ESTH_ERR_INVALID, "math function %d", fn);
    if ((fn == 4 || fn == 5 || fn == 6) && !x2) return esth_fail(ESTH_ERR_INVALID, "function %d needs x2", fn);
    hipLaunchKernelGGL(k_math, dim3(grid_for((uint64_t)n, 256)), dim3(256), 0, 0, fn, x, x2, out, n);
    ESTH_HIP(hipGetLastError());
    ESTH_HIP(hipDeviceSynchronize());
    return ESTH_OK;
}
