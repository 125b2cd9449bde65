/*
 * esth_filter.hip - one island-model filter on one GPU: set-up and the step
 * of particle_filter() (src/partfilt/opencl/pf_generic.c:245-887) as a
 * sequence of kernels on the filter's stream, with no host synchronisation
 * inside a step (the reference blocks three times per step: pf_generic.c:231,
 * 614, 806).
 */
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "esth_internal.h"

#define ESTH_FRAME_RING 16
#define ESTH_TIMING_STEPS 4096

struct esth_filter {
    esth_engine* e = nullptr;
    esth_filter_desc d{};
    int S = 0, V = 0, K = 0;
    uint32_t m = 0, G = 0, t = 0;
    size_t N = 0;
    uint64_t needed = 0, size = 0, words = 0;
    int dimx = 0, dimy = 0;
    unsigned char params[sizeof(esth_unicycle_params) > sizeof(esth_dh6_params) ? sizeof(esth_unicycle_params)
                                                                                 : sizeof(esth_dh6_params)];
    hipStream_t stream = nullptr;
    hipStream_t rng_stream = nullptr;   /* MTGP of the next step runs here, overlapped */
    /* device */
    float *state = nullptr, *tmp = nullptr, *w = nullptr, *bt_w = nullptr, *bt_s = nullptr,
          *part = nullptr, *frames = nullptr;
    /* random words: R[k & 1] holds step k's words, generated from status[k & 1]
     * into status[(k + 1) & 1] */
    float* R[2] = {nullptr, nullptr};
    uint32_t* status[2] = {nullptr, nullptr};
    bool rng_ready = false;             /* R[steps & 1] already holds the next step's words */
    bool prefetch = true;
    bool fuse_sample_sort = false;      /* sample+weight and sort in one launch (k_sample_sort) */
    hipEvent_t ev_rng_done[2], ev_consumed[2], ev_fork;
    float2 *winner = nullptr, *result = nullptr;
    double* est = nullptr;
    uint32_t* top = nullptr;
    esth_mtgp_stream* streams = nullptr;
    std::vector<std::pair<void*, size_t>> allocs;
    /* host */
    float* h_frames = nullptr;   /* pinned ring */
    hipEvent_t frame_ev[ESTH_FRAME_RING];   /* graph replay: per parity slot */
    bool frame_used[ESTH_FRAME_RING] = {};
    hipEvent_t ring_ev;                     /* eager: after the last copy of a pass over the ring */
    bool ring_pending = false;
    int slot = 0;
    hipEvent_t ev[7];
    int timing = 0;   /* 0 off, 1 sample+weight only, 2 every stage */
    int timing_every = 1;   /* time one step in this many */
    uint64_t tcount = 0;
    /* per-step stage events while timing is on: 7 per step, up to ESTH_TIMING_STEPS steps */
    std::vector<hipEvent_t> tev;
    int tsteps = 0;
    uint64_t steps = 0;
    /* graph replay */
    bool use_graph = false;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int gslot[2] = {-1, -1};
};

static int falloc(esth_filter* f, size_t bytes, void** p) {
    ESTH_TRY(esth_engine_alloc(f->e, bytes, p));
    f->allocs.push_back({*p, bytes});
    return ESTH_OK;
}

static void filter_free(esth_filter* f) {
    if (!f) return;
    hipSetDevice(f->e->device);
    if (f->stream) hipStreamSynchronize(f->stream);
    for (int i = 0; i < 2; ++i)
        if (f->gexec[i]) hipGraphExecDestroy(f->gexec[i]);
    for (auto& a : f->allocs) esth_engine_release(f->e, a.first, a.second);
    f->allocs.clear();
    if (f->h_frames) hipHostFree(f->h_frames);
    for (int i = 0; i < ESTH_FRAME_RING; ++i) hipEventDestroy(f->frame_ev[i]);
    hipEventDestroy(f->ring_ev);
    for (int i = 0; i < 7; ++i) hipEventDestroy(f->ev[i]);
    for (int i = 0; i < 2; ++i) { hipEventDestroy(f->ev_rng_done[i]); hipEventDestroy(f->ev_consumed[i]); }
    hipEventDestroy(f->ev_fork);
    if (f->rng_stream) hipStreamDestroy(f->rng_stream);
    for (auto ev : f->tev) hipEventDestroy(ev);
    if (f->stream) hipStreamDestroy(f->stream);
    delete f;
}

extern "C" int esth_filter_create(esth_engine* e, const esth_filter_desc* d, const void* params, size_t params_size,
                                  esth_filter** out) {
    if (!e || !out) return esth_fail(ESTH_ERR_INVALID, "null engine/out");
    int S, V;
    ESTH_TRY(esth_check_desc(d, &S, &V));
    size_t need = 0;
    ESTH_TRY(esth_params_size(d->model, &need));
    if (need && (!params || params_size != need))
        return esth_fail(ESTH_ERR_INVALID, "model %d needs %zu bytes of params, got %zu", d->model, need, params_size);
    ESTH_HIP(hipSetDevice(e->device));
    esth_filter* f = new esth_filter();
    f->e = e;
    f->d = *d;
    f->S = S;
    f->V = V;
    f->m = d->num_particles;
    f->G = d->num_blocks;
    f->t = (d->topology == ESTH_TOPO_NONE) ? 0 : d->num_transfer;
    f->N = (size_t)f->m * f->G;
    f->fuse_sample_sort = esth_sample_sort_pays(d->model, d->state_layout, f->m);
    memset(f->params, 0, sizeof(f->params));
    if (need) memcpy(f->params, params, need);
    f->needed = esth_rng_needed(V, d->resampler, d->adaptive, f->m, f->G);
    f->K = esth_rng_streams(d->rng, f->needed);
    f->size = esth_rng_stream_size(f->K, f->needed);
    f->words = f->size * (uint64_t)f->K;
    f->dimx = d->torus_dimx;
    f->dimy = d->torus_dimy;
    if (d->topology == ESTH_TOPO_TORUS && (f->dimx <= 0 || f->dimy <= 0)) esth_torus_dims(f->G, &f->dimx, &f->dimy);
    for (int i = 0; i < ESTH_FRAME_RING; ++i) hipEventCreateWithFlags(&f->frame_ev[i], hipEventDisableTiming);
    hipEventCreateWithFlags(&f->ring_ev, hipEventDisableTiming);
    for (int i = 0; i < 7; ++i) hipEventCreate(&f->ev[i]);
    for (int i = 0; i < 2; ++i) {
        hipEventCreateWithFlags(&f->ev_rng_done[i], hipEventDisableTiming);
        hipEventCreateWithFlags(&f->ev_consumed[i], hipEventDisableTiming);
    }
    hipEventCreateWithFlags(&f->ev_fork, hipEventDisableTiming);
    int rc = ESTH_OK;
    const size_t SN = (size_t)S * f->N;
    const size_t tG = (size_t)(f->t ? f->t : 1) * f->G;
#define FA(ptr, bytes) if (rc == ESTH_OK) rc = falloc(f, (bytes), (void**)&(ptr))
    FA(f->state, sizeof(float) * SN);
    FA(f->tmp, sizeof(float) * SN);
    FA(f->w, sizeof(float) * f->N);
    FA(f->R[0], sizeof(float) * f->words);
    FA(f->R[1], sizeof(float) * f->words);
    FA(f->bt_w, sizeof(float) * tG);
    FA(f->bt_s, sizeof(float) * tG * S);
    FA(f->part, sizeof(float) * (size_t)f->G * (2 + S));
    FA(f->frames, sizeof(float) * ESTH_FRAME_FLOATS * ESTH_FRAME_RING);
    FA(f->winner, sizeof(float2) * f->G);
    FA(f->result, sizeof(float2));
    FA(f->est, sizeof(double) * (4 + 2 * ESTH_MAX_STATE_DIM));
    FA(f->status[0], sizeof(uint32_t) * ESTH_MTGP_N * (size_t)f->K);
    FA(f->status[1], sizeof(uint32_t) * ESTH_MTGP_N * (size_t)f->K);
    FA(f->top, sizeof(uint32_t) * 64);

    FA(f->streams, sizeof(esth_mtgp_stream) * (size_t)f->K);
#undef FA
    if (rc != ESTH_OK) { filter_free(f); return rc; }
    hipError_t err = hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&f->rng_stream, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipHostMalloc((void**)&f->h_frames, sizeof(float) * ESTH_FRAME_FLOATS * ESTH_FRAME_RING, 0);
    if (err != hipSuccess) { filter_free(f); return esth_fail(ESTH_ERR_DEVICE, "stream/pinned: %s", hipGetErrorString(err)); }
    /* RNG status: make_kernel_data32 (mtgp32-opencl.c:31-60), stream j seeded seed + j + 1 */
    std::vector<uint32_t> st((size_t)ESTH_MTGP_N * f->K);
    std::vector<esth_mtgp_stream> sd(f->K);
    for (int j = 0; j < f->K; ++j) {
        esth_mtgp_seed_state(j, d->seed + (uint32_t)j + 1u, st.data() + (size_t)j * ESTH_MTGP_N);
        sd[j] = {(uint32_t)j, (uint32_t)f->size, (uint64_t)j * f->size, (uint64_t)j};
    }
    err = hipMemcpy(f->status[0], st.data(), st.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(f->streams, sd.data(), sd.size() * sizeof(esth_mtgp_stream), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemset(f->w, 0, sizeof(float) * f->N);

    if (err == hipSuccess) err = hipMemset(f->est, 0, sizeof(double) * (4 + 2 * ESTH_MAX_STATE_DIM));
    if (err != hipSuccess) { filter_free(f); return esth_fail(ESTH_ERR_DEVICE, "init upload: %s", hipGetErrorString(err)); }
    *out = f;
    return ESTH_OK;
}

extern "C" int esth_filter_destroy(esth_filter* f) {
    filter_free(f);
    return ESTH_OK;
}

extern "C" int esth_filter_init_uniform(esth_filter* f, const float* s) {
    if (!f || !s) return esth_fail(ESTH_ERR_INVALID, "null filter/state");
    std::vector<float> h((size_t)f->S * f->N);
    for (int c = 0; c < f->S; ++c) std::fill(h.begin() + (size_t)c * f->N, h.begin() + (size_t)(c + 1) * f->N, s[c]);
    return esth_filter_init(f, h.data());
}

/* the device state from / to host SOA planes, through the filter's layout */
static int upload_state(esth_filter* f, const float* soa) {
    const size_t n = (size_t)f->S * f->N;
    if (f->d.state_layout == ESTH_LAYOUT_SOA) {
        ESTH_HIP(hipMemcpy(f->state, soa, sizeof(float) * n, hipMemcpyHostToDevice));
        return ESTH_OK;
    }
    std::vector<float> h(n);
    ESTH_TRY(esth_state_convert(f->S, f->N, ESTH_LAYOUT_SOA, soa, f->d.state_layout, h.data()));
    ESTH_HIP(hipMemcpy(f->state, h.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    return ESTH_OK;
}

static int download_state(esth_filter* f, float* soa) {
    const size_t n = (size_t)f->S * f->N;
    if (f->d.state_layout == ESTH_LAYOUT_SOA) {
        ESTH_HIP(hipMemcpy(soa, f->state, sizeof(float) * n, hipMemcpyDeviceToHost));
        return ESTH_OK;
    }
    std::vector<float> h(n);
    ESTH_HIP(hipMemcpy(h.data(), f->state, sizeof(float) * n, hipMemcpyDeviceToHost));
    return esth_state_convert(f->S, f->N, f->d.state_layout, h.data(), ESTH_LAYOUT_SOA, soa);
}

extern "C" int esth_filter_init(esth_filter* f, const float* soa) {
    if (!f || !soa) return esth_fail(ESTH_ERR_INVALID, "null filter/state");
    ESTH_HIP(hipSetDevice(f->e->device));
    ESTH_HIP(hipStreamSynchronize(f->stream));
    ESTH_TRY(upload_state(f, soa));
    if (f->d.carry_weights) {   /* a new particle set starts from equal weights */
        std::vector<float> w0(f->N, f->d.weights == ESTH_WEIGHTS_LOG ? 0.0f : 1.0f);
        ESTH_HIP(hipMemcpy(f->w, w0.data(), sizeof(float) * f->N, hipMemcpyHostToDevice));
    }
    return ESTH_OK;
}

/* MTGP words for step k into R[k & 1] */
static int launch_rng(esth_filter* f, uint64_t k, hipStream_t s) {
    const int p = (int)(k & 1);
    return esth_launch_mtgp(f->e->d_params, f->status[p], f->status[p ^ 1], f->streams, f->K, f->R[p], 1, s);
}

/* RNG side of step n = f->steps, before its kernels: if the step's words are
 * not ready they are generated first on the main stream; with prefetch on,
 * step n+1's words are generated on the RNG stream while step n runs (their
 * buffer was last read by step n-1: the ev_consumed event orders that).
 * With `frame` (eager steps) the step's frame is copied here too: on the RNG
 * stream ahead of the event the main stream waits for anyway, so the copy and
 * the words cost the main stream one wait; *copy_stream tells where it went. */
static int rng_before_step(esth_filter* f, float* dfr, const float* hf, bool frame, hipStream_t* copy_stream) {
    hipStream_t s = f->stream;
    const uint64_t n = f->steps;
    const int p = (int)(n & 1);
    const size_t fbytes = sizeof(float) * ESTH_FRAME_FLOATS;
    bool forked = false;
    *copy_stream = s;
    if (!f->rng_ready) {
        if (frame) ESTH_HIP(hipMemcpyAsync(dfr, hf, fbytes, hipMemcpyHostToDevice, s));
        ESTH_TRY(launch_rng(f, n, s));
        if (f->prefetch) {   /* step n+1's words continue from this status */
            ESTH_HIP(hipEventRecord(f->ev_fork, s));
            ESTH_HIP(hipStreamWaitEvent(f->rng_stream, f->ev_fork, 0));
            forked = true;
        }
    } else {
        if (frame) {
            ESTH_HIP(hipMemcpyAsync(dfr, hf, fbytes, hipMemcpyHostToDevice, f->rng_stream));
            ESTH_HIP(hipEventRecord(f->ev_rng_done[p], f->rng_stream));
            *copy_stream = f->rng_stream;
        }
        ESTH_HIP(hipStreamWaitEvent(s, f->ev_rng_done[p], 0));
    }
    if (f->prefetch) {
        if (!forked && n > 0) ESTH_HIP(hipStreamWaitEvent(f->rng_stream, f->ev_consumed[p ^ 1], 0));
        ESTH_TRY(launch_rng(f, n + 1, f->rng_stream));
        /* an eager step n+1 records it after its frame copy instead */
        if (!frame) ESTH_HIP(hipEventRecord(f->ev_rng_done[p ^ 1], f->rng_stream));
    }
    return ESTH_OK;
}

static int rng_after_step(esth_filter* f) {
    if (f->prefetch) ESTH_HIP(hipEventRecord(f->ev_consumed[f->steps & 1], f->stream));
    f->rng_ready = f->prefetch;
    return ESTH_OK;
}

/* The main-stream kernels of step n = f->steps, reading the frame at d_frame
 * and the step's words in R[n & 1].  ev (optional): timing events. */
static int enqueue_step(esth_filter* f, const float* d_frame, hipEvent_t* ev) {
    const bool timing = ev != nullptr;
    const bool all = timing && f->timing >= 2;
    hipStream_t s = f->stream;
    const esth_filter_desc& d = f->d;
    const int log_w = d.weights == ESTH_WEIGHTS_LOG;
    const uint32_t N = (uint32_t)f->N;
    const int p = (int)(f->steps & 1);
    float* R = f->R[p];
    if (timing) ESTH_HIP(hipEventRecord(ev[1], s));
    if (!all && f->fuse_sample_sort) {
        /* sampling_importance (pf_generic.c:521-533) and block_sort
         * (pf_generic.c:572-585) in one launch: the propagated state stays in
         * LDS between them.  Per-stage timing keeps the two launches below
         * (their results are bit-identical). */
        /* a sub-filter the resampling skips keeps its propagated state unless
         * weights are carried (then it takes the sorted one): the scan and
         * the plain alias gate on the uniform rnd[N + g] (pf_generic.cl:403),
         * the adaptive alias on the weights themselves */
        const bool carry = d.carry_weights != 0;
        const int keep_all = !carry && d.resampler == ESTH_RESAMPLE_ALIAS && d.adaptive;
        const float keep_thr = carry ? INFINITY : d.resampling_threshold;
        ESTH_TRY(esth_launch_sample_sort(d.model, f->m, f->state, N, f->G, R, d_frame, log_w, f->w, d.carry_weights,
                                         d.state_layout, f->tmp, f->winner, f->t, f->bt_w, f->bt_s, f->part,
                                         R + (size_t)N * f->V + N, keep_thr, keep_all, s));
        if (timing) ESTH_HIP(hipEventRecord(ev[2], s));
    } else {
        /* sampling_importance with Box-Muller fused (pf_generic.c:521-533) */
        ESTH_TRY(esth_launch_sample_weight(d.model, f->state, N, R, d_frame, log_w, f->w, d.carry_weights,
                                           d.state_layout, s));
        if (timing) ESTH_HIP(hipEventRecord(ev[2], s));
        /* block_sort (pf_generic.c:572-585) */
        ESTH_TRY(esth_launch_sort(f->S, f->m, f->w, f->state, f->tmp, N, f->G, f->winner, nullptr, f->t, f->bt_w,
                                  f->bt_s, f->part, log_w, d.state_layout, s));
    }
    if (all) ESTH_HIP(hipEventRecord(ev[3], s));
    /* reduce (pf_generic.c:604-615), result kept on the device.  With a scan
     * resampler, or the one-wave alias (m >= 64), it rides in the resampling
     * launch (below); its best particle is rank 0 of the winning sub-filter,
     * read from the sort's compact best-t copy, which the exchange does not
     * touch. */
    const bool fused = d.resampler != ESTH_RESAMPLE_ALIAS || f->m >= 64;   /* one wave per sub-filter */
    if (!fused) ESTH_TRY(esth_launch_argmax(f->S, f->winner, f->G, log_w, f->part, f->tmp, N, f->m, f->result, f->est, s));
    if (all) ESTH_HIP(hipEventRecord(ev[4], s));
    /* exchange (pf_generic.c:702-715): in the fused launches its receiving
     * side is applied by the resampling as it reads each sub-filter */
    if (!fused)
        ESTH_TRY(esth_launch_exchange(f->tmp, f->w, f->S, N, f->m, f->G, f->t, d.topology, f->dimx, f->dimy, f->bt_w,
                                      f->bt_s, f->top, s));
    if (all) ESTH_HIP(hipEventRecord(ev[5], s));
    /* resampling (pf_generic.c:732-757), uniforms at rand_offset = N*V */
    if (fused) {
        const bool bt = f->t > 0 && d.topology != ESTH_TOPO_NONE;
        if (d.resampler == ESTH_RESAMPLE_ALIAS)
            ESTH_TRY(esth_launch_alias_argmax(f->S, f->state, f->tmp, f->w, R + (size_t)N * f->V, N, f->m, f->G,
                                              d.resampling_threshold, d.adaptive, log_w, f->winner, f->part,
                                              bt ? f->bt_s : f->tmp, bt ? (size_t)f->G * f->t : (size_t)N,
                                              bt ? f->t : f->m, f->result, f->est, f->t, d.topology, f->dimx,
                                              f->dimy, f->bt_w, f->bt_s, f->top, d.carry_weights, d.state_layout,
                                              s));
        else
            ESTH_TRY(esth_launch_resample_argmax(d.resampler, f->S, f->state, f->tmp, f->w, R + (size_t)N * f->V, N,
                                                 f->m, f->G, d.resampling_threshold, log_w, f->winner, f->part,
                                                 bt ? f->bt_s : f->tmp, bt ? (size_t)f->G * f->t : (size_t)N,
                                                 bt ? f->t : f->m, f->result, f->est, f->t, d.topology, f->dimx,
                                                 f->dimy, f->bt_w, f->bt_s, f->top, d.carry_weights,
                                                 d.state_layout, s));
    } else {
        ESTH_TRY(esth_launch_resample(d.resampler, f->S, f->state, f->tmp, f->w, R + (size_t)N * f->V, N, f->m, f->G,
                                      d.resampling_threshold, d.adaptive, log_w, nullptr, d.carry_weights,
                                      d.state_layout, s));
    }
    if (all) ESTH_HIP(hipEventRecord(ev[6], s));
    return ESTH_OK;
}

/* Capture the main-stream part of the step of parity `slot`: the frame upload
 * from the pinned slot and the step's kernels (the R buffer is fixed by the
 * parity).  The RNG runs outside the graph on its own stream. */
static int capture_graph(esth_filter* f, int slot, hipGraphExec_t* out) {
    hipGraph_t g = nullptr;
    ESTH_HIP(hipStreamBeginCapture(f->stream, hipStreamCaptureModeThreadLocal));
    /* the kernels only: the frame reaches slot `slot` by a copy ahead of each
     * replay, on the RNG stream (as in eager steps) */
    const float* dfr = f->frames + (size_t)slot * ESTH_FRAME_FLOATS;
    int rc = enqueue_step(f, dfr, nullptr);
    hipError_t e2 = hipStreamEndCapture(f->stream, &g);
    if (rc != ESTH_OK) { if (g) hipGraphDestroy(g); return rc; }
    if (e2 != hipSuccess) return esth_fail(ESTH_ERR_DEVICE, "end capture: %s", hipGetErrorString(e2));
    hipError_t e3 = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e3 != hipSuccess) return esth_fail(ESTH_ERR_DEVICE, "instantiate: %s", hipGetErrorString(e3));
    return ESTH_OK;
}

extern "C" int esth_filter_step(esth_filter* f, const void* control, const void* measurement, float dt) {
    if (!f) return esth_fail(ESTH_ERR_INVALID, "null filter");
    ESTH_HIP(hipSetDevice(f->e->device));
    /* frame slot: graphs use the step parity, eager steps a ring */
    int slot = f->use_graph ? (int)(f->steps & 1) : f->slot;
    if (f->use_graph) {
        if (f->frame_used[slot]) ESTH_HIP(hipEventSynchronize(f->frame_ev[slot]));
    } else if (slot == 0 && f->ring_pending) {   /* the previous pass's copies are done */
        ESTH_HIP(hipEventSynchronize(f->ring_ev));
        f->ring_pending = false;
    }
    float* hf = f->h_frames + (size_t)slot * ESTH_FRAME_FLOATS;
    if (esth_build_frame_impl(f->d.model, f->params, control, measurement, dt, hf))
        return esth_fail(ESTH_ERR_INVALID, "bad control/measurement for model %d", f->d.model);
    if (f->use_graph && !f->gexec[slot]) ESTH_TRY(capture_graph(f, slot, &f->gexec[slot]));
    hipEvent_t* ev = nullptr;
    if (!f->use_graph && f->timing && f->tsteps < ESTH_TIMING_STEPS && (f->tcount++ % (uint64_t)f->timing_every) == 0) {
        if ((size_t)(f->tsteps + 1) * 7 > f->tev.size()) {
            for (int i = 0; i < 7; ++i) {
                hipEvent_t e;
                ESTH_HIP(hipEventCreate(&e));
                f->tev.push_back(e);
            }
        }
        ev = &f->tev[(size_t)f->tsteps * 7];
        f->tsteps++;
    }
    float* dfr = f->frames + (size_t)slot * ESTH_FRAME_FLOATS;
    if (ev && f->timing >= 2) ESTH_HIP(hipEventRecord(ev[0], f->stream));
    hipStream_t copy_stream = f->stream;
    ESTH_TRY(rng_before_step(f, dfr, hf, true, &copy_stream));
    if (f->use_graph) {
        ESTH_HIP(hipGraphLaunch(f->gexec[slot], f->stream));
    } else {
        ESTH_TRY(enqueue_step(f, dfr, ev));
        f->slot = (f->slot + 1) % ESTH_FRAME_RING;
    }
    ESTH_TRY(rng_after_step(f));
    if (f->use_graph) {   /* the host slot is free again once its copy has run */
        ESTH_HIP(hipEventRecord(f->frame_ev[slot], copy_stream));
        f->frame_used[slot] = true;
    } else if (slot == ESTH_FRAME_RING - 1) {
        ESTH_HIP(hipEventRecord(f->ring_ev, copy_stream));
        f->ring_pending = true;
    }
    f->steps++;
    return ESTH_OK;
}

extern "C" int esth_filter_use_graph(esth_filter* f, int enable) {
    if (!f) return esth_fail(ESTH_ERR_INVALID, "null filter");
    ESTH_HIP(hipStreamSynchronize(f->stream));
    ESTH_HIP(hipStreamSynchronize(f->rng_stream));
    f->use_graph = enable != 0;
    return ESTH_OK;
}

/* Overlap the next step's MTGP words with the current step (default on). */
extern "C" int esth_filter_set_fusion(esth_filter* f, int enable, int* active) {
    if (!f) return esth_fail(ESTH_ERR_INVALID, "null filter");
    ESTH_HIP(hipSetDevice(f->e->device));
    ESTH_HIP(hipStreamSynchronize(f->stream));
    for (int i = 0; i < 2; ++i)
        if (f->gexec[i]) { hipGraphExecDestroy(f->gexec[i]); f->gexec[i] = nullptr; }
    f->fuse_sample_sort = enable != 0 && esth_sample_sort_pays(f->d.model, f->d.state_layout, f->m);
    if (active) *active = f->fuse_sample_sort ? 1 : 0;
    return ESTH_OK;
}

extern "C" int esth_filter_set_prefetch(esth_filter* f, int enable) {
    if (!f) return esth_fail(ESTH_ERR_INVALID, "null filter");
    ESTH_HIP(hipStreamSynchronize(f->stream));
    ESTH_HIP(hipStreamSynchronize(f->rng_stream));
    for (int i = 0; i < 2; ++i)
        if (f->gexec[i]) { hipGraphExecDestroy(f->gexec[i]); f->gexec[i] = nullptr; }
    f->prefetch = enable != 0;
    return ESTH_OK;
}

extern "C" int esth_filter_sync(esth_filter* f) {
    if (!f) return esth_fail(ESTH_ERR_INVALID, "null filter");
    ESTH_HIP(hipStreamSynchronize(f->stream));
    ESTH_HIP(hipStreamSynchronize(f->rng_stream));
    return ESTH_OK;
}

extern "C" int esth_filter_stream(esth_filter* f, void** s) {
    if (!f || !s) return esth_fail(ESTH_ERR_INVALID, "null");
    *s = (void*)f->stream;
    return ESTH_OK;
}

extern "C" int esth_filter_set_timing(esth_filter* f, int on) {
    if (!f) return esth_fail(ESTH_ERR_INVALID, "null filter");
    ESTH_HIP(hipStreamSynchronize(f->stream));
    f->timing = on < 0 ? 0 : (on > 2 ? 2 : on);
    f->tsteps = 0;
    f->tcount = 0;
    return ESTH_OK;
}

extern "C" int esth_filter_set_timing_every(esth_filter* f, int every) {
    if (!f || every < 1) return esth_fail(ESTH_ERR_INVALID, "timing stride must be >= 1");
    f->timing_every = every;
    f->tcount = 0;
    return ESTH_OK;
}

/* mean per-stage device time over the steps recorded since timing was switched on */
extern "C" int esth_filter_stage_times(esth_filter* f, float us[6]) {
    if (!f || !us) return esth_fail(ESTH_ERR_INVALID, "null");
    if (f->tsteps == 0) return esth_fail(ESTH_ERR_STATE, "no timed steps (timing off, or graph replay)");
    ESTH_HIP(hipStreamSynchronize(f->stream));
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < f->tsteps; ++k) {
        hipEvent_t* ev = &f->tev[(size_t)k * 7];
        for (int i = 0; i < 6; ++i) {
            if (f->timing < 2 && i != 1) continue;
            float ms = 0.0f;
            ESTH_HIP(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
            acc[i] += ms;
        }
    }
    for (int i = 0; i < 6; ++i)
        us[i] = (f->timing < 2 && i != 1) ? NAN : (float)(acc[i] * 1000.0 / f->tsteps);
    return ESTH_OK;
}

extern "C" int esth_filter_estimate(esth_filter* f, esth_estimate* out) {
    if (!f || !out) return esth_fail(ESTH_ERR_INVALID, "null");
    ESTH_HIP(hipStreamSynchronize(f->stream));
    double h[4 + 2 * ESTH_MAX_STATE_DIM];
    ESTH_HIP(hipMemcpy(h, f->est, sizeof(double) * (4 + 2 * (size_t)f->S), hipMemcpyDeviceToHost));
    memset(out, 0, sizeof(*out));
    const int S = f->S;
    out->state_dim = S;
    out->log_weight_max = h[0];
    out->weight_sum = h[1];
    for (int c = 0; c < S; ++c) {
        out->mean[c] = h[1] > 0.0 ? h[2 + c] / h[1] : NAN;
        out->best[c] = (float)h[4 + S + c];
    }
    out->best_weight = (float)h[2 + S];
    out->best_block = (int)h[3 + S];
    out->step = f->steps;
    return ESTH_OK;
}

extern "C" int esth_filter_read_particles(esth_filter* f, float* soa, float* w) {
    if (!f) return esth_fail(ESTH_ERR_INVALID, "null");
    ESTH_HIP(hipStreamSynchronize(f->stream));
    if (soa) ESTH_TRY(download_state(f, soa));
    if (w) ESTH_HIP(hipMemcpy(w, f->w, sizeof(float) * f->N, hipMemcpyDeviceToHost));
    return ESTH_OK;
}

/* snapshot: magic, S, N, K, steps, model, m, G | state (SOA planes, whatever
 * the filter's layout) | w | status */
#define ESTH_SNAP_MAGIC 0x45535448u
extern "C" int esth_filter_snapshot_size(esth_filter* f, size_t* bytes) {
    if (!f || !bytes) return esth_fail(ESTH_ERR_INVALID, "null");
    *bytes = 8 * sizeof(uint64_t) + sizeof(float) * ((size_t)f->S * f->N + f->N) + sizeof(uint32_t) * ESTH_MTGP_N * (size_t)f->K;
    return ESTH_OK;
}

extern "C" int esth_filter_snapshot(esth_filter* f, void* buf, size_t bytes) {
    size_t need = 0;
    ESTH_TRY(esth_filter_snapshot_size(f, &need));
    if (!buf || bytes < need) return esth_fail(ESTH_ERR_INVALID, "snapshot buffer %zu < %zu", bytes, need);
    ESTH_HIP(hipStreamSynchronize(f->stream));
    ESTH_HIP(hipStreamSynchronize(f->rng_stream));
    uint64_t* hdr = (uint64_t*)buf;
    hdr[0] = ESTH_SNAP_MAGIC; hdr[1] = (uint64_t)f->S; hdr[2] = f->N; hdr[3] = (uint64_t)f->K; hdr[4] = f->steps;
    hdr[5] = f->d.model; hdr[6] = f->m; hdr[7] = f->G;
    char* p = (char*)buf + 8 * sizeof(uint64_t);
    ESTH_TRY(download_state(f, (float*)p));
    p += sizeof(float) * (size_t)f->S * f->N;
    ESTH_HIP(hipMemcpy(p, f->w, sizeof(float) * f->N, hipMemcpyDeviceToHost));
    p += sizeof(float) * f->N;
    /* the status that generates the next step's words (any prefetched words are regenerated on restore) */
    ESTH_HIP(hipMemcpy(p, f->status[f->steps & 1], sizeof(uint32_t) * ESTH_MTGP_N * (size_t)f->K, hipMemcpyDeviceToHost));
    return ESTH_OK;
}

extern "C" int esth_filter_restore(esth_filter* f, const void* buf, size_t bytes) {
    size_t need = 0;
    ESTH_TRY(esth_filter_snapshot_size(f, &need));
    if (!buf || bytes < need) return esth_fail(ESTH_ERR_INVALID, "snapshot buffer %zu < %zu", bytes, need);
    const uint64_t* hdr = (const uint64_t*)buf;
    if (hdr[0] != ESTH_SNAP_MAGIC || hdr[1] != (uint64_t)f->S || hdr[2] != f->N || hdr[3] != (uint64_t)f->K ||
        hdr[5] != (uint64_t)f->d.model || hdr[6] != f->m || hdr[7] != f->G)
        return esth_fail(ESTH_ERR_INVALID, "snapshot does not match this filter");
    ESTH_HIP(hipStreamSynchronize(f->stream));
    ESTH_HIP(hipStreamSynchronize(f->rng_stream));
    const char* p = (const char*)buf + 8 * sizeof(uint64_t);
    ESTH_TRY(upload_state(f, (const float*)p));
    p += sizeof(float) * (size_t)f->S * f->N;
    ESTH_HIP(hipMemcpy(f->w, p, sizeof(float) * f->N, hipMemcpyHostToDevice));
    p += sizeof(float) * f->N;
    f->steps = hdr[4];
    ESTH_HIP(hipMemcpy(f->status[f->steps & 1], p, sizeof(uint32_t) * ESTH_MTGP_N * (size_t)f->K, hipMemcpyHostToDevice));
    f->rng_ready = false;
    return ESTH_OK;
}

/* ================================================= trace I/O, reference */
/* read_trace, pf_model_host.h:30-54: 20 whitespace-separated floats per row */
extern "C" int esth_trace_load_robotarm(const char* path, int max_rows, esth_robotarm5_measurement* meas,
                                        esth_robotarm5_control* ctrl, float* dt, float* truth, int* rows) {
    if (!path || !rows) return esth_fail(ESTH_ERR_INVALID, "null");
    FILE* fh = fopen(path, "r");
    if (!fh) return esth_fail(ESTH_ERR_IO, "could not open %s", path);
    int n = 0;
    while (n < max_rows) {
        float v[20];
        int k = 0;
        for (; k < 20; ++k)
            if (fscanf(fh, "%f", &v[k]) != 1) break;
        if (k < 20) break;   /* a short or garbled row ends the trace, as read_trace's -1 */
        if (meas) {
            meas[n].x = v[0]; meas[n].y = v[1];
            for (int i = 0; i < 5; ++i) meas[n].angles[i] = v[2 + i];
        }
        if (ctrl) for (int i = 0; i < 5; ++i) ctrl[n].angles[i] = v[7 + i];
        if (dt) dt[n] = v[12];
        if (truth) {
            float* tr = truth + (size_t)n * 9;
            for (int i = 0; i < 5; ++i) tr[i] = v[15 + i];
            tr[5] = v[13]; tr[6] = v[14]; tr[7] = 0.0f; tr[8] = 0.0f;
        }
        ++n;
    }
    fclose(fh);
    *rows = n;
    return ESTH_OK;
}

static void write_floats(FILE* fh, const float* v, int n) {
    for (int k = 0; k < n; ++k) fprintf(fh, k ? " %.9g" : "%.9g", (double)v[k]);
    fputc('\n', fh);
}

extern "C" int esth_trace_write_robotarm(const char* path, int rows, const esth_robotarm5_measurement* meas,
                                         const esth_robotarm5_control* ctrl, const float* dt, const float* truth) {
    if (!path || rows < 0 || (rows > 0 && (!meas || !ctrl || !dt || !truth))) return esth_fail(ESTH_ERR_INVALID, "null");
    FILE* fh = fopen(path, "w");
    if (!fh) return esth_fail(ESTH_ERR_IO, "could not open %s", path);
    for (int n = 0; n < rows; ++n) {
        const float* tr = truth + (size_t)n * 9;
        float v[20];
        v[0] = meas[n].x; v[1] = meas[n].y;
        for (int i = 0; i < 5; ++i) { v[2 + i] = meas[n].angles[i]; v[7 + i] = ctrl[n].angles[i]; v[15 + i] = tr[i]; }
        v[12] = dt[n]; v[13] = tr[5]; v[14] = tr[6];
        write_floats(fh, v, 20);
    }
    if (fclose(fh)) return esth_fail(ESTH_ERR_IO, "could not write %s", path);
    return ESTH_OK;
}

/* estimate_error, pf_model_host.h:56-69 */
extern "C" float esth_robotarm5_estimate_error(const float* est, const float* act) {
    float error = 0.0f;
    error += (est[5] - act[5]) * (est[5] - act[5]) * ESTH_RA5_NOISE_CAMERA_X;
    error += (est[6] - act[6]) * (est[6] - act[6]) * ESTH_RA5_NOISE_CAMERA_Y;
    for (int i = 0; i < 5; ++i) error += (est[i] - act[i]) * (est[i] - act[i]) * ESTH_RA5_NOISE_ANGLE_MEASUREMENTS;
    return error;
}

static int count_lines(const char* path, int* lines);

/* particle_filter() for the robot arm over a trace (pf_generic.c:245-887) */
extern "C" int esth_particle_filter(uint32_t m, uint32_t G, uint32_t t, float thr, const char* trace, uint32_t seed,
                                    int resampler, double* mean_error, int* steps) {
    const int max_rows = 1 << 20;
    std::vector<esth_robotarm5_measurement> meas;
    std::vector<esth_robotarm5_control> ctrl;
    std::vector<float> dt, truth;
    int rows = 0;
    ESTH_TRY(count_lines(trace, &rows));
    rows = std::min(rows, max_rows);
    meas.resize(rows); ctrl.resize(rows); dt.resize(rows); truth.resize((size_t)rows * 9);
    ESTH_TRY(esth_trace_load_robotarm(trace, rows, meas.data(), ctrl.data(), dt.data(), truth.data(), &rows));
    esth_engine* e = nullptr;
    ESTH_TRY(esth_engine_create(0, 0, &e));
    esth_filter_desc d;
    esth_filter_desc_default(&d, ESTH_MODEL_ROBOTARM5);
    d.num_particles = m; d.num_blocks = G; d.num_transfer = t; d.resampling_threshold = thr;
    d.resampler = resampler; d.seed = seed; d.topology = t ? ESTH_TOPO_RING : ESTH_TOPO_NONE;
    esth_filter* f = nullptr;
    int rc = esth_filter_create(e, &d, nullptr, 0, &f);
    if (rc != ESTH_OK) { esth_engine_destroy(e); return rc; }
    /* initial particles, pf_generic.c:346-361 */
    const float a0 = (float)(0.75f * M_PI);
    float init[9] = {a0, a0, a0, a0, a0, 7.0f, 0.0f, 0.0f, 0.0f};
    rc = esth_filter_init_uniform(f, init);
    double err_sum = 0.0;
    for (int i = 0; rc == ESTH_OK && i < rows; ++i) {
        rc = esth_filter_step(f, &ctrl[i], &meas[i], dt[i]);
        esth_estimate est;
        if (rc == ESTH_OK) rc = esth_filter_estimate(f, &est);
        if (rc == ESTH_OK) {
            float best[9];
            for (int c = 0; c < 9; ++c) best[c] = est.best[c];
            best[7] = 0.0f; best[8] = 0.0f;
            err_sum += esth_robotarm5_estimate_error(best, &truth[(size_t)i * 9]);
        }
    }
    esth_filter_destroy(f);
    esth_engine_destroy(e);
    if (rc != ESTH_OK) return rc;
    if (mean_error) *mean_error = rows ? err_sum / rows : 0.0;
    if (steps) *steps = rows;
    return ESTH_OK;
}

extern "C" int esth_trace_write_unicycle10(const char* path, int rows, const esth_unicycle10_measurement* meas,
                                           const esth_unicycle_control* ctrl, const float* dt, const float* truth) {
    if (!path || rows < 0 || (rows > 0 && (!meas || !ctrl || !dt || !truth))) return esth_fail(ESTH_ERR_INVALID, "null");
    FILE* fh = fopen(path, "w");
    if (!fh) return esth_fail(ESTH_ERR_IO, "could not open %s", path);
    for (int n = 0; n < rows; ++n) {
        float v[16];
        for (int i = 0; i < 10; ++i) v[i] = meas[n].range[i];
        v[10] = ctrl[n].v; v[11] = ctrl[n].omega; v[12] = dt[n];
        for (int i = 0; i < 3; ++i) v[13 + i] = truth[(size_t)n * 3 + i];
        write_floats(fh, v, 16);
    }
    if (fclose(fh)) return esth_fail(ESTH_ERR_IO, "could not write %s", path);
    return ESTH_OK;
}

/* the reference unicycle's input rows (particle_filter.cu:756-772): 16 floats */
extern "C" int esth_trace_load_unicycle10(const char* path, int max_rows, esth_unicycle10_measurement* meas,
                                          esth_unicycle_control* ctrl, float* dt, float* truth, int* rows) {
    if (!path || !rows) return esth_fail(ESTH_ERR_INVALID, "null");
    FILE* fh = fopen(path, "r");
    if (!fh) return esth_fail(ESTH_ERR_IO, "could not open %s", path);
    int n = 0;
    while (n < max_rows) {
        float v[16];
        int k = 0;
        for (; k < 16; ++k)
            if (fscanf(fh, "%f", &v[k]) != 1) break;
        if (k < 16) break;   /* the reference loop runs while fscanf reads all 16 */
        if (meas) for (int i = 0; i < 10; ++i) meas[n].range[i] = v[i];
        if (ctrl) { ctrl[n].v = v[10]; ctrl[n].omega = v[11]; }
        if (dt) dt[n] = v[12];
        if (truth) { truth[(size_t)n * 3] = v[13]; truth[(size_t)n * 3 + 1] = v[14]; truth[(size_t)n * 3 + 2] = v[15]; }
        ++n;
    }
    fclose(fh);
    *rows = n;
    return ESTH_OK;
}

static int count_lines(const char* path, int* lines) {
    FILE* fh = fopen(path, "r");
    if (!fh) return esth_fail(ESTH_ERR_IO, "could not open %s", path);
    int n = 0, ch;
    while ((ch = fgetc(fh)) != EOF) n += ch == '\n';
    fclose(fh);
    *lines = n + 1;
    return ESTH_OK;
}

/* particle_filter.cu:729-1157 with the engine's step: one filter over the
 * trace, error = |best particle - truth| in the plane, averaged over steps */
extern "C" int esth_unicycle10_filter(uint32_t m, uint32_t G, uint32_t t, float thr, const char* trace, uint32_t seed,
                                      int resampler, double* mean_error, int* steps) {
    if (!trace) return esth_fail(ESTH_ERR_INVALID, "null trace");
    int rows = 0;
    ESTH_TRY(count_lines(trace, &rows));
    rows = std::min(rows, 1 << 20);
    std::vector<esth_unicycle10_measurement> meas(rows);
    std::vector<esth_unicycle_control> ctrl(rows);
    std::vector<float> dt(rows), truth((size_t)rows * 3);
    ESTH_TRY(esth_trace_load_unicycle10(trace, rows, meas.data(), ctrl.data(), dt.data(), truth.data(), &rows));
    esth_unicycle10_params prm;
    esth_unicycle10_params_default(&prm);
    esth_engine* e = nullptr;
    ESTH_TRY(esth_engine_create(0, 0, &e));
    esth_filter_desc d;
    esth_filter_desc_default(&d, ESTH_MODEL_UNICYCLE10);
    d.num_particles = m; d.num_blocks = G; d.num_transfer = t; d.resampling_threshold = thr;
    d.resampler = resampler; d.seed = seed; d.topology = t ? ESTH_TOPO_RING : ESTH_TOPO_NONE;
    esth_filter* f = nullptr;
    int rc = esth_filter_create(e, &d, &prm, sizeof(prm), &f);
    if (rc != ESTH_OK) { esth_engine_destroy(e); return rc; }
    const float init[3] = {20.0f, 40.0f, (float)(-M_PI / 2)};
    rc = esth_filter_init_uniform(f, init);
    double err_sum = 0.0;
    for (int i = 0; rc == ESTH_OK && i < rows; ++i) {
        rc = esth_filter_step(f, &ctrl[i], &meas[i], dt[i]);
        esth_estimate est;
        if (rc == ESTH_OK) rc = esth_filter_estimate(f, &est);
        if (rc == ESTH_OK) {
            const double ex = (double)est.best[0] - truth[(size_t)i * 3], ey = (double)est.best[1] - truth[(size_t)i * 3 + 1];
            err_sum += std::sqrt(ex * ex + ey * ey);
        }
    }
    esth_filter_destroy(f);
    esth_engine_destroy(e);
    if (rc != ESTH_OK) return rc;
    if (mean_error) *mean_error = rows ? err_sum / rows : 0.0;
    if (steps) *steps = rows;
    return ESTH_OK;
}

/* ============================================================ scenarios
 * Seeded host generators: splitmix64 words, uniform doubles in [0,1), normals
 * by the polar-free Box-Muller in double. */
struct esth_hostrng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        double u1 = 1.0 - uniform(), u2 = uniform();
        return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    }
};

static double wrap_pi(double a) {
    a = fmod(a + M_PI, 2.0 * M_PI);
    if (a <= 0.0) a += 2.0 * M_PI;
    return a - M_PI;
}

extern "C" int esth_unicycle_scenario(uint32_t seed, int L, float box, int steps, float dt, esth_unicycle_params* p,
                                      esth_unicycle_control* ctrl, esth_unicycle_measurement* meas, float* truth) {
    if (!p || L < 0 || L > ESTH_UNICYCLE_MAX_LANDMARKS || steps < 0) return esth_fail(ESTH_ERR_INVALID, "scenario args");
    esth_hostrng rng{0x5eed000000000000ull ^ (uint64_t)seed};
    memset(p, 0, sizeof(*p));
    p->num_landmarks = L;
    p->sigma_x = 0.05f; p->sigma_y = 0.05f; p->sigma_theta = 0.01f;
    p->sigma_range = 0.1f; p->sigma_bearing = 0.02f;
    if (L == 4 && box == 0.0f) {
        const float lx[4] = {10.0f, -10.0f, -10.0f, 10.0f}, ly[4] = {10.0f, 10.0f, -10.0f, -10.0f};
        for (int j = 0; j < 4; ++j) { p->lx[j] = lx[j]; p->ly[j] = ly[j]; }
    } else {
        for (int j = 0; j < L; ++j) {
            p->lx[j] = (float)(-box + 2.0 * box * rng.uniform());
            p->ly[j] = (float)(-box + 2.0 * box * rng.uniform());
        }
    }
    double x = 0.0, y = 0.0, th = 0.0;
    for (int k = 0; k < steps; ++k) {
        const double v = 1.0, om = 0.2 * sin(0.05 * (double)k);
        double nth = th;
        x += v * dt * cos(nth) + p->sigma_x * rng.normal();
        y += v * dt * sin(nth) + p->sigma_y * rng.normal();
        th += om * dt + p->sigma_theta * rng.normal();
        if (ctrl) { ctrl[k].v = (float)v; ctrl[k].omega = (float)om; }
        if (truth) { truth[3 * k] = (float)x; truth[3 * k + 1] = (float)y; truth[3 * k + 2] = (float)th; }
        if (meas) {
            memset(&meas[k], 0, sizeof(meas[k]));
            for (int j = 0; j < L; ++j) {
                double dx = p->lx[j] - x, dy = p->ly[j] - y;
                meas[k].range[j] = (float)(sqrt(dx * dx + dy * dy) + p->sigma_range * rng.normal());
                meas[k].bearing[j] = (float)wrap_pi(atan2(dy, dx) - th + p->sigma_bearing * rng.normal());
            }
        }
    }
    return ESTH_OK;
}

extern "C" int esth_dh6_scenario(uint32_t seed, int steps, float dt, esth_dh6_params* p, esth_dh6_measurement* meas,
                                 float* truth) {
    if (!p || steps < 0) return esth_fail(ESTH_ERR_INVALID, "scenario args");
    esth_hostrng rng{0xd6a0000000000000ull ^ (uint64_t)seed};
    memset(p, 0, sizeof(*p));
    const double alpha[6] = {M_PI / 2, 0.0, 0.0, M_PI / 2, -M_PI / 2, 0.0};
    for (int i = 0; i < 6; ++i) {
        p->a[i] = (float)(0.2 + 0.4 * rng.uniform());
        p->d[i] = (float)(-0.1 + 0.2 * rng.uniform());
        p->alpha[i] = (float)alpha[i];
    }
    p->sigma_acc = 0.5f; p->sigma_pos = 0.01f; p->sigma_enc = 0.005f;
    double q[6], qd[6];
    for (int i = 0; i < 6; ++i) { q[i] = -0.5 + rng.uniform(); qd[i] = -0.5 + rng.uniform(); }
    /* truth row 0 holds the initial state; rows 1..steps the states after each step */
    if (truth) for (int i = 0; i < 6; ++i) { truth[i] = (float)q[i]; truth[6 + i] = (float)qd[i]; }
    for (int k = 0; k < steps; ++k) {
        for (int i = 0; i < 6; ++i) {
            double a = p->sigma_acc * rng.normal();
            q[i] += qd[i] * dt + 0.5 * dt * dt * a;
            qd[i] += a * dt;
        }
        double T[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
        for (int i = 0; i < 6; ++i) {
            double cq = cos(q[i]), sq = sin(q[i]), ca = cos((double)p->alpha[i]), sa = sin((double)p->alpha[i]);
            double A[4][4] = {{cq, -sq * ca, sq * sa, p->a[i] * cq}, {sq, cq * ca, -cq * sa, p->a[i] * sq}, {0, sa, ca, p->d[i]}, {0, 0, 0, 1}};
            double R[4][4];
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c) {
                    double acc = 0;
                    for (int kk = 0; kk < 4; ++kk) acc += T[r][kk] * A[kk][c];
                    R[r][c] = acc;
                }
            memcpy(T, R, sizeof(T));
        }
        if (meas) {
            for (int i = 0; i < 3; ++i) meas[k].p[i] = (float)(T[i][3] + p->sigma_pos * rng.normal());
            for (int i = 0; i < 6; ++i) meas[k].q[i] = (float)(q[i] + p->sigma_enc * rng.normal());
        }
        if (truth) {
            float* tr = truth + (size_t)(k + 1) * 12;
            for (int i = 0; i < 6; ++i) { tr[i] = (float)q[i]; tr[6 + i] = (float)qd[i]; }
        }
    }
    return ESTH_OK;
}

/* the base's pixel position for an arm state (pinhole + distortion of the
 * robotarm5_camera model), in double */
static void camera_project(const double* q, double x, double y, const esth_camera_params* p, double* u, double* v) {
    double cx = x * cos(q[0]) + y * sin(q[0]);
    double cy = -x * sin(q[0]) + y * cos(q[0]);
    double cz = 0.0;
    for (int i = 1; i < 5; ++i) {
        const double nb = cy * cos(q[i]) - cz * sin(q[i]);
        const double nc = cy * sin(q[i]) + cz * cos(q[i]);
        cy = nb;
        cz = nc - 3.0;
    }
    const double xn = cx / cz, yn = cy / cz, r2 = xn * xn + yn * yn;
    const double radial = 1.0 + p->kc[0] * r2 + p->kc[1] * r2 * r2;
    const double xd = xn * radial + 2.0 * p->kc[2] * xn * yn + p->kc[3] * (r2 + 2.0 * xn * xn);
    const double yd = yn * radial + p->kc[2] * (r2 + 2.0 * yn * yn) + 2.0 * p->kc[3] * xn * yn;
    *u = p->fc[0] * xd + p->cc[0];
    *v = p->fc[1] * yd + p->cc[1];
}

extern "C" int esth_robotarm5_camera_scenario(uint32_t seed, int steps, float dt, const esth_camera_params* params,
                                              esth_robotarm5_control* ctrl, esth_robotarm5_measurement* meas,
                                              float* truth) {
    if (steps < 0 || !(dt > 0.0f)) return esth_fail(ESTH_ERR_INVALID, "scenario args");
    esth_camera_params dflt;
    esth_camera_params_default(&dflt);
    const esth_camera_params* p = params ? params : &dflt;
    esth_hostrng rng{0xca3e000000000000ull ^ (uint64_t)seed};
    double q[5], x = 0.0, y = 0.0, vx = 0.0, vy = 0.0, ph[5];
    for (int i = 0; i < 5; ++i) { q[i] = 0.75 * M_PI; ph[i] = 2.0 * M_PI * rng.uniform(); }
    if (truth) {
        for (int i = 0; i < 5; ++i) truth[i] = (float)q[i];
        truth[5] = truth[6] = truth[7] = truth[8] = 0.0f;
    }
    for (int k = 0; k < steps; ++k) {
        for (int i = 0; i < 5; ++i) {
            const double rate = 0.2 * sin(0.05 * (double)k + ph[i]);
            if (ctrl) ctrl[k].angles[i] = (float)rate;
            q[i] += rate * dt + 0.01 * rng.normal();
        }
        vx = fmin(0.5, fmax(-0.5, vx + 0.1 * rng.normal()));
        vy = fmin(0.5, fmax(-0.5, vy + 0.1 * rng.normal()));
        x += vx * dt;
        y += vy * dt;
        if (fabs(x) > 1.0) { x = copysign(2.0, x) - x; vx = -vx; }
        if (fabs(y) > 1.0) { y = copysign(2.0, y) - y; vy = -vy; }
        if (meas) {
            double u, v;
            camera_project(q, x, y, p, &u, &v);
            for (int i = 0; i < 5; ++i) meas[k].angles[i] = (float)(q[i] + 0.05 * rng.normal());
            meas[k].x = (float)(u + rng.normal());
            meas[k].y = (float)(v + rng.normal());
        }
        if (truth) {
            float* tr = truth + (size_t)(k + 1) * 9;
            for (int i = 0; i < 5; ++i) tr[i] = (float)q[i];
            tr[5] = (float)x; tr[6] = (float)y; tr[7] = (float)vx; tr[8] = (float)vy;
        }
    }
    return ESTH_OK;
}
