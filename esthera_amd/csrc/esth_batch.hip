/*
 * esth_batch.hip - many independent filters stepped together.
 *
 * All filters of a batch live in shared device buffers, each in its own
 * single-filter layout (S state planes of N_f, N_f weights, its own MTGP32
 * stream set of K_f streams seeded seed_f + j + 1).  One step is:
 *   RNG stream : one MTGP32 launch over every filter's streams (next step)
 *   main stream: frame upload, sample+weight, sort (one launch per sub-filter
 *                size class), argmax (one work-group per filter), exchange,
 *                resampling (per size class)
 * and the main-stream part is captured once per step parity into a HIP graph
 * and replayed.  Because every kernel runs the single-filter per-sub-filter
 * code on each filter's own buffers and random words, a filter's particles are
 * bit-identical to the same filter stepped alone with ESTH_RNG_MTGP_PER_FILTER.
 */
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <thread>

#include "esth_internal.h"
#include "esth_kernels.h"
#include "esth_dispatch.h"

#define ESTH_BATCH_TIMING_STEPS 8192

/* A few host threads that build the step's per-filter frames in parallel
 * (10^4 filters take milliseconds on one core, as long as the device step).
 * Every worker takes part in every generation, so no worker can still be in
 * one generation's job when the next is posted. */
struct esth_pool {
    std::vector<std::thread> th;
    std::mutex mu;
    std::condition_variable cv, done_cv;
    std::function<void(int)> job;
    int njobs = 0, finished = 0, nth = 0;
    std::atomic<int> next{0};
    uint64_t gen = 0;
    bool stop = false;
    void work() {
        for (int j = next.fetch_add(1); j < njobs; j = next.fetch_add(1)) job(j);
    }
    void start(int n) {
        nth = n;
        for (int i = 0; i < n; ++i)
            th.emplace_back([this] {
                uint64_t seen = 0;
                for (;;) {
                    {
                        std::unique_lock<std::mutex> lk(mu);
                        cv.wait(lk, [&] { return stop || gen != seen; });
                        if (stop) return;
                        seen = gen;
                    }
                    work();
                    std::lock_guard<std::mutex> lk(mu);
                    if (++finished == nth) done_cv.notify_all();
                }
            });
    }
    void run(int n, std::function<void(int)> f) {
        {
            std::lock_guard<std::mutex> lk(mu);
            job = std::move(f);
            njobs = n;
            next = 0;
            finished = 0;
            ++gen;
        }
        cv.notify_all();
        work();
        std::unique_lock<std::mutex> lk(mu);
        done_cv.wait(lk, [&] { return finished == nth; });
    }
    ~esth_pool() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        for (auto& t : th) t.join();
    }
};

struct esth_bclass {
    int logm = 0;
    int lanes = 64;       /* the resampling's lanes per sub-filter (esth_kb_resample_lanes) */
    uint32_t n = 0;
    uint2* d_sub = nullptr;
    bool fused = false;   /* sample+weight and sort in one launch (kb_sample_sort) */
    uint2* d_units = nullptr;   /* m = 64 fused: pairs of sub-filters (esth_pair_units) */
    uint32_t nunits = 0;
};

size_t esth_pair_units(const uint2* cls, size_t n, uint2* units) {
    size_t k = 0;
    for (size_t i = 0; i < n;) {
        if (i + 1 < n && cls[i + 1].x == cls[i].x && cls[i + 1].y == cls[i].y + 1) {
            units[k++] = cls[i];
            i += 2;
        } else {
            units[k++] = make_uint2(cls[i].x, cls[i].y | ESTH_UNIT_SINGLE);
            i += 1;
        }
    }
    return k;
}

/* kb_sample_sort<MODEL, LOGM, S> for size class lm: 0 it cannot run (not
 * compiled for the model, or it spills), 1 it runs with fewer waves resident
 * per SIMD than kb_sample_weight<MODEL>, 2 with at least as many */
template <int MODEL, int S>
static int kb_fuse_grade_m(int lm) {
    if constexpr (esth_model_traits<MODEL>::S != S) {
        return 0;
    } else {
        const void* fn = nullptr;
        switch (lm) {
        case 6: fn = (const void*)&kb_sample_sort64<MODEL, S>; break;
        case 7: fn = (const void*)&kb_sample_sort<MODEL, 7, S>; break;
        case 8: fn = (const void*)&kb_sample_sort<MODEL, 8, S>; break;
        case 9: fn = (const void*)&kb_sample_sort<MODEL, 9, S>; break;
        case 10: fn = (const void*)&kb_sample_sort<MODEL, 10, S>; break;
        default: return 0;
        }
        const int waves = (lm >= 9 || lm == 6) ? 4 : (1 << lm) / 128;
        const int fused = esth_waves_per_simd(fn, ESTH_WAVE * waves);
        if (fused <= 0) return 0;
        return fused >= esth_waves_per_simd((const void*)&kb_sample_weight<MODEL>, 256) ? 2 : 1;
    }
}

/* grades cached per (model, class): grade + 1, 0 unknown */
static std::atomic<int> g_kb_fuse[16][11];

static int kb_fuse_grade(int model, int S, int lm) {
    if (model < 0 || model >= 16 || lm < 0 || lm > 10) return 0;
    const int c = g_kb_fuse[model][lm].load(std::memory_order_relaxed);
    if (c) return c - 1;
    const int grade = esth_with_model_or(model, 0, [&](auto id) {
        constexpr int MODEL = decltype(id)::value;
        return esth_model_traits<MODEL>::S == S ? kb_fuse_grade_m<MODEL, esth_model_traits<MODEL>::S>(lm) : 0;
    });
    g_kb_fuse[model][lm].store(grade + 1, std::memory_order_relaxed);
    return grade;
}

bool esth_kb_fuse_with(int mode, int model, int S, int lm) {
    const int grade = kb_fuse_grade(model, S, lm);
    if (grade == 0 || mode == ESTH_FUSION_OFF) return false;
    return mode == ESTH_FUSION_ON || grade == 2;
}

int esth_kb_fuse_mode(const esth_engine* e, int lm) {
    return (e && lm >= 0 && lm <= 10) ? e->batch_fusion[lm].load(std::memory_order_relaxed) : ESTH_FUSION_AUTO;
}

bool esth_kb_fuse_pays(const esth_engine* e, int model, int S, int lm) {
    return esth_kb_fuse_with(esth_kb_fuse_mode(e, lm), model, S, lm);
}

static int fusion_class(uint32_t m) {
    for (int lm = 6; lm <= 10; ++lm)
        if (m == (1u << lm)) return lm;
    return -1;
}

extern "C" int esth_engine_set_batch_fusion(esth_engine* e, uint32_t m, int mode) {
    const int lm = fusion_class(m);
    if (!e || lm < 0 || (mode != ESTH_FUSION_AUTO && mode != ESTH_FUSION_OFF && mode != ESTH_FUSION_ON))
        return esth_fail(ESTH_ERR_INVALID, "batch fusion: m %u, mode %d", m, mode);
    e->batch_fusion[lm].store(mode, std::memory_order_relaxed);
    return ESTH_OK;
}

static bool lanes_valid(int lm, int lanes) {
    return lanes == 0 || lanes == 64 || (lanes == 32 && lm == 6) || (lanes == 128 && lm >= 8);
}

int esth_kb_resample_lanes(const esth_engine* e, int lm) {
    const int l = (e && lm >= 0 && lm <= 10) ? e->batch_lanes[lm].load(std::memory_order_relaxed) : 0;
    return (l && lanes_valid(lm, l)) ? l : 64;
}

extern "C" int esth_engine_set_batch_resample_lanes(esth_engine* e, uint32_t m, int lanes) {
    const int lm = fusion_class(m);
    if (!e || lm < 0 || !lanes_valid(lm, lanes))
        return esth_fail(ESTH_ERR_INVALID, "batch resampling: m %u with %d lanes per sub-filter", m, lanes);
    e->batch_lanes[lm].store(lanes, std::memory_order_relaxed);
    return ESTH_OK;
}

extern "C" int esth_engine_batch_resample_lanes(const esth_engine* e, uint32_t m, int* lanes) {
    const int lm = fusion_class(m);
    if (!e || !lanes || lm < 0) return esth_fail(ESTH_ERR_INVALID, "batch resampling: m %u", m);
    *lanes = esth_kb_resample_lanes(e, lm);
    return ESTH_OK;
}

extern "C" int esth_engine_batch_fusion(const esth_engine* e, uint32_t m, int* mode) {
    const int lm = fusion_class(m);
    if (!e || !mode || lm < 0) return esth_fail(ESTH_ERR_INVALID, "batch fusion: m %u", m);
    *mode = e->batch_fusion[lm].load(std::memory_order_relaxed);
    return ESTH_OK;
}

struct esth_batch {
    esth_engine* e = nullptr;
    int n = 0;
    esth_filter_desc d{};   /* shared settings (model, topology, t, resampler, ...) */
    int S = 0, V = 0;
    int nq = 1;
    size_t frame_stride = 0;
    std::vector<esth_filter_desc> descs;
    std::vector<std::vector<unsigned char>> params;
    std::vector<esth_bfilter> hf;   /* host copy of the table */
    std::vector<uint64_t> soff, woff, stream0;
    std::vector<int> K;
    std::vector<uint64_t> rsize;
    uint64_t Ntot = 0, words = 0;
    int nstreams = 0;
    uint32_t nsub = 0;
    hipStream_t stream = nullptr, rng_stream = nullptr;
    /* device */
    esth_bfilter* d_F = nullptr;
    uint2* d_sub_all = nullptr;
    uint2* d_chunks = nullptr;   /* (filter, first pair) per 64 particle pairs */
    uint32_t nchunks = 0;
    std::vector<esth_bclass> cls;
    float *state = nullptr, *tmp = nullptr, *w = nullptr, *part = nullptr, *bt_w = nullptr, *bt_s = nullptr;
    float* R[2] = {nullptr, nullptr};
    float* frames[2] = {nullptr, nullptr};
    uint32_t* status[2] = {nullptr, nullptr};
    esth_mtgp_stream* streams = nullptr;
    float2 *winner = nullptr, *result = nullptr;
    double* est = nullptr;
    uint32_t* top = nullptr;
    std::vector<std::pair<void*, size_t>> allocs;
    /* host */
    float* h_frames[2] = {nullptr, nullptr};
    hipEvent_t frame_ev[2];
    bool frame_used[2] = {false, false};
    bool rng_ready = false, prefetch = true;
    hipEvent_t ev_rng_done[2], ev_consumed[2], ev_fork;
    bool use_graph = true;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    uint64_t steps = 0;
    /* step latency: an event pair per step */
    std::vector<hipEvent_t> tev;
    int tsteps = 0;
    std::unique_ptr<esth_pool> pool;
};

static int balloc(esth_batch* b, size_t bytes, void** p) {
    ESTH_TRY(esth_engine_alloc(b->e, bytes ? bytes : 4, p));
    b->allocs.push_back({*p, bytes ? bytes : 4});
    return ESTH_OK;
}

static void batch_free(esth_batch* b) {
    if (!b) return;
    hipSetDevice(b->e->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->rng_stream) hipStreamSynchronize(b->rng_stream);
    for (int i = 0; i < 2; ++i) {
        if (b->gexec[i]) hipGraphExecDestroy(b->gexec[i]);
        if (b->h_frames[i]) hipHostFree(b->h_frames[i]);
        hipEventDestroy(b->frame_ev[i]);
        hipEventDestroy(b->ev_rng_done[i]);
        hipEventDestroy(b->ev_consumed[i]);
    }
    hipEventDestroy(b->ev_fork);
    for (auto ev : b->tev) hipEventDestroy(ev);
    for (auto& a : b->allocs) esth_engine_release(b->e, a.first, a.second);
    if (b->stream) hipStreamDestroy(b->stream);
    if (b->rng_stream) hipStreamDestroy(b->rng_stream);
    delete b;
}

static size_t frame_floats(int model, const void* params) { return (size_t)esth_model_frame_floats(model, params); }

extern "C" int esth_batch_create(esth_engine* e, int n, const esth_filter_desc* descs, const void* params,
                                 size_t params_stride, esth_batch** out) {
    if (!e || !descs || n <= 0 || !out) return esth_fail(ESTH_ERR_INVALID, "batch: bad arguments");
    const esth_filter_desc& d0 = descs[0];
    int S, V;
    size_t psize = 0;
    ESTH_TRY(esth_params_size(d0.model, &psize));
    if (psize && (!params || params_stride < psize)) return esth_fail(ESTH_ERR_INVALID, "batch: params stride %zu < %zu", params_stride, psize);
    for (int i = 0; i < n; ++i) {
        const esth_filter_desc& d = descs[i];
        ESTH_TRY(esth_check_desc(&d, &S, &V));
        if (d.model != d0.model || d.topology != d0.topology || d.num_transfer != d0.num_transfer ||
            d.resampler != d0.resampler || d.resampling_threshold != d0.resampling_threshold ||
            d.adaptive != d0.adaptive || d.weights != d0.weights)
            return esth_fail(ESTH_ERR_INVALID, "batch: filter %d differs from filter 0 in model/topology/t/resampler/weights", i);
        if (!esth_rng_is_per_filter(d.rng))
            return esth_fail(ESTH_ERR_INVALID, "batch: filter %d must use a per-filter RNG mode", i);
        if (d.state_layout != ESTH_LAYOUT_SOA)
            return esth_fail(ESTH_ERR_UNSUPPORTED, "batch: filter %d: batches keep their state in ESTH_LAYOUT_SOA", i);
        if (d.num_particles < 64) return esth_fail(ESTH_ERR_UNSUPPORTED, "batch: sub-filters need m >= 64 (filter %d)", i);
    }
    ESTH_HIP(hipSetDevice(e->device));
    esth_batch* b = new esth_batch();
    b->e = e;
    b->n = n;
    b->d = d0;
    b->S = S;
    b->V = V;
    b->nq = d0.topology == ESTH_TOPO_TORUS ? 4 : 1;
    for (int i = 0; i < 2; ++i) {
        hipEventCreateWithFlags(&b->frame_ev[i], hipEventDisableTiming);
        hipEventCreateWithFlags(&b->ev_rng_done[i], hipEventDisableTiming);
        hipEventCreateWithFlags(&b->ev_consumed[i], hipEventDisableTiming);
    }
    hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
    b->descs.assign(descs, descs + n);
    b->params.resize(n);
    b->hf.resize(n);
    b->soff.resize(n);
    b->woff.resize(n);
    b->stream0.resize(n);
    b->K.resize(n);
    b->rsize.resize(n);
    /* layout */
    uint64_t Gtot = 0, tGtot = 0, fl = 0;
    for (int i = 0; i < n; ++i) {
        const esth_filter_desc& d = descs[i];
        const unsigned char* pp = psize ? (const unsigned char*)params + (size_t)i * params_stride : nullptr;
        if (psize) b->params[i].assign(pp, pp + psize);
        b->frame_stride = std::max(b->frame_stride, frame_floats(d.model, pp));
        const uint64_t N = (uint64_t)d.num_particles * d.num_blocks;
        const uint64_t need = esth_rng_needed(V, d.resampler, d.adaptive, d.num_particles, d.num_blocks);
        b->K[i] = esth_rng_streams(descs[i].rng, need);
        b->rsize[i] = esth_rng_stream_size(b->K[i], need);
        b->soff[i] = b->Ntot * (uint64_t)S;
        b->woff[i] = b->Ntot;
        b->stream0[i] = (uint64_t)b->nstreams;
        esth_bfilter& f = b->hf[i];
        memset(&f, 0, sizeof(f));
        f.N = (uint32_t)N;
        f.G = d.num_blocks;
        f.m = d.num_particles;
        f.t = d.topology == ESTH_TOPO_NONE ? 0 : d.num_transfer;
        f.dimx = d.torus_dimx;
        f.dimy = d.torus_dimy;
        if (d.topology == ESTH_TOPO_TORUS && (f.dimx <= 0 || f.dimy <= 0)) esth_torus_dims(f.G, &f.dimx, &f.dimy);
        if (f.dimx <= 0) f.dimx = 1;
        if (f.dimy <= 0) f.dimy = 1;
        f.uoff = (uint32_t)(N * (uint64_t)V);
        f.flags = d.carry_weights ? 1u : 0u;
        f.r_off = b->words;
        b->Ntot += N;
        b->words += b->rsize[i] * (uint64_t)b->K[i];
        b->nstreams += b->K[i];
        Gtot += d.num_blocks;
        tGtot += (uint64_t)(f.t ? f.t : 1) * d.num_blocks;
        (void)fl;
    }
    for (int i = 0; i < n; ++i) b->hf[i].f_off = (uint64_t)i * b->frame_stride;
    b->nsub = (uint32_t)Gtot;
    int rc = ESTH_OK;
#define BA(ptr, bytes) if (rc == ESTH_OK) rc = balloc(b, (bytes), (void**)&(ptr))
    BA(b->state, sizeof(float) * b->Ntot * S);
    BA(b->tmp, sizeof(float) * b->Ntot * S);
    BA(b->w, sizeof(float) * b->Ntot);
    BA(b->R[0], sizeof(float) * b->words);
    BA(b->R[1], sizeof(float) * b->words);
    BA(b->status[0], sizeof(uint32_t) * ESTH_MTGP_N * (size_t)b->nstreams);
    BA(b->status[1], sizeof(uint32_t) * ESTH_MTGP_N * (size_t)b->nstreams);
    BA(b->streams, sizeof(esth_mtgp_stream) * (size_t)b->nstreams);
    BA(b->part, sizeof(float) * Gtot * (2 + S));
    BA(b->winner, sizeof(float2) * Gtot);
    BA(b->bt_w, sizeof(float) * tGtot);
    BA(b->bt_s, sizeof(float) * tGtot * S);
    BA(b->result, sizeof(float2) * n);
    BA(b->est, sizeof(double) * (4 + 2 * ESTH_MAX_STATE_DIM) * n);
    BA(b->top, sizeof(uint32_t) * 64 * n);
    BA(b->frames[0], sizeof(float) * b->frame_stride * n);
    BA(b->frames[1], sizeof(float) * b->frame_stride * n);
    BA(b->d_F, sizeof(esth_bfilter) * n);
    BA(b->d_sub_all, sizeof(uint2) * Gtot);
    /* size classes whose sub-filters sample and sort in one launch; the
     * other filters' particles go through kb_sample_weight's chunks */
    bool fuse[11] = {};
    for (int lm = 6; lm <= 10; ++lm) fuse[lm] = esth_kb_fuse_pays(e, b->d.model, S, lm);
    auto class_of = [](uint32_t m) { int k = 0; while ((1u << k) < m) ++k; return k; };
    for (int i = 0; i < n; ++i)
        if (!fuse[class_of(b->hf[i].m)]) b->nchunks += (b->hf[i].N / 2 + ESTH_WAVE - 1) / ESTH_WAVE;
    if (b->nchunks) BA(b->d_chunks, sizeof(uint2) * b->nchunks);
#undef BA
    if (rc != ESTH_OK) { batch_free(b); return rc; }
    /* filter table, sub-filter tables, RNG streams */
    uint64_t goff = 0, tgoff = 0;
    std::vector<std::vector<uint2>> by_class(11);
    std::vector<uint2> all;
    all.reserve(Gtot);
    std::vector<esth_mtgp_stream> hs((size_t)b->nstreams);
    std::vector<uint32_t> st((size_t)ESTH_MTGP_N * b->nstreams);
    for (int i = 0; i < n; ++i) {
        esth_bfilter& f = b->hf[i];
        f.state = b->state + b->soff[i];
        f.tmp = b->tmp + b->soff[i];
        f.w = b->w + b->woff[i];
        f.part = b->part + goff * (2 + S);
        f.winner = (void*)(b->winner + goff);
        f.bt_w = b->bt_w + tgoff;
        f.bt_s = b->bt_s + tgoff * S;
        f.result = (void*)(b->result + i);
        f.est = b->est + (size_t)i * (4 + 2 * ESTH_MAX_STATE_DIM);
        f.top = b->top + (size_t)i * 64;
        int lm = 0;
        while ((1u << lm) < f.m) ++lm;
        for (uint32_t g = 0; g < f.G; ++g) {
            by_class[lm].push_back(make_uint2((uint32_t)i, g));
            all.push_back(make_uint2((uint32_t)i, g));
        }
        for (int j = 0; j < b->K[i]; ++j) {
            const size_t sj = (size_t)b->stream0[i] + (size_t)j;
            hs[sj] = {(uint32_t)j, (uint32_t)b->rsize[i], f.r_off + (uint64_t)j * b->rsize[i], (uint64_t)sj};
            esth_mtgp_seed_state(j, descs[i].seed + (uint32_t)j + 1u, st.data() + sj * ESTH_MTGP_N);
        }
        goff += f.G;
        tgoff += (uint64_t)(f.t ? f.t : 1) * f.G;
    }
    hipError_t err = hipMemcpy(b->d_F, b->hf.data(), sizeof(esth_bfilter) * n, hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(b->d_sub_all, all.data(), sizeof(uint2) * all.size(), hipMemcpyHostToDevice);
    std::vector<uint2> chunks;
    chunks.reserve(b->nchunks);
    for (int i = 0; i < n; ++i)
        if (!fuse[class_of(b->hf[i].m)])
            for (uint32_t q = 0; q < b->hf[i].N / 2; q += ESTH_WAVE) chunks.push_back(make_uint2((uint32_t)i, q));
    if (err == hipSuccess && !chunks.empty())
        err = hipMemcpy(b->d_chunks, chunks.data(), sizeof(uint2) * chunks.size(), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(b->streams, hs.data(), sizeof(esth_mtgp_stream) * hs.size(), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(b->status[0], st.data(), sizeof(uint32_t) * st.size(), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemset(b->w, 0, sizeof(float) * b->Ntot);
    if (err == hipSuccess) err = hipMemset(b->est, 0, sizeof(double) * (4 + 2 * ESTH_MAX_STATE_DIM) * n);
    if (err != hipSuccess) { batch_free(b); return esth_fail(ESTH_ERR_DEVICE, "batch upload: %s", hipGetErrorString(err)); }
    for (int lm = 0; lm <= 10; ++lm) {
        if (by_class[lm].empty()) continue;
        esth_bclass c;
        c.logm = lm;
        c.n = (uint32_t)by_class[lm].size();
        c.fused = fuse[lm];
        c.lanes = esth_kb_resample_lanes(e, lm);
        rc = balloc(b, sizeof(uint2) * c.n, (void**)&c.d_sub);
        if (rc == ESTH_OK && hipMemcpy(c.d_sub, by_class[lm].data(), sizeof(uint2) * c.n, hipMemcpyHostToDevice) != hipSuccess)
            rc = esth_fail(ESTH_ERR_DEVICE, "batch class upload");
        if (rc == ESTH_OK && c.fused && lm == 6) {
            std::vector<uint2> units(c.n);
            c.nunits = (uint32_t)esth_pair_units(by_class[lm].data(), c.n, units.data());
            rc = balloc(b, sizeof(uint2) * c.nunits, (void**)&c.d_units);
            if (rc == ESTH_OK && hipMemcpy(c.d_units, units.data(), sizeof(uint2) * c.nunits, hipMemcpyHostToDevice) != hipSuccess)
                rc = esth_fail(ESTH_ERR_DEVICE, "batch unit upload");
        }
        if (rc != ESTH_OK) { batch_free(b); return rc; }
        b->cls.push_back(c);
    }
    err = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&b->rng_stream, hipStreamNonBlocking);
    for (int i = 0; i < 2 && err == hipSuccess; ++i)
        err = hipHostMalloc((void**)&b->h_frames[i], sizeof(float) * b->frame_stride * n, 0);
    if (err != hipSuccess) { batch_free(b); return esth_fail(ESTH_ERR_DEVICE, "batch streams: %s", hipGetErrorString(err)); }
    *out = b;
    return ESTH_OK;
}

extern "C" int esth_batch_destroy(esth_batch* b) {
    batch_free(b);
    return ESTH_OK;
}

extern "C" int esth_batch_init_uniform(esth_batch* b, const float* states) {
    if (!b || !states) return esth_fail(ESTH_ERR_INVALID, "null");
    ESTH_HIP(hipSetDevice(b->e->device));
    ESTH_HIP(hipStreamSynchronize(b->stream));
    std::vector<float> h((size_t)b->Ntot * b->S);
    for (int i = 0; i < b->n; ++i) {
        const esth_bfilter& f = b->hf[i];
        for (int c = 0; c < b->S; ++c)
            std::fill(h.begin() + b->soff[i] + (size_t)c * f.N, h.begin() + b->soff[i] + (size_t)(c + 1) * f.N,
                      states[(size_t)i * b->S + c]);
    }
    ESTH_HIP(hipMemcpy(b->state, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    /* carried weights restart equal (0 in log mode, 1 linear) */
    std::vector<float> w0((size_t)b->Ntot, 0.0f);
    for (int i = 0; i < b->n; ++i)
        if (b->descs[i].carry_weights && b->d.weights != ESTH_WEIGHTS_LOG)
            std::fill(w0.begin() + b->woff[i], w0.begin() + b->woff[i] + b->hf[i].N, 1.0f);
    ESTH_HIP(hipMemcpy(b->w, w0.data(), sizeof(float) * w0.size(), hipMemcpyHostToDevice));
    return ESTH_OK;
}

/* ------------------------------------------------------------ launches */
template <int LOGM, int S>
static void launch_kb_sort(esth_batch* b, const esth_bclass& c, int log_w) {
    hipLaunchKernelGGL((kb_sort<LOGM, S>), dim3((c.n + 3) / 4), dim3(256), 0, b->stream, b->d_F, c.d_sub, c.n, log_w);
}
template <int LOGM, int S>
static void launch_kb_alias(esth_batch* b, const esth_bclass& c, const float* R, int log_w) {
    constexpr int WPB = esth_alias_wpb<LOGM>::value;
    hipLaunchKernelGGL((kb_alias_w<LOGM, S>), dim3((c.n + WPB - 1) / WPB), dim3(WPB * ESTH_WAVE), 0, b->stream, b->d_F,
                       c.d_sub, c.n, R, b->d.resampling_threshold, b->d.adaptive, log_w);
}
/* am: work-groups running every filter's argmax ahead of the resampling
 * work-groups (the server's first resampling launch) */
template <int LOGM, int S>
static void launch_kb_resample_t(int lanes, const esth_bfilter* F, const uint2* sub, uint32_t n, const float* R,
                                 float thr, int sys, int log_w, int topo, uint32_t am, hipStream_t st) {
    if constexpr (LOGM == 6) {
        if (lanes == 32) {   /* two entries per wave, four waves per work-group */
            hipLaunchKernelGGL((kb_resample64x2<S>), dim3(am + (n + 7) / 8), dim3(256), 0, st, F, sub, n, R, thr, sys,
                               log_w, topo, am);
            return;
        }
    }
    if constexpr (LOGM >= 8) {
        if (lanes == 128) {   /* two waves per entry, two entries per work-group */
            hipLaunchKernelGGL((kb_resample2<LOGM, S>), dim3(am + (n + 1) / 2), dim3(256), 0, st, F, sub, n, R, thr,
                               sys, log_w, topo, am);
            return;
        }
    }
    hipLaunchKernelGGL((kb_resample<LOGM, S>), dim3(am + (n + 3) / 4), dim3(256), 0, st, F, sub, n, R, thr, sys, log_w,
                       topo, am);
}

int esth_launch_kb_resample(int S, int lm, int lanes, const esth_bfilter* F, const uint2* sub, uint32_t n,
                            const float* R, float thr, int systematic, int log_w, int topo, uint32_t argmax_blocks,
                            hipStream_t st) {
    if (n == 0 && argmax_blocks == 0) return ESTH_OK;
    return esth_with_model_state_dim(S, [&](auto sv) {
        constexpr int SV = decltype(sv)::value;
        switch (lm) {
        case 6: launch_kb_resample_t<6, SV>(lanes, F, sub, n, R, thr, systematic, log_w, topo, argmax_blocks, st); break;
        case 7: launch_kb_resample_t<7, SV>(lanes, F, sub, n, R, thr, systematic, log_w, topo, argmax_blocks, st); break;
        case 8: launch_kb_resample_t<8, SV>(lanes, F, sub, n, R, thr, systematic, log_w, topo, argmax_blocks, st); break;
        case 9: launch_kb_resample_t<9, SV>(lanes, F, sub, n, R, thr, systematic, log_w, topo, argmax_blocks, st); break;
        case 10: launch_kb_resample_t<10, SV>(lanes, F, sub, n, R, thr, systematic, log_w, topo, argmax_blocks, st); break;
        default: return esth_fail(ESTH_ERR_INVALID, "m = 2^%d", lm);
        }
        return ESTH_OK;
    });
}

/* sample+weight and sort of a fused size class (m >= 64) of n sub-filters
 * (filter, sub-filter) in `sub`, filters from F */
template <int MODEL, int S>
static int launch_kbss_m(int lm, const esth_bfilter* F, const uint2* sub, uint32_t n, const float* R,
                         const float* frames, int log_w, float thr, int keep_all, hipStream_t st) {
    if constexpr (esth_model_traits<MODEL>::S != S) {
        return esth_fail(ESTH_ERR_INVALID, "state dimension");
    } else {
#define ESTH_KBSS(LM)                                                                                         \
    hipLaunchKernelGGL((kb_sample_sort<MODEL, LM, S>), dim3(n), dim3(ESTH_WAVE * esth_ss_waves<LM>()), 0, st, F, \
                       sub, R, frames, log_w, thr, keep_all)
        switch (lm) {
        case 6:   /* sub: n units (esth_pair_units), one per wave, four waves per work-group */
            hipLaunchKernelGGL((kb_sample_sort64<MODEL, S>), dim3((n + 3) / 4), dim3(256), 0, st, F, sub, n, R, frames,
                               log_w, thr, keep_all);
            break;
        case 7: ESTH_KBSS(7); break;
        case 8: ESTH_KBSS(8); break;
        case 9: ESTH_KBSS(9); break;
        case 10: ESTH_KBSS(10); break;
        default: return esth_fail(ESTH_ERR_INVALID, "fused class m");
        }
#undef ESTH_KBSS
        return ESTH_OK;
    }
}

int esth_launch_kb_sample_sort(int model, int S, int lm, const esth_bfilter* F, const uint2* sub, uint32_t n,
                               const float* R, const float* frames, int log_w, float thr, int keep_all,
                               hipStream_t st) {
    if (n == 0) return ESTH_OK;
    return esth_with_model(model, [&](auto id) {
        constexpr int MODEL = decltype(id)::value;
        if (esth_model_traits<MODEL>::S != S) return esth_fail(ESTH_ERR_UNSUPPORTED, "state dimension %d", S);
        return launch_kbss_m<MODEL, esth_model_traits<MODEL>::S>(lm, F, sub, n, R, frames, log_w, thr, keep_all, st);
    });
}

#define ESTH_BCLASS(FN, S, ...)                        \
    switch (c.logm) {                                  \
    case 6: FN<6, S>(__VA_ARGS__); break;              \
    case 7: FN<7, S>(__VA_ARGS__); break;              \
    case 8: FN<8, S>(__VA_ARGS__); break;              \
    case 9: FN<9, S>(__VA_ARGS__); break;              \
    case 10: FN<10, S>(__VA_ARGS__); break;            \
    default: return esth_fail(ESTH_ERR_INVALID, "m"); \
    }

template <int S>
static int batch_kernels_s(esth_batch* b, int p) {
    const int log_w = b->d.weights == ESTH_WEIGHTS_LOG;
    const float* R = b->R[p];
    const float* frames = b->frames[p];
    hipStream_t s = b->stream;
    const uint32_t nsub = b->nsub;
    const int grid_w = (int)(((uint64_t)b->nchunks * ESTH_WAVE + 255) / 256);
    if (b->nchunks)
        ESTH_TRY(esth_with_model(b->d.model, [&](auto id) {
            hipLaunchKernelGGL(kb_sample_weight<decltype(id)::value>, dim3(grid_w), dim3(256), 0, s, b->d_F,
                               b->d_chunks, b->nchunks, R, frames, log_w);
            return ESTH_OK;
        }));
    for (const esth_bclass& c : b->cls) {
        if (c.fused)
            ESTH_TRY(esth_launch_kb_sample_sort(b->d.model, S, c.logm, b->d_F, c.logm == 6 ? c.d_units : c.d_sub,
                                                c.logm == 6 ? c.nunits : c.n, R, frames, log_w,
                                                b->d.resampling_threshold,
                                                b->d.resampler == ESTH_RESAMPLE_ALIAS && b->d.adaptive, s));
        else
            ESTH_BCLASS(launch_kb_sort, S, b, c, log_w);
    }
    const uint32_t t = b->d.topology == ESTH_TOPO_NONE ? 0 : b->d.num_transfer;
    if (t > 0 && b->d.topology == ESTH_TOPO_ALL2ALL) hipLaunchKernelGGL(kb_topt, dim3(b->n), dim3(1024), 0, s, b->d_F);
    if (b->d.resampler == ESTH_RESAMPLE_ALIAS) {
        hipLaunchKernelGGL(kb_argmax<S>, dim3(b->n), dim3(256), 0, s, b->d_F, log_w);
        if (t > 0) {
            const uint32_t nslots = (uint32_t)b->nq * t;
            const uint64_t total = (uint64_t)nsub * nslots;
            hipLaunchKernelGGL(kb_exchange, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, b->d_F, b->d_sub_all,
                               nsub, S, nslots, b->d.topology);
        }
        for (const esth_bclass& c : b->cls) ESTH_BCLASS(launch_kb_alias, S, b, c, R, log_w);
    } else {
        /* the exchange's receiving side runs inside the resampling */
        hipLaunchKernelGGL(kb_argmax<S>, dim3(b->n), dim3(256), 0, s, b->d_F, log_w);
        for (const esth_bclass& c : b->cls)
            ESTH_TRY(esth_launch_kb_resample(S, c.logm, c.lanes, b->d_F, c.d_sub, c.n, R, b->d.resampling_threshold,
                                             b->d.resampler == ESTH_RESAMPLE_SYSTEMATIC ? 1 : 0, log_w, b->d.topology,
                                             0u, b->stream));
    }
    ESTH_HIP(hipGetLastError());
    return ESTH_OK;
}

static int batch_kernels(esth_batch* b, int p) {
    return esth_with_model_state_dim(b->S, [&](auto sv) { return batch_kernels_s<decltype(sv)::value>(b, p); });
}

static int launch_rng_b(esth_batch* b, uint64_t k, hipStream_t s) {
    const int p = (int)(k & 1);
    return esth_launch_mtgp(b->e->d_params, b->status[p], b->status[p ^ 1], b->streams, b->nstreams, b->R[p], 1, s);
}

static int capture_batch(esth_batch* b, int p) {
    hipGraph_t g = nullptr;
    ESTH_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    ESTH_HIP(hipMemcpyAsync(b->frames[p], b->h_frames[p], sizeof(float) * b->frame_stride * b->n, hipMemcpyHostToDevice,
                            b->stream));
    int rc = batch_kernels(b, p);
    hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (rc != ESTH_OK) { if (g) hipGraphDestroy(g); return rc; }
    if (e2 != hipSuccess) return esth_fail(ESTH_ERR_DEVICE, "batch capture: %s", hipGetErrorString(e2));
    hipError_t e3 = hipGraphInstantiate(&b->gexec[p], g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e3 != hipSuccess) return esth_fail(ESTH_ERR_DEVICE, "batch instantiate: %s", hipGetErrorString(e3));
    return ESTH_OK;
}

extern "C" int esth_batch_step(esth_batch* b, const void* controls, size_t control_stride, const void* measurements,
                               size_t meas_stride, const float* dts) {
    if (!b || !measurements || !dts) return esth_fail(ESTH_ERR_INVALID, "null");
    ESTH_HIP(hipSetDevice(b->e->device));
    const int p = (int)(b->steps & 1);
    if (b->frame_used[p]) ESTH_HIP(hipEventSynchronize(b->frame_ev[p]));
    float* hf = b->h_frames[p];
    std::atomic<int> bad{-1};
    auto build = [&](int i0, int i1) {
        for (int i = i0; i < i1; ++i) {
            const void* c = controls ? (const unsigned char*)controls + (size_t)i * control_stride : nullptr;
            const void* m = (const unsigned char*)measurements + (size_t)i * meas_stride;
            const void* pp = b->params[i].empty() ? nullptr : b->params[i].data();
            static const esth_dh6_control none = {0.0f};
            if (b->d.model == ESTH_MODEL_DH6 && !c) c = &none;
            if (esth_build_frame_n(b->d.model, pp, c, m, dts[i], hf + (size_t)i * b->frame_stride, (int)b->frame_stride))
                bad = i;
        }
    };
    const int chunk = 256;
    if (b->n >= 2 * chunk) {
        if (!b->pool) {
            b->pool.reset(new esth_pool());
            const unsigned hc = std::thread::hardware_concurrency();
            b->pool->start((int)std::min<unsigned>(7u, hc > 1 ? hc - 1 : 1u));
        }
        b->pool->run((b->n + chunk - 1) / chunk, [&](int j) { build(j * chunk, std::min(b->n, (j + 1) * chunk)); });
    } else {
        build(0, b->n);
    }
    if (bad >= 0) return esth_fail(ESTH_ERR_INVALID, "batch: bad control/measurement for filter %d", bad.load());
    if (b->use_graph && !b->gexec[p]) ESTH_TRY(capture_batch(b, p));
    hipEvent_t* ev = nullptr;
    if (b->tsteps < ESTH_BATCH_TIMING_STEPS) {
        if ((size_t)(b->tsteps + 1) * 2 > b->tev.size()) {
            for (int i = 0; i < 2; ++i) {
                hipEvent_t e;
                ESTH_HIP(hipEventCreate(&e));
                b->tev.push_back(e);
            }
        }
        ev = &b->tev[(size_t)b->tsteps * 2];
        b->tsteps++;
    }
    /* RNG: this step's words (if not prefetched) and the next step's on the RNG stream */
    hipStream_t s = b->stream;
    const uint64_t n = b->steps;
    bool forked = false;
    if (!b->rng_ready) {
        ESTH_TRY(launch_rng_b(b, n, s));
        if (b->prefetch) {
            ESTH_HIP(hipEventRecord(b->ev_fork, s));
            ESTH_HIP(hipStreamWaitEvent(b->rng_stream, b->ev_fork, 0));
            forked = true;
        }
    } else {
        ESTH_HIP(hipStreamWaitEvent(s, b->ev_rng_done[p], 0));
    }
    if (b->prefetch) {
        if (!forked && n > 0) ESTH_HIP(hipStreamWaitEvent(b->rng_stream, b->ev_consumed[p ^ 1], 0));
        ESTH_TRY(launch_rng_b(b, n + 1, b->rng_stream));
        ESTH_HIP(hipEventRecord(b->ev_rng_done[p ^ 1], b->rng_stream));
    }
    if (ev) ESTH_HIP(hipEventRecord(ev[0], s));
    if (b->use_graph) {
        ESTH_HIP(hipGraphLaunch(b->gexec[p], s));
    } else {
        ESTH_HIP(hipMemcpyAsync(b->frames[p], hf, sizeof(float) * b->frame_stride * b->n, hipMemcpyHostToDevice, s));
        ESTH_TRY(batch_kernels(b, p));
    }
    if (ev) ESTH_HIP(hipEventRecord(ev[1], s));
    if (b->prefetch) ESTH_HIP(hipEventRecord(b->ev_consumed[p], s));
    b->rng_ready = b->prefetch;
    ESTH_HIP(hipEventRecord(b->frame_ev[p], s));
    b->frame_used[p] = true;
    b->steps++;
    return ESTH_OK;
}

extern "C" int esth_batch_sync(esth_batch* b) {
    if (!b) return esth_fail(ESTH_ERR_INVALID, "null");
    ESTH_HIP(hipStreamSynchronize(b->stream));
    ESTH_HIP(hipStreamSynchronize(b->rng_stream));
    return ESTH_OK;
}

extern "C" int esth_batch_use_graph(esth_batch* b, int enable) {
    if (!b) return esth_fail(ESTH_ERR_INVALID, "null");
    ESTH_TRY(esth_batch_sync(b));
    b->use_graph = enable != 0;
    return ESTH_OK;
}

extern "C" int esth_batch_estimate(esth_batch* b, int i, esth_estimate* out) {
    if (!b || !out || i < 0 || i >= b->n) return esth_fail(ESTH_ERR_INVALID, "batch estimate: bad filter %d", i);
    ESTH_HIP(hipStreamSynchronize(b->stream));
    const int S = b->S;
    double h[4 + 2 * ESTH_MAX_STATE_DIM];
    ESTH_HIP(hipMemcpy(h, b->est + (size_t)i * (4 + 2 * ESTH_MAX_STATE_DIM), sizeof(double) * (4 + 2 * (size_t)S),
                       hipMemcpyDeviceToHost));
    memset(out, 0, sizeof(*out));
    out->state_dim = S;
    out->log_weight_max = h[0];
    out->weight_sum = h[1];
    for (int c = 0; c < S; ++c) {
        out->mean[c] = h[1] > 0.0 ? h[2 + c] / h[1] : NAN;
        out->best[c] = (float)h[4 + S + c];
    }
    out->best_weight = (float)h[2 + S];
    out->best_block = (int)h[3 + S];
    out->step = b->steps;
    return ESTH_OK;
}

extern "C" int esth_batch_read_particles(esth_batch* b, int i, float* soa, float* w) {
    if (!b || i < 0 || i >= b->n) return esth_fail(ESTH_ERR_INVALID, "batch read: bad filter %d", i);
    ESTH_HIP(hipStreamSynchronize(b->stream));
    const esth_bfilter& f = b->hf[i];
    if (soa) ESTH_HIP(hipMemcpy(soa, f.state, sizeof(float) * (size_t)f.N * b->S, hipMemcpyDeviceToHost));
    if (w) ESTH_HIP(hipMemcpy(w, f.w, sizeof(float) * f.N, hipMemcpyDeviceToHost));
    return ESTH_OK;
}

/* device time of the steps since the last call (or creation), microseconds */
extern "C" int esth_batch_step_times(esth_batch* b, float* us, int max, int* count) {
    if (!b || !count) return esth_fail(ESTH_ERR_INVALID, "null");
    ESTH_HIP(hipStreamSynchronize(b->stream));
    int k = std::min(b->tsteps, max);
    for (int i = 0; i < k; ++i) {
        float ms = 0.0f;
        ESTH_HIP(hipEventElapsedTime(&ms, b->tev[(size_t)i * 2], b->tev[(size_t)i * 2 + 1]));
        if (us) us[i] = ms * 1000.0f;
    }
    *count = k;
    b->tsteps = 0;
    return ESTH_OK;
}

extern "C" int esth_batch_step_time(esth_batch* b, float* us) {
    if (!b || !us) return esth_fail(ESTH_ERR_INVALID, "null");
    if (b->tsteps == 0) return esth_fail(ESTH_ERR_STATE, "no timed step");
    ESTH_HIP(hipStreamSynchronize(b->stream));
    float ms = 0.0f;
    ESTH_HIP(hipEventElapsedTime(&ms, b->tev[(size_t)(b->tsteps - 1) * 2], b->tev[(size_t)(b->tsteps - 1) * 2 + 1]));
    *us = ms * 1000.0f;
    return ESTH_OK;
}

extern "C" int esth_batch_info(esth_batch* b, uint64_t* particles, int* streams, uint32_t* sub_filters) {
    if (!b) return esth_fail(ESTH_ERR_INVALID, "null");
    if (particles) *particles = b->Ntot;
    if (streams) *streams = b->nstreams;
    if (sub_filters) *sub_filters = b->nsub;
    return ESTH_OK;
}

/* ============================================================ autotuner
 * The batch path itself, timed per sub-filter size: for m = 64 ... 1024 a
 * one-class batch of 64 filters x 65,536 particles (2^22) with base's model
 * and settings takes 3 untimed steps and 10 timed ones (device time of the
 * graph-replayed step, median), once with sample+weight and the sort as two
 * launches and once fused where the fused kernel fits; then, at the faster
 * form, once more with the resampling's other tiling (32 lanes per
 * sub-filter at m = 64, 128 at m >= 256).  us_per_mparticle[k] is the best
 * cost per million particles, and the engine's batch fusion and resampling
 * tiling for that m are set to it.  A size the descriptor rejects (e.g. t
 * too large for m) costs +inf. */
static int time_batch_class(esth_engine* e, const esth_filter_desc* base, const void* params, size_t psize,
                            const void* control, const void* measurement, float dt, uint32_t m, float* us_per_m) {
    const int nf = 64;
    const uint32_t P = 1u << 16;
    std::vector<esth_filter_desc> descs(nf, *base);
    for (int i = 0; i < nf; ++i) {
        descs[i].num_particles = m;
        descs[i].num_blocks = P / m;
        descs[i].torus_dimx = 0;
        descs[i].torus_dimy = 0;
        descs[i].rng = esth_rng_is_per_filter(base->rng) ? base->rng : ESTH_RNG_MTGP_PER_FILTER;
        descs[i].state_layout = ESTH_LAYOUT_SOA;
        descs[i].seed = base->seed + 7919u * (uint32_t)i;
    }
    int S, V;
    if (esth_check_desc(&descs[0], &S, &V) != ESTH_OK) {
        *us_per_m = INFINITY;
        return ESTH_OK;
    }
    std::vector<unsigned char> pbuf(psize * nf);
    for (int i = 0; i < nf && psize; ++i) memcpy(pbuf.data() + psize * i, params, psize);
    esth_batch* b = nullptr;
    ESTH_TRY(esth_batch_create(e, nf, descs.data(), psize ? pbuf.data() : nullptr, psize, &b));
    std::vector<float> zero((size_t)nf * S, 0.0f), dts(nf, dt);
    int rc = esth_batch_init_uniform(b, zero.data());
    if (rc == ESTH_OK) rc = esth_batch_use_graph(b, 1);
    for (int k = 0; rc == ESTH_OK && k < 3; ++k) rc = esth_batch_step(b, control, 0, measurement, 0, dts.data());
    float us[16];
    int cnt = 0;
    if (rc == ESTH_OK) rc = esth_batch_step_times(b, us, 16, &cnt);
    for (int k = 0; rc == ESTH_OK && k < 10; ++k) rc = esth_batch_step(b, control, 0, measurement, 0, dts.data());
    if (rc == ESTH_OK) rc = esth_batch_step_times(b, us, 16, &cnt);
    esth_batch_destroy(b);
    ESTH_TRY(rc);
    if (cnt <= 0) return esth_fail(ESTH_ERR_STATE, "autotune: no timed steps");
    std::sort(us, us + cnt);
    *us_per_m = us[cnt / 2] / ((float)nf * (float)P * 1e-6f);
    return ESTH_OK;
}

extern "C" int esth_autotune(esth_engine* e, const esth_filter_desc* base, const void* params, size_t params_size,
                             const void* control, const void* measurement, float dt, float us_per_mparticle[5]) {
    if (!e || !base || !us_per_mparticle || !measurement) return esth_fail(ESTH_ERR_INVALID, "null");
    size_t psize = 0;
    ESTH_TRY(esth_params_size(base->model, &psize));
    if (psize && (!params || params_size < psize))
        return esth_fail(ESTH_ERR_INVALID, "autotune: params of %zu bytes, model needs %zu", params_size, psize);
    int S, V;
    ESTH_TRY(esth_model_dims(base->model, &S, &V) ? esth_fail(ESTH_ERR_INVALID, "model") : ESTH_OK);
    for (int k = 0; k < 5; ++k) {
        const uint32_t m = 64u << k;
        const int lm = 6 + k;
        const int saved = esth_kb_fuse_mode(e, lm);
        float c_off = INFINITY, c_on = INFINITY;
        ESTH_TRY(esth_engine_set_batch_fusion(e, m, ESTH_FUSION_OFF));
        int rc = time_batch_class(e, base, params, psize, control, measurement, dt, m, &c_off);
        if (rc == ESTH_OK && kb_fuse_grade(base->model, S, lm) > 0) {
            rc = esth_engine_set_batch_fusion(e, m, ESTH_FUSION_ON);
            if (rc == ESTH_OK) rc = time_batch_class(e, base, params, psize, control, measurement, dt, m, &c_on);
        }
        if (rc != ESTH_OK) {
            esth_engine_set_batch_fusion(e, m, saved);
            return rc;
        }
        us_per_mparticle[k] = std::min(c_off, c_on);
        ESTH_TRY(esth_engine_set_batch_fusion(e, m, c_on < c_off ? ESTH_FUSION_ON : ESTH_FUSION_OFF));
        /* the resampling's other tiling at the chosen form: two sub-filters per
         * wave (m = 64) or two waves per sub-filter (m >= 256) */
        const int alt = lm == 6 ? 32 : (lm >= 8 ? 128 : 0);
        ESTH_TRY(esth_engine_set_batch_resample_lanes(e, m, 64));
        if (alt && us_per_mparticle[k] < INFINITY) {
            float c_alt = INFINITY;
            ESTH_TRY(esth_engine_set_batch_resample_lanes(e, m, alt));
            rc = time_batch_class(e, base, params, psize, control, measurement, dt, m, &c_alt);
            if (rc != ESTH_OK) {
                esth_engine_set_batch_resample_lanes(e, m, 64);
                return rc;
            }
            if (c_alt < us_per_mparticle[k])
                us_per_mparticle[k] = c_alt;
            else
                ESTH_TRY(esth_engine_set_batch_resample_lanes(e, m, 64));
        }
    }
    return ESTH_OK;
}

/* Device time per step of `base`'s filter in each state layout: 3 warm-up
 * steps, then 10 whole steps bracketed by events on the filter's stream
 * (sample+weight and the sort fused as the filter runs them). */
extern "C" int esth_autotune_layout(esth_engine* e, const esth_filter_desc* base, const void* params,
                                    size_t params_size, const void* control, const void* measurement, float dt,
                                    float us_per_step[3], int* best) {
    if (!e || !base || !us_per_step) return esth_fail(ESTH_ERR_INVALID, "null");
    for (int L = ESTH_LAYOUT_SOA; L <= ESTH_LAYOUT_PACKED; ++L) {
        esth_filter_desc d = *base;
        d.state_layout = L;
        esth_filter* f = nullptr;
        ESTH_TRY(esth_filter_create(e, &d, params, params_size, &f));
        int S, V;
        esth_model_dims(d.model, &S, &V);
        std::vector<float> zero(S, 0.0f);
        int rc = esth_filter_init_uniform(f, zero.data());
        for (int i = 0; rc == ESTH_OK && i < 3; ++i) rc = esth_filter_step(f, control, measurement, dt);
        if (rc == ESTH_OK) rc = esth_filter_time_steps(f, 10, control, measurement, dt, &us_per_step[L]);
        esth_filter_destroy(f);
        ESTH_TRY(rc);
    }
    if (best) {
        int b = ESTH_LAYOUT_SOA;
        for (int L = ESTH_LAYOUT_AOS; L <= ESTH_LAYOUT_PACKED; ++L)
            if (us_per_step[L] < us_per_step[b]) b = L;
        *best = b;
    }
    return ESTH_OK;
}

/* The sample+weight arithmetic alone (k_hook_floor) over n particles, the host
 * frame a kernel argument: one untimed launch, then the median device time of
 * `reps` launches, each bracketed by events on a stream of its own. */
extern "C" int esth_engine_time_hook(esth_engine* e, int model, const float* frame, uint64_t n, int reps,
                                     float* us) {
    if (!e || !frame || !us) return esth_fail(ESTH_ERR_INVALID, "null");
    if (n < 2 || n % 2 || n > (1ull << 32) || reps < 1 || reps > 1000)
        return esth_fail(ESTH_ERR_INVALID, "n must be even in [2, 2^32], reps in [1, 1000]");
    ESTH_TRY(esth_with_model(model, [](auto) { return ESTH_OK; }));
    ESTH_HIP(hipSetDevice(e->device));
    const uint32_t npairs = (uint32_t)(n / 2);
    const size_t bytes = (size_t)npairs * sizeof(float);
    float* out = nullptr;
    ESTH_TRY(esth_engine_alloc(e, bytes, (void**)&out));
    esth_frame_arg fa;
    memcpy(fa.v, frame, sizeof(fa.v));
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<float> t;
    auto run = [&]() -> int {
        ESTH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        ESTH_HIP(hipEventCreate(&e0));
        ESTH_HIP(hipEventCreate(&e1));
        const int grid = (int)((npairs + 255u) / 256u);
        for (int i = 0; i <= reps; ++i) {
            ESTH_HIP(hipEventRecord(e0, s));
            ESTH_TRY(esth_with_model(model, [&](auto id) {
                hipLaunchKernelGGL((k_hook_floor<decltype(id)::value>), dim3(grid), dim3(256), 0, s, out, npairs, fa);
                return ESTH_OK;
            }));
            ESTH_HIP(hipGetLastError());
            ESTH_HIP(hipEventRecord(e1, s));
            ESTH_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            ESTH_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (i) t.push_back(ms * 1e3f);
        }
        return ESTH_OK;
    };
    const int rc = run();
    if (e1) hipEventDestroy(e1);
    if (e0) hipEventDestroy(e0);
    if (s) hipStreamDestroy(s);
    esth_engine_release(e, out, bytes);
    ESTH_TRY(rc);
    std::sort(t.begin(), t.end());
    *us = t[t.size() / 2];
    return ESTH_OK;
}

/* the sub-filter size for a filter of `particles` particles: among the powers
 * of two in [max(64, min_m), 1024] that divide it, the one with the lowest
 * measured cost (ties: the larger); if none is >= min_m, the largest divisor. */
extern "C" int esth_autotune_pick(const float us_per_mparticle[5], uint32_t particles, uint32_t min_m, uint32_t* m_out) {
    if (!us_per_mparticle || !m_out || particles % 64) return esth_fail(ESTH_ERR_INVALID, "particles must be a multiple of 64");
    uint32_t best = 0;
    float bc = 0.0f;
    uint32_t largest = 64;
    for (int k = 0; k < 5; ++k) {
        const uint32_t m = 64u << k;
        if (particles % m) continue;
        largest = m;
        if (m < min_m) continue;
        if (best == 0 || us_per_mparticle[k] <= bc) { best = m; bc = us_per_mparticle[k]; }
    }
    *m_out = best ? best : largest;
    return ESTH_OK;
}
