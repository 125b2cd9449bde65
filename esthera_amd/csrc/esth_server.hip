/*
 * esth_server.hip - serving filters by continuous batching.
 *
 * Requests (open a track, update it with a control and measurement, close it)
 * are queued per track.  esth_server_pump() forms a launch wave from the head
 * update of every admitted track that has one, and launches it while earlier
 * waves are still on the device:
 *   RNG stream : wave tables up (one copy), start-up of newly admitted tracks,
 *                one MTGP32 launch over the streams of every track in the wave
 *   main stream: per group of tracks with equal settings, the batched kernels
 *                (sample+weight, sort per sub-filter size, argmax, exchange,
 *                resampling), then the wave's estimates down (one copy)
 * so the RNG of wave w overlaps the step kernels of wave w-1, and requests that
 * arrive meanwhile form wave w+1.  A track's particles, weights and MTGP32
 * statuses stay resident in one block of the server's HBM arena from admission
 * to close; admission waits for a free track slot and arena space (FIFO), and a
 * closed track's block is reused only after the last wave that touched it has
 * completed.  Every kernel runs the single-filter per-sub-filter code on the
 * track's own buffers and random words, so a track's results are bit-identical
 * to the same filter stepped alone with ESTH_RNG_MTGP_PER_FILTER, whatever else
 * is in its waves.
 *
 * The reference has no serving layer; the step each wave runs is the step
 * loop of pf_generic.c:438-808 per track.
 */
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <random>
#include <unordered_map>

#include "esth_internal.h"
#include "esth_kernels.h"
#include "esth_dispatch.h"

#define ESTH_SRV_SLOTS 4
#define ESTH_SRV_EST (4 + 2 * ESTH_MAX_STATE_DIM)
#define ESTH_SRV_ALIGN 256u

static double srv_now_us() {
    using namespace std::chrono;
    return (double)duration_cast<nanoseconds>(steady_clock::now().time_since_epoch()).count() * 1e-3;
}

static inline uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

/* ------------------------------------------------------------ arena
 * Best-fit allocator over one device allocation, offsets only (host side). */
struct esth_arena {
    uint64_t size = 0, used = 0, peak = 0;
    std::map<uint64_t, uint64_t> by_off;          /* free blocks: offset -> size */
    std::multimap<uint64_t, uint64_t> by_size;    /* size -> offset */
    void init(uint64_t n) {
        size = n;
        by_off.clear();
        by_size.clear();
        by_off[0] = n;
        by_size.insert({n, 0});
    }
    void erase_size(uint64_t sz, uint64_t off) {
        auto r = by_size.equal_range(sz);
        for (auto it = r.first; it != r.second; ++it)
            if (it->second == off) { by_size.erase(it); return; }
    }
    bool alloc(uint64_t n, uint64_t* off) {
        n = align_up(n, ESTH_SRV_ALIGN);
        auto it = by_size.lower_bound(n);
        if (it == by_size.end()) return false;
        const uint64_t bsz = it->first, boff = it->second;
        by_size.erase(it);
        by_off.erase(boff);
        if (bsz > n) {
            by_off[boff + n] = bsz - n;
            by_size.insert({bsz - n, boff + n});
        }
        *off = boff;
        used += n;
        peak = std::max(peak, used);
        return true;
    }
    void release(uint64_t off, uint64_t n) {
        n = align_up(n, ESTH_SRV_ALIGN);
        used -= n;
        auto next = by_off.lower_bound(off);
        if (next != by_off.end() && off + n == next->first) {
            erase_size(next->second, next->first);
            n += next->second;
            next = by_off.erase(next);
        }
        if (next != by_off.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                erase_size(prev->second, prev->first);
                off = prev->first;
                n += prev->second;
                by_off.erase(prev);
            }
        }
        by_off[off] = n;
        by_size.insert({n, off});
    }
};

/* ------------------------------------------------------------ tracks */
struct esth_sreq {
    uint64_t id = 0;
    int kind = 1;   /* 0 open, 1 update, 2 close */
    double submit_us = 0.0;
    std::vector<float> frame;
};

struct esth_strack {
    uint64_t id = 0;
    esth_filter_desc d{};
    int S = 0, V = 0, group = 0, logm = 0;
    std::vector<unsigned char> params;
    float init[ESTH_MAX_STATE_DIM] = {0};
    uint64_t open_req = 0;
    double open_submit = 0.0;
    int phase = 0;            /* 0 waiting for admission, 1 admitted, 2 closed */
    bool need_init = false;
    bool in_ready = false;
    bool closing = false;     /* a close request has been queued */
    uint64_t off = 0, bytes = 0;
    uint64_t soff = 0;        /* statuses, from the arena base, in 351-word units */
    esth_bfilter bf{};
    int K = 0;
    uint32_t rsize = 0;
    uint64_t words = 0;
    size_t frame_floats = 0;
    uint64_t last_wave = 0;   /* 1 + index of the last wave that used the block */
    uint64_t steps = 0;
    std::deque<esth_sreq> q;
};

struct esth_sgroup_key {
    int model, topology, t, resampler, adaptive, weights;
    float thr;
    bool operator==(const esth_sgroup_key& o) const {
        return model == o.model && topology == o.topology && t == o.t && resampler == o.resampler &&
               adaptive == o.adaptive && weights == o.weights && thr == o.thr;
    }
};

struct esth_wave_upd {
    esth_strack* t;          /* valid until the wave is launched */
    uint64_t req;
    double submit_us;
    uint64_t step;
    uint64_t track_id;
    int S;
};

struct esth_wave_init {
    uint64_t track_id, req;
    double submit_us;
};

/* one group of a wave: tracks with equal settings, contiguous in the table */
struct esth_wave_group {
    int key = 0;
    uint32_t f0 = 0, nf = 0;        /* filters [f0, f0 + nf) of the table */
    uint64_t sub0 = 0, nsub = 0;    /* their sub-filters in sub_all */
    uint64_t chk0 = 0, nchk = 0;    /* their 64-pair chunks (sample+weight) */
    uint64_t cls0[11] = {0}, ncls[11] = {0};   /* per log2(m) class, in sub_cls */
    uint64_t unit0 = 0, nunit = 0;  /* the fused m = 64 class's units (esth_pair_units), in units */
    uint32_t fuse_bits = 0;         /* bit lm: class lm samples inside kb_sample_sort */
    int lanes[11] = {0};            /* the resampling's lanes per sub-filter, per class */
};

struct esth_wslot {
    unsigned char* h = nullptr;   /* pinned staging of the wave tables */
    size_t hcap = 0;
    unsigned char* dtab = nullptr;
    size_t dcap = 0;
    double* dest = nullptr;       /* device estimates, ESTH_SRV_EST doubles per track */
    float2* dres = nullptr;
    double* hest = nullptr;       /* pinned copy of dest */
    size_t ecap = 0;              /* tracks */
    hipEvent_t ev_rng, ev_done;
};

struct esth_wave {
    uint64_t idx = 0;
    int slot = 0;
    std::vector<esth_wave_upd> upd;
    std::vector<esth_wave_init> inits;
};

struct esth_server {
    esth_engine* e = nullptr;
    int max_tracks = 0;
    int depth = 2;
    int max_wave = 1 << 30;
    esth_arena arena;
    unsigned char* base = nullptr;
    uint64_t base_bytes = 0;
    /* the arena's memory: a reserved virtual range whose chunks are backed by
     * physical HBM when a track's block first needs them (vmm), so the
     * budget costs nothing until tracks occupy it; else one hipMalloc */
    bool vmm = false;
    uint64_t chunk = 0, reserved = 0, mapped_bytes = 0;
    std::vector<hipMemGenericAllocationHandle_t> chunks;
    hipStream_t stream = nullptr, rng_stream = nullptr;
    esth_wslot slot[ESTH_SRV_SLOTS];
    float* R[2] = {nullptr, nullptr};
    uint64_t rcap = 0;            /* words per R buffer */
    std::vector<esth_sgroup_key> keys;
    std::unordered_map<uint64_t, std::unique_ptr<esth_strack>> tracks;
    std::deque<esth_strack*> admit_q;
    std::deque<esth_strack*> ready;
    std::deque<esth_wave> inflight;
    std::vector<std::pair<uint64_t, std::pair<uint64_t, uint64_t>>> deferred;   /* (wave+1, (off, bytes)) */
    std::deque<esth_result> results;
    uint64_t next_track = 1, next_req = 1, waves = 0, completed = 0, wave_updates = 0;
    int live = 0;
    uint64_t queued = 0;
    uint64_t last_slot_wave[ESTH_SRV_SLOTS] = {0, 0, 0, 0};   /* 1 + wave index last launched in a slot */
    std::vector<unsigned char> scratch;
};

#define ESTH_SRV_CHUNK (1ull << 30)

/* the arena's backing: reserve base_bytes of virtual addresses (chunks mapped
 * on demand) where the device supports it, else allocate it whole */
static int arena_create(esth_server* s) {
    int dev = s->e->device, vmm = 0;
    if (hipDeviceGetAttribute(&vmm, hipDeviceAttributeVirtualMemoryManagementSupported, dev) != hipSuccess) vmm = 0;
    if (vmm) {
        hipMemAllocationProp prop = {};
        prop.type = hipMemAllocationTypePinned;
        prop.location.type = hipMemLocationTypeDevice;
        prop.location.id = dev;
        size_t gran = 0;
        if (hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityMinimum) != hipSuccess || !gran)
            vmm = 0;
        else {
            s->chunk = align_up(ESTH_SRV_CHUNK, gran);
            s->reserved = align_up(s->base_bytes, s->chunk);
            void* p = nullptr;
            if (hipMemAddressReserve(&p, s->reserved, s->chunk, nullptr, 0) != hipSuccess || !p) {
                (void)hipGetLastError();
                vmm = 0;
            } else {
                s->base = (unsigned char*)p;
            }
        }
    }
    if (vmm) {
        ESTH_TRY(esth_engine_account(s->e, s->base_bytes));
        s->vmm = true;
        s->chunks.assign(s->reserved / s->chunk, nullptr);
        return ESTH_OK;
    }
    return esth_engine_alloc(s->e, s->base_bytes, (void**)&s->base);
}

/* back every chunk of [off, off + n) with HBM (vmm arenas) */
static int arena_map(esth_server* s, uint64_t off, uint64_t n) {
    if (!s->vmm || n == 0) return ESTH_OK;
    hipMemAllocationProp prop = {};
    prop.type = hipMemAllocationTypePinned;
    prop.location.type = hipMemLocationTypeDevice;
    prop.location.id = s->e->device;
    hipMemAccessDesc acc = {};
    acc.location.type = hipMemLocationTypeDevice;
    acc.location.id = s->e->device;
    acc.flags = hipMemAccessFlagsProtReadWrite;
    for (uint64_t c = off / s->chunk; c <= (off + n - 1) / s->chunk; ++c) {
        if (s->chunks[c]) continue;
        hipMemGenericAllocationHandle_t h = nullptr;
        void* at = s->base + c * s->chunk;
        hipError_t err = hipMemCreate(&h, s->chunk, &prop, 0);
        if (err == hipSuccess) {
            err = hipMemMap(at, s->chunk, 0, h, 0);
            if (err == hipSuccess) {
                err = hipMemSetAccess(at, s->chunk, &acc, 1);
                if (err != hipSuccess) (void)hipMemUnmap(at, s->chunk);
            }
            if (err != hipSuccess) (void)hipMemRelease(h);
        }
        if (err != hipSuccess) return esth_fail(ESTH_ERR_NOMEM, "server arena chunk %llu: %s", (unsigned long long)c,
                                                hipGetErrorString(err));
        s->chunks[c] = h;
        s->mapped_bytes += s->chunk;
    }
    return ESTH_OK;
}

static void arena_destroy(esth_server* s) {
    if (!s->base) return;
    if (s->vmm) {
        for (size_t c = 0; c < s->chunks.size(); ++c)
            if (s->chunks[c]) {
                (void)hipMemUnmap(s->base + c * s->chunk, s->chunk);
                (void)hipMemRelease(s->chunks[c]);
            }
        (void)hipMemAddressFree(s->base, s->reserved);
        esth_engine_account(s->e, -(int64_t)s->base_bytes);
    } else {
        esth_engine_release(s->e, s->base, s->base_bytes);
    }
    s->base = nullptr;
}

static int srv_free_buffers(esth_server* s) {
    for (auto& w : s->slot) {
        if (w.h) hipHostFree(w.h);
        if (w.hest) hipHostFree(w.hest);
        if (w.dtab) esth_engine_release(s->e, w.dtab, w.dcap);
        if (w.dest) esth_engine_release(s->e, w.dest, w.ecap * (sizeof(double) * ESTH_SRV_EST + sizeof(float2)));
        w.h = nullptr;
        w.hest = nullptr;
        w.dtab = nullptr;
        w.dest = nullptr;
        w.dres = nullptr;
        w.hcap = w.dcap = w.ecap = 0;
    }
    for (int p = 0; p < 2; ++p) {
        if (s->R[p]) esth_engine_release(s->e, s->R[p], s->rcap * sizeof(float));
        s->R[p] = nullptr;
    }
    s->rcap = 0;
    return ESTH_OK;
}

extern "C" int esth_server_create(esth_engine* e, int max_tracks, size_t hbm_bytes, esth_server** out) {
    if (!e || !out || max_tracks <= 0) return esth_fail(ESTH_ERR_INVALID, "server: bad arguments");
    if (hbm_bytes < (1u << 20)) return esth_fail(ESTH_ERR_INVALID, "server: HBM arena of %zu bytes is too small", hbm_bytes);
    ESTH_HIP(hipSetDevice(e->device));
    std::unique_ptr<esth_server> s(new esth_server());
    s->e = e;
    s->max_tracks = max_tracks;
    s->base_bytes = align_up(hbm_bytes, ESTH_SRV_ALIGN);
    ESTH_TRY(arena_create(s.get()));
    s->arena.init(s->base_bytes);
    hipError_t err = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&s->rng_stream, hipStreamNonBlocking);
    for (auto& w : s->slot) {
        if (err == hipSuccess) err = hipEventCreateWithFlags(&w.ev_rng, hipEventDisableTiming);
        if (err == hipSuccess) err = hipEventCreateWithFlags(&w.ev_done, hipEventDisableTiming);
    }
    if (err != hipSuccess) {
        arena_destroy(s.get());
        return esth_fail(ESTH_ERR_DEVICE, "server streams: %s", hipGetErrorString(err));
    }
    *out = s.release();
    return ESTH_OK;
}

extern "C" int esth_server_destroy(esth_server* s) {
    if (!s) return ESTH_OK;
    hipSetDevice(s->e->device);
    hipStreamSynchronize(s->stream);
    hipStreamSynchronize(s->rng_stream);
    srv_free_buffers(s);
    for (auto& w : s->slot) {
        hipEventDestroy(w.ev_rng);
        hipEventDestroy(w.ev_done);
    }
    hipStreamDestroy(s->stream);
    hipStreamDestroy(s->rng_stream);
    arena_destroy(s);
    delete s;
    return ESTH_OK;
}

extern "C" int esth_server_configure(esth_server* s, int max_inflight_waves, int max_wave_tracks) {
    if (!s || max_inflight_waves < 1 || max_inflight_waves > ESTH_SRV_SLOTS - 1 || max_wave_tracks < 1)
        return esth_fail(ESTH_ERR_INVALID, "server: waves in flight must be 1..%d, tracks per wave >= 1", ESTH_SRV_SLOTS - 1);
    s->depth = max_inflight_waves;
    s->max_wave = max_wave_tracks;
    return ESTH_OK;
}

/* the track's block: state, tmp, w, part, winner, bt_w, bt_s, top, statuses */
static uint64_t track_layout(esth_strack* t, uint64_t off, bool assign, unsigned char* base) {
    const esth_filter_desc& d = t->d;
    const uint64_t N = (uint64_t)d.num_particles * d.num_blocks, G = d.num_blocks, S = (uint64_t)t->S;
    const uint64_t tt = (d.topology == ESTH_TOPO_NONE ? 0 : d.num_transfer);
    const uint64_t tg = (tt ? tt : 1) * G;
    uint64_t p = 0;
    auto take = [&](uint64_t bytes) { uint64_t q = p; p = align_up(p + bytes, ESTH_SRV_ALIGN); return q; };
    const uint64_t o_state = take(4 * N * S), o_tmp = take(4 * N * S), o_w = take(4 * N);
    const uint64_t o_part = take(4 * G * (2 + S)), o_win = take(8 * G), o_btw = take(4 * tg), o_bts = take(4 * tg * S);
    const uint64_t o_top = take(4 * 64);
    /* statuses start at an arena offset that is a multiple of 351 words */
    const uint64_t sw = 4 * ESTH_MTGP_N;
    const uint64_t o_st = align_up(off + p, sw) - off;
    p = o_st + sw * (uint64_t)t->K;
    if (assign) {
        esth_bfilter& f = t->bf;
        memset(&f, 0, sizeof(f));
        f.state = (float*)(base + off + o_state);
        f.tmp = (float*)(base + off + o_tmp);
        f.w = (float*)(base + off + o_w);
        f.part = (float*)(base + off + o_part);
        f.winner = (void*)(base + off + o_win);
        f.bt_w = (float*)(base + off + o_btw);
        f.bt_s = (float*)(base + off + o_bts);
        f.top = (uint32_t*)(base + off + o_top);
        f.N = (uint32_t)N;
        f.G = (uint32_t)G;
        f.m = d.num_particles;
        f.t = (uint32_t)tt;
        f.dimx = d.torus_dimx;
        f.dimy = d.torus_dimy;
        if (d.topology == ESTH_TOPO_TORUS && (f.dimx <= 0 || f.dimy <= 0)) esth_torus_dims(f.G, &f.dimx, &f.dimy);
        if (f.dimx <= 0) f.dimx = 1;
        if (f.dimy <= 0) f.dimy = 1;
        f.uoff = (uint32_t)(N * (uint64_t)t->V);
        f.flags = d.carry_weights ? 1u : 0u;
        t->soff = (off + o_st) / sw;
    }
    return p;
}

static size_t srv_frame_floats(int model, const void* params) { return (size_t)esth_model_frame_floats(model, params); }

static void push_result(esth_server* s, uint64_t req, uint64_t track, int status, int kind, double submit,
                        const esth_estimate* est) {
    esth_result r;
    memset(&r, 0, sizeof(r));
    r.request_id = req;
    r.track_id = track;
    r.status = status;
    r.kind = kind;
    r.submit_us = submit;
    r.done_us = srv_now_us();
    if (est) r.estimate = *est;
    s->results.push_back(r);
}

extern "C" int esth_server_open(esth_server* s, const esth_filter_desc* desc, const void* params, size_t params_size,
                                const float* init_state, uint64_t* track_id, uint64_t* request_id) {
    if (!s || !desc || !init_state) return esth_fail(ESTH_ERR_INVALID, "server open: null argument");
    int S, V;
    ESTH_TRY(esth_check_desc(desc, &S, &V));
    if (!esth_rng_is_per_filter(desc->rng))
        return esth_fail(ESTH_ERR_INVALID, "server open: tracks use a per-filter RNG mode");
    if (desc->num_particles < 64) return esth_fail(ESTH_ERR_UNSUPPORTED, "server open: sub-filters need m >= 64");
    if (desc->state_layout != ESTH_LAYOUT_SOA)
        return esth_fail(ESTH_ERR_UNSUPPORTED, "server open: tracks keep their state in ESTH_LAYOUT_SOA");
    size_t psize = 0;
    ESTH_TRY(esth_params_size(desc->model, &psize));
    if (psize && (!params || params_size < psize))
        return esth_fail(ESTH_ERR_INVALID, "server open: params of %zu bytes, model needs %zu", params_size, psize);
    if (desc->model == ESTH_MODEL_UNICYCLE) {
        const int L = ((const esth_unicycle_params*)params)->num_landmarks;
        if (L < 0 || L > ESTH_UNICYCLE_MAX_LANDMARKS) return esth_fail(ESTH_ERR_INVALID, "server open: %d landmarks", L);
    }
    std::unique_ptr<esth_strack> t(new esth_strack());
    t->id = s->next_track++;
    t->d = *desc;
    t->S = S;
    t->V = V;
    if (psize) t->params.assign((const unsigned char*)params, (const unsigned char*)params + psize);
    memcpy(t->init, init_state, sizeof(float) * (size_t)S);
    while ((1u << t->logm) < desc->num_particles) ++t->logm;
    const uint64_t need = esth_rng_needed(V, desc->resampler, desc->adaptive, desc->num_particles, desc->num_blocks);
    t->K = esth_rng_streams(desc->rng, need);
    t->rsize = (uint32_t)esth_rng_stream_size(t->K, need);
    t->words = (uint64_t)t->rsize * (uint64_t)t->K;
    t->frame_floats = align_up(srv_frame_floats(desc->model, psize ? t->params.data() : nullptr), 4);
    t->bytes = track_layout(t.get(), 0, false, nullptr) + 4 * ESTH_MTGP_N + ESTH_SRV_ALIGN;
    if (t->bytes > s->base_bytes)
        return esth_fail(ESTH_ERR_BUDGET, "server open: track needs %llu bytes, arena has %llu",
                         (unsigned long long)t->bytes, (unsigned long long)s->base_bytes);
    esth_sgroup_key k{desc->model, desc->topology, (int)(desc->topology == ESTH_TOPO_NONE ? 0 : desc->num_transfer),
                      desc->resampler, desc->adaptive, desc->weights, desc->resampling_threshold};
    int g = -1;
    for (size_t i = 0; i < s->keys.size(); ++i)
        if (s->keys[i] == k) g = (int)i;
    if (g < 0) {
        s->keys.push_back(k);
        g = (int)s->keys.size() - 1;
    }
    t->group = g;
    t->open_req = s->next_req++;
    t->open_submit = srv_now_us();
    if (track_id) *track_id = t->id;
    if (request_id) *request_id = t->open_req;
    s->admit_q.push_back(t.get());
    s->queued++;
    s->tracks[t->id] = std::move(t);
    return ESTH_OK;
}

static esth_strack* find_track(esth_server* s, uint64_t id) {
    auto it = s->tracks.find(id);
    return it == s->tracks.end() ? nullptr : it->second.get();
}

static void make_ready(esth_server* s, esth_strack* t) {
    if (t->phase == 1 && !t->in_ready && !t->q.empty()) {
        t->in_ready = true;
        s->ready.push_back(t);
    }
}

extern "C" int esth_server_update(esth_server* s, uint64_t track_id, const void* control, const void* measurement,
                                  float dt, uint64_t* request_id) {
    if (!s || !measurement) return esth_fail(ESTH_ERR_INVALID, "server update: null argument");
    esth_strack* t = find_track(s, track_id);
    if (!t || t->closing) return esth_fail(ESTH_ERR_STATE, "server update: track %llu is not open", (unsigned long long)track_id);
    esth_sreq r;
    r.kind = 1;
    r.frame.resize(ESTH_FRAME_FLOATS);
    static const esth_dh6_control none = {0.0f};
    if (t->d.model == ESTH_MODEL_DH6 && !control) control = &none;
    if (esth_build_frame_impl(t->d.model, t->params.empty() ? nullptr : t->params.data(), control, measurement, dt,
                              r.frame.data()))
        return esth_fail(ESTH_ERR_INVALID, "server update: bad control/measurement for track %llu", (unsigned long long)track_id);
    r.frame.resize(t->frame_floats);
    r.id = s->next_req++;
    r.submit_us = srv_now_us();
    if (request_id) *request_id = r.id;
    t->q.push_back(std::move(r));
    s->queued++;
    make_ready(s, t);
    return ESTH_OK;
}

extern "C" int esth_server_close(esth_server* s, uint64_t track_id, uint64_t* request_id) {
    if (!s) return esth_fail(ESTH_ERR_INVALID, "server close: null server");
    esth_strack* t = find_track(s, track_id);
    if (!t || t->closing) return esth_fail(ESTH_ERR_STATE, "server close: track %llu is not open", (unsigned long long)track_id);
    esth_sreq r;
    r.kind = 2;
    r.id = s->next_req++;
    r.submit_us = srv_now_us();
    if (request_id) *request_id = r.id;
    t->closing = true;
    t->q.push_back(std::move(r));
    s->queued++;
    make_ready(s, t);
    return ESTH_OK;
}

/* ------------------------------------------------------------ waves */
static void estimate_from(const double* h, int S, uint64_t step, esth_estimate* out) {
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
    out->step = step;
}

static void release_block(esth_server* s, esth_strack* t) {
    if (t->last_wave > s->completed) s->deferred.push_back({t->last_wave, {t->off, t->bytes}});
    else s->arena.release(t->off, t->bytes);
}

static void finish_wave(esth_server* s, esth_wave& w) {
    const esth_wslot& sl = s->slot[w.slot];
    for (const esth_wave_init& t : w.inits) push_result(s, t.req, t.track_id, ESTH_OK, 0, t.submit_us, nullptr);
    for (size_t i = 0; i < w.upd.size(); ++i) {
        esth_estimate est;
        estimate_from(sl.hest + i * ESTH_SRV_EST, w.upd[i].S, w.upd[i].step, &est);
        push_result(s, w.upd[i].req, w.upd[i].track_id, ESTH_OK, 1, w.upd[i].submit_us, &est);
    }
    s->completed = w.idx + 1;
    for (size_t i = 0; i < s->deferred.size();) {
        if (s->deferred[i].first <= s->completed) {
            s->arena.release(s->deferred[i].second.first, s->deferred[i].second.second);
            s->deferred[i] = s->deferred.back();
            s->deferred.pop_back();
        } else {
            ++i;
        }
    }
}

/* collect finished waves; with `block`, wait for the oldest one */
static int collect(esth_server* s, bool block) {
    while (!s->inflight.empty()) {
        esth_wave& w = s->inflight.front();
        hipEvent_t ev = s->slot[w.slot].ev_done;
        if (block) {
            ESTH_HIP(hipEventSynchronize(ev));
            block = false;
        } else {
            hipError_t q = hipEventQuery(ev);
            if (q == hipErrorNotReady) break;
            if (q != hipSuccess) return esth_fail(ESTH_ERR_DEVICE, "server wave: %s", hipGetErrorString(q));
        }
        finish_wave(s, w);
        s->inflight.pop_front();
    }
    return ESTH_OK;
}

static void admit(esth_server* s, std::vector<esth_strack*>& inits) {
    while (!s->admit_q.empty() && s->live < s->max_tracks) {
        esth_strack* t = s->admit_q.front();
        uint64_t off = 0;
        if (!s->arena.alloc(t->bytes, &off)) break;
        if (arena_map(s, off, t->bytes) != ESTH_OK) {   /* no HBM to back it now: wait for closes */
            s->arena.release(off, t->bytes);
            break;
        }
        s->admit_q.pop_front();
        s->queued--;
        t->off = off;
        track_layout(t, off, true, s->base);
        t->phase = 1;
        t->need_init = true;
        s->live++;
        inits.push_back(t);
        make_ready(s, t);
    }
}

static int grow_slot(esth_server* s, esth_wslot& w, size_t hbytes, size_t ntracks) {
    if (hbytes > w.hcap || hbytes > w.dcap) {
        const size_t cap = std::max(hbytes, (size_t)1 << 20) * 2;
        if (w.h) hipHostFree(w.h);
        if (w.dtab) esth_engine_release(s->e, w.dtab, w.dcap);
        w.h = nullptr;
        w.dtab = nullptr;
        w.hcap = w.dcap = 0;
        ESTH_HIP(hipHostMalloc((void**)&w.h, cap, 0));
        w.hcap = cap;
        ESTH_TRY(esth_engine_alloc(s->e, cap, (void**)&w.dtab));
        w.dcap = cap;
    }
    if (ntracks > w.ecap) {
        const size_t cap = std::max(ntracks, (size_t)256) * 2;
        const size_t per = sizeof(double) * ESTH_SRV_EST + sizeof(float2);
        if (w.hest) hipHostFree(w.hest);
        if (w.dest) esth_engine_release(s->e, w.dest, w.ecap * per);
        w.hest = nullptr;
        w.dest = nullptr;
        w.ecap = 0;
        ESTH_HIP(hipHostMalloc((void**)&w.hest, sizeof(double) * ESTH_SRV_EST * cap, 0));
        ESTH_TRY(esth_engine_alloc(s->e, cap * per, (void**)&w.dest));
        w.ecap = cap;
        w.dres = (float2*)(w.dest + ESTH_SRV_EST * cap);
    }
    return ESTH_OK;
}

/* wait until nothing is in flight (buffers about to be reallocated) */
static int quiesce(esth_server* s) {
    while (!s->inflight.empty()) ESTH_TRY(collect(s, true));
    ESTH_HIP(hipStreamSynchronize(s->rng_stream));
    ESTH_HIP(hipStreamSynchronize(s->stream));
    return ESTH_OK;
}

template <int LOGM, int S>
static void srv_sort(const esth_bfilter* F, const uint2* sub, uint32_t n, int log_w, hipStream_t st) {
    hipLaunchKernelGGL((kb_sort<LOGM, S>), dim3((n + 3) / 4), dim3(256), 0, st, F, sub, n, log_w);
}

#define ESTH_SRV_LOGM(FN, S, LM, ...)                  \
    switch (LM) {                                      \
    case 6: FN<6, S>(__VA_ARGS__); break;              \
    case 7: FN<7, S>(__VA_ARGS__); break;              \
    case 8: FN<8, S>(__VA_ARGS__); break;              \
    case 9: FN<9, S>(__VA_ARGS__); break;              \
    case 10: FN<10, S>(__VA_ARGS__); break;            \
    default: return esth_fail(ESTH_ERR_INVALID, "m"); \
    }

template <int LOGM, int S>
static void srv_alias(const esth_bfilter* F, const uint2* sub, uint32_t n, const float* R, float thr, int adaptive,
                      int log_w, hipStream_t st) {
    constexpr int WPB = esth_alias_wpb<LOGM>::value;
    hipLaunchKernelGGL((kb_alias_w<LOGM, S>), dim3((n + WPB - 1) / WPB), dim3(WPB * ESTH_WAVE), 0, st, F, sub, n, R, thr,
                       adaptive, log_w);
}

template <int S>
static int group_kernels(esth_server* s, const esth_sgroup_key& k, const esth_wave_group& g, const esth_bfilter* F,
                         const uint2* sub_all, const uint2* sub_cls, const uint2* units, const uint2* chunks,
                         const float* R, const float* frames) {
    hipStream_t st = s->stream;
    const int log_w = k.weights == ESTH_WEIGHTS_LOG;
    const esth_bfilter* Fg = F + g.f0;
    const uint32_t nsub = (uint32_t)g.nsub;
    const uint2* sa = sub_all + g.sub0;
    const uint2* ch = chunks + g.chk0;
    const uint32_t nch = (uint32_t)g.nchk;
    const int grid_w = (int)(((uint64_t)nch * ESTH_WAVE + 255) / 256);
    if (nch)
        ESTH_TRY(esth_with_model(k.model, [&](auto id) {
            hipLaunchKernelGGL(kb_sample_weight<decltype(id)::value>, dim3(grid_w), dim3(256), 0, st, Fg, ch, nch, R,
                               frames, log_w);
            return ESTH_OK;
        }));
    for (int lm = 6; lm <= 10; ++lm) {
        if (!g.ncls[lm]) continue;
        if ((g.fuse_bits >> lm) & 1u)
            ESTH_TRY(esth_launch_kb_sample_sort(k.model, S, lm, Fg, lm == 6 ? units + g.unit0 : sub_cls + g.cls0[lm],
                                                (uint32_t)(lm == 6 ? g.nunit : g.ncls[lm]), R, frames, log_w, k.thr,
                                                k.resampler == ESTH_RESAMPLE_ALIAS && k.adaptive, st));
        else
            ESTH_SRV_LOGM(srv_sort, S, lm, Fg, sub_cls + g.cls0[lm], (uint32_t)g.ncls[lm], log_w, st);
    }
    if (k.t > 0 && k.topology == ESTH_TOPO_ALL2ALL) hipLaunchKernelGGL(kb_topt, dim3(g.nf), dim3(1024), 0, st, Fg);
    if (k.resampler == ESTH_RESAMPLE_ALIAS) {
        hipLaunchKernelGGL(kb_argmax<S>, dim3(g.nf), dim3(256), 0, st, Fg, log_w);
        if (k.t > 0) {
            const uint32_t nslots = (uint32_t)((k.topology == ESTH_TOPO_TORUS ? 4 : 1) * k.t);
            const uint64_t total = (uint64_t)nsub * nslots;
            hipLaunchKernelGGL(kb_exchange, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, Fg, sa, nsub, S,
                               nslots, k.topology);
        }
        for (int lm = 6; lm <= 10; ++lm)
            if (g.ncls[lm])
                ESTH_SRV_LOGM(srv_alias, S, lm, Fg, sub_cls + g.cls0[lm], (uint32_t)g.ncls[lm], R, k.thr, k.adaptive,
                              log_w, st);
        return ESTH_OK;
    }
    /* exchange inside the resampling, argmax in the first resampling launch */
    uint32_t am = g.nf;
    for (int lm = 6; lm <= 10; ++lm) {
        if (!g.ncls[lm]) continue;
        ESTH_TRY(esth_launch_kb_resample(S, lm, g.lanes[lm], Fg, sub_cls + g.cls0[lm], (uint32_t)g.ncls[lm], R, k.thr,
                                         k.resampler == ESTH_RESAMPLE_SYSTEMATIC ? 1 : 0, log_w, k.topology, am, st));
        am = 0;
    }
    return ESTH_OK;
}

/* form and launch one wave; *launched = false when there was nothing to do */
static int launch_wave(esth_server* s, bool* launched) {
    *launched = false;
    std::vector<esth_strack*> inits;
    admit(s, inits);
    /* the head request of every ready track: closes are done here, updates join the wave */
    std::vector<esth_wave_upd> upd;
    const size_t nready = s->ready.size();
    for (size_t i = 0; i < nready && (int)upd.size() < s->max_wave; ++i) {
        esth_strack* t = s->ready.front();
        s->ready.pop_front();
        t->in_ready = false;
        while (!t->q.empty() && t->q.front().kind == 2) {
            esth_sreq& r = t->q.front();
            push_result(s, r.id, t->id, ESTH_OK, 2, r.submit_us, nullptr);
            s->queued--;
            t->q.pop_front();
            t->phase = 2;
        }
        if (t->phase == 2) {
            s->live--;
            if (t->need_init) {   /* admitted but never launched: nothing on the device uses it */
                inits.erase(std::remove(inits.begin(), inits.end(), t), inits.end());
                t->need_init = false;
                push_result(s, t->open_req, t->id, ESTH_OK, 0, t->open_submit, nullptr);
            }
            release_block(s, t);
            s->tracks.erase(t->id);
            continue;
        }
        if (t->q.empty()) continue;
        esth_sreq& r = t->q.front();
        upd.push_back({t, r.id, r.submit_us, t->steps + 1, t->id, t->S});
    }
    if (upd.empty() && inits.empty()) return ESTH_OK;

    /* order the wave's tracks by group */
    std::stable_sort(upd.begin(), upd.end(), [](const esth_wave_upd& a, const esth_wave_upd& b) {
        return a.t->group < b.t->group;
    });
    const size_t nw = upd.size();
    /* the engine's fusion setting, read once for the whole wave */
    int fmode[11];
    for (int lm = 0; lm <= 10; ++lm) fmode[lm] = esth_kb_fuse_mode(s->e, lm);
    auto fused = [&](const esth_strack* t) { return esth_kb_fuse_with(fmode[t->logm], t->d.model, t->S, t->logm); };
    int flanes[11];
    for (int lm = 0; lm <= 10; ++lm) flanes[lm] = esth_kb_resample_lanes(s->e, lm);
    uint64_t words = 0, nsub = 0, nstreams = 0, frames = 0, nchk = 0;
    for (auto& u : upd) {
        words += u.t->words;
        nsub += u.t->d.num_blocks;
        /* fused size classes sample inside kb_sample_sort, without chunks */
        if (!fused(u.t)) nchk += (u.t->bf.N / 2 + ESTH_WAVE - 1) / ESTH_WAVE;
        nstreams += (uint64_t)u.t->K;
        frames += u.t->frame_floats;
    }
    const size_t o_F = 0;
    const size_t o_all = align_up(o_F + sizeof(esth_bfilter) * nw, ESTH_SRV_ALIGN);
    const size_t o_cls = align_up(o_all + sizeof(uint2) * nsub, ESTH_SRV_ALIGN);
    const size_t o_unt = align_up(o_cls + sizeof(uint2) * nsub, ESTH_SRV_ALIGN);
    const size_t o_chk = align_up(o_unt + sizeof(uint2) * nsub, ESTH_SRV_ALIGN);
    const size_t o_str = align_up(o_chk + sizeof(uint2) * nchk, ESTH_SRV_ALIGN);
    const size_t o_frm = align_up(o_str + sizeof(esth_mtgp_stream) * nstreams, ESTH_SRV_ALIGN);
    const size_t o_ini = align_up(o_frm + sizeof(float) * frames, ESTH_SRV_ALIGN);
    const size_t total = align_up(o_ini + sizeof(esth_sinit) * inits.size(), ESTH_SRV_ALIGN);

    const uint64_t widx = s->waves;
    const int si = (int)(widx % ESTH_SRV_SLOTS);
    esth_wslot& sl = s->slot[si];
    /* the slot's previous wave must have been collected */
    while (s->last_slot_wave[si] > s->completed) ESTH_TRY(collect(s, true));
    if (total > sl.hcap || nw > sl.ecap || words > s->rcap) {
        ESTH_TRY(quiesce(s));
        ESTH_TRY(grow_slot(s, sl, total, nw));
        if (words > s->rcap) {
            for (int p = 0; p < 2; ++p) {
                if (s->R[p]) esth_engine_release(s->e, s->R[p], s->rcap * sizeof(float));
                s->R[p] = nullptr;
            }
            const uint64_t cap = std::max<uint64_t>(words * 2, (uint64_t)1 << 22);
            s->rcap = 0;
            for (int p = 0; p < 2; ++p) ESTH_TRY(esth_engine_alloc(s->e, cap * sizeof(float), (void**)&s->R[p]));
            s->rcap = cap;
        }
    }
    const int p = (int)(widx & 1);
    unsigned char* h = sl.h;
    unsigned char* dt = sl.dtab;
    esth_bfilter* hF = (esth_bfilter*)(h + o_F);
    uint2* hall = (uint2*)(h + o_all);
    uint2* hcls = (uint2*)(h + o_cls);
    uint2* hunt = (uint2*)(h + o_unt);
    uint2* hchk = (uint2*)(h + o_chk);
    esth_mtgp_stream* hst = (esth_mtgp_stream*)(h + o_str);
    float* hfr = (float*)(h + o_frm);
    esth_sinit* hin = (esth_sinit*)(h + o_ini);

    std::vector<esth_wave_group> groups;
    uint64_t r_off = 0, f_off = 0, sub_pos = 0, str_pos = 0, chk_pos = 0;
    for (size_t i = 0; i < nw; ++i) {
        esth_strack* t = upd[i].t;
        if (groups.empty() || groups.back().key != t->group) {
            esth_wave_group g;
            g.key = t->group;
            g.f0 = (uint32_t)i;
            g.sub0 = sub_pos;
            g.chk0 = chk_pos;
            for (int lm = 0; lm <= 10; ++lm) g.lanes[lm] = flanes[lm];
            groups.push_back(g);
        }
        esth_wave_group& g = groups.back();
        esth_bfilter f = t->bf;
        f.r_off = r_off;
        f.f_off = f_off;
        f.est = sl.dest + i * ESTH_SRV_EST;
        f.result = (void*)(sl.dres + i);
        hF[i] = f;
        const uint32_t li = (uint32_t)(i - g.f0);
        for (uint32_t b = 0; b < t->d.num_blocks; ++b) hall[sub_pos + b] = make_uint2(li, b);
        sub_pos += t->d.num_blocks;
        if (fused(t))
            g.fuse_bits |= 1u << t->logm;
        else
            for (uint32_t q = 0; q < t->bf.N / 2; q += ESTH_WAVE) hchk[chk_pos++] = make_uint2(li, q);
        g.nchk = chk_pos - g.chk0;
        g.nf++;
        g.nsub += t->d.num_blocks;
        g.ncls[t->logm] += t->d.num_blocks;
        for (int j = 0; j < t->K; ++j)
            hst[str_pos++] = {(uint32_t)j, t->rsize, r_off + (uint64_t)j * t->rsize, t->soff + (uint64_t)j};
        memcpy(hfr + f_off, t->q.front().frame.data(), sizeof(float) * t->frame_floats);
        r_off += t->words;
        f_off += t->frame_floats;
    }
    /* per-group class lists */
    uint64_t cpos = 0;
    for (esth_wave_group& g : groups) {
        uint64_t c0[11];
        for (int lm = 0; lm <= 10; ++lm) {
            g.cls0[lm] = cpos;
            c0[lm] = cpos;
            cpos += g.ncls[lm];
        }
        for (uint64_t j = g.sub0; j < g.sub0 + g.nsub; ++j) {
            const esth_strack* t = upd[g.f0 + hall[j].x].t;
            hcls[c0[t->logm]++] = hall[j];
        }
    }
    uint64_t upos = 0;
    for (esth_wave_group& g : groups) {
        g.unit0 = upos;
        if (g.ncls[6] && ((g.fuse_bits >> 6) & 1u))
            g.nunit = esth_pair_units(hcls + g.cls0[6], g.ncls[6], hunt + upos);
        upos += g.nunit;
    }
    for (size_t i = 0; i < inits.size(); ++i) {
        esth_strack* t = inits[i];
        esth_sinit& r = hin[i];
        memset(&r, 0, sizeof(r));
        r.state = t->bf.state;
        r.w = t->bf.w;
        r.status = (uint32_t*)s->base + t->soff * ESTH_MTGP_N;
        r.N = t->bf.N;
        r.S = (uint32_t)t->S;
        r.K = (uint32_t)t->K;
        r.seed = t->d.seed;
        r.w0 = (t->d.carry_weights && t->d.weights != ESTH_WEIGHTS_LOG) ? 1.0f : 0.0f;
        memcpy(r.init, t->init, sizeof(float) * (size_t)t->S);
        t->need_init = false;
    }
    /* RNG stream: tables up, start-up, MTGP32 */
    if (widx >= 2) {
        const uint64_t prev = widx - 2;   /* last reader of R[p] */
        if (prev >= s->completed) ESTH_HIP(hipStreamWaitEvent(s->rng_stream, s->slot[prev % ESTH_SRV_SLOTS].ev_done, 0));
    }
    ESTH_HIP(hipMemcpyAsync(dt, h, total, hipMemcpyHostToDevice, s->rng_stream));
    if (!inits.empty()) {
        uint32_t maxN = 0;
        for (esth_strack* t : inits) maxN = std::max(maxN, t->bf.N);
        const unsigned bx = std::min<unsigned>(64u, (maxN + 255) / 256);
        hipLaunchKernelGGL(ks_init, dim3(bx, (unsigned)inits.size()), dim3(256), 0, s->rng_stream,
                           (const esth_sinit*)(dt + o_ini), (const uint32_t*)s->e->d_params);
    }
    if (nstreams)
        ESTH_TRY(esth_launch_mtgp(s->e->d_params, (const uint32_t*)s->base, (uint32_t*)s->base,
                                  (const esth_mtgp_stream*)(dt + o_str), (int)nstreams, s->R[p], 1, s->rng_stream));
    ESTH_HIP(hipEventRecord(sl.ev_rng, s->rng_stream));
    /* main stream: the step of every track in the wave */
    ESTH_HIP(hipStreamWaitEvent(s->stream, sl.ev_rng, 0));
    const esth_bfilter* dF = (const esth_bfilter*)(dt + o_F);
    for (const esth_wave_group& g : groups) {
        const esth_sgroup_key& k = s->keys[g.key];
        const uint2* sa = (const uint2*)(dt + o_all);
        const uint2* sc = (const uint2*)(dt + o_cls);
        const uint2* su = (const uint2*)(dt + o_unt);
        const uint2* ck = (const uint2*)(dt + o_chk);
        const float* fr = (const float*)(dt + o_frm);
        ESTH_TRY(esth_with_model_state_dim(upd[g.f0].t->S, [&](auto sv) {
            return group_kernels<decltype(sv)::value>(s, k, g, dF, sa, sc, su, ck, s->R[p], fr);
        }));
    }
    ESTH_HIP(hipGetLastError());
    if (nw) ESTH_HIP(hipMemcpyAsync(sl.hest, sl.dest, sizeof(double) * ESTH_SRV_EST * nw, hipMemcpyDeviceToHost, s->stream));
    ESTH_HIP(hipEventRecord(sl.ev_done, s->stream));

    esth_wave w;
    w.idx = widx;
    w.slot = si;
    for (esth_strack* t : inits) w.inits.push_back({t->id, t->open_req, t->open_submit});
    for (auto& u : upd) {
        esth_strack* t = u.t;
        t->q.pop_front();
        s->queued--;
        t->steps++;
        t->last_wave = widx + 1;
        make_ready(s, t);
    }
    for (esth_strack* t : inits) t->last_wave = std::max(t->last_wave, widx + 1);
    w.upd = std::move(upd);
    s->inflight.push_back(std::move(w));
    s->last_slot_wave[si] = widx + 1;
    s->waves++;
    s->wave_updates += nw;
    *launched = true;
    return ESTH_OK;
}

extern "C" int esth_server_pump(esth_server* s) {
    if (!s) return esth_fail(ESTH_ERR_INVALID, "null server");
    ESTH_HIP(hipSetDevice(s->e->device));
    ESTH_TRY(collect(s, false));
    while ((int)s->inflight.size() < s->depth) {
        bool launched = false;
        ESTH_TRY(launch_wave(s, &launched));
        if (!launched) break;
    }
    return ESTH_OK;
}

extern "C" int esth_server_poll(esth_server* s, esth_result* out, int max_results, int* n_out) {
    if (!s || !n_out || (max_results > 0 && !out)) return esth_fail(ESTH_ERR_INVALID, "server poll: null argument");
    int n = 0;
    while (n < max_results && !s->results.empty()) {
        out[n++] = s->results.front();
        s->results.pop_front();
    }
    *n_out = n;
    return ESTH_OK;
}

extern "C" int esth_server_drain(esth_server* s) {
    if (!s) return esth_fail(ESTH_ERR_INVALID, "null server");
    ESTH_HIP(hipSetDevice(s->e->device));
    for (;;) {
        ESTH_TRY(esth_server_pump(s));
        if (s->inflight.empty()) {
            /* nothing launched: whatever is still queued waits for admission */
            bool pending = false;
            for (auto* t : s->ready) pending |= !t->q.empty();
            if (!pending) break;
        } else {
            ESTH_TRY(collect(s, true));
        }
    }
    return ESTH_OK;
}

extern "C" int esth_server_stats(esth_server* s, size_t* hbm_used, int* open_tracks, int* queued) {
    if (!s) return esth_fail(ESTH_ERR_INVALID, "null server");
    if (hbm_used) *hbm_used = (size_t)s->arena.used;
    if (open_tracks) *open_tracks = s->live;
    if (queued) *queued = (int)s->queued;
    return ESTH_OK;
}

extern "C" int esth_server_wave_stats(esth_server* s, uint64_t* waves, uint64_t* wave_updates, size_t* hbm_peak) {
    if (!s) return esth_fail(ESTH_ERR_INVALID, "null server");
    if (waves) *waves = s->waves;
    if (wave_updates) *wave_updates = s->wave_updates;
    if (hbm_peak) *hbm_peak = (size_t)s->arena.peak;
    return ESTH_OK;
}

extern "C" int esth_server_read_particles(esth_server* s, uint64_t track_id, float* state_soa, float* weights) {
    if (!s) return esth_fail(ESTH_ERR_INVALID, "null server");
    esth_strack* t = find_track(s, track_id);
    if (!t || t->phase != 1) return esth_fail(ESTH_ERR_STATE, "server read: track %llu is not admitted", (unsigned long long)track_id);
    ESTH_HIP(hipSetDevice(s->e->device));
    ESTH_TRY(quiesce(s));
    if (t->need_init) return esth_fail(ESTH_ERR_STATE, "server read: track %llu has not started", (unsigned long long)track_id);
    if (state_soa) ESTH_HIP(hipMemcpy(state_soa, t->bf.state, sizeof(float) * (size_t)t->bf.N * t->S, hipMemcpyDeviceToHost));
    if (weights) ESTH_HIP(hipMemcpy(weights, t->bf.w, sizeof(float) * t->bf.N, hipMemcpyDeviceToHost));
    return ESTH_OK;
}

/* ============================================================ workload driver
 * A request stream for the served form: Poisson arrivals at `rate`, a share
 * `open_fraction` opening tracks (unicycle with probability
 * `unicycle_fraction`, else the robot arm; particle counts log-uniform in
 * [min_particles, max_particles], multiples of 64), the rest updating an open
 * track chosen uniformly; a track closes after a geometric number of updates
 * of mean `mean_lifetime`.  With rate == 0 the stream is closed-loop instead:
 * `initial_tracks` tracks, each with one update always queued.  Unicycle
 * tracks replay one of eight seeded scenarios (esth_unicycle_scenario), robot
 * arm tracks the caller's trace.  `plain` serves the same stream with one
 * esth_filter per track and per-request launches. */
namespace {

struct sim_track {
    uint64_t id = 0;          /* server track id */
    int model = 0, scen = 0;
    uint32_t particles = 0;
    uint64_t step = 0;
    int64_t remaining = -1;   /* updates before close; < 0: never */
    esth_filter* f = nullptr; /* plain mode */
};

struct sim_pool {
    int L = 0;
    int steps = 256;
    std::vector<esth_unicycle_params> up;
    std::vector<std::vector<esth_unicycle_control>> uc;
    std::vector<std::vector<esth_unicycle_measurement>> um;
};

struct sim {
    esth_engine* e;
    const esth_serve_workload* w;
    sim_pool pool;
    std::mt19937_64 rng;
    std::vector<sim_track> open;   /* open from the client's view */
    uint64_t next_seed = 1;
    std::vector<double> lat;
    uint64_t opens = 0, updates = 0, closes = 0, failed = 0, requests = 0;
    double particle_updates = 0.0;
    explicit sim(esth_engine* e_, const esth_serve_workload* w_) : e(e_), w(w_), rng(w_->seed) {}

    double unif() { return std::uniform_real_distribution<double>(0.0, 1.0)(rng); }

    uint32_t draw_particles() {
        const double lo = std::log((double)w->min_particles), hi = std::log((double)w->max_particles);
        uint64_t p = (uint64_t)std::llround(std::exp(lo + (hi - lo) * unif()) / 64.0) * 64u;
        p = std::max<uint64_t>(p, std::max<uint32_t>(64u, (w->min_particles + 63u) / 64u * 64u));
        return (uint32_t)p;
    }

    int desc_for(int model, uint32_t particles, esth_filter_desc* d) {
        esth_filter_desc_default(d, model);
        uint32_t m = 0;
        if (w->us_per_mparticle) {
            ESTH_TRY(esth_autotune_pick(w->us_per_mparticle, particles, 64, &m));
        } else {
            m = 64;
            while (m < 256 && particles % (2 * m) == 0) m *= 2;
        }
        d->num_particles = m;
        d->num_blocks = particles / m;
        d->topology = w->topology;
        d->num_transfer = (uint32_t)w->num_transfer;
        d->resampler = w->resampler;
        d->weights = w->weights;
        d->resampling_threshold = 1.0f;
        d->rng = ESTH_RNG_MTGP_PER_FILTER;
        d->seed = (uint32_t)(w->seed * 2654435761u + 7919u * (uint32_t)next_seed++);
        if (d->topology == ESTH_TOPO_TORUS) {   /* counts whose grid does not factor run as a ring */
            int dx = 0, dy = 0;
            esth_torus_dims(d->num_blocks, &dx, &dy);
            if ((uint64_t)dx * (uint64_t)dy != d->num_blocks) d->topology = ESTH_TOPO_RING;
        }
        if (d->topology == ESTH_TOPO_NONE || d->num_blocks < 2) { d->topology = ESTH_TOPO_NONE; d->num_transfer = 0; }
        return ESTH_OK;
    }

    void init_state(int model, float* s) {
        memset(s, 0, sizeof(float) * ESTH_MAX_STATE_DIM);
        if (model == ESTH_MODEL_ROBOTARM5) {
            for (int i = 0; i < 5; ++i) s[i] = 0.75f * 3.14159265358979f;
            s[5] = 7.0f;
        }
    }

    int build_pool() {
        if (w->unicycle_fraction > 0.0) {
            pool.L = w->landmarks;
            pool.up.resize(8);
            pool.uc.assign(8, std::vector<esth_unicycle_control>(pool.steps));
            pool.um.assign(8, std::vector<esth_unicycle_measurement>(pool.steps));
            std::vector<float> truth((size_t)pool.steps * 3);
            for (int i = 0; i < 8; ++i)
                ESTH_TRY(esth_unicycle_scenario(w->seed * 31u + (uint32_t)i, pool.L, 50.0f, pool.steps, 0.1f, &pool.up[i],
                                                pool.uc[i].data(), pool.um[i].data(), truth.data()));
        }
        if (w->unicycle_fraction < 1.0 && (!w->ra_controls || !w->ra_measurements || !w->ra_dt || w->ra_steps <= 0))
            return esth_fail(ESTH_ERR_INVALID, "serve workload: robot-arm tracks need a trace");
        return ESTH_OK;
    }

    const void* params_of(const sim_track& t, size_t* sz) {
        if (t.model == ESTH_MODEL_UNICYCLE) { *sz = sizeof(esth_unicycle_params); return &pool.up[t.scen]; }
        *sz = 0;
        return nullptr;
    }
    void inputs_of(const sim_track& t, const void** c, const void** m, float* dt) {
        if (t.model == ESTH_MODEL_UNICYCLE) {
            const int k = (int)(t.step % (uint64_t)pool.steps);
            *c = &pool.uc[t.scen][k];
            *m = &pool.um[t.scen][k];
            *dt = 0.1f;
        } else {
            const int k = (int)(t.step % (uint64_t)w->ra_steps);
            *c = &w->ra_controls[k];
            *m = &w->ra_measurements[k];
            *dt = w->ra_dt[k];
        }
    }
    sim_track new_track() {
        sim_track t;
        t.model = unif() < w->unicycle_fraction ? ESTH_MODEL_UNICYCLE : ESTH_MODEL_ROBOTARM5;
        t.scen = (int)(rng() % 8u);
        t.particles = draw_particles();
        if (w->mean_lifetime > 0.0) {
            std::geometric_distribution<int64_t> g(1.0 / w->mean_lifetime);
            t.remaining = 1 + g(rng);
        }
        return t;
    }
};

double pct(std::vector<double>& v, double q) {
    if (v.empty()) return NAN;
    const size_t k = std::min(v.size() - 1, (size_t)std::floor(q * (double)(v.size() - 1) + 0.5));
    std::nth_element(v.begin(), v.begin() + (long)k, v.end());
    return v[k];
}

}  // namespace

static int sim_served(sim& S, esth_serve_report* out) {
    const esth_serve_workload* w = S.w;
    esth_server* srv = nullptr;
    ESTH_TRY(esth_server_create(S.e, w->max_tracks, w->hbm_bytes, &srv));
    std::unique_ptr<esth_server, int (*)(esth_server*)> guard(srv, esth_server_destroy);
    ESTH_TRY(esth_server_configure(srv, w->max_inflight > 0 ? w->max_inflight : 2, 1 << 30));
    std::unordered_map<uint64_t, double> sched;      /* update request -> arrival (us, absolute) */
    std::unordered_map<uint64_t, size_t> by_track;    /* server track id -> index into S.open */
    std::unordered_map<uint64_t, uint32_t> parts;
    auto open_track = [&](sim_track t) -> int {
        esth_filter_desc d;
        ESTH_TRY(S.desc_for(t.model, t.particles, &d));
        size_t psz = 0;
        const void* pp = S.params_of(t, &psz);
        float init[ESTH_MAX_STATE_DIM];
        S.init_state(t.model, init);
        uint64_t rq;
        ESTH_TRY(esth_server_open(srv, &d, pp, psz, init, &t.id, &rq));
        S.opens++;
        S.requests++;
        by_track[t.id] = S.open.size();
        parts[t.id] = t.particles;
        S.open.push_back(t);
        return ESTH_OK;
    };
    auto update_track = [&](size_t i, double arrival) -> int {
        sim_track& t = S.open[i];
        const void *c, *m;
        float dt;
        S.inputs_of(t, &c, &m, &dt);
        uint64_t rq;
        ESTH_TRY(esth_server_update(srv, t.id, c, m, dt, &rq));
        sched[rq] = arrival;
        t.step++;
        S.requests++;
        if (t.remaining > 0 && --t.remaining == 0) {
            ESTH_TRY(esth_server_close(srv, t.id, &rq));
            S.requests++;
            S.closes++;
            by_track.erase(t.id);
            if (i + 1 != S.open.size()) {
                S.open[i] = S.open.back();
                by_track[S.open[i].id] = i;
            }
            S.open.pop_back();
        }
        return ESTH_OK;
    };
    for (int i = 0; i < w->initial_tracks; ++i) ESTH_TRY(open_track(S.new_track()));
    ESTH_TRY(esth_server_drain(srv));
    std::vector<esth_result> res(4096);
    int n = 0;
    do ESTH_TRY(esth_server_poll(srv, res.data(), (int)res.size(), &n)); while (n > 0);

    const bool closed_loop = w->rate <= 0.0;
    const double t0 = srv_now_us();
    const double t_end = t0 + w->duration_s * 1e6;
    double next = t0;
    std::exponential_distribution<double> gap(closed_loop ? 1.0 : w->rate * 1e-6);
    if (closed_loop) {
        for (size_t i = 0; i < S.open.size(); ++i) {
            S.open[i].remaining = -1;
            ESTH_TRY(update_track(i, t0));
        }
    } else {
        next += gap(S.rng);
    }
    double last_done = t0;
    bool arrivals = true;
    for (;;) {
        double now = srv_now_us();
        if (!closed_loop) {
            while (arrivals && next <= now) {
                if (next >= t_end) { arrivals = false; break; }
                if (S.open.empty() || S.unif() < w->open_fraction) ESTH_TRY(open_track(S.new_track()));
                else ESTH_TRY(update_track((size_t)(S.rng() % S.open.size()), next));
                next += gap(S.rng);
            }
        } else if (now >= t_end) {
            arrivals = false;
        }
        ESTH_TRY(esth_server_pump(srv));
        ESTH_TRY(esth_server_poll(srv, res.data(), (int)res.size(), &n));
        for (int i = 0; i < n; ++i) {
            const esth_result& r = res[(size_t)i];
            if (r.status != ESTH_OK) { S.failed++; continue; }
            if (r.kind != 1) continue;
            auto it = sched.find(r.request_id);
            if (it == sched.end()) continue;
            S.lat.push_back(r.done_us - it->second);
            sched.erase(it);
            S.updates++;
            S.particle_updates += parts[r.track_id];
            last_done = std::max(last_done, r.done_us);
            if (closed_loop && arrivals) {
                auto bt = by_track.find(r.track_id);
                if (bt != by_track.end()) ESTH_TRY(update_track(bt->second, r.done_us));
            }
        }
        /* after the window: stop once nothing is left that can still run (requests
         * of tracks that never got admitted stay in the backlog) */
        if (!arrivals && (sched.empty() || srv->inflight.empty())) break;
    }
    out->wall_s = (last_done - t0) * 1e-6;
    out->backlog = sched.size();
    uint64_t waves = 0, wu = 0;
    size_t peak = 0;
    esth_server_wave_stats(srv, &waves, &wu, &peak);
    out->waves = waves;
    out->mean_wave_updates = waves ? (double)wu / (double)waves : 0.0;
    out->hbm_peak = peak;
    out->peak_tracks = srv->live;
    return ESTH_OK;
}

static int sim_plain(sim& S, esth_serve_report* out) {
    const esth_serve_workload* w = S.w;
    struct closer {
        std::vector<sim_track>* v;
        ~closer() { for (auto& t : *v) if (t.f) esth_filter_destroy(t.f); }
    } cl{&S.open};
    auto open_track = [&](sim_track t) -> int {
        esth_filter_desc d;
        ESTH_TRY(S.desc_for(t.model, t.particles, &d));
        size_t psz = 0;
        const void* pp = S.params_of(t, &psz);
        float init[ESTH_MAX_STATE_DIM];
        S.init_state(t.model, init);
        ESTH_TRY(esth_filter_create(S.e, &d, pp, psz, &t.f));
        int rc = esth_filter_set_prefetch(t.f, 0);
        if (rc == ESTH_OK) rc = esth_filter_init_uniform(t.f, init);
        if (rc != ESTH_OK) { esth_filter_destroy(t.f); return rc; }
        S.opens++;
        S.requests++;
        S.open.push_back(t);
        return ESTH_OK;
    };
    auto step_track = [&](sim_track& t) -> int {
        const void *c, *m;
        float dt;
        S.inputs_of(t, &c, &m, &dt);
        ESTH_TRY(esth_filter_step(t.f, c, m, dt));
        t.step++;
        S.requests++;
        return ESTH_OK;
    };
    for (int i = 0; i < w->initial_tracks; ++i) ESTH_TRY(open_track(S.new_track()));
    ESTH_TRY(esth_dev_sync());
    const bool closed_loop = w->rate <= 0.0;
    const double t0 = srv_now_us();
    const double t_end = t0 + w->duration_s * 1e6;
    double last_done = t0;
    esth_estimate est;
    if (closed_loop) {
        for (auto& t : S.open) t.remaining = -1;
        while (srv_now_us() < t_end) {
            const double r0 = srv_now_us();
            for (auto& t : S.open) ESTH_TRY(step_track(t));
            for (auto& t : S.open) {
                ESTH_TRY(esth_filter_estimate(t.f, &est));
                const double d = srv_now_us();
                S.lat.push_back(d - r0);
                S.updates++;
                S.particle_updates += t.particles;
                last_done = d;
            }
        }
    } else {
        std::exponential_distribution<double> gap(w->rate * 1e-6);
        double next = t0 + gap(S.rng);
        while (next < t_end) {
            while (srv_now_us() < next) {}
            if (S.open.empty() || S.unif() < w->open_fraction) {
                ESTH_TRY(open_track(S.new_track()));
            } else {
                const size_t i = (size_t)(S.rng() % S.open.size());
                sim_track& t = S.open[i];
                ESTH_TRY(step_track(t));
                ESTH_TRY(esth_filter_estimate(t.f, &est));
                const double d = srv_now_us();
                S.lat.push_back(d - next);
                S.updates++;
                S.particle_updates += t.particles;
                last_done = d;
                if (t.remaining > 0 && --t.remaining == 0) {
                    esth_filter_destroy(t.f);
                    S.closes++;
                    S.requests++;
                    S.open[i] = S.open.back();
                    S.open.pop_back();
                }
            }
            next += gap(S.rng);
        }
    }
    out->wall_s = (last_done - t0) * 1e-6;
    out->peak_tracks = (int)S.open.size();
    return ESTH_OK;
}

extern "C" int esth_serve_simulate(esth_engine* e, const esth_serve_workload* w, esth_serve_report* out) {
    if (!e || !w || !out) return esth_fail(ESTH_ERR_INVALID, "serve simulate: null argument");
    if (w->duration_s <= 0.0 || w->min_particles < 64 || w->max_particles < w->min_particles || w->max_tracks <= 0)
        return esth_fail(ESTH_ERR_INVALID, "serve simulate: bad workload");
    if (w->rate <= 0.0 && w->initial_tracks <= 0) return esth_fail(ESTH_ERR_INVALID, "serve simulate: closed loop needs tracks");
    memset(out, 0, sizeof(*out));
    sim S(e, w);
    ESTH_TRY(S.build_pool());
    ESTH_TRY(w->plain ? sim_plain(S, out) : sim_served(S, out));
    out->requests = S.requests;
    out->opens = S.opens;
    out->updates = S.updates;
    out->closes = S.closes;
    out->failed = S.failed;
    out->updates_per_s = out->wall_s > 0.0 ? (double)S.updates / out->wall_s : 0.0;
    out->particle_updates_per_s = out->wall_s > 0.0 ? S.particle_updates / out->wall_s : 0.0;
    out->mean_particles = S.updates ? S.particle_updates / (double)S.updates : 0.0;
    double sum = 0.0, mx = 0.0;
    for (double v : S.lat) { sum += v; mx = std::max(mx, v); }
    out->mean_us = S.lat.empty() ? NAN : sum / (double)S.lat.size();
    out->max_us = mx;
    out->p50_us = pct(S.lat, 0.50);
    out->p99_us = pct(S.lat, 0.99);
    return ESTH_OK;
}
