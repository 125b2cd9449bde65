/*
 * esth_kernels.h - gfx950 kernels of the filter step.
 *
 * Layout in HBM (one filter, N = m*G particles):
 *   state, tmp : S float planes of N (structure of arrays, plane c at c*N)
 *   w          : N floats, linear or log weights
 *   R          : the step's MTGP32 words, K streams x size (float in [1,2))
 *   winner     : G float2 (best weight, block id)
 *   bt_w, bt_s : best t of every sub-filter, compact (G*t, S planes of G*t)
 *   part       : G x (2+S) floats: block max, sum w, sum w*x_c (posterior)
 *
 * Execution model: one 64-lane wave per sub-filter for sort and resampling
 * (E = m/64 elements per lane, lane-major, e = lane*E + r; for m < 64 a wave
 * holds 64/m sub-filters with one element per lane).  Strides below E are
 * register compare-exchanges, larger ones cross lanes by xor shuffles, so the
 * sort needs no LDS and no barrier.  Vose's alias build keeps the reference's
 * work-group-per-sub-filter shape because its pairing rounds are
 * data-dependent and share worklists.
 */
#ifndef ESTH_KERNELS_H
#define ESTH_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "esth_layout.h"
#include "esth_models.h"

#define ESTH_WAVE 64

/* ------------------------------------------------------------ wave ops */
__device__ __forceinline__ float esth_shfl_xor(float v, int mask) {
    int idx = ((int)__lane_id() ^ mask) << 2;
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(idx, __builtin_bit_cast(int, v)));
}
__device__ __forceinline__ int esth_shfl_xor_i(int v, int mask) {
    int idx = ((int)__lane_id() ^ mask) << 2;
    return __builtin_amdgcn_ds_bpermute(idx, v);
}
__device__ __forceinline__ float esth_shfl_idx(float v, int src_lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(src_lane << 2, __builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double esth_shfl_xor_d(double v, int mask) {
    uint64_t u = __builtin_bit_cast(uint64_t, v);
    int lo = esth_shfl_xor_i((int)(uint32_t)u, mask);
    int hi = esth_shfl_xor_i((int)(uint32_t)(u >> 32), mask);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

/* the halves of a 32- or 16-lane xor exchange (MASK 32 or 16), from one
 * v_permlane32_swap / v_permlane16_swap of v with itself: lo[i] is the value
 * of the lane of the pair {i, i ^ MASK} whose MASK bit is 0, hi[i] that of the
 * lane whose bit is 1, in both lanes of the pair.  No LDS access. */
template <int MASK>
__device__ __forceinline__ void esth_xor_halves(int v, int& lo, int& hi) {
    static_assert(MASK == 16 || MASK == 32, "permlane swaps pair lanes 16 or 32 apart");
    if constexpr (MASK == 32) {
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, true, false);
        lo = (int)r[0];
        hi = (int)r[1];
    } else {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, true, false);
        lo = (int)r[0];
        hi = (int)r[1];
    }
}

/* xor-lane exchange with a compile-time mask: DPP quad_perm (1, 2), DPP
 * row_ror:8 (8), ds_swizzle SWAP mode (4), a permlane swap (16, 32). */
template <int MASK>
__device__ __forceinline__ int esth_xor_lane(int v) {
    if constexpr (MASK == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);
    else if constexpr (MASK == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);
    else if constexpr (MASK == 4) return __builtin_amdgcn_ds_swizzle(v, 0x101F);
    else if constexpr (MASK == 8) return __builtin_amdgcn_mov_dpp(v, 0x128, 0xF, 0xF, false);
    else if constexpr (MASK == 16 || MASK == 32) {
        int lo, hi;
        esth_xor_halves<MASK>(v, lo, hi);
        return ((int)__lane_id() & MASK) ? lo : hi;
    } else return esth_shfl_xor_i(v, MASK);
}
template <int MASK>
__device__ __forceinline__ float esth_xor_lane_f(float v) {
    return __builtin_bit_cast(float, esth_xor_lane<MASK>(__builtin_bit_cast(int, v)));
}
/* butterfly sum over lane groups of size 2^LOG (fixed pairing order) */
template <int K, int LIM>
__device__ __forceinline__ float esth_bfly_sum(float v) {
    if constexpr (K < LIM) return esth_bfly_sum<2 * K, LIM>(v + esth_xor_lane_f<K>(v));
    else return v;
}
template <int K, int LIM>
__device__ __forceinline__ float esth_bfly_max(float v) {
    if constexpr (K < LIM) {
        float o = esth_xor_lane_f<K>(v);
        return esth_bfly_max<2 * K, LIM>(o > v ? o : v);
    } else {
        return v;
    }
}

/* ================================================================= MTGP32
 * One work-group of 256 per stream, 768-word status ring in LDS; each round
 * produces 768 words in three 256-wide phases, each one the recursion
 * x[n] = rec(x[n-351], x[n-350], x[n-351+pos]) for 256 consecutive n
 * (mtgp32-opencl.cl:186-242).  Output is the sequential stream of the
 * stream's parameter set, continuing across launches through the status.
 *
 * One launch generates `nsteps` consecutive steps' words: step s's `size`
 * words of stream j go to out + s * out_step + out_off(j), and the status
 * after step s (the 351 words that generate step s + 1) to status_out +
 * s * status_step + 351 * status(j).  status_in may equal status_out when
 * nsteps is 1: each work-group reads its status into LDS before the first
 * barrier and writes it back after the last.
 *
 * Every value a phase reads is older than the phase: x[n-351+pos] with
 * pos <= 93 < 96 for all 200 parameter sets, and the tempering's
 * x[n-352+pos] before it, so the four LDS reads are issued together ahead
 * of the phase's write. */
/* stream j of a launch, run by one 256-thread work-group; st: 800 words of
 * LDS (the 768-word ring and the two 16-word tables) */
__device__ __forceinline__ void esth_mtgp_wg(const uint32_t* __restrict__ params, uint32_t mask,
                                             const uint32_t* status_in, uint32_t* status_out,
                                             const esth_mtgp_stream* __restrict__ streams, uint32_t j,
                                             uint32_t* __restrict__ out, int float_mode, int nsteps,
                                             uint64_t out_step, uint64_t status_step, uint32_t* st) {
    uint32_t* tbl = st + ESTH_MTGP_LARGE;
    uint32_t* ttbl = tbl + 16;
    const int tid = threadIdx.x;
    const esth_mtgp_stream sd = streams[j];
    const uint32_t* p = params + (size_t)sd.set * ESTH_MTGP_PARAM_WORDS;
    const int pos = (int)p[0];
    const uint32_t sh1 = p[1], sh2 = p[2];
    if (tid < 16) {
        tbl[tid] = p[4 + tid];
        ttbl[tid] = float_mode ? p[36 + tid] : p[20 + tid];
    }
    const uint32_t* gin = status_in + (size_t)sd.status * ESTH_MTGP_N;
    st[ESTH_MTGP_LARGE - ESTH_MTGP_N + tid] = gin[tid];
    if (tid < ESTH_MTGP_N - ESTH_MTGP_THREADS)
        st[ESTH_MTGP_LARGE - ESTH_MTGP_N + ESTH_MTGP_THREADS + tid] = gin[ESTH_MTGP_THREADS + tid];
    __syncthreads();
    const int size = (int)sd.size;
    /* start offsets (mod 768) of x[n-351] for the three phases */
    const int b0 = ESTH_MTGP_LARGE - ESTH_MTGP_N;          /* 417 */
    const int b1 = 4 * ESTH_MTGP_THREADS - ESTH_MTGP_N;    /* 673 */
    const int b2 = 2 * ESTH_MTGP_THREADS - ESTH_MTGP_N;    /* 161 */
    for (int step = 0; step < nsteps; ++step) {
        uint32_t* o = out + (size_t)step * out_step + sd.out_off;
        for (int i = 0; i < size; i += ESTH_MTGP_LARGE) {
#pragma unroll
            for (int ph = 0; ph < 3; ++ph) {
                const int b = ph == 0 ? b0 : (ph == 1 ? b1 : b2);
                int i0 = b + tid, i1 = b + tid + 1, ip = b + tid + pos, it = b + tid + pos - 1;
                if (ph == 1) { i0 %= ESTH_MTGP_LARGE; i1 %= ESTH_MTGP_LARGE; ip %= ESTH_MTGP_LARGE; it %= ESTH_MTGP_LARGE; }
                const uint32_t x0 = st[i0], x1 = st[i1], xp = st[ip];
                uint32_t T = st[it];
                uint32_t X = (x0 & mask) ^ x1;
                X ^= X << sh1;
                uint32_t Y = X ^ (xp >> sh2);
                uint32_t r = Y ^ tbl[Y & 0x0f];
                st[ph * ESTH_MTGP_THREADS + tid] = r;
                T ^= T >> 16;
                T ^= T >> 8;
                uint32_t v = float_mode ? ((r >> 9) ^ ttbl[T & 0x0f]) : (r ^ ttbl[T & 0x0f]);
                o[i + ph * ESTH_MTGP_THREADS + tid] = v;
                __syncthreads();
            }
        }
        /* the status after this step: the last 351 words of the ring, whose
         * newest word sits at 767 after a whole number of rounds */
        uint32_t* gst = status_out + (size_t)step * status_step + (size_t)sd.status * ESTH_MTGP_N;
        gst[tid] = st[ESTH_MTGP_LARGE - ESTH_MTGP_N + tid];
        if (tid < ESTH_MTGP_N - ESTH_MTGP_THREADS)
            gst[ESTH_MTGP_THREADS + tid] = st[4 * ESTH_MTGP_THREADS - ESTH_MTGP_N + tid];
    }
}

/* streams j = blockIdx.x, blockIdx.x + gridDim.x, ...: a launch over many
 * short streams (a batch's per-filter streams) keeps a resident grid instead
 * of one work-group launch per stream */
static __global__ __launch_bounds__(256) void k_mtgp(const uint32_t* __restrict__ params, uint32_t mask,
                                              const uint32_t* status_in, uint32_t* status_out,
                                              const esth_mtgp_stream* __restrict__ streams, uint32_t nstreams,
                                              uint32_t* __restrict__ out, int float_mode, int nsteps,
                                              uint64_t out_step, uint64_t status_step) {
    __shared__ uint32_t st[ESTH_MTGP_LARGE + 32];
    for (uint32_t j = blockIdx.x; j < nstreams; j += gridDim.x) {
        esth_mtgp_wg(params, mask, status_in, status_out, streams, j, out, float_mode, nsteps, out_step, status_step,
                     st);
        __syncthreads();   /* the ring and tables are reloaded for the next stream */
    }
}

/* one step of stream j for a sampling launch, kept out of line so that its
 * registers do not add to the sampling path's (a call per work-group) */
__device__ __noinline__ void esth_mtgp_wg_call(const esth_mtgp_inline mi, uint32_t j, uint32_t* st) {
    esth_mtgp_wg(mi.params, mi.mask, mi.status_in, mi.status_out, mi.streams, j, mi.out, 1, 1, 0, 0, st);
}

/* the next step's words from the first mi.n work-groups of a 256-thread
 * sampling launch: true for those work-groups, which do nothing else */
__device__ __forceinline__ bool esth_mtgp_inline_wg(const esth_mtgp_inline& mi, uint32_t* st) {
    if ((int)blockIdx.x >= mi.n) return false;
    esth_mtgp_wg_call(mi, blockIdx.x, st);
    return true;
}

/* ============================================================ Box-Muller */
static __global__ void k_boxmuller(float* __restrict__ d, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (2 * i < n) {
        float z0, z1;
        esth_boxmuller_pair(d[2 * i], d[2 * i + 1], &z0, &z1);
        d[2 * i] = z0;
        d[2 * i + 1] = z1;
    }
}

/* ========================================================= sample+weight
 * One thread per particle pair (p, p+1), p even.  The reference transforms
 * uniform pairs (2j, 2j+1) of the buffer in place and particle i reads the
 * normal at k*N + i; with N even the pair at k*N + p gives particle p its
 * cosine half and particle p+1 its sine half, so the transform is fused here
 * and the normals never touch HBM. */
template <int MODEL>
struct esth_model_traits;
/* one specialisation per line of esth_model_list.h: state floats S, normals
 * V, and the model's one- and two-particle hooks */
#define ESTH_X(ID, NAME, PFX, S_, V_, PSZ, CSZ, MSZ)                                                        \
    template <>                                                                                            \
    struct esth_model_traits<ID> {                                                                         \
        static constexpr int S = S_, V = V_;                                                               \
        __device__ static float iw(float* s, const float* z, const float* fr) {                            \
            return esth_##PFX##_sampling_iw(s, z, fr);                                                     \
        }                                                                                                  \
        __device__ static void iw_pair(float* s0, float* s1, const float* z0, const float* z1, const float* fr, \
                                       float* lw0, float* lw1) {                                           \
            esth_##PFX##_sampling_iw_pair(s0, s1, z0, z1, fr, lw0, lw1);                                    \
        }                                                                                                  \
    };
ESTH_MODEL_LIST(ESTH_X)
#undef ESTH_X

/* particles p (s0) and p+1 (s1) given their uniform pairs (u0[k], u1[k]) =
 * (R[k N + p], R[k N + p + 1]): Box-Muller, propagation and log-weight */
template <int MODEL>
__device__ __forceinline__ void esth_sample_pair_regs(float (&s0)[esth_model_traits<MODEL>::S],
                                                      float (&s1)[esth_model_traits<MODEL>::S],
                                                      const float (&u0)[esth_model_traits<MODEL>::V],
                                                      const float (&u1)[esth_model_traits<MODEL>::V],
                                                      const float* __restrict__ frame, float& lw0, float& lw1) {
    using T = esth_model_traits<MODEL>;
    float z0[T::V], z1[T::V];
#pragma unroll
    for (int k = 0; k < T::V; ++k) esth_boxmuller_pair(u0[k], u1[k], &z0[k], &z1[k]);
    T::iw_pair(s0, s1, z0, z1, frame, &lw0, &lw1);
}

/* particles p and p+1 (p even) of a state in layout L: SOA one float2 per
 * plane; AOS the pair's 2S consecutive floats as float2s; PACKED, per group
 * of w planes, the pair's 2w consecutive floats (float4s where w is 2 or 4:
 * a group starts at a multiple of 4N floats and the pair at 2pw) */
template <int S, int L>
__device__ __forceinline__ void esth_load_pair(const float* __restrict__ st, size_t N, uint32_t p, float (&s0)[S],
                                               float (&s1)[S]) {
    if constexpr (L == ESTH_LAYOUT_SOA) {
#pragma unroll
        for (int c = 0; c < S; ++c) {
            const float2 v = *reinterpret_cast<const float2*>(st + (size_t)c * N + p);
            s0[c] = v.x;
            s1[c] = v.y;
        }
    } else if constexpr (L == ESTH_LAYOUT_AOS) {
        const float* b = st + (size_t)p * S;
        float run[2 * S];
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const float2 v = *reinterpret_cast<const float2*>(b + 2 * k);
            run[2 * k] = v.x;
            run[2 * k + 1] = v.y;
        }
#pragma unroll
        for (int c = 0; c < S; ++c) { s0[c] = run[c]; s1[c] = run[S + c]; }
    } else {
#pragma unroll
        for (int q = 0; q < S; q += 4) {
            const int w = S - q < 4 ? S - q : 4;
            const float* b = st + (size_t)q * N + (size_t)p * w;
            float run[8];
            if (w == 4 || w == 2) {
#pragma unroll
                for (int k = 0; k < w / 2; ++k) {
                    const float4 v = *reinterpret_cast<const float4*>(b + 4 * k);
                    run[4 * k] = v.x; run[4 * k + 1] = v.y; run[4 * k + 2] = v.z; run[4 * k + 3] = v.w;
                }
            } else {
#pragma unroll
                for (int k = 0; k < w; ++k) {
                    const float2 v = *reinterpret_cast<const float2*>(b + 2 * k);
                    run[2 * k] = v.x;
                    run[2 * k + 1] = v.y;
                }
            }
#pragma unroll
            for (int j = 0; j < w; ++j) { s0[q + j] = run[j]; s1[q + j] = run[w + j]; }
        }
    }
}

template <int S, int L>
__device__ __forceinline__ void esth_store_pair(float* __restrict__ st, size_t N, uint32_t p, const float (&s0)[S],
                                                const float (&s1)[S]) {
    if constexpr (L == ESTH_LAYOUT_SOA) {
#pragma unroll
        for (int c = 0; c < S; ++c) *reinterpret_cast<float2*>(st + (size_t)c * N + p) = make_float2(s0[c], s1[c]);
    } else if constexpr (L == ESTH_LAYOUT_AOS) {
        float run[2 * S];
#pragma unroll
        for (int c = 0; c < S; ++c) { run[c] = s0[c]; run[S + c] = s1[c]; }
        float* b = st + (size_t)p * S;
#pragma unroll
        for (int k = 0; k < S; ++k) *reinterpret_cast<float2*>(b + 2 * k) = make_float2(run[2 * k], run[2 * k + 1]);
    } else {
#pragma unroll
        for (int q = 0; q < S; q += 4) {
            const int w = S - q < 4 ? S - q : 4;
            float run[8];
#pragma unroll
            for (int j = 0; j < w; ++j) { run[j] = s0[q + j]; run[w + j] = s1[q + j]; }
            float* b = st + (size_t)q * N + (size_t)p * w;
            if (w == 4 || w == 2) {
#pragma unroll
                for (int k = 0; k < w / 2; ++k)
                    *reinterpret_cast<float4*>(b + 4 * k) =
                        make_float4(run[4 * k], run[4 * k + 1], run[4 * k + 2], run[4 * k + 3]);
            } else {
#pragma unroll
                for (int k = 0; k < w; ++k) *reinterpret_cast<float2*>(b + 2 * k) = make_float2(run[2 * k], run[2 * k + 1]);
            }
        }
    }
}

template <int MODEL, int L = ESTH_LAYOUT_SOA>
__device__ __forceinline__ void esth_sample_weight_pair(float* __restrict__ state, uint32_t N,
                                                        const float* __restrict__ R,
                                                        const float* __restrict__ frame, int log_w,
                                                        float* __restrict__ w, uint32_t p, int carry = 0) {
    using T = esth_model_traits<MODEL>;
    float s0[T::S], s1[T::S], u0[T::V], u1[T::V];
    esth_load_pair<T::S, L>(state, N, p, s0, s1);
#pragma unroll
    for (int k = 0; k < T::V; ++k) {
        float2 u = *reinterpret_cast<const float2*>(R + (size_t)k * N + p);
        u0[k] = u.x;
        u1[k] = u.y;
    }
    float lw0, lw1;
    esth_sample_pair_regs<MODEL>(s0, s1, u0, u1, frame, lw0, lw1);
    esth_store_pair<T::S, L>(state, N, p, s0, s1);
    float2 wo = log_w ? make_float2(lw0, lw1) : make_float2(esth_expf(lw0), esth_expf(lw1));
    if (carry) {
        const float2 wp = *reinterpret_cast<const float2*>(w + p);
        wo = log_w ? make_float2(wp.x + wo.x, wp.y + wo.y) : make_float2(wp.x * wo.x, wp.y * wo.y);
    }
    *reinterpret_cast<float2*>(w + p) = wo;
}

/* The step's frame reaches the sampling kernels either in device memory
 * (frame_ptr: graph replay, whose captured arguments are fixed) or as a
 * kernel argument (FA: eager steps; the frame then travels with the launch
 * and needs no copy ordered before it).  Either way the kernels read it with
 * scalar loads. */
struct esth_frame_arg {
    float v[ESTH_FRAME_FLOATS];
};

template <int MODEL, int L, bool FA>
__global__ __launch_bounds__(256) void k_sample_weight(float* __restrict__ state, uint32_t N,
                                                       const float* __restrict__ R,
                                                       const float* __restrict__ frame_ptr,
                                                       const esth_frame_arg fa, int log_w,
                                                       float* __restrict__ w, int carry,
                                                       const esth_mtgp_inline mi) {
    if (mi.n) {
        __shared__ uint32_t mst[ESTH_MTGP_LARGE + 32];
        if (esth_mtgp_inline_wg(mi, mst)) return;
    }
    const float* __restrict__ frame = FA ? fa.v : frame_ptr;
    const uint32_t p = 2u * ((blockIdx.x - (uint32_t)mi.n) * blockDim.x + threadIdx.x);
    if (p < N) esth_sample_weight_pair<MODEL, L>(state, N, R, frame, log_w, w, p, carry);
}

/* Calibration: the sample+weight arithmetic of k_sample_weight (Box-Muller,
 * propagation, log-weight of a particle pair) with no memory access but one
 * 4-byte store per pair.  States in [-0.5, 0.5) and uniforms in [1, 2) are
 * made from the pair index with shifts, xors and adds (about 3 operations
 * per value, against ~10^3 per pair in the hooks). */
template <int MODEL>
__global__ __launch_bounds__(256) void k_hook_floor(float* __restrict__ out, uint32_t npairs,
                                                    const esth_frame_arg fa) {
    using T = esth_model_traits<MODEL>;
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npairs) return;
    uint32_t h = q ^ 0x9E3779B9u;
    h ^= h << 13;
    h ^= h >> 17;
    h ^= h << 5;
    float s0[T::S], s1[T::S], u0[T::V], u1[T::V];
#pragma unroll
    for (int c = 0; c < T::S; ++c) {
        s0[c] = esth_u2f(0x3f800000u | ((h + 0x1F123u * (uint32_t)c) & 0x7fffffu)) - 1.5f;
        s1[c] = esth_u2f(0x3f800000u | ((h ^ (0x2E571u * (uint32_t)c + 0x5A5u)) & 0x7fffffu)) - 1.5f;
    }
#pragma unroll
    for (int k = 0; k < T::V; ++k) {
        u0[k] = esth_u2f(0x3f800000u | ((h + 0x3B9ACu * (uint32_t)k + 0x11u) & 0x7fffffu));
        u1[k] = esth_u2f(0x3f800000u | ((h ^ (0x4C4B4u * (uint32_t)k + 0x7Fu)) & 0x7fffffu));
    }
    float lw0, lw1;
    esth_sample_pair_regs<MODEL>(s0, s1, u0, u1, fa.v, lw0, lw1);
    out[q] = lw0 + lw1;
}

/* ============================================================ block sort
 * The reference's descending bitonic network (bitonic_sort.cl:28-61 with
 * dir = 0): for size = 2..m, stride = size/2..1, pair (pos, pos+stride) with
 * pos & stride == 0 swaps iff (w[pos] > w[pos+stride]) == ddd, where
 * ddd = (pos & size) != 0 for size < m and 0 in the final merge.  Keys equal
 * to each other therefore swap in descending segments, exactly as there. */
__host__ __device__ constexpr int esth_ilog2(int x) { return x <= 1 ? 0 : 1 + esth_ilog2(x / 2); }

/* a sub-filter of M = 2^LOGM particles held by LANES lanes (by default one
 * wave, or all M lanes when M < 64), E = M / LANES consecutive positions each */
template <int LOGM, int LANES = ((1 << LOGM) >= ESTH_WAVE ? ESTH_WAVE : (1 << LOGM))>
struct esth_tile {
    static_assert(LANES >= 1 && LANES <= ESTH_WAVE && (1 << LOGM) % LANES == 0, "lanes per sub-filter");
    static constexpr int M = 1 << LOGM;
    static constexpr int LPS = LANES;                  /* lanes per sub-filter */
    static constexpr int E = M / LANES;                /* elements per lane */
    static constexpr int LOGL = esth_ilog2(LANES);
    static constexpr int LOGE = LOGM - LOGL;
};

/* a lane's E consecutive floats at p (p aligned to 4 * min(E, 4) bytes) */
template <int E>
__device__ __forceinline__ void esth_store_run(float* __restrict__ p, const float (&v)[E]) {
    if constexpr (E % 4 == 0) {
#pragma unroll
        for (int r = 0; r < E; r += 4)
            *reinterpret_cast<float4*>(p + r) = make_float4(v[r], v[r + 1], v[r + 2], v[r + 3]);
    } else if constexpr (E == 2) {
        *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
    } else {
#pragma unroll
        for (int r = 0; r < E; ++r) p[r] = v[r];
    }
}

/* a lane's E consecutive floats from p (p aligned like esth_store_run's) */
template <int E>
__device__ __forceinline__ void esth_load_run(const float* __restrict__ p, float (&v)[E]) {
    if constexpr (E % 4 == 0) {
#pragma unroll
        for (int r = 0; r < E; r += 4) {
            const float4 q = *reinterpret_cast<const float4*>(p + r);
            v[r] = q.x; v[r + 1] = q.y; v[r + 2] = q.z; v[r + 3] = q.w;
        }
    } else if constexpr (E == 2) {
        const float2 q = *reinterpret_cast<const float2*>(p);
        v[0] = q.x; v[1] = q.y;
    } else {
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = p[r];
    }
}

/* orders this wave's LDS accesses before and after it (the wave's lanes
 * exchange data through its own LDS region) */
__device__ __forceinline__ void esth_wave_lds_fence() {
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
}

/* v[r] = plane[pos[r]] over one sub-filter's M floats, through LDS: the
 * sub-filter's plane is read coalesced (each lane its E consecutive floats,
 * vector loads) into `buf`, the sub-filter's M-float region of this wave's
 * LDS, and the permuted reads are served there instead of as scattered
 * global loads, each of which would touch a different cache line. */
template <int M, int E>
__device__ __forceinline__ void esth_permute_lds(const float (&in)[E], int le, float* buf, const int (&pos)[E],
                                                 float (&v)[E]) {
    esth_wave_lds_fence();   /* earlier reads of buf are done */
    esth_store_run<E>(buf + le, in);
    esth_wave_lds_fence();
#pragma unroll
    for (int r = 0; r < E; ++r) v[r] = buf[pos[r] & (M - 1)];
}

template <int M, int E>
__device__ __forceinline__ void esth_gather_lds(const float* __restrict__ plane, int le, float* buf,
                                                const int (&pos)[E], float (&v)[E]) {
    float in[E];
    esth_load_run<E>(plane + le, in);
    esth_permute_lds<M, E>(in, le, buf, pos, v);
}

/* a[0] + a[1], then (a[0] + a[1]) + (a[2] + a[3]), ...: the order of a
 * butterfly sum over K adjacent lanes */
template <int K>
__device__ __forceinline__ float esth_tree_sum(float (&a)[K]) {
    if constexpr (K == 1) {
        return a[0];
    } else {
        float h[K / 2];
#pragma unroll
        for (int j = 0; j < K / 2; ++j) h[j] = a[2 * j] + a[2 * j + 1];
        return esth_tree_sum<K / 2>(h);
    }
}

template <int LOGM, int SIZE, int STRIDE, int LANES = esth_tile<LOGM>::LPS>
__device__ __forceinline__ void esth_bitonic_stage(float (&key)[(esth_tile<LOGM, LANES>::E)],
                                                   int (&idx)[(esth_tile<LOGM, LANES>::E)], int le, int laneg) {
    using TL = esth_tile<LOGM, LANES>;
    constexpr int M = TL::M, E = TL::E;
    /* compare-exchanges as selects: a swap under a divergent branch costs an
     * exec-mask save/restore and register moves per pair */
    if constexpr (STRIDE < E) {
#pragma unroll
        for (int r = 0; r < E; ++r) {
            if ((r & STRIDE) == 0) {
                const int a = r, b = r + STRIDE;
                const bool d = (SIZE == M) ? false : (((le + a) & SIZE) != 0);
                const float ka = key[a], kb = key[b];
                const int ia = idx[a], ib = idx[b];
                const bool sw = (ka > kb) == d;
                key[a] = sw ? kb : ka;
                key[b] = sw ? ka : kb;
                idx[a] = sw ? ib : ia;
                idx[b] = sw ? ia : ib;
            }
        }
    } else if constexpr (STRIDE / E == 16 || STRIDE / E == 32) {
        /* the pair's lower and upper keys straight from the permlane swap; the
         * lower lane keeps the upper key iff it swaps, the upper lane iff not */
        constexpr int MASK = STRIDE / E;
        const bool lower = (laneg & MASK) == 0;
#pragma unroll
        for (int r = 0; r < E; ++r) {
            int kl_i, ku_i, il, iu;
            esth_xor_halves<MASK>(__builtin_bit_cast(int, key[r]), kl_i, ku_i);
            esth_xor_halves<MASK>(idx[r], il, iu);
            const float kl = __builtin_bit_cast(float, kl_i), ku = __builtin_bit_cast(float, ku_i);
            const bool d = (SIZE == M) ? false : (((le + r) & SIZE) != 0);
            const bool sw = (kl > ku) == d;
            const bool take_u = sw == lower;
            key[r] = take_u ? ku : kl;
            idx[r] = take_u ? iu : il;
        }
    } else {
        constexpr int MASK = STRIDE / E;
        const bool lower = (laneg & MASK) == 0;
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const float other = esth_xor_lane_f<MASK>(key[r]);
            const int oidx = esth_xor_lane<MASK>(idx[r]);
            const bool d = (SIZE == M) ? false : (((le + r) & SIZE) != 0);
            const float kl = lower ? key[r] : other;
            const float ku = lower ? other : key[r];
            const bool sw = (kl > ku) == d;
            key[r] = sw ? other : key[r];
            idx[r] = sw ? oidx : idx[r];
        }
    }
}

template <int LOGM, int SIZE, int STRIDE, int LANES = esth_tile<LOGM>::LPS>
__device__ __forceinline__ void esth_bitonic_strides(float (&key)[(esth_tile<LOGM, LANES>::E)],
                                                     int (&idx)[(esth_tile<LOGM, LANES>::E)], int le, int laneg) {
    if constexpr (STRIDE > 0) {
        esth_bitonic_stage<LOGM, SIZE, STRIDE, LANES>(key, idx, le, laneg);
        esth_bitonic_strides<LOGM, SIZE, STRIDE / 2, LANES>(key, idx, le, laneg);
    }
}

template <int LOGM, int SIZE, int LANES = esth_tile<LOGM>::LPS>
__device__ __forceinline__ void esth_bitonic_sizes(float (&key)[(esth_tile<LOGM, LANES>::E)],
                                                   int (&idx)[(esth_tile<LOGM, LANES>::E)], int le, int laneg) {
    if constexpr (SIZE <= (1 << LOGM)) {
        esth_bitonic_strides<LOGM, SIZE, SIZE / 2, LANES>(key, idx, le, laneg);
        esth_bitonic_sizes<LOGM, 2 * SIZE, LANES>(key, idx, le, laneg);
    }
}

/* the network over one sub-filter held by LANES lanes, E positions each
 * (lane-major: lane l holds positions l E .. l E + E - 1); the same
 * compare-exchanges on the same positions whatever LANES is */
template <int LOGM, int LANES = esth_tile<LOGM>::LPS>
__device__ __forceinline__ void esth_bitonic(float (&key)[(esth_tile<LOGM, LANES>::E)],
                                             int (&idx)[(esth_tile<LOGM, LANES>::E)], int le) {
    esth_bitonic_sizes<LOGM, 2, LANES>(key, idx, le, le / esth_tile<LOGM, LANES>::E);
}

/* ============================================================== argmax
 * reduce (reduce.cl:19-55), one work-group of 128 work-items: a strided pass
 * (the later element wins ties), then a tree (`other` wins ties).  Here the
 * 128 work-items are the first two waves of a wider group; each walks its
 * strided elements in order, loading them 16 at a time so the loads overlap.
 * The whole group then combines the sub-filters' posterior partial sums in
 * double and copies the best particle (rank 0 of the winning sub-filter,
 * read before any exchange).
 * est (doubles): [0] max log-weight, [1] sum w, [2..2+S) sum w x, [2+S] best w,
 * [3+S] best g, [4+S..4+2S) best particle. */
template <int S, int NT>
__device__ __forceinline__ void esth_argmax_block(const float2* __restrict__ winner, uint32_t G, int log_w,
                                                  const float* __restrict__ part,
                                                  const float* __restrict__ best_src, size_t plane_stride,
                                                  uint32_t g_stride, float2* __restrict__ result,
                                                  double* __restrict__ est, float2* sc, float* smax, double* sd) {
    constexpr int NW = NT / 64;
    const int l = threadIdx.x;
    const int lane = l & 63, wid = l >> 6;
    if (l < 128) {
        float2 acc = make_float2(log_w ? -INFINITY : -1.0f, -1.0f);
        for (uint32_t g0 = (uint32_t)l; g0 < G; g0 += 128u * 16u) {
            float2 e[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                uint32_t gi = g0 + 128u * (uint32_t)j;
                e[j] = gi < G ? winner[gi] : make_float2(0.0f, 0.0f);
            }
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (g0 + 128u * (uint32_t)j < G) acc = (acc.x > e[j].x) ? acc : e[j];
        }
        sc[l] = acc;
    }
    float bmax = -INFINITY;
    if (log_w) {   /* 8 loads in flight per work-item */
        for (uint32_t g0 = (uint32_t)l; g0 < G; g0 += 8u * (uint32_t)NT) {
            float pm[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t gi = g0 + (uint32_t)(j * NT);
                pm[j] = gi < G ? part[(size_t)gi * (2 + S)] : -INFINITY;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) bmax = pm[j] > bmax ? pm[j] : bmax;
        }
        bmax = esth_bfly_max<1, 64>(bmax);
        if (lane == 0) smax[wid] = bmax;
    }
    __syncthreads();
    if (l < 64) {   /* the reference tree over the 128 strided results */
        for (int off = 64; off > 0; off >>= 1) {
            if (l < off) {
                float2 other = sc[l + off];
                float2 mine = sc[l];
                sc[l] = (mine.x > other.x) ? mine : other;
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        }
    }
    float gmax = 0.0f;
    if (log_w) {
        gmax = smax[0];
        for (int i = 1; i < NW; ++i) gmax = smax[i] > gmax ? smax[i] : gmax;
        gmax = esth_max_shift(gmax);
    }
    double a[2 + S];
#pragma unroll
    for (int c = 0; c < 2 + S; ++c) a[c] = 0.0;
    /* the same per-work-item order of sub-filters, U of them loaded at once */
    constexpr int U = (2 + S) <= 5 ? 4 : 2;
    for (uint32_t g0 = (uint32_t)l; g0 < G; g0 += (uint32_t)U * (uint32_t)NT) {
        float pg[U][2 + S];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const uint32_t gi = g0 + (uint32_t)(j * NT);
#pragma unroll
            for (int c = 0; c < 2 + S; ++c) pg[j][c] = gi < G ? part[(size_t)gi * (2 + S) + c] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < U; ++j) {
            if (g0 + (uint32_t)(j * NT) >= G) break;
            const double scale = log_w ? (double)esth_expf(pg[j][0] - gmax) : 1.0;
#pragma unroll
            for (int c = 1; c < 2 + S; ++c) a[c] += (double)pg[j][c] * scale;
        }
    }
#pragma unroll
    for (int c = 1; c < 2 + S; ++c) {
        double v = a[c];
        for (int k = 32; k > 0; k >>= 1) v += esth_shfl_xor_d(v, k);
        if (lane == 0) sd[wid * (2 + S) + c] = v;
    }
    __syncthreads();
    if (l == 0) {
        result[0] = sc[0];
        est[0] = (double)gmax;
#pragma unroll
        for (int c = 1; c < 2 + S; ++c) {
            double v = 0.0;
            for (int i = 0; i < NW; ++i) v += sd[i * (2 + S) + c];
            est[c] = v;
        }
        est[2 + S] = (double)sc[0].x;
        est[3 + S] = (double)sc[0].y;
    }
    if (best_src && l < S) {   /* rank 0 of the winning sub-filter, component l */
        int gb = (int)sc[0].y;
        if (gb >= 0 && (uint32_t)gb < G) est[4 + S + l] = (double)best_src[(size_t)l * plane_stride + (size_t)gb * g_stride];
    }
}

template <int S>
__global__ __launch_bounds__(1024) void k_argmax(const float2* __restrict__ winner, uint32_t G,
                                                 int log_w, const float* __restrict__ part,
                                                 const float* __restrict__ sorted, uint32_t N,
                                                 uint32_t m, float2* __restrict__ result,
                                                 double* __restrict__ est) {
    __shared__ float2 sc[128];
    __shared__ float smax[16];
    __shared__ double sd[16 * (2 + S)];
    esth_argmax_block<S, 1024>(winner, G, log_w, part, sorted, N, m, result, est, sc, smax, sd);
}

/* Sorted weights in place, gathered state to `sorted`, winner (w[0], g),
 * compact best t, and the posterior partial sums of the sub-filter.  (The
 * argmax stays a separate launch: folding it into the last work-group needs an
 * agent-scope release per work-group, which on the 8-XCD part writes back the
 * L2 and costs more than the launch it saves.) */
/* One sub-filter (one wave, or 64/m sub-filters per wave): sort its weights
 * (block_sort, pf_generic.cl:252-291) and write the sorted weights, the
 * gathered state, the compact best t, the winner and the posterior partials. */
/* E consecutive particles i0.. of plane c in a state of any layout: SOA
 * through esth_load_run's vector loads, AOS / PACKED one float each at the
 * layout's particle stride (S, or the plane's group width) */
template <int E, int S>
__device__ __forceinline__ void esth_load_plane_run(const float* __restrict__ st, int layout, size_t N, int c,
                                                    size_t i0, float (&out)[E]) {
    if (layout == ESTH_LAYOUT_SOA) {
        esth_load_run<E>(st + (size_t)c * N + i0, out);
        return;
    }
    const size_t b = esth_state_index(layout, S, N, c, i0);
    const size_t stride = esth_state_index(layout, S, N, c, i0 + 1) - b;
#pragma unroll
    for (int r = 0; r < E; ++r) out[r] = st[b + (size_t)r * stride];
}

template <int E, int S>
__device__ __forceinline__ void esth_store_plane_run(float* __restrict__ st, int layout, size_t N, int c, size_t i0,
                                                     const float (&v)[E]) {
    if (layout == ESTH_LAYOUT_SOA) {
        esth_store_run<E>(st + (size_t)c * N + i0, v);
        return;
    }
    const size_t b = esth_state_index(layout, S, N, c, i0);
    const size_t stride = esth_state_index(layout, S, N, c, i0 + 1) - b;
#pragma unroll
    for (int r = 0; r < E; ++r) st[b + (size_t)r * stride] = v[r];
}

/* FROM_LDS: the sub-filter's weights and planes are already in LDS at `src`
 * (src[0..M) the weights, src[(1 + c) M ..) plane c, one sub-filter per
 * wave): the keys are read from there and the gather permutes there, with
 * no global reads */
/* v[r] to particle i0 + r * stride of plane c (a strided run: with stride
 * 64, each r is one coalesced store across the wave) */
template <int E, int S>
__device__ __forceinline__ void esth_store_plane_strided(float* __restrict__ st, int layout, size_t N, int c,
                                                         size_t i0, int stride, const float (&v)[E]) {
    if (layout == ESTH_LAYOUT_SOA) {
        float* p = st + (size_t)c * N + i0;
#pragma unroll
        for (int r = 0; r < E; ++r) p[(size_t)r * stride] = v[r];
        return;
    }
#pragma unroll
    for (int r = 0; r < E; ++r) st[esth_state_index(layout, S, N, c, i0 + (size_t)r * stride)] = v[r];
}

template <int LOGM, int S, bool HOIST = false, bool FROM_LDS = false, int LANES = esth_tile<LOGM>::LPS>
__device__ __forceinline__ void esth_sort_sub(float* __restrict__ w, const float* __restrict__ st_in,
                                              float* __restrict__ st_out, uint32_t N, uint32_t G, uint32_t g,
                                              bool valid, float2* __restrict__ winner, int* __restrict__ perm,
                                              uint32_t t, float* __restrict__ bt_w, float* __restrict__ bt_s,
                                              float* __restrict__ part, int log_w, float* stage_w,
                                              int layout = ESTH_LAYOUT_SOA, const float* src = nullptr) {
    using TL = esth_tile<LOGM, LANES>;
    constexpr int M = TL::M, E = TL::E, LPS = TL::LPS;
    static_assert(!FROM_LDS || M >= ESTH_WAVE, "an LDS-resident sub-filter of at least one wave's positions");
    const int lane = threadIdx.x & (ESTH_WAVE - 1);
    const int le = (lane % LPS) * E;
    float* stage = FROM_LDS ? nullptr : stage_w + (lane - lane % LPS) * E;   /* this sub-filter's M floats */
    const size_t base = (size_t)g * M;
    float key[E];
    int idx[E];
    if constexpr (FROM_LDS)
        esth_load_run<E>(src + le, key);
    else
        esth_load_run<E>(w + base + le, key);
    /* the particles' planes do not depend on the order: with HOIST, when
     * they fit in registers, their loads are issued here and land while the
     * network runs (a gain for many small sub-filters; on a few large ones
     * the early burst delays the other waves' keys, so there they follow the
     * network) */
    constexpr int PB = FROM_LDS ? 1 : (S * E <= 64) ? S : 1;
    constexpr bool EARLY = HOIST && !FROM_LDS && PB == S;
    float in[PB][E];
    if constexpr (EARLY) {
#pragma unroll
        for (int j = 0; j < S; ++j) esth_load_plane_run<E, S>(st_in, layout, N, j, base + le, in[j]);
    }
#pragma unroll
    for (int r = 0; r < E; ++r) idx[r] = le + r;
    esth_bitonic<LOGM, LANES>(key, idx, le);
    /* block max M_g = key of position 0 */
    const int lead = lane - (lane % LPS);
    const float mx = esth_shfl_idx(key[0], lead);
    /* the partial sums follow the one-wave tiling's order whatever LANES is:
     * each run of ED positions (that tiling's positions per lane) is chained
     * in a lane, the runs then join pairwise as the lanes' butterfly does */
    constexpr int ED = esth_tile<LOGM>::E, K = E / ED;
    static_assert(E % ED == 0, "a narrower tiling holds whole runs of the one-wave tiling");
    float wt[E];
    float swg[K];
#pragma unroll
    for (int q = 0; q < K; ++q) swg[q] = 0.0f;
#pragma unroll
    for (int r = 0; r < E; ++r) {
        wt[r] = log_w ? esth_expf(key[r] - esth_max_shift(mx)) : key[r];
        swg[r / ED] = swg[r / ED] + wt[r];
    }
    float sw = esth_tree_sum<K>(swg);
    float sx[S];
    if (valid) {
#pragma unroll
        for (int r = 0; r < E; ++r) w[base + le + r] = key[r];
        if (perm) {
#pragma unroll
            for (int r = 0; r < E; ++r) perm[base + le + r] = idx[r];
        }
    }
    /* gather (particle_set of block_sort, pf_generic.cl:281-283) and partial
     * sums; each lane's E consecutive outputs leave as vector stores; unless
     * loaded early, the planes' loads are issued together, PB at a time */
#pragma unroll
    for (int c = 0; c < S; ++c) {
        float* dst = st_out + (size_t)c * N + base;
        float accg[K];
#pragma unroll
        for (int q = 0; q < K; ++q) accg[q] = 0.0f;
        float v[E];
        if constexpr (FROM_LDS) {
#pragma unroll
            for (int r = 0; r < E; ++r) v[r] = src[(1 + c) * M + idx[r]];
        } else {
            if (!EARLY && c % PB == 0) {
#pragma unroll
                for (int j = 0; j < PB; ++j) esth_load_plane_run<E, S>(st_in, layout, N, c + j, base + le, in[j]);
            }
            esth_permute_lds<M, E>(in[c % PB], le, stage, idx, v);
        }
#pragma unroll
        for (int r = 0; r < E; ++r) accg[r / ED] = esth_fma(wt[r], v[r], accg[r / ED]);
        const float acc = esth_tree_sum<K>(accg);
        if (valid) esth_store_run<E>(dst + le, v);
        if (valid && t > 0) {
#pragma unroll
            for (int r = 0; r < E; ++r)
                if ((uint32_t)(le + r) < t) bt_s[(size_t)c * G * t + (size_t)g * t + le + r] = v[r];
        }
        sx[c] = acc;
    }
    if (valid && t > 0) {
#pragma unroll
        for (int r = 0; r < E; ++r)
            if ((uint32_t)(le + r) < t) bt_w[(size_t)g * t + le + r] = key[r];
    }
    sw = esth_bfly_sum<1, LPS>(sw);
#pragma unroll
    for (int c = 0; c < S; ++c) sx[c] = esth_bfly_sum<1, LPS>(sx[c]);
    if (valid && le == 0) {
        winner[g] = make_float2(key[0], (float)g);
        float* pg = part + (size_t)g * (2 + S);
        pg[0] = log_w ? mx : 0.0f;
        pg[1] = sw;
#pragma unroll
        for (int c = 0; c < S; ++c) pg[2 + c] = sx[c];
    }
}

template <int LOGM, int S>
__global__ __launch_bounds__(256) void k_sort(float* __restrict__ w, const float* __restrict__ st_in,
                                              float* __restrict__ st_out, uint32_t N, uint32_t G,
                                              float2* __restrict__ winner, int* __restrict__ perm,
                                              uint32_t t, float* __restrict__ bt_w,
                                              float* __restrict__ bt_s, float* __restrict__ part,
                                              int log_w, int layout) {
    constexpr int LPS = esth_tile<LOGM>::LPS;
    const int lane = threadIdx.x & (ESTH_WAVE - 1);
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / ESTH_WAVE;
    uint32_t g = wave * (ESTH_WAVE / LPS) + (uint32_t)(lane / LPS);
    const bool valid = g < G;   /* a whole sub-filter is valid or not */
    if (!valid) g = G - 1;
    __shared__ float stage[4][ESTH_WAVE * esth_tile<LOGM>::E];
    esth_sort_sub<LOGM, S>(w, st_in, st_out, N, G, g, valid, winner, perm, t, bt_w, bt_s, part, log_w,
                           stage[threadIdx.x / ESTH_WAVE], layout);
}

/* waves per work-group of k_sample_sort (W = min(M / 128, 4)): each thread
 * samples M / (128 W) pairs.  Four waves beat one or two at m = 512 (90 µs
 * against 93 and 95 for the 2M-particle unicycle): more sampling waves
 * resident around each sorting one. */
template <int LOGM>
__host__ __device__ constexpr int esth_ss_waves() {
    return (1 << LOGM) / 128 < 4 ? (1 << LOGM) / 128 : 4;
}

/* sample+weight and block sort of one sub-filter in one work-group of
 * esth_ss_waves waves: each thread samples and weights its pairs (as
 * k_sample_weight, same uniforms, same arithmetic) into the work-group's LDS,
 * then wave 0
 * sorts the sub-filter there (esth_sort_sub, FROM_LDS) and writes what
 * k_sort writes.  The propagated unsorted state is read only by the sort,
 * except in a sub-filter the resampling will skip without carried weights,
 * which keeps it (pf_generic.cl:403): it is stored only for those, known
 * here from the gate's uniform (gate[g] - 1 > keep_thr) or, for the
 * data-dependent alias gate, always (keep_all).  The sort's wave runs beside
 * other work-groups' sampling instead of after all of them. */
template <int MODEL, int L, int LOGM, bool FA>
__global__ __launch_bounds__(ESTH_WAVE * esth_ss_waves<LOGM>()) void k_sample_sort(float* __restrict__ state, uint32_t N,
                                                                 uint32_t G, const float* __restrict__ R,
                                                                 const float* __restrict__ frame_ptr,
                                                                 const esth_frame_arg fa, int log_w,
                                                                 float* __restrict__ w, int carry,
                                                                 float* __restrict__ st_out,
                                                                 float2* __restrict__ winner, uint32_t t,
                                                                 float* __restrict__ bt_w,
                                                                 float* __restrict__ bt_s,
                                                                 float* __restrict__ part,
                                                                 const float* __restrict__ gate, float keep_thr,
                                                                 int keep_all, const esth_mtgp_inline mi) {
    using T = esth_model_traits<MODEL>;
    constexpr int M = 1 << LOGM, S = T::S;
    static_assert(LOGM >= 7 && LOGM <= 10, "one work-group per sub-filter of 128..1024 particles");
    const float* __restrict__ frame = FA ? fa.v : frame_ptr;
    __shared__ __align__(16) float sub[(1 + S) * M];   /* weights, then plane c at (1 + c) M */
    constexpr int NT = ESTH_WAVE * esth_ss_waves<LOGM>();
    /* the next step's MTGP32 work-groups come first (the launcher adds them
     * only where a work-group has 256 threads and the LDS to hold a ring) */
    constexpr bool MT = NT == ESTH_MTGP_THREADS && (1 + S) * M >= ESTH_MTGP_LARGE + 32;
    if constexpr (MT) {
        if (mi.n && esth_mtgp_inline_wg(mi, reinterpret_cast<uint32_t*>(sub))) return;
    }
    const uint32_t g = blockIdx.x - (MT ? (uint32_t)mi.n : 0u);
    if (g >= G) return;
    const bool keep = keep_all || (gate[g] - 1.0f) > keep_thr;
#pragma unroll 1
    for (int q = 2 * (int)threadIdx.x; q < M; q += 2 * NT) {
        const uint32_t p = g * (uint32_t)M + (uint32_t)q;
        float s0[S], s1[S], u0[T::V], u1[T::V];
        esth_load_pair<S, L>(state, N, p, s0, s1);
#pragma unroll
        for (int k = 0; k < T::V; ++k) {
            const float2 u = *reinterpret_cast<const float2*>(R + (size_t)k * N + p);
            u0[k] = u.x;
            u1[k] = u.y;
        }
        float lw0, lw1;
        esth_sample_pair_regs<MODEL>(s0, s1, u0, u1, frame, lw0, lw1);
        if (keep) esth_store_pair<S, L>(state, N, p, s0, s1);
        float2 wo = log_w ? make_float2(lw0, lw1) : make_float2(esth_expf(lw0), esth_expf(lw1));
        if (carry) {
            const float2 wp = *reinterpret_cast<const float2*>(w + p);
            wo = log_w ? make_float2(wp.x + wo.x, wp.y + wo.y) : make_float2(wp.x * wo.x, wp.y * wo.y);
        }
        *reinterpret_cast<float2*>(sub + q) = wo;
#pragma unroll
        for (int c = 0; c < S; ++c) *reinterpret_cast<float2*>(sub + (1 + c) * M + q) = make_float2(s0[c], s1[c]);
    }
    __syncthreads();
    if (threadIdx.x >= ESTH_WAVE) return;
    esth_sort_sub<LOGM, S, false, true>(w, nullptr, st_out, N, G, g, true, winner, nullptr, t, bt_w, bt_s, part,
                                        log_w, nullptr, ESTH_LAYOUT_SOA, sub);
}

/* ============================================================= exchange */
/* compact best-t index feeding received slot (q, k) of sub-filter b */
__device__ __forceinline__ size_t esth_recv_src(uint32_t b, uint32_t q, uint32_t k, uint32_t G, uint32_t t, int topo,
                                                int dimx, int dimy, const uint32_t* __restrict__ top) {
    if (topo == ESTH_TOPO_RING) return (size_t)((b + G - 1) % G) * t + k;
    if (topo == ESTH_TOPO_ALL2ALL) return top[k];
    int x = (int)(b % (uint32_t)dimx), y = (int)(b / (uint32_t)dimx);
    int sx = x, sy = y;
    if (q == 0) sx = (x + 1) % dimx;
    else if (q == 1) sx = (x - 1 + dimx) % dimx;
    else if (q == 2) sy = (y + 1) % dimy;
    else sy = (y - 1 + dimy) % dimy;
    return (size_t)(sy * dimx + sx) * t + k;
}

/*
 * Receiver-side gather from the compact best-t buffer, so every read sees
 * the pre-exchange sorted state whatever the launch order:
 *  ring  : slot m-1-k <- block (b-1) mod G rank k (pf_generic.cl:368-383)
 *  torus : slot m-1-k-q*t <- rank k of the right (q=0), left (1), lower (2),
 *          upper (3) neighbour: the receiving side of exchange_2dtorus
 *          (cuda/pf_generic.cu:352-393) with in-block offsets
 *  all2all: slot m-1-k <- global top-t candidate k */
static __global__ void k_exchange(float* __restrict__ st, float* __restrict__ w, int S, uint32_t N,
                           uint32_t m, uint32_t G, uint32_t t, int topo, int dimx, int dimy,
                           const float* __restrict__ bt_w, const float* __restrict__ bt_s,
                           const uint32_t* __restrict__ top) {
    const uint32_t nq = (topo == ESTH_TOPO_TORUS) ? 4u : 1u;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = (uint64_t)G * nq * t;
    if (i >= total) return;
    const uint32_t k = (uint32_t)(i % t);
    const uint32_t q = (uint32_t)((i / t) % nq);
    const uint32_t b = (uint32_t)(i / ((uint64_t)t * nq));
    const size_t src = esth_recv_src(b, q, k, G, t, topo, dimx, dimy, top);
    const size_t dst = (size_t)b * m + (m - 1 - k - q * t);
    w[dst] = bt_w[src];
    const size_t Gt = (size_t)G * t;
    for (int c = 0; c < S; ++c) st[(size_t)c * N + dst] = bt_s[(size_t)c * Gt + src];
}

/* Global top-t of the G*t candidates by (weight desc, index asc): t rounds,
 * each selecting the best candidate strictly after the previous one. */
static __global__ __launch_bounds__(1024) void k_topt(const float* __restrict__ bt_w, uint32_t n, uint32_t t,
                                               uint32_t* __restrict__ top) {
    __shared__ float sw[1024];
    __shared__ uint32_t si[1024];
    const int l = threadIdx.x;
    float pw = INFINITY;
    uint32_t pi = 0;
    bool have_prev = false;
    for (uint32_t k = 0; k < t; ++k) {
        float bw = -INFINITY;
        uint32_t bi = 0xffffffffu;
        for (uint32_t c = (uint32_t)l; c < n; c += 1024) {
            float v = bt_w[c];
            bool after = !have_prev || (v < pw) || (v == pw && c > pi);
            bool better = (bi == 0xffffffffu) || (v > bw) || (v == bw && c < bi);
            if (after && better) { bw = v; bi = c; }
        }
        sw[l] = bw;
        si[l] = bi;
        __syncthreads();
        for (int off = 512; off > 0; off >>= 1) {
            if (l < off) {
                float ow = sw[l + off];
                uint32_t oi = si[l + off];
                bool take = (oi != 0xffffffffu) &&
                            ((si[l] == 0xffffffffu) || ow > sw[l] || (ow == sw[l] && oi < si[l]));
                if (take) { sw[l] = ow; si[l] = oi; }
            }
            __syncthreads();
        }
        pw = sw[0];
        pi = si[0];
        have_prev = true;
        if (l == 0) top[k] = pi;
        __syncthreads();
    }
}

/* cross-lane levels of the scan tree: lx[k] = sibling node at level k */
template <int K, int LOGL>
__device__ __forceinline__ void esth_scan_levels(float& node, float* lx) {
    if constexpr (K < LOGL) {
        float sib = esth_xor_lane_f<(1 << K)>(node);
        lx[K] = sib;
        node = node + sib;
        esth_scan_levels<K + 1, LOGL>(node, lx);
    }
}

/* ============================================ resampling (prefix sum)
 * resampling (pf_generic.cl:385-432): gate, exclusive scan of the sub-filter's
 * weights, selection, binary search, gather.  The scan reproduces the
 * Blelloch tree's additions exactly: the up-sweep node over [j 2^k, (j+1) 2^k)
 * is the sum of its two halves, and the exclusive prefix of element e is
 * 0 + L_top + ... + L_0 taken top-down over the levels k where bit k of e is
 * set, L_k being the left sibling's node.  Levels inside a lane come from a
 * register tree, levels across lanes from xor butterflies (a + b == b + a in
 * IEEE arithmetic, so both lanes of a pair hold the same node bits). */
/* The receiving side of the exchange, applied while the resampling reads its
 * sub-filter: slot m-1-k-q*t takes rank k of sender q's compact best t
 * (esth_recv_src), in the weights the scan sees, in the states the gather
 * draws from, and in the stored sorted weights (w_out).  The exchanged
 * states are only ever read by the resampling, so they are not stored. */
struct esth_xch {
    const float* bt_w;
    const float* bt_s;
    const uint32_t* top;
    float* w_out;
    uint32_t G, t, nq;
    int topo, dimx, dimy;
};

/* one sub-filter per wave (or 64/m per wave); every wave of the work-group
 * must call it (one barrier inside); cdf_w: this wave's 64*E floats of LDS;
 * recv: apply the exchange xch on the way in */
template <int LOGM, int S, bool HOIST = false, int LANES = esth_tile<LOGM>::LPS>
__device__ __forceinline__ void esth_resample_sub(float* __restrict__ st_main,
                                                  const float* __restrict__ st_tmp,
                                                  const float* w,   /* xch.w_out may write it */
                                                  const float* __restrict__ rnd, uint32_t N, uint32_t g,
                                                  bool valid, float thr, int systematic, int log_w,
                                                  int* __restrict__ sel, float* cdf_w,
                                                  bool recv = false, esth_xch xch = esth_xch{},
                                                  int carry = 0, float* w_mut = nullptr,
                                                  int layout = ESTH_LAYOUT_SOA) {
    using TL = esth_tile<LOGM, LANES>;
    constexpr int M = TL::M, E = TL::E, LPS = TL::LPS, LOGE = TL::LOGE, LOGL = TL::LOGL;
    const int lane = threadIdx.x & (ESTH_WAVE - 1);
    const int laneg = lane % LPS;
    const int le = laneg * E;
    const size_t base = (size_t)g * M;
    const bool go = valid && !((rnd[N + g] - 1.0f) > thr);   /* gate, pf_generic.cl:403 */
    int rsrc[E];   /* compact best-t index of a received slot, or -1 */
#pragma unroll
    for (int r = 0; r < E; ++r) rsrc[r] = -1;
    recv = recv && xch.t > 0;
    if (recv) {
        const uint32_t t = xch.t, nrecv = xch.nq * t;
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const uint32_t pos = (uint32_t)(le + r);
            if (pos >= (uint32_t)M - nrecv) {
                const uint32_t j = (uint32_t)M - 1u - pos, q = j / t, k = j - q * t;
                rsrc[r] = (int)esth_recv_src(g, q, k, xch.G, t, xch.topo, xch.dimx, xch.dimy, xch.top);
            }
        }
    }
    float x[E];
    esth_load_run<E>(w + base + le, x);
    /* the selection's uniforms and, with HOIST when they fit in registers,
     * the sorted particles' planes do not depend on the scan: their loads
     * are issued here and land while it runs */
    constexpr int PB = (S * E <= 64) ? S : 1;
    constexpr bool EARLY = HOIST && PB == S;
    float inh[PB][E];
    if constexpr (EARLY) {
#pragma unroll
        for (int c = 0; c < S; ++c) esth_load_run<E>(st_tmp + (size_t)c * N + base + le, inh[c]);
    }
    /* output r of this lane is particle laneg + LPS r of the sub-filter:
     * lanes take consecutive outputs, so at every level of the binary
     * searches the wave's probes fall on consecutive (or equal) CDF words,
     * not E words apart, and the stores of each r are one coalesced run */
    float uu[E];
    if (!systematic) {
#pragma unroll
        for (int r = 0; r < E; ++r) uu[r] = rnd[base + laneg + LPS * r];
    }
    if (recv) {
#pragma unroll
        for (int r = 0; r < E; ++r)
            if (rsrc[r] >= 0) {
                x[r] = xch.bt_w[rsrc[r]];
                if (valid) xch.w_out[base + le + r] = x[r];
            }
        if constexpr (EARLY) {
            const size_t Gt = (size_t)xch.G * xch.t;
#pragma unroll
            for (int c = 0; c < S; ++c)
#pragma unroll
                for (int r = 0; r < E; ++r)
                    if (rsrc[r] >= 0) inh[c][r] = xch.bt_s[(size_t)c * Gt + (size_t)rsrc[r]];
        }
    }
    if (log_w) {
        float mx = x[0];
#pragma unroll
        for (int r = 1; r < E; ++r) mx = x[r] > mx ? x[r] : mx;
        mx = esth_bfly_max<1, LPS>(mx);
#pragma unroll
        for (int r = 0; r < E; ++r) x[r] = esth_expf(x[r] - esth_max_shift(mx));
    }
    /* register tree: lvl[k][j], k = 0..LOGE */
    float lvl[LOGE + 1][E];
#pragma unroll
    for (int r = 0; r < E; ++r) lvl[0][r] = x[r];
#pragma unroll
    for (int k = 1; k <= LOGE; ++k)
#pragma unroll
        for (int j = 0; j < (E >> k); ++j) lvl[k][j] = lvl[k - 1][2 * j] + lvl[k - 1][2 * j + 1];
    float node = lvl[LOGE][0];
    float lx[LOGL > 0 ? LOGL : 1];
    esth_scan_levels<0, LOGL>(node, lx);
    const float total = node;
    float acc = 0.0f;
#pragma unroll
    for (int k = LOGL - 1; k >= 0; --k)
        if ((laneg >> k) & 1) acc = acc + lx[k];
    float* cdf = cdf_w + (lane - laneg) * E;
#pragma unroll
    for (int r = 0; r < E; ++r) {
        float a = acc;
#pragma unroll
        for (int k = LOGE - 1; k >= 0; --k)
            if ((r >> k) & 1) a = a + lvl[k][(r >> k) - 1];
        cdf[le + r] = a;
    }
    __syncthreads();
    const float u0 = rnd[base] - 1.0f;
    const float step = total / (float)M;
    int pick[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        float s;
        if (systematic) s = ((float)(laneg + LPS * r) + u0) * step;
        else s = (uu[r] - 1.0f) * total;
        /* binary_search, pf_generic.cl:182-201: on a power-of-two range each
         * round halves it exactly, so its do-while runs LOGM rounds; unrolled,
         * the E searches of a lane interleave */
        int mn = 0, mxi = M - 1;
#pragma unroll
        for (int it = 0; it < LOGM; ++it) {
            const int cur = mn + (mxi - mn + 1) / 2;
            const bool left = cdf[cur] > s;
            mxi = left ? cur - 1 : mxi;
            mn = left ? mn : cur;
        }
        pick[r] = mn & (M - 1);
    }
    /* gather the selected particles through the cdf's LDS region, plane by
     * plane (every lane takes part; only resampled sub-filters store); unless
     * loaded early, the planes' loads are issued together, PB at a time */
#pragma unroll
    for (int c0 = 0; c0 < S; c0 += PB) {
        float in[PB][E];
        if constexpr (EARLY) {
#pragma unroll
            for (int c = 0; c < PB; ++c)
#pragma unroll
                for (int r = 0; r < E; ++r) in[c][r] = inh[c][r];
        } else {
#pragma unroll
            for (int c = 0; c < PB; ++c) esth_load_run<E>(st_tmp + (size_t)(c0 + c) * N + base + le, in[c]);
        }
        if (!EARLY && recv) {
            const size_t Gt = (size_t)xch.G * xch.t;
#pragma unroll
            for (int c = 0; c < PB; ++c)
#pragma unroll
                for (int r = 0; r < E; ++r)
                    if (rsrc[r] >= 0) in[c][r] = xch.bt_s[(size_t)(c0 + c) * Gt + (size_t)rsrc[r]];
        }
#pragma unroll
        for (int c = 0; c < PB; ++c) {
            float v[E];
            esth_permute_lds<M, E>(in[c], le, cdf, pick, v);
            if (go) esth_store_plane_strided<E, S>(st_main, layout, N, c0 + c, base + laneg, LPS, v);
            /* carried weights: a skipped sub-filter keeps its sorted, exchanged particles */
            else if (carry && valid) esth_store_plane_run<E, S>(st_main, layout, N, c0 + c, base + le, in[c]);
        }
    }
    if (carry && go) {   /* resampled: equal weights again */
        float w0[E];
#pragma unroll
        for (int r = 0; r < E; ++r) w0[r] = log_w ? 0.0f : 1.0f;
        esth_store_run<E>(w_mut + base + le, w0);
    }
    if (sel && valid) {
#pragma unroll
        for (int r = 0; r < E; ++r) sel[base + laneg + LPS * r] = go ? pick[r] : -1;
    }
}

template <int LOGM, int S>
__global__ __launch_bounds__(256) void k_resample(float* __restrict__ st_main,
                                                  const float* __restrict__ st_tmp,
                                                  const float* __restrict__ w,
                                                  const float* __restrict__ rnd, uint32_t N, uint32_t G,
                                                  float thr, int systematic, int log_w,
                                                  int* __restrict__ sel, int layout) {
    using TL = esth_tile<LOGM>;
    __shared__ float cdf_s[4][ESTH_WAVE * TL::E];
    const int lane = threadIdx.x & (ESTH_WAVE - 1);
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / ESTH_WAVE;
    uint32_t g = wave * (ESTH_WAVE / TL::LPS) + (uint32_t)(lane / TL::LPS);
    const bool valid = g < G;
    if (!valid) g = G - 1;   /* keep the lane busy; its writes are dropped */
    esth_resample_sub<LOGM, S>(st_main, st_tmp, w, rnd, N, g, valid, thr, systematic, log_w, sel,
                               cdf_s[threadIdx.x / ESTH_WAVE], false, esth_xch{}, 0, nullptr, layout);
}

/* The step's argmax + posterior (work-group 0) beside the scan resampling
 * (work-groups 1..): the two read disjoint data (the argmax: winners,
 * partials and the compact best-t copy of rank 0; the resampling: weights,
 * uniforms and the exchanged sorted state) and neither writes what the other
 * reads, so one launch does both and the argmax's single work-group runs in
 * the shadow of the resampling instead of on its own before the exchange. */
template <int LOGM, int S>
__global__ __launch_bounds__(256) void k_resample_argmax(float* __restrict__ st_main,
                                                         const float* __restrict__ st_tmp,
                                                         const float* w,
                                                         const float* __restrict__ rnd, uint32_t N, uint32_t G,
                                                         float thr, int systematic, int log_w,
                                                         const float2* __restrict__ winner,
                                                         const float* __restrict__ part,
                                                         const float* __restrict__ best_src, size_t plane_stride,
                                                         uint32_t g_stride, float2* __restrict__ result,
                                                         double* __restrict__ est, esth_xch xch, int carry,
                                                         int layout) {
    using TL = esth_tile<LOGM>;
    __shared__ float cdf_s[4][ESTH_WAVE * TL::E];
    __shared__ float2 sc[128];
    __shared__ float smax[4];
    __shared__ double sd[4 * (2 + S)];
    if (blockIdx.x == 0) {
        esth_argmax_block<S, 256>(winner, G, log_w, part, best_src, plane_stride, g_stride, result, est, sc, smax, sd);
        return;
    }
    const int lane = threadIdx.x & (ESTH_WAVE - 1);
    const uint32_t wave = ((blockIdx.x - 1) * blockDim.x + threadIdx.x) / ESTH_WAVE;
    uint32_t g = wave * (ESTH_WAVE / TL::LPS) + (uint32_t)(lane / TL::LPS);
    const bool valid = g < G;
    if (!valid) g = G - 1;
    esth_resample_sub<LOGM, S>(st_main, st_tmp, w, rnd, N, g, valid, thr, systematic, log_w, nullptr,
                               cdf_s[threadIdx.x / ESTH_WAVE], true, xch, carry, const_cast<float*>(w), layout);
}

/* the sum of the wave-level nodes n[a .. a + K), as the position tree forms
 * it: (n[a] + n[a+1]) + (n[a+2] + n[a+3]) for K = 4 */
template <int K>
__device__ __forceinline__ float esth_wave_node(const float* n, int a) {
    if constexpr (K == 1) return n[a];
    else return esth_wave_node<K / 2>(n, a) + esth_wave_node<K / 2>(n, a + K / 2);
}

/* The same resampling of one sub-filter by W waves (W = 2: m >= 256; W = 4:
 * m >= 512): wave h holds particles h M/W .. (h + 1) M/W - 1, E = M/(64 W)
 * per lane, lane-major.  Every value is the one-wave version's: the scan
 * tree is the same binary tree over positions (register levels, cross-lane
 * levels, then the wave levels, whose left nodes a wave adds top-down as the
 * high bits of its positions ask), the block max is a max, the search and the
 * selections are per output.  W times the waves per sub-filter cover W
 * times the latency.  cdf: the sub-filter's M floats of LDS; stage: S M
 * floats (the gather's planes); xl: 2 W floats (the waves' maxima and
 * totals).  All W waves (and every wave of the work-group) must call it:
 * three work-group barriers inside. */
template <int LOGM, int S, bool TO_LDS = false, int W = 2>
__device__ __forceinline__ void esth_resample_sub2(float* __restrict__ st_main, const float* __restrict__ st_tmp,
                                                   const float* w, const float* __restrict__ rnd, uint32_t N,
                                                   uint32_t g, bool valid, float thr, int systematic, int log_w,
                                                   float* cdf, float* stage, float* xl, int h, esth_xch xch,
                                                   int carry, float* w_mut, int layout, float* res = nullptr) {
    static_assert(W == 2 || W == 4, "two or four waves per sub-filter");
    constexpr int LOGW = W == 4 ? 2 : 1;
    constexpr int M = 1 << LOGM, E = M / (64 * W), LOGE = LOGM - 6 - LOGW;
    static_assert(LOGE >= 1, "at least two particles per lane");
    const int lane = threadIdx.x & (ESTH_WAVE - 1);
    const int le = h * (M / W) + lane * E;   /* first position of this lane */
    const size_t base = (size_t)g * M;
    const bool go = valid && !((rnd[N + g] - 1.0f) > thr);   /* gate, pf_generic.cl:403 */
    int rsrc[E];
#pragma unroll
    for (int r = 0; r < E; ++r) rsrc[r] = -1;
    const bool recv = xch.t > 0;
    if (recv) {
        const uint32_t t = xch.t, nrecv = xch.nq * t;
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const uint32_t pos = (uint32_t)(le + r);
            if (pos >= (uint32_t)M - nrecv) {
                const uint32_t j = (uint32_t)M - 1u - pos, q = j / t, k = j - q * t;
                rsrc[r] = (int)esth_recv_src(g, q, k, xch.G, t, xch.topo, xch.dimx, xch.dimy, xch.top);
            }
        }
    }
    float x[E];
    esth_load_run<E>(w + base + le, x);
    /* the sorted particles' planes, loaded here so that their latency runs
     * beside the block max and the scan instead of after them */
    float in[S][E];
#pragma unroll
    for (int c = 0; c < S; ++c) esth_load_run<E>(st_tmp + (size_t)c * N + base + le, in[c]);
    /* outputs: particle h M/W + lane + 64 r */
    const int ob = h * (M / W) + lane;
    float uu[E];
    if (!systematic) {
#pragma unroll
        for (int r = 0; r < E; ++r) uu[r] = rnd[base + ob + ESTH_WAVE * r];
    }
    if (recv) {
#pragma unroll
        for (int r = 0; r < E; ++r)
            if (rsrc[r] >= 0) {
                x[r] = xch.bt_w[rsrc[r]];
                if (valid) xch.w_out[base + le + r] = x[r];
            }
    }
    if (log_w) {
        float mx = x[0];
#pragma unroll
        for (int r = 1; r < E; ++r) mx = x[r] > mx ? x[r] : mx;
        mx = esth_bfly_max<1, ESTH_WAVE>(mx);
        if (lane == 0) xl[h] = mx;
        __syncthreads();
        mx = xl[0];
#pragma unroll
        for (int i = 1; i < W; ++i) mx = xl[i] > mx ? xl[i] : mx;
#pragma unroll
        for (int r = 0; r < E; ++r) x[r] = esth_expf(x[r] - esth_max_shift(mx));
    } else {
        __syncthreads();
    }
    float lvl[LOGE + 1][E];
#pragma unroll
    for (int r = 0; r < E; ++r) lvl[0][r] = x[r];
#pragma unroll
    for (int k = 1; k <= LOGE; ++k)
#pragma unroll
        for (int j = 0; j < (E >> k); ++j) lvl[k][j] = lvl[k - 1][2 * j] + lvl[k - 1][2 * j + 1];
    float node = lvl[LOGE][0];
    float lx[6];
    esth_scan_levels<0, 6>(node, lx);
    if (lane == 0) xl[W + h] = node;   /* the wave's node: [h M/W, (h + 1) M/W) */
    __syncthreads();
    const float* wn = xl + W;
    const float total = esth_wave_node<W>(wn, 0);
    float acc = 0.0f;
    if constexpr (W == 4) {
        if (h >= 2) acc = acc + esth_wave_node<2>(wn, 0);
    }
    if (h & 1) acc = acc + wn[h - 1];
#pragma unroll
    for (int k = 5; k >= 0; --k)
        if ((lane >> k) & 1) acc = acc + lx[k];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        float a = acc;
#pragma unroll
        for (int k = LOGE - 1; k >= 0; --k)
            if ((r >> k) & 1) a = a + lvl[k][(r >> k) - 1];
        cdf[le + r] = a;
    }
    /* the sorted particles (with the received ones), staged for the gather */
#pragma unroll
    for (int c = 0; c < S; ++c) {
        if (recv) {
            const size_t Gt = (size_t)xch.G * xch.t;
#pragma unroll
            for (int r = 0; r < E; ++r)
                if (rsrc[r] >= 0) in[c][r] = xch.bt_s[(size_t)c * Gt + (size_t)rsrc[r]];
        }
        esth_store_run<E>(stage + c * M + le, in[c]);
        /* carried weights: a skipped sub-filter keeps its sorted, exchanged particles */
        if (!go && carry && valid) esth_store_plane_run<E, S>(st_main, layout, N, c, base + le, in[c]);
    }
    __syncthreads();
    const float u0 = rnd[base] - 1.0f;
    const float step = total / (float)M;
    int pick[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        float sv;
        if (systematic) sv = ((float)(ob + ESTH_WAVE * r) + u0) * step;
        else sv = (uu[r] - 1.0f) * total;
        int mn = 0, mxi = M - 1;
#pragma unroll
        for (int it = 0; it < LOGM; ++it) {
            const int cur = mn + (mxi - mn + 1) / 2;
            const bool lft = cdf[cur] > sv;
            mxi = lft ? cur - 1 : mxi;
            mn = lft ? mn : cur;
        }
        pick[r] = mn & (M - 1);
    }
    if constexpr (TO_LDS) {
        /* the sub-filter's next particles into res (S planes of M): the
         * selections, or, for a sub-filter the gate skips, the propagated
         * state the previous launch kept for it in st_main */
#pragma unroll
        for (int c = 0; c < S; ++c)
#pragma unroll
            for (int r = 0; r < E; ++r)
                res[c * M + ob + ESTH_WAVE * r] =
                    go ? stage[c * M + pick[r]] : st_main[(size_t)c * N + base + ob + ESTH_WAVE * r];
    } else if (go) {
#pragma unroll
        for (int c = 0; c < S; ++c) {
            float v[E];
#pragma unroll
            for (int r = 0; r < E; ++r) v[r] = stage[c * M + pick[r]];
            esth_store_plane_strided<E, S>(st_main, layout, N, c, base + ob, ESTH_WAVE, v);
        }
    }
    if (carry && go) {   /* resampled: equal weights again */
        float w0[E];
#pragma unroll
        for (int r = 0; r < E; ++r) w0[r] = log_w ? 0.0f : 1.0f;
        esth_store_run<E>(w_mut + base + le, w0);
    }
}

/* k_resample_argmax with two waves per sub-filter (m >= 256): work-group 0
 * the argmax, each other work-group of four waves two sub-filters */
template <int LOGM, int S>
__global__ __launch_bounds__(256) void k_resample2_argmax(float* __restrict__ st_main,
                                                          const float* __restrict__ st_tmp, const float* w,
                                                          const float* __restrict__ rnd, uint32_t N, uint32_t G,
                                                          float thr, int systematic, int log_w,
                                                          const float2* __restrict__ winner,
                                                          const float* __restrict__ part,
                                                          const float* __restrict__ best_src, size_t plane_stride,
                                                          uint32_t g_stride, float2* __restrict__ result,
                                                          double* __restrict__ est, esth_xch xch, int carry,
                                                          int layout) {
    constexpr int M = 1 << LOGM;
    __shared__ __align__(16) float cdf_s[2][M];
    __shared__ __align__(16) float stage_s[2][S * M];
    __shared__ float xl_s[2][4];
    __shared__ float2 sc[128];
    __shared__ float smax[4];
    __shared__ double sd[4 * (2 + S)];
    if (blockIdx.x == 0) {
        esth_argmax_block<S, 256>(winner, G, log_w, part, best_src, plane_stride, g_stride, result, est, sc, smax, sd);
        return;
    }
    const int wv = threadIdx.x / ESTH_WAVE;   /* 0..3: sub-filter wv / 2, half wv % 2 */
    const int j = wv >> 1;
    uint32_t g = (blockIdx.x - 1) * 2u + (uint32_t)j;
    const bool valid = g < G;
    if (!valid) g = G - 1;
    esth_resample_sub2<LOGM, S>(st_main, st_tmp, w, rnd, N, g, valid, thr, systematic, log_w, cdf_s[j], stage_s[j],
                                xl_s[j], wv & 1, xch, carry, const_cast<float*>(w), layout);
}

/* One launch per step across the step boundary: work-group b resamples
 * sub-filter b of step k (all four waves, esth_resample_sub2 with W = 4 and
 * the particles kept in LDS), then samples, weights and sorts it for step k + 1
 * (four waves, as k_sample_sort) from there; work-group mi.n runs step k's
 * argmax and posterior, the first mi.n work-groups step k + 2's MTGP32 words.
 * Between the steps a resampled particle never leaves LDS.  Step k's
 * exchange reads the senders' best t of step k while this launch writes
 * step k + 1's into the other buffer of a pair (bt_*_k / bt_*_k1, and the
 * winners and partials the same way); w and tmp of sub-filter b are read and
 * then rewritten by its own work-group only.  For the scan resamplers on a
 * ring or torus with t > 0, SOA state and no carried weights. */
template <int MODEL, int LOGM, bool FA>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_step_x(
    float* tmp, float* w, const float* __restrict__ rnd_k, float thr, int systematic, const esth_xch xk,
    const float2* __restrict__ winner_k, const float* __restrict__ part_k, const float* __restrict__ best_k,
    float2* __restrict__ result, double* __restrict__ est, float* __restrict__ state, uint32_t N, uint32_t G,
    const float* __restrict__ R_k1, const float* __restrict__ frame_ptr, const esth_frame_arg fa, int log_w,
    float2* __restrict__ winner_k1, float* __restrict__ bt_w_k1, float* __restrict__ bt_s_k1,
    float* __restrict__ part_k1, const float* __restrict__ gate_k1, float keep_thr, const esth_mtgp_inline mi) {
    using T = esth_model_traits<MODEL>;
    constexpr int M = 1 << LOGM, S = T::S;
    static_assert(LOGM >= 9 && LOGM <= 10, "four resampling and sampling waves per sub-filter");
    /* res: S planes of M; then the resampling's cdf (M), staged planes (S M)
     * and xl (8), whose space the sampling's weights and planes ((1 + S) M)
     * take over */
    __shared__ __align__(16) float lds[S * M + (1 + S) * M + 8];
    __shared__ float2 sc[128];
    __shared__ float smax[4];
    __shared__ double sd[4 * (2 + S)];
    float* res = lds;
    float* cdf = lds + S * M;
    float* stage = cdf + M;
    float* xl = stage + S * M;
    float* sub = lds + S * M;
    if (mi.n && esth_mtgp_inline_wg(mi, reinterpret_cast<uint32_t*>(lds))) return;
    const uint32_t b0 = (uint32_t)mi.n;
    if (blockIdx.x == b0) {
        esth_argmax_block<S, 256>(winner_k, G, log_w, part_k, best_k, (size_t)G * xk.t, xk.t, result, est, sc, smax,
                                  sd);
        return;
    }
    const uint32_t g = blockIdx.x - b0 - 1u;
    if (g >= G) return;
    const int wv = threadIdx.x / ESTH_WAVE;
    esth_resample_sub2<LOGM, S, true, 4>(state, tmp, w, rnd_k, N, g, true, thr, systematic, log_w, cdf, stage, xl, wv,
                                         xk, 0, w, ESTH_LAYOUT_SOA, res);
    __syncthreads();   /* res complete; cdf and stage free */
    const float* __restrict__ frame = FA ? fa.v : frame_ptr;
    const bool keep = (gate_k1[g] - 1.0f) > keep_thr;
#pragma unroll 1
    for (int q = 2 * (int)threadIdx.x; q < M; q += 2 * 256) {
        const uint32_t p = g * (uint32_t)M + (uint32_t)q;
        float s0[S], s1[S], u0[T::V], u1[T::V];
#pragma unroll
        for (int c = 0; c < S; ++c) {
            const float2 v = *reinterpret_cast<const float2*>(res + c * M + q);
            s0[c] = v.x;
            s1[c] = v.y;
        }
#pragma unroll
        for (int k = 0; k < T::V; ++k) {
            const float2 u = *reinterpret_cast<const float2*>(R_k1 + (size_t)k * N + p);
            u0[k] = u.x;
            u1[k] = u.y;
        }
        float lw0, lw1;
        esth_sample_pair_regs<MODEL>(s0, s1, u0, u1, frame, lw0, lw1);
        if (keep) esth_store_pair<S, ESTH_LAYOUT_SOA>(state, N, p, s0, s1);
        const float2 wo = log_w ? make_float2(lw0, lw1) : make_float2(esth_expf(lw0), esth_expf(lw1));
        *reinterpret_cast<float2*>(sub + q) = wo;
#pragma unroll
        for (int c = 0; c < S; ++c) *reinterpret_cast<float2*>(sub + (1 + c) * M + q) = make_float2(s0[c], s1[c]);
    }
    __syncthreads();
    if (threadIdx.x >= ESTH_WAVE) return;
    esth_sort_sub<LOGM, S, false, true>(w, nullptr, tmp, N, G, g, true, winner_k1, nullptr, xk.t, bt_w_k1, bt_s_k1,
                                        part_k1, log_w, nullptr, ESTH_LAYOUT_SOA, sub);
}

/* =========================================================== Vose alias
 * resampling_valias (resampling-valias.cl:214-304): one work-group of m per
 * sub-filter.  Normalisation sums with sum_reduce4's exact tree (:85-104);
 * the worklists are filled in ascending work-item order (ballot ranks), the
 * canonical order of the reference's local atomics (:124-126,175-178). */
__device__ __forceinline__ float esth_sum_reduce4(float* arr, uint32_t n, int l) {
    uint32_t offset = n;
    while (offset > 2) {
        offset >>= 2;
        if ((uint32_t)l < offset)
            arr[l] += arr[l + offset] + arr[l + offset + offset] + arr[l + offset + offset + offset];
        __syncthreads();
    }
    if (offset > 1) {
        if (l == 0) arr[0] += arr[1];
        __syncthreads();
    }
    return arr[0];
}

/* rank of this lane among lanes with `pred` in ascending local id, and the
 * total, over the whole work-group (cnt: LDS scratch of 16 uint). */
__device__ __forceinline__ uint32_t esth_block_rank(bool pred, uint32_t* cnt, int l, int nw,
                                                    uint32_t* total) {
    const uint64_t bal = __ballot(pred);
    const int lane = l & 63, wid = l >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t r = (uint32_t)__popcll(bal & lt);
    if (lane == 0) cnt[wid] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t before = 0, tot = 0;
    for (int i = 0; i < nw; ++i) {
        uint32_t c = cnt[i];
        if (i < wid) before += c;
        tot += c;
    }
    __syncthreads();
    *total = tot;
    return before + r;
}

/* one work-group of M work-items on sub-filter g; carry: a skipped
 * sub-filter keeps its sorted, exchanged particles (and weights), a
 * resampled one restarts from equal weights */
template <int S>
__device__ __forceinline__ void esth_alias_sub(float* __restrict__ st_main, const float* __restrict__ st_tmp,
                                               const float* w, const float* __restrict__ rnd,
                                               uint32_t N, uint32_t M, uint32_t g, float thr, int adaptive,
                                               int log_w, int* __restrict__ sel, float* __restrict__ prob_out,
                                               uint32_t* __restrict__ alias_out, int carry = 0,
                                               int layout = ESTH_LAYOUT_SOA) {
    __shared__ float prob[ESTH_MAX_NUM_PARTICLES];
    __shared__ uint32_t alias[ESTH_MAX_NUM_PARTICLES];
    __shared__ uint32_t sl[ESTH_MAX_NUM_PARTICLES];
    __shared__ float scratch[ESTH_MAX_NUM_PARTICLES];
    __shared__ uint32_t cnt[16];
    __shared__ float red[16];
    const int l = threadIdx.x;
    const int nw = (int)((M + 63) / 64);
    const size_t idx = (size_t)g * M + (size_t)l;
    const bool table_only = prob_out != nullptr;
    if (!table_only && !adaptive) {
        if ((rnd[N + g] - 1.0f) > thr) {
            if (sel) sel[idx] = -1;
#pragma unroll
            for (int c = 0; c < S; ++c)
                if (carry) st_main[esth_state_index(layout, S, N, c, idx)] = st_tmp[(size_t)c * N + idx];
            return;
        }
    }
    float p = w[idx];
    if (log_w) {
        float mx = p;
        for (int k = 1; k < 64; k <<= 1) {
            float o = esth_shfl_xor(mx, k);
            if ((l ^ k) < (int)M) mx = o > mx ? o : mx;
        }
        if ((l & 63) == 0) red[l >> 6] = mx;
        __syncthreads();
        mx = red[0];
        for (int i = 1; i < nw; ++i) mx = red[i] > mx ? red[i] : mx;
        p = esth_expf(p - esth_max_shift(mx));
    }
    scratch[l] = p;
    __syncthreads();
    float sum = esth_sum_reduce4(scratch, M, l);
    if (sum > 0.0f) p = p / sum;
    else p = 1.0f / (float)M;
    prob[l] = p;
    __syncthreads();   /* every item has read the sum before scratch is reused */
    if (adaptive && !table_only) {
        scratch[l] = p * p;
        __syncthreads();
        float recip_ess = esth_sum_reduce4(scratch, M, l);
        if (!(recip_ess > thr)) {
            if (sel) sel[idx] = -1;
#pragma unroll
            for (int c = 0; c < S; ++c)
                if (carry) st_main[esth_state_index(layout, S, N, c, idx)] = st_tmp[(size_t)c * N + idx];
            return;
        }
    }
    /* vose_alias_init, resampling-valias.cl:111-195 */
    float pl = prob[l] * (float)M;
    prob[l] = pl;
    alias[l] = (uint32_t)l;
    uint32_t nsmall;
    const bool small = pl <= 1.0f;
    uint32_t rk = esth_block_rank(small, cnt, l, nw, &nsmall);
    uint32_t slot = small ? rk : (M - 1u - ((uint32_t)l - rk));
    sl[slot] = (uint32_t)l;
    uint32_t small_end = nsmall, small_begin = 0;
    /* Pairing rounds.  The pairs of a round never outnumber those of the
     * round before (a round of n = min(small, large) pairs leaves at most n of
     * the smaller side), so once a round fits in one wave, wave 0 runs every
     * remaining round alone with wave-level ordering instead of work-group
     * barriers: the same rounds in the same order, hence the same table. */
    for (;;) {
        __syncthreads();
        uint32_t a = small_end - small_begin, b = M - small_end;
        uint32_t nactive = a < b ? a : b;
        if (nactive <= (uint32_t)ESTH_WAVE) break;
        const bool act = (uint32_t)l < nactive;
        uint32_t si = 0, li = 0;
        float pS = 0.0f, pL = 0.0f;
        if (act) {
            si = sl[small_begin + l];
            li = sl[small_end + l];
            pS = prob[si];
            alias[si] = li;
            pL = prob[li];
        }
        __syncthreads();
        bool bs = false;
        if (act) {
            pL = (pL + pS) - 1.0f;
            bs = pL <= 1.0f;
        }
        uint32_t nbs;
        uint32_t rs = esth_block_rank(act && bs, cnt, l, nw, &nbs);
        if (act) {
            uint32_t s2 = bs ? (small_end + rs) : (small_end - 1u + nactive - ((uint32_t)l - rs));
            sl[s2] = li;
            prob[li] = pL;
        }
        small_end += nbs;
        small_begin += nactive;
    }
    if (l < ESTH_WAVE) {
        const uint32_t lane = (uint32_t)l;
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        for (;;) {
            const uint32_t a = small_end - small_begin, b = M - small_end;
            const uint32_t nactive = a < b ? a : b;
            if (nactive == 0) break;
            const bool act = lane < nactive;
            uint32_t si = 0, li = 0;
            float pS = 0.0f, pL = 0.0f;
            if (act) {
                si = sl[small_begin + lane];
                li = sl[small_end + lane];
                pS = prob[si];
                alias[si] = li;
                pL = prob[li];
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            bool bs = false;
            if (act) {
                pL = (pL + pS) - 1.0f;
                bs = pL <= 1.0f;
            }
            const uint64_t bal = __ballot(act && bs);
            const uint32_t rs = (uint32_t)__popcll(bal & below), nbs = (uint32_t)__popcll(bal);
            if (act) {
                const uint32_t s2 = bs ? (small_end + rs) : (small_end - 1u + nactive - (lane - rs));
                sl[s2] = li;
                prob[li] = pL;
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            small_end += nbs;
            small_begin += nactive;
        }
    }
    __syncthreads();
    if (table_only) {
        prob_out[idx] = prob[l];
        alias_out[idx] = alias[l];
        return;
    }
    /* vose_alias_generate_next (:198-202) on d_random2[idx] - 1 (:288-295) */
    const float r0 = rnd[2 * idx] - 1.0f, r1 = rnd[2 * idx + 1] - 1.0f;
    const uint32_t col = (uint32_t)floorf(r0 * (float)M);
    const uint32_t pick = (r1 < prob[col]) ? col : alias[col];
    const size_t sb = (size_t)g * M;
#pragma unroll
    for (int c = 0; c < S; ++c) st_main[esth_state_index(layout, S, N, c, idx)] = st_tmp[(size_t)c * N + sb + pick];
    if (sel) sel[idx] = (int)pick;
    /* every item has read its weight (before the first barrier) */
    if (carry) const_cast<float*>(w)[idx] = log_w ? 0.0f : 1.0f;
}

template <int S>
__global__ __launch_bounds__(1024) void k_alias(float* __restrict__ st_main,
                                                const float* __restrict__ st_tmp,
                                                const float* w,
                                                const float* __restrict__ rnd, uint32_t N, uint32_t M,
                                                float thr, int adaptive, int log_w,
                                                int* __restrict__ sel, float* __restrict__ prob_out,
                                                uint32_t* __restrict__ alias_out, int carry, int layout) {
    esth_alias_sub<S>(st_main, st_tmp, w, rnd, N, M, blockIdx.x, thr, adaptive, log_w, sel, prob_out, alias_out,
                      carry, layout);
}


/* ==================================================== Vose alias, one wave
 * The same resampling for M >= 64 with one wave per sub-filter (E = M/64
 * items per lane, item i = lane E + r): the work-group algorithm above with
 * wave-level ordering in place of its barriers.  The sums follow
 * sum_reduce4's tree, the worklists the same canonical order (ascending item
 * index) and the pairing rounds the same reads-then-writes steps, so the
 * tables are identical; but no waves idle at barriers while the pairing
 * tail runs on one of them, and a sub-filter costs one wave, not M/64. */
__device__ __forceinline__ int esth_wave_excl_scan_i(int x, int lane, int* total) {
    int v = x;
#pragma unroll
    for (int d = 1; d < ESTH_WAVE; d <<= 1) {
        const int o = __builtin_amdgcn_ds_bpermute((lane - d) << 2, v);
        v += lane >= d ? o : 0;
    }
    *total = __builtin_amdgcn_readlane(v, ESTH_WAVE - 1);
    return v - x;
}

/* sum_reduce4 (resampling-valias.cl:85-104) over arr[0..M) by one wave */
template <int M>
__device__ __forceinline__ float esth_wave_sum_reduce4(float* arr, int lane) {
    int offset = M;
#pragma unroll
    for (int it = 0; it < 6 && offset > 2; ++it) {
        offset >>= 2;
        for (int l = lane; l < offset; l += ESTH_WAVE) {
            const float rhs = (arr[l + offset] + arr[l + offset + offset]) + arr[l + offset + offset + offset];
            arr[l] = arr[l] + rhs;
        }
        esth_wave_lds_fence();
    }
    if (offset > 1) {
        if (lane == 0) arr[0] = arr[0] + arr[1];
        esth_wave_lds_fence();
    }
    return arr[0];
}

template <int LOGM, int S>
__device__ __forceinline__ void esth_alias_wave(float* __restrict__ st_main, const float* __restrict__ st_tmp,
                                                const float* w, const float* __restrict__ rnd, uint32_t N,
                                                uint32_t g, float thr, int adaptive, int log_w,
                                                int* __restrict__ sel, float* __restrict__ prob_out,
                                                uint32_t* __restrict__ alias_out, int carry, int layout,
                                                float* lds, bool recv = false, esth_xch xch = esth_xch{}) {
    static_assert(LOGM >= 6, "one wave per sub-filter needs M >= 64");
    constexpr int M = 1 << LOGM, E = M / ESTH_WAVE;
    constexpr int J = (M / 2 + ESTH_WAVE - 1) / ESTH_WAVE;   /* a round pairs at most M/2 */
    const int lane = threadIdx.x & (ESTH_WAVE - 1);
    const int le = lane * E;
    const size_t base = (size_t)g * M;
    float* prob = lds;
    uint32_t* alias = reinterpret_cast<uint32_t*>(lds + M);
    float* scratch = lds + 2 * M;                           /* sums, then the worklists */
    uint32_t* sl = reinterpret_cast<uint32_t*>(lds + 2 * M);
    uint32_t* rtab = reinterpret_cast<uint32_t*>(lds + 3 * M);   /* received slots' sources */
    const bool table_only = prob_out != nullptr;
    /* the exchange's receiving side (as in esth_resample_sub): slot M-1-j
     * holds rank k of neighbour q, j = q t + k, read from the compact best-t
     * copy (its index into it tabulated once, rtab[j]); its weight is written
     * back to w */
    recv = recv && xch.t > 0;
    const uint32_t nrecv = recv ? xch.nq * xch.t : 0u;
    const size_t Gt = (size_t)xch.G * xch.t;
    if (recv) {
        for (uint32_t j = (uint32_t)lane; j < nrecv; j += ESTH_WAVE) {
            const uint32_t q = j / xch.t, k = j - q * xch.t;
            rtab[j] = (uint32_t)esth_recv_src(g, q, k, xch.G, xch.t, xch.topo, xch.dimx, xch.dimy, xch.top);
        }
        esth_wave_lds_fence();
    }
    bool skip = !table_only && !adaptive && (rnd[N + g] - 1.0f) > thr;
    float p[E];
    esth_load_run<E>(w + base + le, p);
    if (recv) {
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const uint32_t pos = (uint32_t)(le + r);
            if (pos >= (uint32_t)M - nrecv) {
                p[r] = xch.bt_w[rtab[(uint32_t)M - 1u - pos]];
                xch.w_out[base + pos] = p[r];
            }
        }
    }
    if (!skip) {
        if (log_w) {
            float mx = p[0];
#pragma unroll
            for (int r = 1; r < E; ++r) mx = p[r] > mx ? p[r] : mx;
            mx = esth_bfly_max<1, ESTH_WAVE>(mx);
#pragma unroll
            for (int r = 0; r < E; ++r) p[r] = esth_expf(p[r] - esth_max_shift(mx));
        }
#pragma unroll
        for (int r = 0; r < E; ++r) scratch[le + r] = p[r];
        esth_wave_lds_fence();
        const float sum = esth_wave_sum_reduce4<M>(scratch, lane);
#pragma unroll
        for (int r = 0; r < E; ++r) p[r] = sum > 0.0f ? p[r] / sum : 1.0f / (float)M;
        esth_wave_lds_fence();   /* every lane has read the sum */
        if (adaptive && !table_only) {
#pragma unroll
            for (int r = 0; r < E; ++r) scratch[le + r] = p[r] * p[r];
            esth_wave_lds_fence();
            const float recip_ess = esth_wave_sum_reduce4<M>(scratch, lane);
            skip = !(recip_ess > thr);
            esth_wave_lds_fence();
        }
    }
    if (skip) {
        if (sel) {
#pragma unroll
            for (int r = 0; r < E; ++r) sel[base + le + r] = -1;
        }
        if (carry) {
#pragma unroll
            for (int c = 0; c < S; ++c) {
                float v[E];
                esth_load_run<E>(st_tmp + (size_t)c * N + base + le, v);
                if (recv) {
#pragma unroll
                    for (int r = 0; r < E; ++r) {
                        const uint32_t pos = (uint32_t)(le + r);
                        if (pos >= (uint32_t)M - nrecv) v[r] = xch.bt_s[(size_t)c * Gt + rtab[(uint32_t)M - 1u - pos]];
                    }
                }
                esth_store_plane_run<E, S>(st_main, layout, N, c, base + le, v);
            }
        }
        return;
    }
    /* vose_alias_init, resampling-valias.cl:111-195 */
    bool small[E];
    int cs = 0;
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const float pl = p[r] * (float)M;
        prob[le + r] = pl;
        alias[le + r] = (uint32_t)(le + r);
        small[r] = pl <= 1.0f;
        cs += small[r] ? 1 : 0;
    }
    int nsmall_i;
    int rk = esth_wave_excl_scan_i(cs, lane, &nsmall_i);
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const uint32_t item = (uint32_t)(le + r);
        const uint32_t slot = small[r] ? (uint32_t)rk : ((uint32_t)M - 1u - (item - (uint32_t)rk));
        sl[slot] = item;
        rk += small[r] ? 1 : 0;
    }
    esth_wave_lds_fence();
    uint32_t small_begin = 0, small_end = (uint32_t)nsmall_i;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (ESTH_WAVE - lane));
    for (;;) {
        const uint32_t a = small_end - small_begin, b = (uint32_t)M - small_end;
        const uint32_t n = a < b ? a : b;
        if (n == 0) break;
        if (n <= (uint32_t)ESTH_WAVE) {   /* the common short round: one pair per lane */
            const uint32_t k = (uint32_t)lane;
            const bool act = k < n;
            uint32_t li = 0;
            float pL = 0.0f;
            if (act) {
                const uint32_t si = sl[small_begin + k];
                li = sl[small_end + k];
                const float pS = prob[si];
                alias[si] = li;
                pL = (prob[li] + pS) - 1.0f;
            }
            esth_wave_lds_fence();
            const bool bs = act && pL <= 1.0f;
            const uint64_t bal = __ballot(bs);
            const uint32_t rs = (uint32_t)__popcll(bal & below);
            if (act) {
                const uint32_t s2 = bs ? (small_end + rs) : (small_end - 1u + n - (k - rs));
                sl[s2] = li;
                prob[li] = pL;
            }
            esth_wave_lds_fence();
            small_end += (uint32_t)__popcll(bal);
            small_begin += n;
            continue;
        }
        uint32_t li[J];
        float pL[J];
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const uint32_t k = (uint32_t)(j * ESTH_WAVE + lane);
            li[j] = 0;
            pL[j] = 0.0f;
            if (k < n) {
                const uint32_t si = sl[small_begin + k];
                li[j] = sl[small_end + k];
                const float pS = prob[si];
                alias[si] = li[j];
                pL[j] = (prob[li[j]] + pS) - 1.0f;
            }
        }
        esth_wave_lds_fence();   /* the round's reads before its writes */
        uint32_t nbs = 0;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            if ((uint32_t)(j * ESTH_WAVE) >= n) break;
            const uint32_t k = (uint32_t)(j * ESTH_WAVE + lane);
            const bool act = k < n;
            const bool bs = act && pL[j] <= 1.0f;
            const uint64_t bal = __ballot(bs);
            const uint32_t rs = nbs + (uint32_t)__popcll(bal & below);
            if (act) {
                const uint32_t s2 = bs ? (small_end + rs) : (small_end - 1u + n - (k - rs));
                sl[s2] = li[j];
                prob[li[j]] = pL[j];
            }
            nbs += (uint32_t)__popcll(bal);
        }
        esth_wave_lds_fence();
        small_end += nbs;
        small_begin += n;
    }
    if (table_only) {
#pragma unroll
        for (int r = 0; r < E; ++r) {
            prob_out[base + le + r] = prob[le + r];
            alias_out[base + le + r] = alias[le + r];
        }
        return;
    }
    /* vose_alias_generate_next (:198-202) on d_random2[idx] - 1 (:288-295) */
    float u[2 * E];
    esth_load_run<2 * E>(rnd + 2 * (base + le), u);
    int pick[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const float r0 = u[2 * r] - 1.0f, r1 = u[2 * r + 1] - 1.0f;
        const uint32_t col = (uint32_t)floorf(r0 * (float)M);
        pick[r] = (int)((r1 < prob[col]) ? col : alias[col]);
    }
    size_t src[E];   /* where each pick's particle is: the sorted copy or a received slot */
    bool rx[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        rx[r] = (uint32_t)pick[r] >= (uint32_t)M - nrecv;
        src[r] = rx[r] ? (size_t)rtab[(uint32_t)M - 1u - (uint32_t)pick[r]] : base + (size_t)pick[r];
    }
#pragma unroll
    for (int c = 0; c < S; ++c) {
        float v[E];
#pragma unroll
        for (int r = 0; r < E; ++r)
            v[r] = rx[r] ? xch.bt_s[(size_t)c * Gt + src[r]] : st_tmp[(size_t)c * N + src[r]];
        esth_store_plane_run<E, S>(st_main, layout, N, c, base + le, v);
    }
    if (sel) {
#pragma unroll
        for (int r = 0; r < E; ++r) sel[base + le + r] = pick[r];
    }
    if (carry) {   /* every lane has read its weights (before the sums) */
        float w0[E];
#pragma unroll
        for (int r = 0; r < E; ++r) w0[r] = log_w ? 0.0f : 1.0f;
        esth_store_run<E>(const_cast<float*>(w) + base + le, w0);
    }
}

/* waves per work-group of the one-wave alias kernels (LDS: 16 M bytes a wave) */
template <int LOGM>
struct esth_alias_wpb {
    static constexpr int value = LOGM >= 10 ? 2 : 4;
};

template <int LOGM, int S>
__global__ __launch_bounds__(256) void k_alias_w(float* __restrict__ st_main, const float* __restrict__ st_tmp,
                                                 const float* w, const float* __restrict__ rnd, uint32_t N,
                                                 uint32_t G, float thr, int adaptive, int log_w,
                                                 int* __restrict__ sel, float* __restrict__ prob_out,
                                                 uint32_t* __restrict__ alias_out, int carry, int layout) {
    constexpr int WPB = esth_alias_wpb<LOGM>::value;
    __shared__ float lds[WPB][4 << LOGM];
    const uint32_t wv = threadIdx.x / ESTH_WAVE;
    const uint32_t g = blockIdx.x * WPB + wv;
    if (g >= G) return;   /* a whole wave: no work-group barrier follows */
    esth_alias_wave<LOGM, S>(st_main, st_tmp, w, rnd, N, g, thr, adaptive, log_w, sel, prob_out, alias_out, carry,
                             layout, lds[wv]);
}

/* The step's argmax + posterior (work-group 0) and the exchange's receiving
 * side beside the one-wave alias resampling, as k_resample_argmax does for
 * the scan resamplers: the argmax reads the winners, partials and best-t
 * copy, the resampling the weights, uniforms, sorted state and best-t copy;
 * neither writes what the other reads. */
template <int LOGM, int S>
__global__ __launch_bounds__(256) void k_alias_argmax_w(float* __restrict__ st_main, const float* __restrict__ st_tmp,
                                                        const float* w, const float* __restrict__ rnd, uint32_t N,
                                                        uint32_t G, float thr, int adaptive, int log_w,
                                                        const float2* __restrict__ winner,
                                                        const float* __restrict__ part,
                                                        const float* __restrict__ best_src, size_t plane_stride,
                                                        uint32_t g_stride, float2* __restrict__ result,
                                                        double* __restrict__ est, esth_xch xch, int carry,
                                                        int layout) {
    constexpr int WPB = esth_alias_wpb<LOGM>::value;
    __shared__ float lds[WPB][4 << LOGM];
    __shared__ float2 sc[128];
    __shared__ float smax[WPB];
    __shared__ double sd[WPB * (2 + S)];
    if (blockIdx.x == 0) {
        esth_argmax_block<S, WPB * ESTH_WAVE>(winner, G, log_w, part, best_src, plane_stride, g_stride, result, est, sc,
                                              smax, sd);
        return;
    }
    const uint32_t wv = threadIdx.x / ESTH_WAVE;
    const uint32_t g = (blockIdx.x - 1) * WPB + wv;
    if (g >= G) return;
    esth_alias_wave<LOGM, S>(st_main, st_tmp, w, rnd, N, g, thr, adaptive, log_w, nullptr, nullptr, nullptr, carry,
                             layout, lds[wv], true, xch);
}

/* ============================================================== batches
 * Kernels over many filters: a device table of esth_bfilter and tables of
 * (filter, sub-filter) pairs, one per sub-filter size class.  Each wave (or
 * work-group for the alias build) runs exactly the per-sub-filter code of the
 * single-filter kernels above on its filter's buffers, so a filter's particles
 * are bit-identical alone and in a batch. */
__device__ __forceinline__ uint32_t esth_wave_id() {
    return __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / ESTH_WAVE);
}

/* one wave per chunk of 64 particle pairs of one filter, chunk = (filter,
 * first pair): sample+weight has no sub-filter structure, so the chunks tile
 * a filter's particles whatever its sub-filter size, and a wave never mixes
 * filters (its frame reads stay uniform) */
template <int MODEL>
__global__ __launch_bounds__(256) void kb_sample_weight(const esth_bfilter* __restrict__ F,
                                                        const uint2* __restrict__ chunks, uint32_t nchunks,
                                                        const float* __restrict__ R,
                                                        const float* __restrict__ frames, int log_w) {
    const uint32_t wv = esth_wave_id();
    if (wv >= nchunks) return;
    const uint2 c = chunks[wv];
    const esth_bfilter& f = F[c.x];
    const uint32_t p = 2u * (c.y + (threadIdx.x & (ESTH_WAVE - 1)));
    if (p < f.N)
        esth_sample_weight_pair<MODEL>(f.state, f.N, R + f.r_off, frames + f.f_off, log_w, f.w, p, (int)(f.flags & 1u));
}

template <int LOGM, int S>
__global__ __launch_bounds__(256) void kb_sort(const esth_bfilter* __restrict__ F, const uint2* __restrict__ sub,
                                               uint32_t nsub, int log_w) {
    const uint32_t wv = esth_wave_id();
    if (wv >= nsub) return;
    const uint2 e = sub[wv];
    const esth_bfilter& f = F[e.x];
    __shared__ float stage[4][ESTH_WAVE * esth_tile<LOGM>::E];
    esth_sort_sub<LOGM, S, true>(f.w, f.state, f.tmp, f.N, f.G, e.y, true, (float2*)f.winner, nullptr, f.t, f.bt_w,
                           f.bt_s, f.part, log_w, stage[threadIdx.x / ESTH_WAVE]);
}

/* k_sample_sort over one size class of a batch (m >= 128): work-group b
 * samples, weights and sorts sub-filter sub[b] = (filter, sub-filter) with
 * that filter's uniforms, frame and buffers.  keep_all: the resampler's gate
 * depends on the weights (adaptive alias), so every sub-filter without
 * carried weights stores its propagated state; otherwise those whose gate
 * uniform exceeds thr do.  The 3-dimensional models are compiled for 8
 * waves per SIMD (at most 64 VGPRs): left free, the compiler spends 80 VGPRs
 * on the unicycle's m = 128 and 256 classes and halves the sampling's
 * occupancy.  The others are left free (a 6-wave target made the arm models
 * spill).  Instantiations that spill, or keep fewer waves resident than
 * kb_sample_weight, are not used. */
template <int MODEL>
__host__ __device__ constexpr int esth_kbss_waves_per_eu() {
    return esth_model_traits<MODEL>::S <= 3 ? 8 : 1;
}

template <int MODEL, int LOGM, int S>
__global__ __launch_bounds__(ESTH_WAVE * esth_ss_waves<LOGM>())
__attribute__((amdgpu_waves_per_eu(esth_kbss_waves_per_eu<MODEL>()))) void kb_sample_sort(
    const esth_bfilter* __restrict__ F, const uint2* __restrict__ sub, const float* __restrict__ R,
    const float* __restrict__ frames, int log_w, float thr, int keep_all) {
    using T = esth_model_traits<MODEL>;
    static_assert(T::S == S, "state dimension of the model");
    constexpr int M = 1 << LOGM;
    constexpr int NT = ESTH_WAVE * esth_ss_waves<LOGM>();
    __shared__ __align__(16) float lds[(1 + S) * M];
    const uint2 e = sub[blockIdx.x];
    const esth_bfilter& f = F[e.x];
    const int carry = (int)(f.flags & 1u);
    const float* Rf = R + f.r_off;
    const bool keep = !carry && (keep_all || (Rf[f.uoff + f.N + e.y] - 1.0f) > thr);
#pragma unroll 1
    for (int q = 2 * (int)threadIdx.x; q < M; q += 2 * NT) {
        const uint32_t p = e.y * (uint32_t)M + (uint32_t)q;
        float s0[S], s1[S], u0[T::V], u1[T::V];
        esth_load_pair<S, ESTH_LAYOUT_SOA>(f.state, f.N, p, s0, s1);
#pragma unroll
        for (int k = 0; k < T::V; ++k) {
            const float2 u = *reinterpret_cast<const float2*>(Rf + (size_t)k * f.N + p);
            u0[k] = u.x;
            u1[k] = u.y;
        }
        float lw0, lw1;
        esth_sample_pair_regs<MODEL>(s0, s1, u0, u1, frames + f.f_off, lw0, lw1);
        if (keep) esth_store_pair<S, ESTH_LAYOUT_SOA>(f.state, f.N, p, s0, s1);
        float2 wo = log_w ? make_float2(lw0, lw1) : make_float2(esth_expf(lw0), esth_expf(lw1));
        if (carry) {
            const float2 wp = *reinterpret_cast<const float2*>(f.w + p);
            wo = log_w ? make_float2(wp.x + wo.x, wp.y + wo.y) : make_float2(wp.x * wo.x, wp.y * wo.y);
        }
        *reinterpret_cast<float2*>(lds + q) = wo;
#pragma unroll
        for (int c = 0; c < S; ++c) *reinterpret_cast<float2*>(lds + (1 + c) * M + q) = make_float2(s0[c], s1[c]);
    }
    __syncthreads();
    if (threadIdx.x >= ESTH_WAVE) return;
    esth_sort_sub<LOGM, S, false, true>(f.w, nullptr, f.tmp, f.N, f.G, e.y, true, (float2*)f.winner, nullptr, f.t,
                                        f.bt_w, f.bt_s, f.part, log_w, nullptr, ESTH_LAYOUT_SOA, lds);
}

/* kb_sample_sort for the m = 64 class, over units of one filter's
 * sub-filters g0 and g0 + 1 (unit = (filter, g0 | single << 31); a single
 * unit has sub-filter g0 only): a wave samples and weights the unit's
 * particles (lanes 0-31 the first sub-filter's 32 pairs, lanes 32-63 the
 * second's) into its own LDS, then sorts both at once from there, 32 lanes
 * and two positions per lane each (esth_sort_sub, FROM_LDS).  A unit never spans two filters, so the model
 * hook's frame reads stay wave-uniform.  Per-particle arithmetic and the sort
 * are those of kb_sample_weight + kb_sort. */
#define ESTH_UNIT_SINGLE 0x80000000u
template <int MODEL, int S>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(esth_kbss_waves_per_eu<MODEL>()))) void kb_sample_sort64(
    const esth_bfilter* __restrict__ F, const uint2* __restrict__ units, uint32_t nunits, const float* __restrict__ R,
    const float* __restrict__ frames, int log_w, float thr, int keep_all) {
    using T = esth_model_traits<MODEL>;
    static_assert(T::S == S, "state dimension of the model");
    constexpr int M = 64;
    constexpr int SUB = (1 + S) * M;   /* LDS floats of one sub-filter: weights, then the planes */
    __shared__ __align__(16) float lds[4][2 * SUB];
    const int lane = threadIdx.x & (ESTH_WAVE - 1);
    const int wv = threadIdx.x / ESTH_WAVE;
    float* L = lds[wv];
    const uint32_t u = blockIdx.x * 4u + (uint32_t)wv;
    const bool valid = u < nunits;
    const uint2 e = units[valid ? u : 0];
    const esth_bfilter& f = F[e.x];
    const uint32_t g0 = e.y & ~ESTH_UNIT_SINGLE;
    const int nsubs = (e.y & ESTH_UNIT_SINGLE) ? 1 : 2;
    const int half = lane / 32;
    if (valid && half < nsubs) {
        const int q = 2 * (lane & 31);
        const uint32_t g = g0 + (uint32_t)half;
        const int carry = (int)(f.flags & 1u);
        const float* Rf = R + f.r_off;
        const bool keep = !carry && (keep_all || (Rf[f.uoff + f.N + g] - 1.0f) > thr);
        const uint32_t p = g * (uint32_t)M + (uint32_t)q;
        float s0[S], s1[S], u0[T::V], u1[T::V];
        esth_load_pair<S, ESTH_LAYOUT_SOA>(f.state, f.N, p, s0, s1);
#pragma unroll
        for (int k = 0; k < T::V; ++k) {
            const float2 uu = *reinterpret_cast<const float2*>(Rf + (size_t)k * f.N + p);
            u0[k] = uu.x;
            u1[k] = uu.y;
        }
        float lw0, lw1;
        esth_sample_pair_regs<MODEL>(s0, s1, u0, u1, frames + f.f_off, lw0, lw1);
        if (keep) esth_store_pair<S, ESTH_LAYOUT_SOA>(f.state, f.N, p, s0, s1);
        float2 wo = log_w ? make_float2(lw0, lw1) : make_float2(esth_expf(lw0), esth_expf(lw1));
        if (carry) {
            const float2 wp = *reinterpret_cast<const float2*>(f.w + p);
            wo = log_w ? make_float2(wp.x + wo.x, wp.y + wo.y) : make_float2(wp.x * wo.x, wp.y * wo.y);
        }
        float* Lh = L + half * SUB;
        *reinterpret_cast<float2*>(Lh + q) = wo;
#pragma unroll
        for (int c = 0; c < S; ++c) *reinterpret_cast<float2*>(Lh + (1 + c) * M + q) = make_float2(s0[c], s1[c]);
    }
    __syncthreads();
    if (!valid) return;
    /* both sub-filters at once, 32 lanes and two positions per lane each */
    esth_sort_sub<6, S, false, true, 32>(f.w, nullptr, f.tmp, f.N, f.G, g0 + (uint32_t)half, half < nsubs,
                                         (float2*)f.winner, nullptr, f.t, f.bt_w, f.bt_s, f.part, log_w, nullptr,
                                         ESTH_LAYOUT_SOA, L + half * SUB);
}

template <int S>
__global__ __launch_bounds__(256) void kb_argmax(const esth_bfilter* __restrict__ F, int log_w) {
    __shared__ float2 sc[128];
    __shared__ float smax[4];
    __shared__ double sd[4 * (2 + S)];
    const esth_bfilter& f = F[blockIdx.x];
    esth_argmax_block<S, 256>((const float2*)f.winner, f.G, log_w, f.part, f.tmp, f.N, f.m, (float2*)f.result,
                              f.est, sc, smax, sd);
}

/* all-to-all: the global top t of each filter's compact best-t buffer */
static __global__ __launch_bounds__(1024) void kb_topt(const esth_bfilter* __restrict__ F) {
    __shared__ float sw[1024];
    __shared__ uint32_t si[1024];
    const esth_bfilter& f = F[blockIdx.x];
    const int l = threadIdx.x;
    const uint32_t n = f.G * f.t;
    float pw = INFINITY;
    uint32_t pi = 0;
    bool have_prev = false;
    for (uint32_t k = 0; k < f.t; ++k) {
        float bw = -INFINITY;
        uint32_t bi = 0xffffffffu;
        for (uint32_t c = (uint32_t)l; c < n; c += 1024) {
            float v = f.bt_w[c];
            bool after = !have_prev || (v < pw) || (v == pw && c > pi);
            bool better = (bi == 0xffffffffu) || (v > bw) || (v == bw && c < bi);
            if (after && better) { bw = v; bi = c; }
        }
        sw[l] = bw;
        si[l] = bi;
        __syncthreads();
        for (int off = 512; off > 0; off >>= 1) {
            if (l < off) {
                float ow = sw[l + off];
                uint32_t oi = si[l + off];
                bool take = (oi != 0xffffffffu) &&
                            ((si[l] == 0xffffffffu) || ow > sw[l] || (ow == sw[l] && oi < si[l]));
                if (take) { sw[l] = ow; si[l] = oi; }
            }
            __syncthreads();
        }
        pw = sw[0];
        pi = si[0];
        have_prev = true;
        if (l == 0) f.top[k] = pi;
        __syncthreads();
    }
}

/* one thread per received slot: nslots = (torus ? 4 : 1) * t slots per sub-filter */
static __global__ void kb_exchange(const esth_bfilter* __restrict__ F, const uint2* __restrict__ sub,
                                   uint32_t nsub, int S, uint32_t nslots, int topo) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (uint64_t)nsub * nslots) return;
    const uint2 e = sub[j / nslots];
    const uint32_t slot = (uint32_t)(j % nslots);
    const esth_bfilter& f = F[e.x];
    const uint32_t t = f.t;
    const uint32_t q = slot / t, k = slot % t;
    const size_t src = esth_recv_src(e.y, q, k, f.G, t, topo, f.dimx, f.dimy, f.top);
    const size_t dst = (size_t)e.y * f.m + (f.m - 1 - k - q * t);
    f.w[dst] = f.bt_w[src];
    const size_t Gt = (size_t)f.G * t;
    for (int c = 0; c < S; ++c) f.tmp[(size_t)c * f.N + dst] = f.bt_s[(size_t)c * Gt + src];
}

template <int LOGM, int S>
__global__ __launch_bounds__(256) void kb_resample(const esth_bfilter* __restrict__ F,
                                                   const uint2* __restrict__ sub, uint32_t nsub,
                                                   const float* __restrict__ R, float thr, int systematic,
                                                   int log_w, int topo, uint32_t argmax_blocks) {
    __shared__ float cdf_s[4][ESTH_WAVE * esth_tile<LOGM>::E];
    __shared__ float2 sc[128];
    __shared__ float smax[4];
    __shared__ double sd[4 * (2 + S)];
    if (blockIdx.x < argmax_blocks) {   /* filter blockIdx.x's argmax + posterior, beside the resampling */
        const esth_bfilter& f = F[blockIdx.x];
        esth_argmax_block<S, 256>((const float2*)f.winner, f.G, log_w, f.part, f.tmp, f.N, f.m, (float2*)f.result,
                                  f.est, sc, smax, sd);
        return;
    }
    const uint32_t wv = __builtin_amdgcn_readfirstlane(((blockIdx.x - argmax_blocks) * blockDim.x + threadIdx.x) /
                                                       ESTH_WAVE);
    const bool valid = wv < nsub;
    const uint2 e = sub[valid ? wv : nsub - 1];
    const esth_bfilter& f = F[e.x];
    const esth_xch x{f.bt_w, f.bt_s, f.top, f.w, f.G, f.t, topo == ESTH_TOPO_TORUS ? 4u : 1u, topo, f.dimx, f.dimy};
    esth_resample_sub<LOGM, S, true>(f.state, f.tmp, f.w, R + f.r_off + f.uoff, f.N, e.y, valid, thr, systematic, log_w,
                               nullptr, cdf_s[threadIdx.x / ESTH_WAVE], topo != ESTH_TOPO_NONE, x,
                               (int)(f.flags & 1u), f.w);
}

/* kb_resample's other tilings of a size class; work-groups below
 * argmax_blocks run filter blockIdx.x's argmax + posterior instead.
 * kb_resample64x2 (m = 64): 32 lanes per sub-filter, two positions each, so
 * a wave resamples entries 2w and 2w + 1 of the class list (which may belong
 * to different filters: nothing here reads a frame).  kb_resample2 (m >= 256):
 * two waves per sub-filter (esth_resample_sub2), a work-group of four waves
 * two entries.  The scan's binary tree over positions, the searches and the
 * selections are the one-wave kernel's, so the particles are the same bit
 * for bit. */
template <int S>
__global__ __launch_bounds__(256) void kb_resample64x2(const esth_bfilter* __restrict__ F,
                                                       const uint2* __restrict__ sub, uint32_t nsub,
                                                       const float* __restrict__ R, float thr, int systematic,
                                                       int log_w, int topo, uint32_t argmax_blocks) {
    __shared__ float cdf_s[4][2 * ESTH_WAVE];
    __shared__ float2 sc[128];
    __shared__ float smax[4];
    __shared__ double sd[4 * (2 + S)];
    if (blockIdx.x < argmax_blocks) {
        const esth_bfilter& f = F[blockIdx.x];
        esth_argmax_block<S, 256>((const float2*)f.winner, f.G, log_w, f.part, f.tmp, f.N, f.m, (float2*)f.result,
                                  f.est, sc, smax, sd);
        return;
    }
    const uint32_t wv = ((blockIdx.x - argmax_blocks) * blockDim.x + threadIdx.x) / ESTH_WAVE;
    const uint32_t i = 2u * wv + (uint32_t)((threadIdx.x & (ESTH_WAVE - 1)) / 32);
    const bool valid = i < nsub;
    const uint2 e = sub[valid ? i : nsub - 1];
    const esth_bfilter& f = F[e.x];
    const esth_xch x{f.bt_w, f.bt_s, f.top, f.w, f.G, f.t, topo == ESTH_TOPO_TORUS ? 4u : 1u, topo, f.dimx, f.dimy};
    esth_resample_sub<6, S, true, 32>(f.state, f.tmp, f.w, R + f.r_off + f.uoff, f.N, e.y, valid, thr, systematic,
                                      log_w, nullptr, cdf_s[threadIdx.x / ESTH_WAVE], topo != ESTH_TOPO_NONE, x,
                                      (int)(f.flags & 1u), f.w);
}

template <int LOGM, int S>
__global__ __launch_bounds__(256) void kb_resample2(const esth_bfilter* __restrict__ F, const uint2* __restrict__ sub,
                                                    uint32_t nsub, const float* __restrict__ R, float thr,
                                                    int systematic, int log_w, int topo, uint32_t argmax_blocks) {
    constexpr int M = 1 << LOGM;
    __shared__ __align__(16) float cdf_s[2][M];
    __shared__ __align__(16) float stage_s[2][S * M];
    __shared__ float xl_s[2][4];
    __shared__ float2 sc[128];
    __shared__ float smax[4];
    __shared__ double sd[4 * (2 + S)];
    if (blockIdx.x < argmax_blocks) {
        const esth_bfilter& f = F[blockIdx.x];
        esth_argmax_block<S, 256>((const float2*)f.winner, f.G, log_w, f.part, f.tmp, f.N, f.m, (float2*)f.result,
                                  f.est, sc, smax, sd);
        return;
    }
    const int wv = threadIdx.x / ESTH_WAVE;   /* 0..3: entry wv / 2, half wv % 2 */
    const int j = wv >> 1;
    const uint32_t i = (blockIdx.x - argmax_blocks) * 2u + (uint32_t)j;
    const bool valid = i < nsub;
    const uint2 e = sub[valid ? i : nsub - 1];
    const esth_bfilter& f = F[e.x];
    const esth_xch x{f.bt_w, f.bt_s, f.top, f.w, f.G, f.t, topo == ESTH_TOPO_TORUS ? 4u : 1u, topo, f.dimx, f.dimy};
    esth_resample_sub2<LOGM, S>(f.state, f.tmp, f.w, R + f.r_off + f.uoff, f.N, e.y, valid, thr, systematic, log_w,
                                cdf_s[j], stage_s[j], xl_s[j], wv & 1, x, (int)(f.flags & 1u), f.w, ESTH_LAYOUT_SOA);
}

/* one wave per (filter, sub-filter) entry of a size class (m >= 64) */
template <int LOGM, int S>
__global__ __launch_bounds__(256) void kb_alias_w(const esth_bfilter* __restrict__ F, const uint2* __restrict__ sub,
                                                  uint32_t nsub, const float* __restrict__ R, float thr, int adaptive,
                                                  int log_w) {
    constexpr int WPB = esth_alias_wpb<LOGM>::value;
    __shared__ float lds[WPB][4 << LOGM];
    const uint32_t wv = threadIdx.x / ESTH_WAVE;
    const uint32_t i = blockIdx.x * WPB + wv;
    if (i >= nsub) return;
    const uint2 e = sub[i];
    const esth_bfilter& f = F[e.x];
    esth_alias_wave<LOGM, S>(f.state, f.tmp, f.w, R + f.r_off + f.uoff, f.N, e.y, thr, adaptive, log_w, nullptr, nullptr,
                             nullptr, (int)(f.flags & 1u), ESTH_LAYOUT_SOA, lds[wv]);
}


/* ============================================================== serving
 * Start-up of tracks admitted in a wave: the state planes set to the initial
 * state, the weights to w0 and the track's K MTGP32 statuses seeded on the
 * device as mtgp32_init_state (mtgp32-fast.c:119-137) does on the host, with
 * stream j of parameter set j seeded seed + j + 1 (mtgp32-opencl.c:47).
 * grid (blocks per track, tracks). */
struct esth_sinit {
    float* state;
    float* w;
    uint32_t* status;   /* K consecutive 351-word statuses */
    uint32_t N, S, K, seed;
    float w0;           /* initial weight: 0, or 1 for carried linear weights */
    float init[ESTH_MAX_STATE_DIM];
};

__device__ __forceinline__ void esth_mtgp_seed_dev(const uint32_t* __restrict__ set_params, uint32_t seed,
                                                   uint32_t* __restrict__ st) {
    const uint32_t hidden = set_params[4 + 4] ^ (set_params[4 + 8] << 16);
    uint32_t tmp = hidden;
    tmp += tmp >> 16;
    tmp += tmp >> 8;
    const uint32_t fill = (tmp & 0xffu) * 0x01010101u;
    uint32_t prev = seed;
    st[0] = seed;
    uint32_t cur = hidden ^ (1812433253u * (prev ^ (prev >> 30)) + 1u);
    /* a[1] starts as `hidden` (not the fill) before the recursion is applied */
    st[1] = cur;
    prev = cur;
    for (int i = 2; i < ESTH_MTGP_N; ++i) {
        cur = fill ^ (1812433253u * (prev ^ (prev >> 30)) + (uint32_t)i);
        st[i] = cur;
        prev = cur;
    }
}

static __global__ __launch_bounds__(256) void ks_init(const esth_sinit* __restrict__ rec,
                                                      const uint32_t* __restrict__ params) {
    const esth_sinit& r = rec[blockIdx.y];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t c = 0; c < r.S; ++c) {
        const float v = r.init[c];
        float* pl = r.state + (size_t)c * r.N;
        for (uint32_t i = i0; i < r.N; i += stride) pl[i] = v;
    }
    for (uint32_t i = i0; i < r.N; i += stride) r.w[i] = r.w0;
    if (blockIdx.x == 0)
        for (uint32_t j = threadIdx.x; j < r.K; j += blockDim.x)
            esth_mtgp_seed_dev(params + (size_t)j * ESTH_MTGP_PARAM_WORDS, r.seed + j + 1u,
                               r.status + (size_t)j * ESTH_MTGP_N);
}

/* ================================================================ copy */
static __global__ __launch_bounds__(256) void k_copy(float4* __restrict__ dst, const float4* __restrict__ src, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

/* ================================================================ math */
static __global__ void k_math(int fn, const float* __restrict__ x, const float* __restrict__ x2,
                       float* __restrict__ out, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = x[i], r = 0.0f, s, c;
    switch (fn) {
    case 0: r = esth_expf(v); break;
    case 1: r = esth_logf(v); break;
    case 2: esth_sincosf(v, &s, &c); r = s; break;
    case 3: esth_sincosf(v, &s, &c); r = c; break;
    case 4: r = esth_atan2f(v, x2[i]); break;
    case 5: esth_polar_abs(v, x2[i], &r, &s); break;
    case 6: esth_polar_abs(v, x2[i], &s, &r); break;
    case 7: r = esth_rcpf(v); break;
    default: r = 0.0f;
    }
    out[i] = r;
}

#endif /* ESTH_KERNELS_H */
