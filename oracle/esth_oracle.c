/*
 * esth_oracle.c - CPU oracle of the island-model filter step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke check and bench.py's CPU
 * baseline leg load this library, as the checker; the engine never does.
 *
 * It restates the reference's OpenCL path (src/partfilt/opencl/) the way a
 * CPU OpenCL runtime executes it: each work-group in turn, and inside a
 * work-group each barrier-delimited phase as a loop over work-items in
 * ascending local id.  Function-by-function citations are given below.
 *
 * Floating point: compiled with -ffp-contract=off, so every expression is the
 * IEEE operation sequence written here.  The per-particle model hooks and the
 * transcendentals come from the engine's shared headers (esth_models.h,
 * esth_math.h) so that this FP32 path and the GPU agree bit for bit; the
 * models are pinned independently by the FP64 twins below (libm double), and
 * the transcendentals by error-bound tests against libm.
 *
 * Parity pins (see tests/): MTGP words against the reference's own vendored
 * C library (golden fixtures); sort by brute force over all tie patterns and
 * invariants; scan by closed form; alias by the Vose invariant; exchange by
 * closed-form index maps; posterior by the FP64 twin.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "esth_oracle.h"
#include "esth_models.h"
#include "esth_layout.h"

/* ================================================================ MTGP32
 * Saito & Matsumoto, "Variants of Mersenne Twister suitable for graphic
 * processors", ACM TOMS 39(2), 2013.  MEXP 11213, N = 351 words.
 * Parameter word layout (52 words per set): pos, sh1, sh2, mask, tbl[16],
 * tmp_tbl[16], flt_tmp_tbl[16]. */

/* mtgp32_init_state, mtgp32-fast.c:119-137 */
void or_mtgp_seed(const uint32_t* p, uint32_t seed, uint32_t* st) {
    const uint32_t* tbl = p + 4;
    uint32_t hidden = tbl[4] ^ (tbl[8] << 16);
    uint32_t tmp = hidden;
    tmp += tmp >> 16;
    tmp += tmp >> 8;
    uint32_t fill = tmp & 0xffu;
    fill |= fill << 8;
    fill |= fill << 16;                 /* memset(array, tmp & 0xff, ...) */
    for (int i = 0; i < ESTH_MTGP_N; ++i) st[i] = fill;
    st[0] = seed;
    st[1] = hidden;
    for (int i = 1; i < ESTH_MTGP_N; ++i)
        st[i] ^= 1812433253u * (st[i - 1] ^ (st[i - 1] >> 30)) + (uint32_t)i;
}

/* Sequential stream x[n] = rec(x[n-351], x[n-350], x[n-351+pos]), output
 * temper(x[n], x[n-351+pos-1]); mtgp32-fast.h:149-218,252-293, which the
 * OpenCL kernel (mtgp32-opencl.cl:186-242) reproduces 256 words at a time. */
void or_mtgp_generate(const uint32_t* p, uint32_t mask, uint32_t* st, int n, uint32_t* out,
                      int float_mode) {
    const int pos = (int)p[0], sh1 = (int)p[1], sh2 = (int)p[2];
    const uint32_t* tbl = p + 4;
    const uint32_t* ttbl = float_mode ? p + 36 : p + 20;
    uint32_t* x = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(ESTH_MTGP_N + n));
    memcpy(x, st, sizeof(uint32_t) * ESTH_MTGP_N);
    for (int k = 0; k < n; ++k) {
        uint32_t X = (x[k] & mask) ^ x[k + 1];
        X ^= X << sh1;
        uint32_t Y = X ^ (x[k + pos] >> sh2);
        uint32_t r = Y ^ tbl[Y & 0x0f];
        x[ESTH_MTGP_N + k] = r;
        uint32_t T = x[k + pos - 1];
        T ^= T >> 16;
        T ^= T >> 8;
        out[k] = float_mode ? ((r >> 9) ^ ttbl[T & 0x0f]) : (r ^ ttbl[T & 0x0f]);
    }
    memcpy(st, x + n, sizeof(uint32_t) * ESTH_MTGP_N);
    free(x);
}

/* The reference buffer layout: stream j (parameter set first_set + j) writes
 * `size` words at out[j*size] (mtgp32-opencl.cl:212,225,237); the mask is the
 * first set's (mtgp32-opencl.c:68). */
void or_mtgp_fill(const uint32_t* params, int nstreams, int first_set, uint32_t* status, int size,
                  uint32_t* out, int float_mode) {
    const uint32_t mask = params[(size_t)first_set * ESTH_MTGP_PARAM_WORDS + 3];
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nstreams; ++j) {
        const uint32_t* p = params + (size_t)(first_set + j) * ESTH_MTGP_PARAM_WORDS;
        or_mtgp_generate(p, mask, status + (size_t)j * ESTH_MTGP_N, size, out + (size_t)j * size,
                         float_mode);
    }
}

/* ============================================================ Box-Muller
 * boxmuller, pf_generic.cl:465-483: pair (2i, 2i+1) in place, for 2i < n. */
void or_boxmuller(float* d, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (n + 1) / 2; ++i) {
        float z0, z1;
        esth_boxmuller_pair(d[2 * i], d[2 * i + 1], &z0, &z1);
        d[2 * i] = z0;
        d[2 * i + 1] = z1;
    }
}

/* ============================================================== models */
int or_build_frame(int model, const void* params, const void* ctrl, const void* meas, float dt,
                   float* frame) {
    return esth_build_frame_impl(model, params, ctrl, meas, dt, frame);
}

/* sampling_importance, pf_generic.cl:206-250: particle idx reads its normals
 * at d_random[k*N + idx] (pf_model.cl:128-158) and stores weight[idx]. */
int or_sample_weight(int model, float* state, uint32_t N, const float* normals, const float* frame,
                     int weights, float* w) {
    int S, V;
    if (esth_model_dims(model, &S, &V)) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)N; ++i) {
        float s[ESTH_MAX_STATE_DIM], z[ESTH_MAX_STATE_DIM];
        for (int c = 0; c < S; ++c) s[c] = state[(size_t)c * N + i];
        for (int k = 0; k < V; ++k) z[k] = normals[(size_t)k * N + i];
        float lw = 0.0f;
        /* the hook of the listed model (every listed model, no hand-kept table) */
#define OR_X(ID, NAME, PFX, S_, V_, PSZ, CSZ, MSZ) \
        if (model == ID) lw = esth_##PFX##_sampling_iw(s, z, frame);
        ESTH_MODEL_LIST(OR_X)
#undef OR_X
        for (int c = 0; c < S; ++c) state[(size_t)c * N + i] = s[c];
        w[i] = (weights == ESTH_WEIGHTS_LOG) ? lw : esth_expf(lw);
    }
    return 0;
}

/* ---- FP64 twins: the models written directly from their definitions with
 * libm double functions, independent of esth_math.h / esth_models.h. */

/* robotarm5's walk and joint chain (pf_model.cl:70-117) in double; za / zp:
 * noise of the joint angles / of x, y, vx, vy.  Returns the angle cost and
 * the camera-frame point in cam. */
static double ra5_chain_f64(double* s, const double* za, const double* zp, const esth_robotarm5_control* c,
                            const esth_robotarm5_measurement* m, double dt, double cam[3]) {
    static const double L[5] = {0.0, 3.0, 3.0, 3.0, 3.0};
    double x = s[5] + s[7] * dt + zp[0] * 0.2;
    double y = s[6] + s[8] * dt + zp[1] * 0.2;
    double vx = s[7] + zp[2] * 2.0;
    double vy = s[8] + zp[3] * 2.0;
    double value = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    for (int i = 0; i < 5; ++i) {
        double a = s[i] + (double)c->angles[i] * dt + za[i] * 0.2;
        double sf = sin(a), cf = cos(a);
        double e = a - (double)m->angles[i];
        value += e * e * (double)0.1f;
        if (i == 0) {
            cx = x * cf + y * sf;
            cy = -x * sf + y * cf;
            cz = 0.0;
        } else {
            double nB = cy * cf - cz * sf;
            double nC = cy * sf + cz * cf;
            cy = nB;
            cz = nC - L[i];
        }
        s[i] = a;
    }
    s[5] = x; s[6] = y; s[7] = vx; s[8] = vy;
    cam[0] = cx; cam[1] = cy; cam[2] = cz;
    return value;
}

/* robotarm5: pf_model.cl:52-131 in double */
static double ra5_f64(double* s, const double* z, const esth_robotarm5_control* c,
                      const esth_robotarm5_measurement* m, double dt) {
    double cam[3];
    double value = ra5_chain_f64(s, z + 4, z, c, m, dt, cam);
    value += (cam[0] - m->x) * (cam[0] - m->x) * (double)0.1f;
    value += (cam[1] - m->y) * (cam[1] - m->y) * (double)0.1f;
    return -value;
}

/* robotarm5_camera (pf_model_soa_packed_code.cuh:57-143) in double, the
 * projection in the Brown-Conrady form with tangential coefficients
 * p1 = kc3, p2 = kc4:
 *   xd = x (1 + k1 r^2 + k2 r^4) + 2 p1 x y + p2 (r^2 + 2 x^2),
 *   yd = y (1 + k1 r^2 + k2 r^4) + p1 (r^2 + 2 y^2) + 2 p2 x y,
 *   (u, v) = (fx xd + cx, fy yd + cy). */
static double ra5c_f64(double* s, const double* z, const esth_camera_params* p, const esth_robotarm5_control* c,
                       const esth_robotarm5_measurement* m, double dt) {
    double cam[3];
    double value = ra5_chain_f64(s, z, z + 5, c, m, dt, cam);
    const double x = cam[0] / cam[2], y = cam[1] / cam[2];
    const double r2 = x * x + y * y, r4 = r2 * r2;
    const double k1 = p->kc[0], k2 = p->kc[1], p1 = p->kc[2], p2 = p->kc[3];
    const double radial = 1.0 + k1 * r2 + k2 * r4;
    const double xd = x * radial + 2.0 * p1 * x * y + p2 * (r2 + 2.0 * x * x);
    const double yd = y * radial + p1 * (r2 + 2.0 * y * y) + 2.0 * p2 * x * y;
    const double u = (double)p->fc[0] * xd + (double)p->cc[0];
    const double v = (double)p->fc[1] * yd + (double)p->cc[1];
    value += (u - m->x) * (u - m->x) * (double)0.1f;
    value += (v - m->y) * (v - m->y) * (double)0.1f;
    return -value;
}

static double wrap_pi_f64(double a) {
    a = fmod(a + M_PI, 2.0 * M_PI);
    if (a <= 0.0) a += 2.0 * M_PI;
    return a - M_PI;
}

static double uni_f64(double* s, const double* z, const esth_unicycle_params* p,
                      const esth_unicycle_control* c, const esth_unicycle_measurement* m,
                      double dt) {
    double th0 = s[2];
    double x = s[0] + (double)c->v * dt * cos(th0) + (double)p->sigma_x * z[0];
    double y = s[1] + (double)c->v * dt * sin(th0) + (double)p->sigma_y * z[1];
    double th = th0 + (double)c->omega * dt + (double)p->sigma_theta * z[2];
    s[0] = x; s[1] = y; s[2] = th;
    double sr = p->sigma_range, sb = p->sigma_bearing, acc = 0.0;
    for (int j = 0; j < p->num_landmarks; ++j) {
        double dx = p->lx[j] - x, dy = p->ly[j] - y;
        double r = sqrt(dx * dx + dy * dy);
        double b = atan2(dy, dx) - th;
        double er = (m->range[j] - r) / sr;
        double eb = wrap_pi_f64((double)m->bearing[j] - b) / sb;
        acc += er * er + eb * eb;
    }
    return -0.5 * acc;
}

/* unicycle10: particle_filter.cu:313-359 in double, the arc as the reference
 * writes it (straight line only where w dt is below double resolution) */
static double u10_f64(double* s, const double* z, const esth_unicycle10_params* p,
                      const esth_unicycle_control* c, const esth_unicycle10_measurement* m, double dt) {
    double v = (double)c->v + (double)p->noise_v * z[0];
    double w = (double)c->omega + (double)p->noise_omega * z[1];
    double g = (double)p->noise_gamma * z[2];
    double th = s[2];
    if (fabs(w * dt) >= 1e-9) {
        s[0] += v / w * (sin(th + w * dt) - sin(th));
        s[1] -= v / w * (cos(th + w * dt) - cos(th));
    } else {
        s[0] += v * dt * cos(th);
        s[1] += v * dt * sin(th);
    }
    s[2] = th + w * dt + g * dt;
    double value = 0.0;
    for (int j = 0; j < 10; ++j) {
        double d1 = s[0] - p->sx[j], d2 = s[1] - p->sy[j], d3 = p->sz[j];
        double e = sqrt(d1 * d1 + d2 * d2 + d3 * d3) - (double)m->range[j];
        value += e * e;
    }
    return log(1e5) - 5.0 * log(2.0 * M_PI) - value;
}

/* randomwalk3 from its definition: p += v dt + sigma_q z, log w = -|m - p|^2 / (2 sigma_r^2) */
static double rw3_f64(double* s, const double* z, const esth_randomwalk3_params* p,
                      const esth_randomwalk3_control* c, const esth_randomwalk3_measurement* m, double dt) {
    double e2 = 0.0;
    for (int i = 0; i < 3; ++i) {
        s[i] += (double)c->v[i] * (double)(float)dt + (double)p->sigma_q * z[i];
        const double e = (double)m->p[i] - s[i];
        e2 += e * e;
    }
    return -0.5 * e2 / ((double)p->sigma_r * p->sigma_r);
}

static double dh6_f64(double* s, const double* z, const esth_dh6_params* p,
                      const esth_dh6_measurement* m, double dt) {
    for (int i = 0; i < 6; ++i) {
        double acc = (double)p->sigma_acc * z[i];
        s[i] = s[i] + s[6 + i] * dt + 0.5 * dt * dt * acc;
        s[6 + i] = s[6 + i] + acc * dt;
    }
    /* T = A_0 A_1 ... A_5 as 4x4 matrices, A_i = Rz(q) Tz(d) Tx(a) Rx(alpha) */
    double T[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    for (int i = 0; i < 6; ++i) {
        double cq = cos(s[i]), sq = sin(s[i]), ca = cos(p->alpha[i]), sa = sin(p->alpha[i]);
        double A[4][4] = {{cq, -sq * ca, sq * sa, p->a[i] * cq},
                          {sq, cq * ca, -cq * sa, p->a[i] * sq},
                          {0, sa, ca, p->d[i]},
                          {0, 0, 0, 1}};
        double R[4][4];
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) {
                double acc = 0;
                for (int k = 0; k < 4; ++k) acc += T[r][k] * A[k][c];
                R[r][c] = acc;
            }
        memcpy(T, R, sizeof(T));
    }
    double ep = 0.0, eq = 0.0;
    for (int i = 0; i < 3; ++i) { double e = m->p[i] - T[i][3]; ep += e * e; }
    for (int i = 0; i < 6; ++i) { double e = m->q[i] - s[i]; eq += e * e; }
    return -0.5 * (ep / ((double)p->sigma_pos * p->sigma_pos) + eq / ((double)p->sigma_enc * p->sigma_enc));
}

/* normals: N*V FP32 normals in the reference layout (already Box-Muller'd) */
int or_sample_weight_f64(int model, const float* state_in, uint32_t N, const float* normals,
                         const void* params, const void* ctrl, const void* meas, float dt,
                         double* state_out, double* lw_out) {
    int S, V;
    if (esth_model_dims(model, &S, &V)) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)N; ++i) {
        double s[ESTH_MAX_STATE_DIM], z[ESTH_MAX_STATE_DIM];
        for (int c = 0; c < S; ++c) s[c] = state_in[(size_t)c * N + i];
        for (int k = 0; k < V; ++k) z[k] = normals[(size_t)k * N + i];
        double lw;
        if (model == ESTH_MODEL_ROBOTARM5)
            lw = ra5_f64(s, z, (const esth_robotarm5_control*)ctrl,
                         (const esth_robotarm5_measurement*)meas, dt);
        else if (model == ESTH_MODEL_UNICYCLE)
            lw = uni_f64(s, z, (const esth_unicycle_params*)params,
                         (const esth_unicycle_control*)ctrl,
                         (const esth_unicycle_measurement*)meas, dt);
        else if (model == ESTH_MODEL_UNICYCLE10)
            lw = u10_f64(s, z, (const esth_unicycle10_params*)params, (const esth_unicycle_control*)ctrl,
                         (const esth_unicycle10_measurement*)meas, dt);
        else if (model == ESTH_MODEL_ROBOTARM5_CAMERA)
            lw = ra5c_f64(s, z, (const esth_camera_params*)params, (const esth_robotarm5_control*)ctrl,
                          (const esth_robotarm5_measurement*)meas, dt);
        else if (model == ESTH_MODEL_RANDOMWALK3)
            lw = rw3_f64(s, z, (const esth_randomwalk3_params*)params, (const esth_randomwalk3_control*)ctrl,
                         (const esth_randomwalk3_measurement*)meas, dt);
        else if (model == ESTH_MODEL_DH6)
            lw = dh6_f64(s, z, (const esth_dh6_params*)params, (const esth_dh6_measurement*)meas, dt);
        else
            lw = NAN;   /* a listed model without a twin: its FP64 posterior reads NaN */
        for (int c = 0; c < S; ++c) state_out[(size_t)c * N + i] = s[c];
        lw_out[i] = lw;
    }
    return 0;
}

/* ========================================================= bitonic sort
 * block_sort (pf_generic.cl:252-291) calling bitonic_sort/comparator
 * (bitonic_sort.cl:11-61) with dir = 0: m/2 work-items, each stage one
 * barrier phase; comparator swaps iff (w[a] > w[b]) == ddd, so equal keys
 * swap when ddd == 0. */
static void bitonic_block(float* w, int* idx, uint32_t m) {
    const uint32_t half = m / 2, dir = 0;
    for (uint32_t size = 2; size < m; size <<= 1) {
        for (uint32_t stride = size / 2; stride > 0; stride >>= 1) {
            for (uint32_t tid = 0; tid < half; ++tid) {
                uint32_t ddd = dir ^ ((tid & (size / 2)) != 0);
                uint32_t pos = 2 * tid - (tid & (stride - 1));
                if ((w[pos] > w[pos + stride]) == (int)ddd) {
                    float t = w[pos]; w[pos] = w[pos + stride]; w[pos + stride] = t;
                    int ti = idx[pos]; idx[pos] = idx[pos + stride]; idx[pos + stride] = ti;
                }
            }
        }
    }
    for (uint32_t stride = m / 2; stride > 0; stride >>= 1) {
        for (uint32_t tid = 0; tid < half; ++tid) {
            uint32_t pos = 2 * tid - (tid & (stride - 1));
            if ((w[pos] > w[pos + stride]) == (int)dir) {
                float t = w[pos]; w[pos] = w[pos + stride]; w[pos + stride] = t;
                int ti = idx[pos]; idx[pos] = idx[pos + stride]; idx[pos + stride] = ti;
            }
        }
    }
}

void or_block_sort(float* w, const float* state_in, float* state_out, int S, uint32_t m,
                   uint32_t G, float* winner, int* perm) {
    const size_t N = (size_t)m * G;
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < (int64_t)G; ++g) {
        float sw[ESTH_MAX_NUM_PARTICLES];
        int si[ESTH_MAX_NUM_PARTICLES];
        const size_t base = (size_t)g * m;
        for (uint32_t i = 0; i < m; ++i) { sw[i] = w[base + i]; si[i] = (int)i; }
        bitonic_block(sw, si, m);
        for (uint32_t i = 0; i < m; ++i) {
            w[base + i] = sw[i];
            for (int c = 0; c < S; ++c)
                state_out[(size_t)c * N + base + i] = state_in[(size_t)c * N + base + si[i]];
            if (perm) perm[base + i] = si[i];
        }
        if (winner) { winner[2 * g] = sw[0]; winner[2 * g + 1] = (float)g; }
    }
}

/* =============================================================== argmax
 * reduce, reduce.cl:19-55, launched as 1 work-group of 128 (pf_generic.c:604-612):
 * strided sequential pass keeping the later element on ties, then a tree in
 * which `other` wins ties.  The accumulator starts at (-1,-1); with
 * log-weights it starts at (-inf,-1). */
void or_argmax(const float* winner, uint32_t G, int weights, float* result) {
    float sx[128], sy[128];
    const float init = (weights == ESTH_WEIGHTS_LOG) ? -INFINITY : -1.0f;
    for (int l = 0; l < 128; ++l) {
        float ax = init, ay = -1.0f;
        for (uint32_t gi = (uint32_t)l; gi < G; gi += 128) {
            float ex = winner[2 * gi], ey = winner[2 * gi + 1];
            if (!(ax > ex)) { ax = ex; ay = ey; }
        }
        sx[l] = ax; sy[l] = ay;
    }
    for (int off = 64; off > 0; off >>= 1) {
        for (int l = 0; l < off; ++l) {
            if (!(sx[l] > sx[l + off])) { sx[l] = sx[l + off]; sy[l] = sy[l + off]; }
        }
    }
    result[0] = sx[0];
    result[1] = sy[0];
}

/* ============================================================= exchange
 * On sorted sub-filters (best first), block b's worst slots receive
 * neighbours' best t particles.
 *  ring  : exchange_ring, pf_generic.cl:368-383:
 *          slot m-1-k <- block (b-1) mod G, rank k.
 *  torus : exchange_2dtorus, cuda/pf_generic.cu:352-393, with the slot
 *          offsets corrected to stay inside the receiving block:
 *          a sender writes its rank-k particle to its left neighbour's slot
 *          m-1-k, right's m-1-k-t, top's m-1-k-2t, bottom's m-1-k-3t;
 *          x = b % dimx, y = b / dimx, wrap-around.
 *  all2all: every block's slot m-1-k <- the k-th best of all blocks' best t
 *          (cuda/pf_generic2_all2all.cu:404-419 after sort_by_key at :1007),
 *          ties by ascending candidate index g*t + k.
 * All reads see the pre-exchange sorted state (the reference's in-place ring
 * reads only ranks < t, which it never writes while t <= m/2). */
static int cmp_cand(const void* a, const void* b) {
    const float* x = (const float*)a;
    const float* y = (const float*)b;
    if (x[0] > y[0]) return -1;
    if (x[0] < y[0]) return 1;
    return (x[1] < y[1]) ? -1 : (x[1] > y[1]) ? 1 : 0;
}

int or_exchange(float* state, float* w, int S, uint32_t m, uint32_t G, uint32_t t, int topo,
                int dimx, int dimy) {
    if (t == 0 || topo == ESTH_TOPO_NONE) return 0;
    const size_t N = (size_t)m * G;
    const uint32_t nq = (topo == ESTH_TOPO_TORUS) ? 4 : 1;
    if ((uint64_t)nq * t > m / 2 && topo != ESTH_TOPO_TORUS) return -1;
    if (topo == ESTH_TOPO_TORUS && 4ull * t > m) return -1;
    if (topo == ESTH_TOPO_TORUS) {
        if (dimx <= 0 || dimy <= 0) esth_torus_dims(G, &dimx, &dimy);
        if ((uint64_t)dimx * (uint64_t)dimy != G) return -1;
    }
    /* snapshot of every block's best t */
    float* bw = (float*)malloc(sizeof(float) * (size_t)G * t);
    float* bs = (float*)malloc(sizeof(float) * (size_t)G * t * S);
    for (uint32_t g = 0; g < G; ++g)
        for (uint32_t k = 0; k < t; ++k) {
            bw[(size_t)g * t + k] = w[(size_t)g * m + k];
            for (int c = 0; c < S; ++c)
                bs[((size_t)c * G + g) * t + k] = state[(size_t)c * N + (size_t)g * m + k];
        }
    uint32_t* top = NULL;
    if (topo == ESTH_TOPO_ALL2ALL) {
        size_t nc = (size_t)G * t;
        float* cand = (float*)malloc(sizeof(float) * 2 * nc);
        for (size_t i = 0; i < nc; ++i) { cand[2 * i] = bw[i]; cand[2 * i + 1] = (float)i; }
        qsort(cand, nc, 2 * sizeof(float), cmp_cand);
        top = (uint32_t*)malloc(sizeof(uint32_t) * t);
        for (uint32_t k = 0; k < t; ++k) top[k] = (uint32_t)cand[2 * k + 1];
        free(cand);
    }
    for (uint32_t b = 0; b < G; ++b) {
        for (uint32_t q = 0; q < nq; ++q) {
            for (uint32_t k = 0; k < t; ++k) {
                size_t src;
                if (topo == ESTH_TOPO_RING) {
                    src = (size_t)((b + G - 1) % G) * t + k;
                } else if (topo == ESTH_TOPO_ALL2ALL) {
                    src = top[k];
                } else {
                    int x = (int)(b % (uint32_t)dimx), y = (int)(b / (uint32_t)dimx);
                    int sx = x, sy = y;
                    if (q == 0) sx = (x + 1) % dimx;            /* from the right neighbour */
                    else if (q == 1) sx = (x - 1 + dimx) % dimx; /* from the left */
                    else if (q == 2) sy = (y + 1) % dimy;        /* from the block below */
                    else sy = (y - 1 + dimy) % dimy;             /* from the block above */
                    src = (size_t)(sy * dimx + sx) * t + k;
                }
                size_t dst = (size_t)b * m + (m - 1 - k - q * t);
                w[dst] = bw[src];
                for (int c = 0; c < S; ++c)
                    state[(size_t)c * N + dst] = bs[((size_t)c * G) * t + src];
            }
        }
    }
    free(bw);
    free(bs);
    free(top);
    return 0;
}

/* ============================================================ resampling */

/* prefix_sum, pf_generic.cl:130-177: exclusive Blelloch scan in place over m
 * elements by m work-items; returns the total from the up-sweep. */
static float prefix_sum(float* el, uint32_t n) {
    uint32_t offset = 1;
    for (uint32_t d = n / 2; d > 0; d /= 2) {
        for (uint32_t tid = 0; tid < d; ++tid) {
            uint32_t ai = offset * (2 * tid + 1) - 1;
            uint32_t bi = offset * (2 * tid + 2) - 1;
            el[bi] += el[ai];
        }
        offset *= 2;
    }
    float cum = el[n - 1];
    el[n - 1] = 0.0f;
    for (uint32_t d = 1; d < n; d *= 2) {
        offset /= 2;
        for (uint32_t tid = 0; tid < d; ++tid) {
            uint32_t ai = offset * (2 * tid + 1) - 1;
            uint32_t bi = offset * (2 * tid + 2) - 1;
            float t = el[ai];
            el[ai] = el[bi];
            el[bi] += t;
        }
    }
    return cum;
}

/* binary_search, pf_generic.cl:182-201: largest i with haystack[i] <= needle
 * (returns 0 if none). */
static int binary_search(const float* h, float needle, int size) {
    int mn = 0, mx = size - 1;
    do {
        int cur = mn + (mx - mn + 1) / 2;
        if (h[cur] > needle) mx = cur - 1;
        else mn = cur;
    } while (mn < mx);
    return mn;
}

static float block_max(const float* w, uint32_t m) {
    float mx = w[0];
    for (uint32_t i = 1; i < m; ++i) mx = (w[i] > mx) ? w[i] : mx;
    return mx;
}

/* sum_reduce4, resampling-valias.cl:85-104 */
static float sum_reduce4(float* arr, uint32_t n) {
    uint32_t offset = n;
    while (offset > 2) {
        offset >>= 2;
        for (uint32_t l = 0; l < offset; ++l)
            arr[l] += arr[l + offset] + arr[l + offset + offset] + arr[l + offset + offset + offset];
    }
    if (offset > 1) arr[0] += arr[1];
    return arr[0];
}

/* vose_alias_init, resampling-valias.cl:111-195, with the local atomics taken
 * in ascending work-item order (the canonical order; the reference's order
 * is whatever its runtime produced). prob holds the normalised p on entry. */
static void vose_init(float* prob, uint32_t* alias, uint32_t m) {
    uint32_t sl[ESTH_MAX_NUM_PARTICLES];
    uint32_t small_end = 0, large_rend = m - 1;
    for (uint32_t l = 0; l < m; ++l) {
        prob[l] *= (float)m;
        alias[l] = l;
    }
    for (uint32_t l = 0; l < m; ++l) {
        if (prob[l] <= 1.0f) sl[small_end++] = l;
        else sl[large_rend--] = l;
    }
    uint32_t small_begin = 0;
    for (;;) {
        uint32_t se = small_end;
        uint32_t a = se - small_begin, b = m - se;
        uint32_t nactive = a < b ? a : b;
        if (nactive == 0) break;
        large_rend += nactive;
        uint32_t sidx[ESTH_MAX_NUM_PARTICLES], lidx[ESTH_MAX_NUM_PARTICLES];
        float pS[ESTH_MAX_NUM_PARTICLES], pL[ESTH_MAX_NUM_PARTICLES];
        for (uint32_t l = 0; l < nactive; ++l) {
            sidx[l] = sl[small_begin + l];
            lidx[l] = sl[se + l];
            pS[l] = prob[sidx[l]];
            alias[sidx[l]] = lidx[l];
            pL[l] = prob[lidx[l]];
        }
        for (uint32_t l = 0; l < nactive; ++l) {
            float probL = (pL[l] + pS[l]) - 1.0f;
            uint32_t old;
            if (probL <= 1.0f) old = small_end++;
            else old = large_rend--;
            sl[old] = lidx[l];
            prob[lidx[l]] = probL;
        }
        small_begin += nactive;
    }
}

/* normalised probabilities of one block (resampling-valias.cl:262-270);
 * log-weights are first mapped to exp(lw - max). */
static void alias_probs(const float* wblk, uint32_t m, int weights, float* prob) {
    float scratch[ESTH_MAX_NUM_PARTICLES], p[ESTH_MAX_NUM_PARTICLES];
    float mx = (weights == ESTH_WEIGHTS_LOG) ? block_max(wblk, m) : 0.0f;
    for (uint32_t l = 0; l < m; ++l) {
        p[l] = (weights == ESTH_WEIGHTS_LOG) ? esth_expf(wblk[l] - esth_max_shift(mx)) : wblk[l];
        scratch[l] = p[l];
    }
    float sum = sum_reduce4(scratch, m);
    for (uint32_t l = 0; l < m; ++l) {
        if (sum > 0.0f) prob[l] = p[l] / sum;
        else prob[l] = 1.0f / (float)m;
    }
}

void or_alias_table(const float* w, uint32_t m, uint32_t G, int weights, float* prob,
                    uint32_t* alias) {
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < (int64_t)G; ++g) {
        alias_probs(w + (size_t)g * m, m, weights, prob + (size_t)g * m);
        vose_init(prob + (size_t)g * m, alias + (size_t)g * m, m);
    }
}

/* resampling (pf_generic.cl:385-432) and resampling_valias
 * (resampling-valias.cl:214-304).  rnd points at the uniform region
 * (d_random + rand_offset), values in [1,2). */
int or_resample(int resampler, float* dst, const float* src, int S, const float* w, const float* rnd,
                uint32_t m, uint32_t G, float thr, int adaptive, int weights, int* sel) {
    const size_t N = (size_t)m * G;
    if (adaptive && resampler != ESTH_RESAMPLE_ALIAS) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < (int64_t)G; ++g) {
        const size_t base = (size_t)g * m;
        if (!(resampler == ESTH_RESAMPLE_ALIAS && adaptive)) {
            if ((rnd[N + g] - 1.0f) > thr) {        /* gate, pf_generic.cl:403 */
                if (sel) for (uint32_t i = 0; i < m; ++i) sel[base + i] = -1;
                continue;
            }
        }
        int pick[ESTH_MAX_NUM_PARTICLES];
        if (resampler == ESTH_RESAMPLE_ALIAS) {
            float prob[ESTH_MAX_NUM_PARTICLES];
            uint32_t alias[ESTH_MAX_NUM_PARTICLES];
            alias_probs(w + base, m, weights, prob);
            if (adaptive) {
                float sq[ESTH_MAX_NUM_PARTICLES];
                for (uint32_t l = 0; l < m; ++l) sq[l] = prob[l] * prob[l];
                float recip_ess = sum_reduce4(sq, m);
                if (!(recip_ess > thr)) {
                    if (sel) for (uint32_t i = 0; i < m; ++i) sel[base + i] = -1;
                    continue;
                }
            }
            vose_init(prob, alias, m);
            for (uint32_t l = 0; l < m; ++l) {
                size_t idx = base + l;
                float r0 = rnd[2 * idx] - 1.0f, r1 = rnd[2 * idx + 1] - 1.0f;
                uint32_t col = (uint32_t)floorf(r0 * (float)m);  /* vose_alias_generate_next, :198-202 */
                pick[l] = (r1 < prob[col]) ? (int)col : (int)alias[col];
            }
        } else {
            float cdf[ESTH_MAX_NUM_PARTICLES];
            float mx = (weights == ESTH_WEIGHTS_LOG) ? block_max(w + base, m) : 0.0f;
            for (uint32_t l = 0; l < m; ++l)
                cdf[l] = (weights == ESTH_WEIGHTS_LOG) ? esth_expf(w[base + l] - esth_max_shift(mx)) : w[base + l];
            float cum = prefix_sum(cdf, m);
            if (resampler == ESTH_RESAMPLE_MULTINOMIAL) {
                for (uint32_t l = 0; l < m; ++l) {
                    float s = (rnd[base + l] - 1.0f) * cum;
                    pick[l] = binary_search(cdf, s, (int)m);
                }
            } else {
                /* systematic: one uniform per sub-filter (the block's first
                 * selection uniform), positions (l + u) * cum / m */
                float u = rnd[base] - 1.0f;
                float step = cum / (float)m;
                for (uint32_t l = 0; l < m; ++l) {
                    float s = ((float)l + u) * step;
                    pick[l] = binary_search(cdf, s, (int)m);
                }
            }
        }
        for (uint32_t l = 0; l < m; ++l) {
            for (int c = 0; c < S; ++c)
                dst[(size_t)c * N + base + l] = src[(size_t)c * N + base + (size_t)pick[l]];
            if (sel) sel[base + l] = pick[l];
        }
    }
    return 0;
}

/* ================================================================= math */
float or_expf(float x) { return esth_expf(x); }
float or_logf(float x) { return esth_logf(x); }
void or_sincosf(float x, float* s, float* c) { esth_sincosf(x, s, c); }
float or_atan2f(float y, float x) { return esth_atan2f(y, x); }
void or_polar_abs(float u, float v, float* r, float* a) { esth_polar_abs(u, v, r, a); }
float or_rcpf(float x) { return esth_rcpf(x); }

/* =========================================================== whole step
 * particle_filter's loop body, pf_generic.c:438-808. */
struct or_filter {
    esth_filter_desc d;
    int S, V, K;
    size_t N;
    uint64_t needed, size, words;
    const uint32_t* params;   /* 200 x 52 */
    uint32_t* status;         /* K x 351 */
    float* R;                 /* words */
    float* state;             /* S x N */
    float* tmp;               /* S x N */
    float* w;                 /* N */
    float winner_best[2];
    float best_state[ESTH_MAX_STATE_DIM];
    double* mean32;
    double* mean64;
    double* st64;
    double* lw64;
    unsigned char mparams[2048];
    int dimx, dimy;
    int f64;
};

or_filter* or_filter_create(const esth_filter_desc* d, const void* params, size_t params_size,
                            const uint32_t* mtgp_params) {
    or_filter* f = (or_filter*)calloc(1, sizeof(or_filter));
    f->d = *d;
    if (esth_model_dims(d->model, &f->S, &f->V)) { free(f); return NULL; }
    f->N = (size_t)d->num_particles * d->num_blocks;
    f->needed = esth_rng_needed(f->V, d->resampler, d->adaptive, d->num_particles, d->num_blocks);
    f->K = esth_rng_streams(d->rng, f->needed);
    f->size = esth_rng_stream_size(f->K, f->needed);
    f->words = f->size * (uint64_t)f->K;
    f->params = mtgp_params;
    f->status = (uint32_t*)malloc(sizeof(uint32_t) * ESTH_MTGP_N * (size_t)f->K);
    for (int j = 0; j < f->K; ++j)   /* make_kernel_data32, mtgp32-opencl.c:46-48 */
        or_mtgp_seed(mtgp_params + (size_t)j * ESTH_MTGP_PARAM_WORDS, d->seed + (uint32_t)j + 1u,
                     f->status + (size_t)j * ESTH_MTGP_N);
    f->R = (float*)malloc(sizeof(float) * f->words);
    f->state = (float*)calloc((size_t)f->S * f->N, sizeof(float));
    f->tmp = (float*)calloc((size_t)f->S * f->N, sizeof(float));
    f->w = (float*)calloc(f->N, sizeof(float));
    f->mean32 = (double*)calloc(ESTH_MAX_STATE_DIM, sizeof(double));
    f->mean64 = (double*)calloc(ESTH_MAX_STATE_DIM, sizeof(double));
    f->st64 = (double*)malloc(sizeof(double) * (size_t)f->S * f->N);
    f->lw64 = (double*)malloc(sizeof(double) * f->N);
    if (params && params_size <= sizeof(f->mparams)) memcpy(f->mparams, params, params_size);
    f->dimx = d->torus_dimx;
    f->dimy = d->torus_dimy;
    if (d->topology == ESTH_TOPO_TORUS && (f->dimx <= 0 || f->dimy <= 0))
        esth_torus_dims(d->num_blocks, &f->dimx, &f->dimy);
    f->f64 = 1;
    return f;
}

void or_filter_destroy(or_filter* f) {
    if (!f) return;
    free(f->status); free(f->R); free(f->state); free(f->tmp); free(f->w);
    free(f->mean32); free(f->mean64); free(f->st64); free(f->lw64);
    free(f);
}

/* with carried weights a new particle set starts from equal weights */
static void reset_weights(or_filter* f) {
    const float w0 = (f->d.carry_weights && f->d.weights != ESTH_WEIGHTS_LOG) ? 1.0f : 0.0f;
    for (size_t i = 0; i < f->N; ++i) f->w[i] = w0;
}

void or_filter_init_uniform(or_filter* f, const float* s) {
    for (int c = 0; c < f->S; ++c)
        for (size_t i = 0; i < f->N; ++i) f->state[(size_t)c * f->N + i] = s[c];
    reset_weights(f);
}

void or_filter_set_state(or_filter* f, const float* soa) {
    memcpy(f->state, soa, sizeof(float) * (size_t)f->S * f->N);
    reset_weights(f);
}

static void posterior_f64(const double* st, const double* lw, size_t N, int S, double* mean) {
    double mx = -INFINITY;
    for (size_t i = 0; i < N; ++i) mx = lw[i] > mx ? lw[i] : mx;
    if (mx == -INFINITY) mx = 0.0;   /* every weight 0 */
    double sw = 0.0, acc[ESTH_MAX_STATE_DIM] = {0};
    for (size_t i = 0; i < N; ++i) {
        double wi = exp(lw[i] - mx);
        sw += wi;
        for (int c = 0; c < S; ++c) acc[c] += wi * st[(size_t)c * N + i];
    }
    for (int c = 0; c < S; ++c) mean[c] = acc[c] / sw;
}

int or_filter_step(or_filter* f, const void* ctrl, const void* meas, float dt) {
    const esth_filter_desc* d = &f->d;
    const uint32_t m = d->num_particles, G = d->num_blocks;
    const size_t N = f->N;
    float frame[ESTH_FRAME_FLOATS];
    if (esth_build_frame_impl(d->model, f->mparams, ctrl, meas, dt, frame)) return -1;
    /* generate_normal_random, pf_generic.c:191-232 */
    or_mtgp_fill(f->params, f->K, 0, f->status, (int)f->size, (uint32_t*)f->R, 1);
    or_boxmuller(f->R, (int64_t)N * f->V);
    const int carry = d->carry_weights != 0;
    const int logw = d->weights == ESTH_WEIGHTS_LOG;
    float* w_prev = NULL;   /* carried weights, aligned with the particles */
    if (carry) {
        w_prev = (float*)malloc(sizeof(float) * N);
        memcpy(w_prev, f->w, sizeof(float) * N);
    }
    /* FP64 twin from the same prior and normals (lock-step posterior) */
    if (f->f64) {
    double* pre = f->st64;
    float* prior = (float*)malloc(sizeof(float) * (size_t)f->S * N);
    memcpy(prior, f->state, sizeof(float) * (size_t)f->S * N);
    or_sample_weight_f64(d->model, prior, (uint32_t)N, f->R, f->mparams, ctrl, meas, dt, pre, f->lw64);
    if (carry)
        for (size_t i = 0; i < N; ++i) f->lw64[i] += logw ? (double)w_prev[i] : log((double)w_prev[i]);
    posterior_f64(pre, f->lw64, N, f->S, f->mean64);
    free(prior);
    }
    /* sampling_importance, pf_generic.c:521-533 */
    or_sample_weight(d->model, f->state, (uint32_t)N, f->R, frame, d->weights, f->w);
    if (carry) {   /* w = w_prev * likelihood (w_prev + log-likelihood) */
        for (size_t i = 0; i < N; ++i) f->w[i] = logw ? w_prev[i] + f->w[i] : w_prev[i] * f->w[i];
        free(w_prev);
    }
    /* FP64 posterior of the FP32 path's particles and weights */
    if (f->f64) {
        double* lw = f->lw64;
        double* st = f->st64;
        for (size_t i = 0; i < N; ++i)
            lw[i] = (d->weights == ESTH_WEIGHTS_LOG) ? (double)f->w[i] : log((double)f->w[i]);
        for (size_t i = 0; i < (size_t)f->S * N; ++i) st[i] = f->state[i];
        posterior_f64(st, lw, N, f->S, f->mean32);
    }
    /* block_sort, pf_generic.c:572-585 */
    float* winner = (float*)malloc(sizeof(float) * 2 * G);
    or_block_sort(f->w, f->state, f->tmp, f->S, m, G, winner, NULL);
    /* reduce, pf_generic.c:604-615 */
    or_argmax(winner, G, d->weights, f->winner_best);
    free(winner);
    if (f->winner_best[1] >= 0.0f)
        for (int c = 0; c < f->S; ++c)
            f->best_state[c] = f->tmp[(size_t)c * N + (size_t)f->winner_best[1] * m];
    /* exchange, pf_generic.c:702-715 */
    if (d->num_transfer > 0 && d->topology != ESTH_TOPO_NONE)
        if (or_exchange(f->tmp, f->w, f->S, m, G, d->num_transfer, d->topology, f->dimx, f->dimy))
            return -1;
    /* resampling, pf_generic.c:732-757, uniforms at rand_offset = Nn */
    int* sel = carry ? (int*)malloc(sizeof(int) * N) : NULL;
    if (or_resample(d->resampler, f->state, f->tmp, f->S, f->w, f->R + N * f->V, m, G,
                    d->resampling_threshold, d->adaptive, d->weights, sel)) {
        free(sel);
        return -1;
    }
    if (carry) {   /* skipped: keep the sorted, exchanged particles and weights; resampled: equal weights */
        for (uint32_t g = 0; g < G; ++g) {
            const size_t base = (size_t)g * m;
            if (sel[base] < 0) {
                for (int c = 0; c < f->S; ++c)
                    memcpy(f->state + (size_t)c * N + base, f->tmp + (size_t)c * N + base, sizeof(float) * m);
            } else {
                for (uint32_t l = 0; l < m; ++l) f->w[base + l] = logw ? 0.0f : 1.0f;
            }
        }
        free(sel);
    }
    return 0;
}

void or_filter_read(or_filter* f, float* state, float* w) {
    if (state) memcpy(state, f->state, sizeof(float) * (size_t)f->S * f->N);
    if (w) memcpy(w, f->w, sizeof(float) * f->N);
}

void or_filter_estimate(or_filter* f, double* m32, double* m64, float* best) {
    if (m32) memcpy(m32, f->mean32, sizeof(double) * f->S);
    if (m64) memcpy(m64, f->mean64, sizeof(double) * f->S);
    if (best) { best[0] = f->winner_best[0]; best[1] = f->winner_best[1]; }
}

/* the best particle of the best sub-filter of the last step: position 0 of
 * its sorted block, as the argmax found it (pf_generic.c:634-668 reads the
 * sorted copy before the exchange) */
void or_filter_best_particle(or_filter* f, float* out) {
    for (int c = 0; c < f->S; ++c) out[c] = f->best_state[c];
}

uint64_t or_filter_rng_words(or_filter* f) { return f->words; }

/* 0: time only the FP32 restatement (the CPU baseline), 1: also the FP64 twin */
void or_filter_set_f64(or_filter* f, int on) { f->f64 = on; }

void or_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}
