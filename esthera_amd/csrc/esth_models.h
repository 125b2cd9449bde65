/*
 * esth_models.h - model plug-ins: the per-particle sample+weight hook and the
 * per-step frame builder, shared by the gfx950 kernels and the host.
 *
 * Hook contract (reference: src/models/robotarm/opencl/pf_model.cl:33-131):
 *   sampling_iw(state, noise, frame) propagates one particle in place with
 *   its standard normals and returns its log importance weight.  The engine
 *   turns that into exp(lw) for the reference's linear weights
 *   (pf_model.cl:129) or keeps it in the log domain.
 *   particle_set (pf_model.cl:33-48) is the engine's generic per-plane copy;
 *   state is structure-of-arrays, one float plane per state variable.
 *
 * Frame: everything a step's hook needs that is uniform over particles
 * (dt, control, measurement, model constants), precomputed once per step on
 * the host into ESTH_FRAME_FLOATS floats, read by the kernels with scalar
 * loads.  Arithmetic is written operation by operation with explicit fmaf
 * where a fused multiply-add is meant, so host and device agree bit for bit.
 */
#ifndef ESTH_MODELS_H
#define ESTH_MODELS_H

#include "esth_math.h"
#include "esthera.h"

/* float offset of state variable c of particle i among N particles of S
 * variables in an ESTH_LAYOUT_* layout */
ESTH_HD size_t esth_state_index(int layout, int S, size_t N, int c, size_t i) {
    if (layout == ESTH_LAYOUT_AOS) return i * (size_t)S + (size_t)c;
    if (layout == ESTH_LAYOUT_PACKED) {
        const int q = c & ~3;
        const int w = S - q < 4 ? S - q : 4;
        return (size_t)q * N + i * (size_t)w + (size_t)(c - q);
    }
    return (size_t)c * N + i;
}

/* ------------------------------------------------------------ robotarm5
 * src/models/robotarm/opencl/pf_model.cl:52-131, pf_model.h:26-32.
 * State planes: 0..4 joint angles, 5 x, 6 y, 7 vx, 8 vy.
 * Noise k (d_random[k*N + idx], pf_model.cl:128-158): 0 x, 1 y, 2 vx, 3 vy,
 * 4..8 joint angles. */
#define ESTH_RA5_S 9
#define ESTH_RA5_V 9
/* frame: [0] dt, [1..5] control angles, [6..10] measured angles, [11] x, [12] y */
#define ESTH_RA5_NOISE_ANGLES 0.2f   /* pf_model.cl:3 */
#define ESTH_RA5_NOISE_X      0.2f
#define ESTH_RA5_NOISE_Y      0.2f
#define ESTH_RA5_NOISE_V_X    2.0f
#define ESTH_RA5_NOISE_V_Y    2.0f
#define ESTH_RA5_NOISE_ANGLE_MEASUREMENTS 0.1f  /* pf_model.h:28-30 */
#define ESTH_RA5_NOISE_CAMERA_X 0.1f
#define ESTH_RA5_NOISE_CAMERA_Y 0.1f

ESTH_HD float esth_ra5_arm_length(int i) { return i == 0 ? 0.0f : 3.0f; } /* pf_model.cl:13 */

/* The base's random walk and the joint chain (pf_model.cl:70-117):
 * zp = noise of x, y, vx, vy; za = noise of the five joint angles.  Leaves
 * the propagated particle in s, the camera-frame point of the base in cam
 * and returns the weighted squared angle-measurement errors. */
ESTH_HD float esth_ra5_chain(float* s, const float* za, const float* zp, const float* fr, float* cam) {
    const float dt = fr[0];
    float value = 0.0f;
    float x = s[5], y = s[6], vx = s[7], vy = s[8];
    x = x + ((vx * dt) + (zp[0] * ESTH_RA5_NOISE_X));
    y = y + ((vy * dt) + (zp[1] * ESTH_RA5_NOISE_Y));
    vx = vx + (zp[2] * ESTH_RA5_NOISE_V_X);
    vy = vy + (zp[3] * ESTH_RA5_NOISE_V_Y);
    float camX, camY, camZ;
    /* joint 0, pf_model.cl:76-90 */
    {
        float a = s[0] + ((fr[1] * dt) + (za[0] * ESTH_RA5_NOISE_ANGLES));
        float sf, cf;
        esth_sincosf(a, &sf, &cf);
        camX = (x * cf) + (y * sf);
        camY = ((-x) * sf) + (y * cf);
        camZ = 0.0f;
        float e = a - fr[6];
        value = value + ((e * e) * ESTH_RA5_NOISE_ANGLE_MEASUREMENTS);
        s[0] = a;
    }
    /* joints 1..4, pf_model.cl:91-117 */
    for (int i = 1; i < 5; ++i) {
        float a = s[i] + ((fr[1 + i] * dt) + (za[i] * ESTH_RA5_NOISE_ANGLES));
        float sf, cf;
        esth_sincosf(a, &sf, &cf);
        float e = a - fr[6 + i];
        value = value + ((e * e) * ESTH_RA5_NOISE_ANGLE_MEASUREMENTS);
        float nB = (camY * cf) - (camZ * sf);
        float nC = (camY * sf) + (camZ * cf);
        camY = nB;
        camZ = nC - esth_ra5_arm_length(i);
        s[i] = a;
    }
    s[5] = x; s[6] = y; s[7] = vx; s[8] = vy;
    cam[0] = camX; cam[1] = camY; cam[2] = camZ;
    return value;
}

ESTH_HD float esth_robotarm5_sampling_iw(float* s, const float* z, const float* fr) {
    float cam[3];
    float value = esth_ra5_chain(s, z + 4, z, fr, cam);
    /* camera error, pf_model.cl:124-125 */
    value = value + (((cam[0] - fr[11]) * (cam[0] - fr[11])) * ESTH_RA5_NOISE_CAMERA_X);
    value = value + (((cam[1] - fr[12]) * (cam[1] - fr[12])) * ESTH_RA5_NOISE_CAMERA_Y);
    return -value;
}

/* ----------------------------------------------------- robotarm5_camera
 * The arm seen through a calibrated camera
 * (src/models/robotarm/cuda/pf_model_soa_packed_code.cuh:33-43,57-143): the
 * same walk and joint chain, noise in the CUDA order (0..4 joint angles,
 * 5 x, 6 y, 7 vx, 8 vy, :79-83), and the camera-frame point projected to
 * pixels by the pinhole + Brown-Conrady model before it is compared with the
 * measurement (:119-131):
 *   xn = X/Z, yn = Y/Z, r2 = xn^2 + yn^2, k = 1 + kc1 r2 + kc2 r2^2,
 *   u = fc_x (k xn + 2 kc3 xn yn + kc4 (r2 + 2 xn^2)) + cc_x,
 *   v = fc_y (k yn + kc3 (r2 + 2 yn^2) + 2 kc4 xn yn) + cc_y,
 *   log w = -(angle terms + 0.1 (u - mu)^2 + 0.1 (v - mv)^2).
 * A point whose projection overflows gives +inf, never NaN: a NaN cost is
 * taken as +inf (weight 0), so no NaN key reaches the sort.
 * State planes and frame as robotarm5, measurement x, y in pixels; frame
 * [13] fc_x, [14] fc_y, [15] cc_x, [16] cc_y, [17..20] kc1..kc4. */
#define ESTH_RA5C_FLOATS 21

ESTH_HD float esth_robotarm5_camera_sampling_iw(float* s, const float* z, const float* fr) {
    float cam[3];
    float value = esth_ra5_chain(s, z, z + 5, fr, cam);
    const float xn = cam[0] / cam[2];
    const float yn = cam[1] / cam[2];
    const float xx = xn * xn, yy = yn * yn, xy = xn * yn;
    const float r2 = xx + yy;
    const float kc1 = fr[17], kc2 = fr[18], kc3 = fr[19], kc4 = fr[20];
    const float radial = esth_fma(esth_fma(kc2, r2, kc1), r2, 1.0f);
    const float tx = esth_fma(2.0f * kc3, xy, kc4 * esth_fma(2.0f, xx, r2));
    const float ty = esth_fma(kc3, esth_fma(2.0f, yy, r2), (2.0f * kc4) * xy);
    const float u = esth_fma(fr[13], esth_fma(radial, xn, tx), fr[15]);
    const float v = esth_fma(fr[14], esth_fma(radial, yn, ty), fr[16]);
    const float eu = u - fr[11], ev = v - fr[12];
    value = value + ((eu * eu) * ESTH_RA5_NOISE_CAMERA_X);
    value = value + ((ev * ev) * ESTH_RA5_NOISE_CAMERA_Y);
    return value == value ? -value : -INFINITY;
}

/* ------------------------------------------------------------- unicycle
 * x += v dt cos(theta), y += v dt sin(theta), theta += omega dt, each plus
 * additive Gaussian noise; measurement: range and bearing to each landmark.
 * log w = -1/2 sum_j [ ((zr_j - r_j)/sigma_r)^2 + (wrap(zb_j - b_j)/sigma_b)^2 ],
 * b_j = atan2(ly_j - y, lx_j - x) - theta.
 * The bearing residual is taken as the angle of the landmark vector rotated
 * by -(theta + zb_j), which lands in [-pi, pi] without a separate wrap.
 * State planes: 0 x, 1 y, 2 theta.  Noise: 0 x, 1 y, 2 theta. */
#define ESTH_UNI_S 3
#define ESTH_UNI_V 3
/* frame: [0] dt, [1] v*dt, [2] omega*dt, [3] sx, [4] sy, [5] stheta,
 * [6] 0.5/sr^2, [7] 0.5/sb^2, [8] L, then at 16+5j: lx, ly, zr, cos zb, sin zb */
#define ESTH_UNI_HDR 16

/* propagation; returns cos and sin of the new heading */
ESTH_HD void esth_unicycle_propagate(float* s, const float* z, const float* fr, float* ctp, float* stp) {
    float ct, st;
    esth_sincosf(s[2], &st, &ct);
    float x = s[0] + fr[1] * ct;
    x = x + fr[3] * z[0];
    float y = s[1] + fr[1] * st;
    y = y + fr[4] * z[1];
    float th = s[2] + fr[2];
    th = th + fr[5] * z[2];
    s[0] = x; s[1] = y; s[2] = th;
    esth_sincosf(th, stp, ctp);
}

/* one landmark's squared range and bearing residuals, accumulated.  The
 * landmark offset is rotated by -(theta + zb): its length is the predicted
 * range and its angle the bearing residual wrapped to (-pi, pi]; only the
 * residual's square enters, so its magnitude suffices (esth_polar_abs). */
ESTH_HD void esth_unicycle_landmark(const float* lm, float x, float y, float ct, float st, float* acc_r,
                                    float* acc_b) {
    float dx = lm[0] - x;
    float dy = lm[1] - y;
    float C = esth_fma(ct, lm[3], -(st * lm[4]));   /* cos(theta + zb) */
    float S = esth_fma(st, lm[3], ct * lm[4]);      /* sin(theta + zb) */
    float u = esth_fma(dx, C, dy * S);
    float v = esth_fma(dy, C, -(dx * S));
    float r, eb;
    esth_polar_abs(u, v, &r, &eb);
    float er = lm[2] - r;
    *acc_r = esth_fma(er, er, *acc_r);
    *acc_b = esth_fma(eb, eb, *acc_b);
}

ESTH_HD float esth_unicycle_finish(float acc_r, float acc_b, const float* fr) {
    return -esth_fma(acc_r, fr[6], acc_b * fr[7]);
}

ESTH_HD float esth_unicycle_sampling_iw(float* s, const float* z, const float* fr) {
    float ct, st;
    esth_unicycle_propagate(s, z, fr, &ct, &st);
    const int L = (int)fr[8];
    float acc_r = 0.0f, acc_b = 0.0f;
    for (int j = 0; j < L; ++j)
        esth_unicycle_landmark(fr + ESTH_UNI_HDR + 5 * j, s[0], s[1], ct, st, &acc_r, &acc_b);
    return esth_unicycle_finish(acc_r, acc_b, fr);
}

/* ------------------------------------------------------------------ dh6
 * 6-joint serial arm, Denavit-Hartenberg links A_i = Rz(q_i) Tz(d_i) Tx(a_i) Rx(alpha_i).
 * Constant-velocity dynamics with Gaussian joint acceleration noise:
 *   acc = sigma_acc z,  q += qd dt + acc dt^2/2,  qd += acc dt.
 * Measurement: end-effector position (FK) with sigma_pos, joint encoders with sigma_enc.
 * State planes: 0..5 q, 6..11 qd.  Noise: 0..5 joint accelerations. */
#define ESTH_DH6_S 12
#define ESTH_DH6_V 6
/* frame: [0] dt, [1] dt^2/2, [2] sigma_acc, [3] 0.5/sp^2, [4] 0.5/se^2,
 * [5..7] measured p, [8..13] measured q, then at 16+4i: a_i, d_i, cos alpha_i, sin alpha_i */
#define ESTH_DH6_HDR 16

ESTH_HD float esth_dh6_sampling_iw(float* s, const float* z, const float* fr) {
    const float dt = fr[0];
    for (int i = 0; i < 6; ++i) {
        float acc = fr[2] * z[i];
        float q = esth_fma(s[6 + i], dt, s[i]);
        s[i] = esth_fma(fr[1], acc, q);
        s[6 + i] = esth_fma(dt, acc, s[6 + i]);
    }
    /* end-effector position: p = A_0 (A_1 ( ... (A_5 0))) evaluated inside out */
    float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f;
    for (int i = 5; i >= 0; --i) {
        const float* lk = fr + ESTH_DH6_HDR + 4 * i;
        float sq, cq;
        esth_sincosf(s[i], &sq, &cq);
        float w = esth_fma(lk[2], v1, -(lk[3] * v2));   /* ca v1 - sa v2 */
        float u = v0 + lk[0];
        float n2 = esth_fma(lk[3], v1, esth_fma(lk[2], v2, lk[1]));
        v0 = esth_fma(cq, u, -(sq * w));
        v1 = esth_fma(sq, u, cq * w);
        v2 = n2;
    }
    float ep0 = fr[5] - v0, ep1 = fr[6] - v1, ep2 = fr[7] - v2;
    float acc_p = esth_fma(ep0, ep0, esth_fma(ep1, ep1, ep2 * ep2));
    float acc_q = 0.0f;
    for (int i = 0; i < 6; ++i) {
        float e = fr[8 + i] - s[i];
        acc_q = esth_fma(e, e, acc_q);
    }
    return -esth_fma(acc_p, fr[3], acc_q * fr[4]);
}

/* ------------------------------------------------------------ unicycle10
 * The reference unicycle (src/models/unicycle/particle_filter.cu): sampling()
 * (:313-333) moves the particle on the arc of the noisy speed and turn rate,
 *   v = v_c + nv z0, w = w_c + nw z1, gamma = ng z2,
 *   x += v/w (sin(th + w dt) - sin th), y -= v/w (cos(th + w dt) - cos th),
 *   th += w dt + gamma dt,
 * and importance_weight() (:339-359) compares the distances to 10 sensors at
 * height sz with the measured ranges: weight = norm exp(-sum (d - range)^2),
 * norm = 1e5 / (2 pi)^5, here returned as its logarithm.  The arc is written
 * in its half-angle form, h = w dt / 2:
 *   x += v dt (sin h / h) cos(th + h),  y += v dt (sin h / h) sin(th + h),
 * the same displacement without the cancellation of the two sines (whose
 * difference loses all digits as w -> 0, and is 0/0 at w = 0); sin h / h is
 * 1 - h^2/6 below |h| = 1e-3. */
#define ESTH_U10_S 3
#define ESTH_U10_V 3
/* frame: [0] dt, [1] v, [2] w, [3] nv, [4] nw, [5] ng, [6] log norm,
 * [16..26) ranges, then at 26+3j: sensor x, y, z */
#define ESTH_U10_HDR 16
#define ESTH_U10_FLOATS (ESTH_U10_HDR + 10 + 30)
#define ESTH_U10_LOG_NORM 2.3235402f   /* log(1e5) - 5 log(2 pi) */

ESTH_HD float esth_unicycle10_sampling_iw(float* s, const float* z, const float* fr) {
    const float dt = fr[0];
    const float v = fr[1] + fr[3] * z[0];
    const float w = fr[2] + fr[4] * z[1];
    const float g = fr[5] * z[2];
    const float th = s[2];
    const float wdt = w * dt;
    const float h = 0.5f * wdt;
    float sh, ch, sm, cm;
    esth_sincosf(h, &sh, &ch);
    esth_sincosf(th + h, &sm, &cm);
    const float sinc = fabsf(h) < 1e-3f ? 1.0f - (h * h) * (1.0f / 6.0f) : sh / h;
    const float chord = (v * dt) * sinc;
    const float x = s[0] + chord * cm;
    const float y = s[1] + chord * sm;
    s[0] = x;
    s[1] = y;
    s[2] = (th + wdt) + g * dt;
    float value = 0.0f;
    for (int j = 0; j < 10; ++j) {
        const float* sp = fr + ESTH_U10_HDR + 10 + 3 * j;
        const float d1 = x - sp[0];
        const float d2 = y - sp[1];
        const float d3 = sp[2];
        const float e = sqrtf(d1 * d1 + d2 * d2 + d3 * d3) - fr[ESTH_U10_HDR + j];
        value += e * e;
    }
    return fr[6] - value;
}

/* ------------------------------------------------------- frame builders
 * esth_<model>_frame fills the model's part of a zeroed frame whose [0] is
 * dt (at most nfloats floats); 0, or -1 for bad inputs.
 * esth_<model>_frame_floats: the floats of the frame its hook reads. */
ESTH_HD int esth_robotarm5_frame(const void* params, const void* control, const void* measurement, float dt,
                                 float* fr, int nfloats) {
    (void)params;
    (void)dt;
    (void)nfloats;
    const esth_robotarm5_control* c = (const esth_robotarm5_control*)control;
    const esth_robotarm5_measurement* m = (const esth_robotarm5_measurement*)measurement;
    if (!c || !m) return -1;
    for (int i = 0; i < 5; ++i) { fr[1 + i] = c->angles[i]; fr[6 + i] = m->angles[i]; }
    fr[11] = m->x;
    fr[12] = m->y;
    return 0;
}
ESTH_HD int esth_robotarm5_frame_floats(const void* params) { (void)params; return 16; }

ESTH_HD int esth_robotarm5_camera_frame(const void* params, const void* control, const void* measurement, float dt,
                                        float* fr, int nfloats) {
    const esth_camera_params* p = (const esth_camera_params*)params;
    if (ESTH_RA5C_FLOATS > nfloats || !p) return -1;
    if (esth_robotarm5_frame(params, control, measurement, dt, fr, nfloats)) return -1;
    fr[13] = p->fc[0];
    fr[14] = p->fc[1];
    fr[15] = p->cc[0];
    fr[16] = p->cc[1];
    for (int i = 0; i < 4; ++i) fr[17 + i] = p->kc[i];
    return 0;
}
ESTH_HD int esth_robotarm5_camera_frame_floats(const void* params) { (void)params; return ESTH_RA5C_FLOATS; }

ESTH_HD int esth_unicycle_frame(const void* params, const void* control, const void* measurement, float dt,
                                float* fr, int nfloats) {
    const esth_unicycle_params* p = (const esth_unicycle_params*)params;
    const esth_unicycle_control* c = (const esth_unicycle_control*)control;
    const esth_unicycle_measurement* m = (const esth_unicycle_measurement*)measurement;
    if (!p || !c || !m) return -1;
    if (p->num_landmarks < 0 || p->num_landmarks > ESTH_UNICYCLE_MAX_LANDMARKS) return -1;
    if (ESTH_UNI_HDR + 5 * p->num_landmarks > nfloats) return -1;
    if (!(p->sigma_range > 0.0f) || !(p->sigma_bearing > 0.0f)) return -1;
    fr[1] = c->v * dt;
    fr[2] = c->omega * dt;
    fr[3] = p->sigma_x;
    fr[4] = p->sigma_y;
    fr[5] = p->sigma_theta;
    fr[6] = 0.5f / (p->sigma_range * p->sigma_range);
    fr[7] = 0.5f / (p->sigma_bearing * p->sigma_bearing);
    fr[8] = (float)p->num_landmarks;
    for (int j = 0; j < p->num_landmarks; ++j) {
        float* lm = fr + ESTH_UNI_HDR + 5 * j;
        float sb, cb;
        esth_sincosf(m->bearing[j], &sb, &cb);
        lm[0] = p->lx[j];
        lm[1] = p->ly[j];
        lm[2] = m->range[j];
        lm[3] = cb;
        lm[4] = sb;
    }
    return 0;
}
ESTH_HD int esth_unicycle_frame_floats(const void* params) {
    return params ? ESTH_UNI_HDR + 5 * ((const esth_unicycle_params*)params)->num_landmarks : ESTH_FRAME_FLOATS;
}

ESTH_HD int esth_unicycle10_frame(const void* params, const void* control, const void* measurement, float dt,
                                  float* fr, int nfloats) {
    (void)dt;
    const esth_unicycle10_params* p = (const esth_unicycle10_params*)params;
    const esth_unicycle_control* c = (const esth_unicycle_control*)control;
    const esth_unicycle10_measurement* m = (const esth_unicycle10_measurement*)measurement;
    if (ESTH_U10_FLOATS > nfloats || !p || !c || !m) return -1;
    fr[1] = c->v;
    fr[2] = c->omega;
    fr[3] = p->noise_v;
    fr[4] = p->noise_omega;
    fr[5] = p->noise_gamma;
    fr[6] = ESTH_U10_LOG_NORM;
    for (int j = 0; j < 10; ++j) {
        fr[ESTH_U10_HDR + j] = m->range[j];
        float* sp = fr + ESTH_U10_HDR + 10 + 3 * j;
        sp[0] = p->sx[j];
        sp[1] = p->sy[j];
        sp[2] = p->sz[j];
    }
    return 0;
}
ESTH_HD int esth_unicycle10_frame_floats(const void* params) { (void)params; return ESTH_U10_FLOATS; }

ESTH_HD int esth_dh6_frame(const void* params, const void* control, const void* measurement, float dt, float* fr,
                           int nfloats) {
    (void)control;
    const esth_dh6_params* p = (const esth_dh6_params*)params;
    const esth_dh6_measurement* m = (const esth_dh6_measurement*)measurement;
    if (ESTH_DH6_HDR + 24 > nfloats || !p || !m) return -1;
    if (!(p->sigma_pos > 0.0f) || !(p->sigma_enc > 0.0f)) return -1;
    fr[1] = 0.5f * (dt * dt);
    fr[2] = p->sigma_acc;
    fr[3] = 0.5f / (p->sigma_pos * p->sigma_pos);
    fr[4] = 0.5f / (p->sigma_enc * p->sigma_enc);
    for (int i = 0; i < 3; ++i) fr[5 + i] = m->p[i];
    for (int i = 0; i < 6; ++i) fr[8 + i] = m->q[i];
    for (int i = 0; i < 6; ++i) {
        float* lk = fr + ESTH_DH6_HDR + 4 * i;
        float sa, ca;
        esth_sincosf(p->alpha[i], &sa, &ca);
        lk[0] = p->a[i];
        lk[1] = p->d[i];
        lk[2] = ca;
        lk[3] = sa;
    }
    return 0;
}
ESTH_HD int esth_dh6_frame_floats(const void* params) { (void)params; return ESTH_DH6_HDR + 24; }

/* two particles through a model's hook, one after the other (a model may
 * define its own esth_<model>_sampling_iw_pair to interleave them) */
#define ESTH_PAIR_HOOK_FROM_SINGLE(PFX)                                                                        \
    ESTH_HD void esth_##PFX##_sampling_iw_pair(float* s0, float* s1, const float* z0, const float* z1,          \
                                               const float* fr, float* lw0, float* lw1) {                       \
        *lw0 = esth_##PFX##_sampling_iw(s0, z0, fr);                                                           \
        *lw1 = esth_##PFX##_sampling_iw(s1, z1, fr);                                                           \
    }
ESTH_PAIR_HOOK_FROM_SINGLE(robotarm5)
ESTH_PAIR_HOOK_FROM_SINGLE(robotarm5_camera)
ESTH_PAIR_HOOK_FROM_SINGLE(dh6)
ESTH_PAIR_HOOK_FROM_SINGLE(unicycle10)

/* the unicycle's particles walk the landmarks together: one scalar load of a
 * landmark feeds two independent chains (the same operations per particle) */
ESTH_HD void esth_unicycle_sampling_iw_pair(float* s0, float* s1, const float* z0, const float* z1, const float* fr,
                                            float* lw0, float* lw1) {
    float c0, n0, c1, n1;
    esth_unicycle_propagate(s0, z0, fr, &c0, &n0);
    esth_unicycle_propagate(s1, z1, fr, &c1, &n1);
    const int L = (int)fr[8];
    float r0 = 0.0f, b0 = 0.0f, r1 = 0.0f, b1 = 0.0f;
#ifdef __HIP_DEVICE_COMPILE__
#pragma unroll 4
#endif
    for (int j = 0; j < L; ++j) {
        const float* lm = fr + ESTH_UNI_HDR + 5 * j;
        esth_unicycle_landmark(lm, s0[0], s0[1], c0, n0, &r0, &b0);
        esth_unicycle_landmark(lm, s1[0], s1[1], c1, n1, &r1, &b1);
    }
    *lw0 = esth_unicycle_finish(r0, b0, fr);
    *lw1 = esth_unicycle_finish(r1, b1, fr);
}

/* ----------------------------------------------------------- the models */
#include "esth_model_list.h"

ESTH_HD int esth_model_dims(int model, int* S, int* V) {
    switch (model) {
#define ESTH_X(ID, NAME, PFX, S_, V_, PSZ, CSZ, MSZ) \
    case ID: *S = S_; *V = V_; return 0;
    ESTH_MODEL_LIST(ESTH_X)
#undef ESTH_X
    default: return -1;
    }
}

/* the frame's first `nfloats` floats (at least the model's own length) */
ESTH_HD int esth_build_frame_n(int model, const void* params, const void* control, const void* measurement,
                               float dt, float* fr, int nfloats) {
    for (int i = 0; i < nfloats; ++i) fr[i] = 0.0f;
    if (nfloats < 16) return -1;
    fr[0] = dt;
    switch (model) {
#define ESTH_X(ID, NAME, PFX, S_, V_, PSZ, CSZ, MSZ) \
    case ID: return esth_##PFX##_frame(params, control, measurement, dt, fr, nfloats);
    ESTH_MODEL_LIST(ESTH_X)
#undef ESTH_X
    default: return -1;
    }
}

/* floats of a model's frame that its sampling_iw reads */
ESTH_HD int esth_model_frame_floats(int model, const void* params) {
    switch (model) {
#define ESTH_X(ID, NAME, PFX, S_, V_, PSZ, CSZ, MSZ) \
    case ID: return esth_##PFX##_frame_floats(params);
    ESTH_MODEL_LIST(ESTH_X)
#undef ESTH_X
    default: return 16;
    }
}

ESTH_HD int esth_build_frame_impl(int model, const void* params, const void* control,
                                  const void* measurement, float dt, float* fr) {
    return esth_build_frame_n(model, params, control, measurement, dt, fr, ESTH_FRAME_FLOATS);
}

/* Box-Muller of one uniform pair in [1,2) (pf_generic.cl:465-483):
 * u1 = 2 - x0 in (0,1], u2 = 2 - x1, r = sqrt(-2 ln u1), phi = 2 pi u2. */
ESTH_HD void esth_boxmuller_pair(float x0, float x1, float* z0, float* z1) {
    float u1 = 2.0f - x0;
    float u2 = 2.0f - x1;
    float r = sqrtf(-2.0f * esth_logf(u1));
    float phi = ESTH_2PI * u2;
    float sp, cp;
    esth_sincosf(phi, &sp, &cp);
    *z0 = r * cp;
    *z1 = r * sp;
}

#endif /* ESTH_MODELS_H */
