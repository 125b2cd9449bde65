/*
 * randomwalk3 - a 3-D random walk observed directly: the smallest model, and
 * the example of adding one.  A model is this one header (its hook and frame
 * builder) plus its #include and one line in esth_model_list.h, and, for
 * callers, its id and input structs in the public header esthera.h; the
 * engine's single filters, batches, servers, stage entry points and the
 * oracle pick it up from that list.
 *
 *   p += v dt + sigma_q z   (z: three standard normals)
 *   log w = -1/2 |m - p|^2 / sigma_r^2
 *
 * State planes: 0 x, 1 y, 2 z.  Noise: 0 x, 1 y, 2 z.  The hook follows the
 * contract of sampling_iw (src/models/robotarm/opencl/pf_model.cl:52-131):
 * propagate one particle in place with its normals, return its log weight.
 */
#ifndef ESTH_MODEL_RANDOMWALK3_H
#define ESTH_MODEL_RANDOMWALK3_H

#define ESTH_RW3_S 3
#define ESTH_RW3_V 3
/* frame: [0] dt, [1..3] v dt, [4] sigma_q, [5] 0.5 / sigma_r^2, [6..8] measured p */

ESTH_HD float esth_randomwalk3_sampling_iw(float* s, const float* z, const float* fr) {
    float acc = 0.0f;
    for (int c = 0; c < 3; ++c) {
        const float x = esth_fma(fr[4], z[c], s[c] + fr[1 + c]);
        s[c] = x;
        const float e = fr[6 + c] - x;
        acc = esth_fma(e, e, acc);
    }
    return -(acc * fr[5]);
}

ESTH_HD int esth_randomwalk3_frame(const void* params, const void* control, const void* measurement, float dt,
                                   float* fr, int nfloats) {
    const esth_randomwalk3_params* p = (const esth_randomwalk3_params*)params;
    const esth_randomwalk3_control* c = (const esth_randomwalk3_control*)control;
    const esth_randomwalk3_measurement* m = (const esth_randomwalk3_measurement*)measurement;
    if (nfloats < 16 || !p || !c || !m || !(p->sigma_r > 0.0f)) return -1;
    fr[0] = dt;
    for (int i = 0; i < 3; ++i) {
        fr[1 + i] = c->v[i] * dt;
        fr[6 + i] = m->p[i];
    }
    fr[4] = p->sigma_q;
    fr[5] = 0.5f / (p->sigma_r * p->sigma_r);
    return 0;
}

ESTH_HD int esth_randomwalk3_frame_floats(const void* params) {
    (void)params;
    return 16;
}

ESTH_PAIR_HOOK_FROM_SINGLE(randomwalk3)

#endif
