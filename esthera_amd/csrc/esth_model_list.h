/*
 * esth_model_list.h - the compiled-in models, one line each:
 *
 *   X(id, name, prefix, state floats, normals, params bytes, control bytes, measurement bytes)
 *
 * A model with prefix P provides, in a header included here:
 *   ESTH_HD float esth_P_sampling_iw(float* state, const float* normals, const float* frame)
 *       propagate one particle in place, return its log importance weight
 *       (the reference hook sampling_iw, pf_model.cl:52-131);
 *   ESTH_HD void esth_P_sampling_iw_pair(s0, s1, z0, z1, frame, &lw0, &lw1)
 *       two particles at once (ESTH_PAIR_HOOK_FROM_SINGLE(P) makes it from
 *       the one-particle hook);
 *   ESTH_HD int esth_P_frame(params, control, measurement, dt, frame, nfloats)
 *       the step's frame from the model's inputs (0, or -1 for bad inputs);
 *   ESTH_HD int esth_P_frame_floats(params)
 *       how many frame floats the hook reads.
 * Every dispatch over models (kernels, filters, batches, servers, stage
 * entry points, the oracle) and over state dimensions is generated from this
 * list.  The id is the model's number in esth_model_lookup / descriptors.
 */
#ifndef ESTH_MODEL_LIST_H
#define ESTH_MODEL_LIST_H

#include "models/randomwalk3.h"

#define ESTH_MODEL_LIST(X)                                                                                           \
    X(ESTH_MODEL_ROBOTARM5, "robotarm5", robotarm5, ESTH_RA5_S, ESTH_RA5_V, 0, sizeof(esth_robotarm5_control),          \
      sizeof(esth_robotarm5_measurement))                                                                            \
    X(ESTH_MODEL_UNICYCLE, "unicycle", unicycle, ESTH_UNI_S, ESTH_UNI_V, sizeof(esth_unicycle_params),                 \
      sizeof(esth_unicycle_control), sizeof(esth_unicycle_measurement))                                              \
    X(ESTH_MODEL_DH6, "dh6", dh6, ESTH_DH6_S, ESTH_DH6_V, sizeof(esth_dh6_params), sizeof(esth_dh6_control),          \
      sizeof(esth_dh6_measurement))                                                                                  \
    X(ESTH_MODEL_UNICYCLE10, "unicycle10", unicycle10, ESTH_U10_S, ESTH_U10_V, sizeof(esth_unicycle10_params),        \
      sizeof(esth_unicycle_control), sizeof(esth_unicycle10_measurement))                                            \
    X(ESTH_MODEL_ROBOTARM5_CAMERA, "robotarm5_camera", robotarm5_camera, ESTH_RA5_S, ESTH_RA5_V,                      \
      sizeof(esth_camera_params), sizeof(esth_robotarm5_control), sizeof(esth_robotarm5_measurement))                \
    X(ESTH_MODEL_RANDOMWALK3, "randomwalk3", randomwalk3, ESTH_RW3_S, ESTH_RW3_V, sizeof(esth_randomwalk3_params),    \
      sizeof(esth_randomwalk3_control), sizeof(esth_randomwalk3_measurement))

#endif
