/*
 * esth_dispatch.h - run-time model id / state dimension -> compile-time
 * constant, generated from esth_model_list.h, for the kernel launchers.
 */
#ifndef ESTH_DISPATCH_H
#define ESTH_DISPATCH_H

#include <type_traits>

#include "esth_internal.h"

/* f(std::integral_constant<int, id>) for the listed model `model` */
template <typename F>
static int esth_with_model(int model, F&& f) {
#define ESTH_X(ID, NAME, PFX, S_, V_, PSZ, CSZ, MSZ) \
    if (model == ID) return f(std::integral_constant<int, ID>{});
    ESTH_MODEL_LIST(ESTH_X)
#undef ESTH_X
    return esth_fail(ESTH_ERR_INVALID, "unknown model %d", model);
}

/* the same returning f's value, or `otherwise` for an unknown model */
template <typename R, typename F>
static R esth_with_model_or(int model, R otherwise, F&& f) {
#define ESTH_X(ID, NAME, PFX, S_, V_, PSZ, CSZ, MSZ) \
    if (model == ID) return f(std::integral_constant<int, ID>{});
    ESTH_MODEL_LIST(ESTH_X)
#undef ESTH_X
    return otherwise;
}

/* f(std::integral_constant<int, S>) for the state dimension of a listed
 * model, or S = 1 (the stage entry points' single plane) */
template <typename F>
static int esth_with_state_dim(int S, F&& f) {
    if (S == 1) return f(std::integral_constant<int, 1>{});
#define ESTH_X(ID, NAME, PFX, S_, V_, PSZ, CSZ, MSZ) \
    if (S == (S_)) return f(std::integral_constant<int, (S_)>{});
    ESTH_MODEL_LIST(ESTH_X)
#undef ESTH_X
    return esth_fail(ESTH_ERR_UNSUPPORTED, "state dimension %d", S);
}

/* the same over the listed models' state dimensions only */
template <typename F>
static int esth_with_model_state_dim(int S, F&& f) {
#define ESTH_X(ID, NAME, PFX, S_, V_, PSZ, CSZ, MSZ) \
    if (S == (S_)) return f(std::integral_constant<int, (S_)>{});
    ESTH_MODEL_LIST(ESTH_X)
#undef ESTH_X
    return esth_fail(ESTH_ERR_UNSUPPORTED, "state dimension %d", S);
}

/* whether some listed model (or the single plane) has state dimension S */
static inline bool esth_state_dim_known(int S) {
    if (S == 1) return true;
#define ESTH_X(ID, NAME, PFX, S_, V_, PSZ, CSZ, MSZ) \
    if (S == (S_)) return true;
    ESTH_MODEL_LIST(ESTH_X)
#undef ESTH_X
    return false;
}

/* the largest params struct of the listed models */
constexpr size_t esth_max_model_params() {
    size_t m = 1;
#define ESTH_X(ID, NAME, PFX, S_, V_, PSZ, CSZ, MSZ) \
    m = (size_t)(PSZ) > m ? (size_t)(PSZ) : m;
    ESTH_MODEL_LIST(ESTH_X)
#undef ESTH_X
    return m;
}

#endif
