"""ctypes binding of oracle/liboracle.so (the CPU oracle).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The engine never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REPO = os.path.dirname(HERE)

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def _opt(p):
    return None if p is None else p.ctypes.data_as(C.c_void_p)


_lib = None


def use_library(path):
    """Load the oracle from another build of oracle/esth_oracle.c (the CPU
    baseline's -O3 -march=native build); later calls use it."""
    global _lib
    _lib = None
    return lib(path)


def lib(path=None):
    global _lib
    if _lib is None:
        path = path or LIB_PATH
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        L = C.CDLL(path)
        L.or_mtgp_seed.argtypes = [_u32p, C.c_uint32, _u32p]
        L.or_mtgp_generate.argtypes = [_u32p, C.c_uint32, _u32p, C.c_int, _u32p, C.c_int]
        L.or_mtgp_fill.argtypes = [_u32p, C.c_int, C.c_int, _u32p, C.c_int, _u32p, C.c_int]
        L.or_boxmuller.argtypes = [_f32p, C.c_int64]
        L.or_build_frame.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, _f32p]
        L.or_sample_weight.argtypes = [C.c_int, _f32p, C.c_uint32, _f32p, _f32p, C.c_int, _f32p]
        L.or_sample_weight_f64.argtypes = [C.c_int, _f32p, C.c_uint32, _f32p, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_float, _f64p, _f64p]
        L.or_block_sort.argtypes = [_f32p, _f32p, _f32p, C.c_int, C.c_uint32, C.c_uint32, C.c_void_p,
                                    C.c_void_p]
        L.or_argmax.argtypes = [_f32p, C.c_uint32, C.c_int, _f32p]
        L.or_exchange.argtypes = [_f32p, _f32p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                  C.c_int, C.c_int]
        L.or_resample.argtypes = [C.c_int, _f32p, _f32p, C.c_int, _f32p, _f32p, C.c_uint32, C.c_uint32,
                                  C.c_float, C.c_int, C.c_int, C.c_void_p]
        L.or_alias_table.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_int, _f32p, _u32p]
        L.or_expf.argtypes = [C.c_float]
        L.or_expf.restype = C.c_float
        L.or_logf.argtypes = [C.c_float]
        L.or_logf.restype = C.c_float
        L.or_atan2f.argtypes = [C.c_float, C.c_float]
        L.or_atan2f.restype = C.c_float
        L.or_sincosf.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.or_polar_abs.argtypes = [C.c_float, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.or_rcpf.argtypes = [C.c_float]
        L.or_rcpf.restype = C.c_float
        L.or_filter_create.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, _u32p]
        L.or_filter_create.restype = C.c_void_p
        L.or_filter_destroy.argtypes = [C.c_void_p]
        L.or_filter_init_uniform.argtypes = [C.c_void_p, _f32p]
        L.or_filter_set_state.argtypes = [C.c_void_p, _f32p]
        L.or_filter_step.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float]
        L.or_filter_read.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_filter_estimate.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_filter_rng_words.argtypes = [C.c_void_p]
        L.or_filter_best_particle.argtypes = [C.c_void_p, _f32p]
        L.or_filter_rng_words.restype = C.c_uint64
        L.or_set_threads.argtypes = [C.c_int]
        L.or_filter_set_f64.argtypes = [C.c_void_p, C.c_int]
        _lib = L
    return _lib


def mtgp_params():
    """The 200 x 52 parameter table from esthera_amd/data (the same data the engine embeds)."""
    rows = []
    with open(os.path.join(REPO, "esthera_amd", "data", "mtgp32dc_11213.txt")) as fh:
        for line in fh:
            if line.startswith("#") or not line.strip():
                continue
            tok = line.split()
            rows.append([int(tok[0]), int(tok[1]), int(tok[2])] + [int(t, 16) for t in tok[3:]])
    arr = np.array(rows, dtype=np.uint32)
    assert arr.shape == (200, 52)
    return arr


def mtgp_seed(params, set_idx, seed):
    st = np.zeros(351, np.uint32)
    lib().or_mtgp_seed(np.ascontiguousarray(params[set_idx]), seed & 0xFFFFFFFF, st)
    return st


def mtgp_generate(params, set_idx, state, n, float_mode=True, mask=None):
    out = np.zeros(n, np.uint32)
    m = int(params[0, 3]) if mask is None else mask
    lib().or_mtgp_generate(np.ascontiguousarray(params[set_idx]), m, state, n, out, 1 if float_mode else 0)
    return out


def mtgp_fill(params, nstreams, status, size, float_mode=True, first_set=0):
    out = np.zeros(nstreams * size, np.uint32)
    lib().or_mtgp_fill(np.ascontiguousarray(params), nstreams, first_set, status, size, out,
                       1 if float_mode else 0)
    return out


def boxmuller(data):
    d = np.array(data, dtype=np.float32, copy=True)
    lib().or_boxmuller(d, d.size)
    return d


def build_frame(model, params, ctrl, meas, dt):
    fr = np.zeros(336, np.float32)
    rc = lib().or_build_frame(model, _ref(params), _ref(ctrl), _ref(meas), dt, fr)
    assert rc == 0
    return fr


def _ref(x):
    return None if x is None else C.cast(C.pointer(x), C.c_void_p)


def sample_weight(model, state, N, normals, frame, weights):
    st = np.array(state, dtype=np.float32, copy=True)
    w = np.zeros(N, np.float32)
    assert lib().or_sample_weight(model, st, N, np.ascontiguousarray(normals, np.float32), frame, weights, w) == 0
    return st, w


def sample_weight_f64(model, state, N, normals, params, ctrl, meas, dt, S):
    so = np.zeros(S * N, np.float64)
    lw = np.zeros(N, np.float64)
    assert lib().or_sample_weight_f64(model, np.ascontiguousarray(state, np.float32), N,
                                      np.ascontiguousarray(normals, np.float32), _ref(params), _ref(ctrl),
                                      _ref(meas), dt, so, lw) == 0
    return so, lw


def block_sort(w, state_in, S, m, G):
    w = np.array(w, dtype=np.float32, copy=True)
    out = np.zeros(S * m * G, np.float32)
    win = np.zeros(2 * G, np.float32)
    perm = np.zeros(m * G, np.int32)
    st = np.ascontiguousarray(state_in, np.float32) if S > 0 else np.zeros(1, np.float32)
    lib().or_block_sort(w, st, out if S > 0 else np.zeros(1, np.float32), S, m, G, _opt(win), _opt(perm))
    return w, out, win, perm


def argmax(winner, G, weights):
    r = np.zeros(2, np.float32)
    lib().or_argmax(np.ascontiguousarray(winner, np.float32), G, weights, r)
    return r


def exchange(state, w, S, m, G, t, topo, dimx=0, dimy=0):
    st = np.array(state, dtype=np.float32, copy=True)
    ww = np.array(w, dtype=np.float32, copy=True)
    rc = lib().or_exchange(st, ww, S, m, G, t, topo, dimx, dimy)
    assert rc == 0
    return st, ww


def resample(resampler, dst, src, S, w, rnd, m, G, thr, adaptive, weights):
    d = np.array(dst, dtype=np.float32, copy=True)
    sel = np.zeros(m * G, np.int32)
    rc = lib().or_resample(resampler, d, np.ascontiguousarray(src, np.float32), S,
                           np.ascontiguousarray(w, np.float32), np.ascontiguousarray(rnd, np.float32),
                           m, G, thr, adaptive, weights, _opt(sel))
    assert rc == 0
    return d, sel


def alias_table(w, m, G, weights):
    prob = np.zeros(m * G, np.float32)
    alias = np.zeros(m * G, np.uint32)
    lib().or_alias_table(np.ascontiguousarray(w, np.float32), m, G, weights, prob, alias)
    return prob, alias


class Filter:
    """Whole-step oracle: FP32 restatement plus FP64 posterior."""

    def __init__(self, desc, params=None, params_size=0):
        self._params = mtgp_params()
        self._desc = desc
        self._mp = params
        self.h = lib().or_filter_create(C.cast(C.pointer(desc), C.c_void_p), _ref(params),
                                        params_size, self._params)
        assert self.h
        self.N = desc.num_particles * desc.num_blocks

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_filter_destroy(self.h)
            self.h = None

    def set_f64(self, on):
        lib().or_filter_set_f64(self.h, 1 if on else 0)

    def init_uniform(self, state):
        lib().or_filter_init_uniform(self.h, np.ascontiguousarray(state, np.float32))

    def set_state(self, soa):
        lib().or_filter_set_state(self.h, np.ascontiguousarray(soa, np.float32))

    def step(self, ctrl, meas, dt):
        rc = lib().or_filter_step(self.h, _ref(ctrl), _ref(meas), dt)
        assert rc == 0

    def read(self, S):
        st = np.zeros(S * self.N, np.float32)
        w = np.zeros(self.N, np.float32)
        lib().or_filter_read(self.h, st.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p))
        return st, w

    def best_particle(self, S):
        out = np.zeros(S, np.float32)
        lib().or_filter_best_particle(self.h, out)
        return out

    def estimate(self, S):
        m32 = np.zeros(S, np.float64)
        m64 = np.zeros(S, np.float64)
        best = np.zeros(2, np.float32)
        lib().or_filter_estimate(self.h, m32.ctypes.data_as(C.c_void_p), m64.ctypes.data_as(C.c_void_p),
                                 best.ctypes.data_as(C.c_void_p))
        return m32, m64, best
