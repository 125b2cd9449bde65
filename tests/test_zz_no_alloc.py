"""The allocation-free block sort (esth_stage_sort_noalloc) against the
allocating one (esth_stage_sort) on the same inputs, through the C ABI: sorted
keys, permutations, gathered planes and winners bit for bit, on consecutive
calls that share the library's device scratch (the second call smaller than
the first, so stale scratch from the first would show)."""
import numpy as np
import pytest

import esthera_amd as ea

pytestmark = pytest.mark.gpu

DB = ea.DeviceBuffer


def _inputs(m, G, S, seed):
    rng = np.random.default_rng(seed)
    N = m * G
    w = rng.integers(0, 7, N).astype(np.float32) * 0.5 - 2.0   # ties
    w[: N // 3] = rng.normal(size=N // 3).astype(np.float32)
    st = rng.normal(size=S * N).astype(np.float32)
    return w, st


def _sort(entry, w, st, m, G, S, with_winner):
    N = m * G
    dw, dst, dout, dperm = DB.from_array(w), DB.from_array(st), DB(S * N * 4), DB(N * 4)
    dwin = DB(G * 8) if with_winner else None
    L = ea.lib()
    if entry == "alloc":
        ea.check(L.esth_stage_sort(dw.ptr, dst.ptr, dout.ptr, S, m, G, dwin.ptr if dwin else None, dperm.ptr))
    else:
        ea.check(L.esth_stage_sort_noalloc(dw.ptr, dst.ptr, dout.ptr, S, m, G, dwin.ptr if dwin else None,
                                           dperm.ptr, None))
        ea.check(L.esth_dev_sync())
    out = [dw.to_array(np.float32, N), dperm.to_array(np.int32, N), dout.to_array(np.float32, S * N)]
    if dwin is not None:
        out.append(dwin.to_array(np.float32, 2 * G))
    return [o.view(np.uint32) if o.dtype == np.float32 else o for o in out]


@pytest.mark.parametrize("calls", [
    [(512, 4096, 3, True), (128, 1000, 3, True)],
    [(256, 2048, 12, False), (64, 333, 9, True)],
])
def test_noalloc_sort_equals_allocating_sort(calls, engine):
    for k, (m, G, S, with_winner) in enumerate(calls):
        w, st = _inputs(m, G, S, 100 * k + m + S)
        ref = _sort("alloc", w, st, m, G, S, with_winner)
        got = _sort("noalloc", w, st, m, G, S, with_winner)
        for r, g in zip(ref, got):
            assert np.array_equal(r, g), f"call {k}: m={m} G={G} S={S}"


def test_noalloc_sort_rejects_sizes_beyond_its_scratch(engine):
    dw = DB(4)
    L = ea.lib()
    rc = L.esth_stage_sort_noalloc(dw.ptr, dw.ptr, dw.ptr, 12, 2, 65536, None, None, None)
    assert rc == ea.ESTH_ERR_BUDGET
    assert "at most 524288" in L.esth_last_error().decode()
