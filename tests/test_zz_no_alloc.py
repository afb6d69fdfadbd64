"""gf_average_noalloc against gf_average through the C ABI: bit-identical
averages on two calls in a row, the second on a smaller context (fewer
partial sums than the first left in the shared device array), for FM, AFM
and FE fields; and its argument errors.

The output buffer is allocated here with the HIP runtime the library links
(the entry point itself allocates nothing)."""
import ctypes
import sys

import numpy as np
import pytest

from goferro import (FIELD_M, FIELD_M_AFM1, FIELD_M_AFM2, FIELD_P, FIELD_Q, Context, GoFerroError, load_library,
                     workloads)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


class DeviceOut:
    """Three doubles of device memory from the library's own HIP runtime."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(24)) == 0
        self.clear()

    def clear(self):
        nan = np.full(3, np.nan)
        assert self.hip.hipMemcpy(self.ptr, nan.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(24), 1) == 0

    def read(self):
        out = np.empty(3)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(24), 2) == 0
        return out

    def close(self):
        self.hip.hipFree(self.ptr)


def noalloc(ctx, fid, dev):
    lib = load_library()
    st = lib.gf_average_noalloc(ctx._h, fid, dev.ptr)
    if st != 0:
        raise GoFerroError(f"gf_average_noalloc failed ({st}): {lib.gf_last_error(ctx._h).decode()}")
    ctx.sync()
    return dev.read()


def fm_ctx(shape, seed, demag):
    prob, m = workloads.fm_film(shape, seed=seed, demag=demag)
    c = Context(prob)
    c.upload(FIELD_M, m)
    if demag:
        c.demag()
    c.step(2)
    return c


def test_noalloc_average_matches_average_on_two_calls():
    dev = DeviceOut()
    # 128x128x16 uses the full 1024 partial sums, 6x5x3 a single one: the
    # second call reads only what its own sum kernel wrote
    for shape, seed, demag in (((128, 128, 16), 3, True), ((6, 5, 3), 4, False)):
        say("case noalloc FM average", shape)
        c = fm_ctx(shape, seed, demag)
        dev.clear()
        got = noalloc(c, FIELD_M, dev)
        ref = np.asarray(c.average(FIELD_M))
        say("  noalloc", got, "average", ref)
        assert np.array_equal(got, ref), (shape, got, ref)
        c.close()
    dev.close()


def test_noalloc_average_afm_and_fe_fields():
    dev = DeviceOut()
    say("case noalloc AFM average (16, 16, 8)")
    prob, m1, m2 = workloads.afm_block((16, 16, 8))
    c = Context(prob)
    c.upload(FIELD_M_AFM1, m1)
    c.upload(FIELD_M_AFM2, m2)
    c.step(3)
    for fid in (FIELD_M_AFM1, FIELD_M_AFM2):
        dev.clear()
        got = noalloc(c, fid, dev)
        assert np.array_equal(got, np.asarray(c.average(fid))), fid
    with pytest.raises(GoFerroError, match="average is defined"):
        noalloc(c, FIELD_M, dev)
    c.close()
    say("case noalloc FE average (4, 4, 96)")
    prob, p, q = workloads.fe_thz_column((4, 4, 96), fe_layers=40)
    c = Context(prob)
    c.upload(FIELD_P, p)
    c.upload(FIELD_Q, q)
    c.step(3)
    dev.clear()
    got = noalloc(c, FIELD_P, dev)
    assert np.array_equal(got, np.asarray(c.average(FIELD_P)))
    with pytest.raises(GoFerroError, match="average is defined"):
        noalloc(c, FIELD_Q, dev)
    c.close()
    dev.close()
