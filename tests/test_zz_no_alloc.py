"""owm_run_tile_device_noalloc: the whole pipeline for one tile with every
device buffer carved from the library's static scratch (nothing allocated),
against owm_run_tile_device on the same inputs, called through the C ABI.

Two calls in a row, the second on a smaller tile (the scratch keeps the first
call's contents, so everything that must start at zero has to be cleared by
the second), then a tile larger than the scratch is refused with a message,
and the entry point still works afterwards.  The context's own device
memory stays at zero bytes throughout: the no-allocation calls never grow its
arenas."""
import ctypes

import re

import numpy as np
import pytest

import pyoracle as O
from fastowm_amd import owm, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _call(fn, ctx, dev, n, p, cap):
    out = np.zeros(cap, np.uint32)
    ns = ctypes.c_uint64()
    st = fn(ctx.h, dev[0].data_ptr(), dev[1].data_ptr(), dev[2].data_ptr(), n, ctypes.byref(p),
            out.ctypes.data, cap, ctypes.byref(ns))
    return st, out[:ns.value].copy(), ns.value


def test_noalloc_equals_device_entry_point():
    import torch
    L = owm.lib()
    big = synth.config_c1(seed=0)  # 1M points, 500 m, Wsize 12 / Bsize 20 / Overlap 0.8
    small_t = synth.scaled(synth.config_c2(seed=4), 300_000)
    cases = [(big, 12, 20, 0.8, 32), (small_t, 9, 15, 0.6, 16)]
    data = []
    for t, ws, bs, ov, leaf in cases:
        x, y, z = t.generate()
        dev = [torch.from_numpy(a).to("cuda:0") for a in (x, y, z)]
        data.append((x, y, z, dev, owm.make_params(ws, bs, ov, leaf), (ws, bs, ov, leaf)))
    torch.cuda.synchronize()
    with owm.Context(0) as nctx, owm.Context(0) as rctx:
        assert nctx.device_bytes() == 0
        got = []
        for rep, (x, y, z, dev, p, args) in enumerate(data):
            say("noalloc call", rep, len(x), "points", args)
            st, seeds, ns = _call(L.owm_run_tile_device_noalloc, nctx, dev, len(x), p, len(x) + 1)
            assert st == 0, nctx.last_error()
            got.append(seeds)
            assert nctx.device_bytes() == 0, "the no-allocation call grew the context's arenas"
        for rep, (x, y, z, dev, p, args) in enumerate(data):
            say("device entry point", rep)
            st, want, _ = _call(L.owm_run_tile_device, rctx, dev, len(x), p, len(x) + 1)
            assert st == 0, rctx.last_error()
            assert np.array_equal(got[rep], want), f"call {rep}: differs from owm_run_tile_device"
            ws, bs, ov, leaf = args
            assert np.array_equal(want, O.run(x, y, z, ws, bs, ov, leaf)[0])
        say("capacity")
        x, y, z, dev, p, _ = data[1]
        st, _, ns = _call(L.owm_run_tile_device_noalloc, nctx, dev, len(x), p, 3)
        assert st == 5 and ns == len(got[1])  # OWM_E_CAPACITY with the size needed
        say("too large for the scratch")
        t = synth.scaled(synth.config_c2(seed=6), 6_000_000)
        bx, by, bz = t.generate()
        bdev = [torch.from_numpy(a).to("cuda:0") for a in (bx, by, bz)]
        torch.cuda.synchronize()
        st, _, _ = _call(L.owm_run_tile_device_noalloc, nctx, bdev, len(bx), owm.make_params(12, 20, 0.8, 32),
                         len(bx) + 1)
        msg = nctx.last_error()
        assert st == 3 and "bytes of scratch" in msg, msg
        need = int(re.search(r"needs (\d+) bytes", msg).group(1))
        say("scratch needed", need, "bytes,", round(need / len(bx), 1), "per point")
        assert need > 256 << 20
        assert nctx.device_bytes() == 0
        say("after the refusal")
        x, y, z, dev, p, _ = data[1]
        st, seeds, _ = _call(L.owm_run_tile_device_noalloc, nctx, dev, len(x), p, len(x) + 1)
        assert st == 0 and np.array_equal(seeds, got[1])
