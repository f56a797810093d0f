"""xme_me_search_static (frame-level ME with its plan in the library's static
device scratch, no device allocation) against the plan-based entry point
(xme_plan_create + xme_me_run) on the same inputs: two calls in a row, the
second with a smaller job, for the full search and the hexagon search; a job
whose plan exceeds the scratch is refused with XME_ERR_BUDGET."""
import ctypes as C

import numpy as np
import pytest

from tests.test_gpu_me import LAM, device_cur
from xavs2_amd import lib as X
from xavs2_amd import synth

pytestmark = pytest.mark.gpu


def static_search(prm, pus, n_ref, dcur, drefs):
    arr = np.ascontiguousarray(pus, dtype=X.PU_DTYPE)
    n = len(arr) * n_ref
    out = X.DeviceBuffer(n * X.RESULT_DTYPE.itemsize)
    refs = (X.RefDesc * n_ref)(*[r.desc() for r in drefs[:n_ref]])
    rc = X.lib().xme_me_search_static(C.byref(prm), X.ptr(arr), len(arr), n_ref, C.c_void_p(dcur.origin),
                                      dcur.stride, refs, C.c_void_p(out.ptr), None)
    return rc, (out.download(n, X.RESULT_DTYPE) if rc == 0 else None)


def plan_search(prm, pus, n_ref, dcur, drefs):
    plan = X.Plan(prm, pus, n_ref)
    plan.run(dcur, drefs[:n_ref])
    return plan.fetch()


@pytest.mark.parametrize("method", [X.ME_FS, X.ME_HEX])
def test_static_search_equals_plan_search_twice(method):
    w, h = 512, 320
    refs = [synth.smooth_field(w, h, seed=300 + r) for r in range(2)]
    cur = synth.multi_ref_current(refs, 16, 24, seed=311)
    drefs = []
    for r in refs:
        d = X.DeviceRef(w, h, frac=2)
        d.load(r)
        drefs.append(d)
    dcur = device_cur(cur)
    prm = X.me_params(w, h, 32, LAM, me_method=method)
    big = X.make_pus([(x, y, pw, ph, 4, -6) for (x, y, pw, ph) in synth.mixed_pus(w, h, 21)])
    small = big[: len(big) // 3]
    for pus, n_ref in ((big, 2), (small, 1)):
        print(f"static search: method {method}, {len(pus)} PUs, {n_ref} refs", flush=True)
        rc, got = static_search(prm, pus, n_ref, dcur, drefs)
        assert rc == 0, X.lib().xme_last_error()
        exp = plan_search(prm, pus, n_ref, dcur, drefs)
        assert np.array_equal(got, exp)


def test_static_search_refuses_a_plan_beyond_the_scratch():
    # 4x4 PUs over 4096x2048 against 8 references: 4.2 M instances, well past 64 MiB of plan
    w, h = 4096, 2048
    ref = X.DeviceRef(w, h, frac=2)
    ref.load(synth.smooth_field(w, h, seed=5))
    dcur = device_cur(synth.smooth_field(w, h, seed=6))
    pus = np.zeros((h // 4) * (w // 4), dtype=X.PU_DTYPE)
    pus["x"] = np.tile(np.arange(0, w, 4), h // 4)
    pus["y"] = np.repeat(np.arange(0, h, 4), w // 4)
    pus["w"] = 4
    pus["h"] = 4
    rc, _ = static_search(X.me_params(w, h, 16, LAM), pus, 8, dcur, [ref] * 8)
    assert rc == X.XME_ERR_BUDGET
    assert b"static scratch" in X.lib().xme_last_error()
