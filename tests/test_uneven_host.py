"""Host-side checks of the uneven-batch plan: argument validation happens
before any device call, and the autotuner's candidate menu obeys the
kernel's limits."""
import ctypes

import numpy as np
import pytest

from tpr_amd import _lib, autotune
from tpr_amd.uneven import schedule_array


def _create(sizes, schedule=None, method=_lib.PTPR_M, arith=0):
    L = _lib.lib()
    h = ctypes.c_void_p()
    sz = np.asarray(sizes, np.int32)
    tab = schedule_array(schedule) if schedule else None
    st = L.tprh_uneven_plan_create(ctypes.byref(h), method, arith, sz.ctypes.data_as(ctypes.c_void_p), sz.size,
                                   tab.ctypes.data_as(ctypes.c_void_p) if tab is not None else None)
    return st, h


@pytest.mark.parametrize("sizes", [[24], [1024, 8], [0], [-16], [1 << 15]])
def test_bad_sizes_rejected(sizes):
    st, h = _create(sizes)
    assert st in (_lib.ERROR_INVALID_VALUE, _lib.ERROR_NOT_SUPPORTED)
    assert h.value is None


def test_bad_method_and_schedule_rejected():
    assert _create([256], method=_lib.PCR_M)[0] == _lib.ERROR_INVALID_VALUE
    # s does not divide n
    assert _create([256], {256: (512, 8, 64)})[0] == _lib.ERROR_INVALID_VALUE
    # shape not in the persistent kernel's menu
    assert _create([4096], {4096: (512, 16, 32)})[0] == _lib.ERROR_NOT_SUPPORTED
    # E * L != s
    assert _create([4096], {4096: (256, 8, 16)})[0] == _lib.ERROR_INVALID_VALUE
    # too many lanes per system: 16384 / 128 slices * 16 lanes > 1024
    assert _create([16384], {16384: (128, 8, 16)})[0] == _lib.ERROR_NOT_SUPPORTED


@pytest.mark.parametrize("method", [_lib.PTPR_M, _lib.TPR_M])
@pytest.mark.parametrize("n", [1024, 4096, 8192])
def test_exact_classes_at_s512_take_a_menu_shape(method, n):
    """A schedule row giving only s = 512: the default 16 x 32 slices have no
    path in the persistent kernel, so fast arithmetic refuses the class;
    exact arithmetic (the same bits with any shape) takes a shape of the
    menu and the plan passes validation (here it then fails at the device
    allocation, without a GPU)."""
    st, h = _create([n, 256], {n: (512, 0, 0)}, method=method, arith=1)
    assert st in (_lib.SUCCESS, _lib.ERROR_HIP), st
    if st == _lib.SUCCESS:
        _lib.lib().tprh_uneven_plan_destroy(h)
    assert _create([n, 256], {n: (512, 0, 0)}, method=method, arith=0)[0] == _lib.ERROR_NOT_SUPPORTED


def test_schedule_array_layout():
    tab = schedule_array({1024: (256, 16, 16), 16: (16, 1, 16)})
    assert tab.shape == (32, 3)
    assert tuple(tab[10]) == (256, 16, 16) and tuple(tab[4]) == (16, 1, 16)
    assert not tab[11].any()
    with pytest.raises(ValueError):
        schedule_array({1000: (8, 1, 8)})


def test_autotune_candidates_respect_kernel_limits():
    for k in range(4, 15):
        n = 1 << k
        cands = autotune.candidates(n)
        assert cands, n
        assert autotune.default_choice(n) in cands or n < 16
        for s, E, L in cands:
            assert E * L == s and n % s == 0 and (n // s) * L <= 1024
            assert (E, L) in autotune.UNEVEN_SHAPES
    assert autotune.candidates(1 << 15) == []


def test_committed_schedule_is_valid():
    import os
    if not os.path.exists(autotune.DEFAULT_PATH):
        pytest.skip("no tuned schedule committed")
    for n, cand in autotune.schedule_of(autotune.load()).items():
        assert tuple(cand) in autotune.candidates(n)


def test_class_choice_keeps_the_default_within_the_margin():
    """A candidate replaces the default s = min(n, 256) shape only when it is
    faster by more than the margin; the committed table follows the rule
    from its own measurements."""
    us = {"128/8x16": 98.0, "256/16x16": 100.0, "64/4x16": 130.0}
    assert autotune.choose_class(1024, us) == [256, 16, 16]          # 2 % lead: noise
    us["128/8x16"] = 90.0
    assert autotune.choose_class(1024, us) == [128, 8, 16]           # 10 % lead
    assert autotune.choose_class(1024, us, margin=0.2) == [256, 16, 16]
    assert autotune.choose_class(128, {"128/8x16": 50.0, "64/4x16": 49.0}) == [128, 8, 16]
    assert autotune.choose_class(64, {"64/4x16": 50.0, "32/2x16": 40.0}) == [32, 2, 16]
    import os
    if os.path.exists(autotune.DEFAULT_PATH):
        table = autotune.load()
        for n, row in table["classes"].items():
            assert row["choice"] == autotune.choose_class(int(n), row["us"], table.get("margin", 0.0))
