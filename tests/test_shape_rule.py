"""The default slice shape of the batched entry points (tprh_slice_shape, no
device call): the rows of the committed tuner table
(tpr_amd/schedules/batched_gfx950.json, compiled into the library from
tpr_amd/csrc/batched_shapes.inc), the nearest tuned batch size for other
batch sizes, and the default shape wherever the table has no row."""
import json
import os

import pytest

from tpr_amd import _lib
from tpr_amd import autotune as A
from tpr_amd.batched import slice_shape

TABLE = A.load(A.BATCHED_PATH)
ROWS = sorted(TABLE["cases"].items())


def test_header_is_the_table():
    """The rows compiled into the library are the committed table's."""
    with open(A.HEADER_PATH) as f:
        assert f.read() == A.header_text(TABLE)


@pytest.mark.parametrize("key,row", ROWS, ids=[k for k, _ in ROWS])
def test_library_takes_the_tuned_shape(key, row):
    method, arith, n, s, nsys = A.parse_batched_key(key)
    assert slice_shape(n, s, nsys, method, arith) == tuple(row["choice"])


@pytest.mark.parametrize("key,row", ROWS, ids=[k for k, _ in ROWS])
def test_choice_keeps_the_answers_and_is_fastest(key, row):
    """The choice gave the default shape's bits and is within the tuner's
    margin of the fastest shape that did; when the default is within that
    margin it is the choice."""
    method, arith, n, s, nsys = A.parse_batched_key(key)
    choice = "%dx%d" % tuple(row["choice"])
    default = "%dx%d" % A.default_batched_shape(s)
    assert row["same_bits"][choice] and row["same_bits"][default]
    same = [k for k, v in row["same_bits"].items() if v]
    best = min(row["us"][k] for k in same)
    margin = TABLE["margin"]
    assert row["us"][choice] <= best * (1 + margin) + 1e-9
    if row["us"][default] <= best * (1 + margin):
        assert choice == default
    if arith == "exact":
        assert all(row["same_bits"].values())  # exact answers never depend on the shape


def test_same_bits_follow_the_arithmetic_form():
    """On the tuning inputs, exact answers were the default shape's for every
    candidate, and fast answers exactly for the candidates that run the
    default's form of the fast stage 1 (tpr_amd.autotune.fast_form); config
    1 (fast PTPR, 4096 x 1024, s = 256) keeps 16 x 16."""
    for key, row in ROWS:
        method, arith, n, s, nsys = A.parse_batched_key(key)
        if arith == "exact":
            assert all(row["same_bits"].values()), key
            continue
        dform = A.fast_form(method, *A.default_batched_shape(s))
        for shape, same in row["same_bits"].items():
            E, L = (int(v) for v in shape.split("x"))
            assert same == (A.fast_form(method, E, L) == dform), (key, shape)
    assert slice_shape(1024, 256, 4096, "ptpr", "fast") == (16, 16)


def _row(method, arith, n, s, nsys):
    return tuple(TABLE["cases"][A.batched_key(method, arith, n, s, nsys)]["choice"])


@pytest.mark.parametrize("nsys,tuned", [(0, 1), (2, 1), (3, 1), (100, 256), (1000, 256), (3000, 4096),
                                        (100000, 4096)])
def test_nearest_tuned_batch_size(nsys, tuned):
    for method in ("ptpr", "tpr"):
        for arith in ("fast", "exact"):
            assert slice_shape(2048, 512, nsys, method, arith) == _row(method, arith, 2048, 512, tuned)


def test_capped_batch_row_for_the_largest_class():
    """n = 8192 was tuned with 2048 systems (the 16 Mi-equation cap)."""
    for method in ("ptpr", "tpr"):
        assert slice_shape(8192, 512, 4096, method, "exact") == _row(method, "exact", 8192, 512, 2048)


@pytest.mark.parametrize("n,s,shape", [(512, 128, (8, 16)), (4096, 1024, (16, 64)), (16, 16, (1, 16)),
                                       (8192, 256, (16, 16)), (32768, 2048, (32, 64))])
def test_default_shape_outside_the_table(n, s, shape):
    for method in ("ptpr", "tpr"):
        for arith in ("fast", "exact"):
            assert slice_shape(n, s, 7, method, arith) == shape


def test_table_covers_every_tuned_case():
    keys = set(TABLE["cases"])
    for method in ("ptpr", "tpr"):
        for arith in ("fast", "exact"):
            for s in (256, 512):
                for m in (1, 2, 4, 8, 16):
                    n = s * m
                    for nsys in (1, 256, min(4096, TABLE["eq_cap"] // n)):
                        assert A.batched_key(method, arith, n, s, nsys) in keys


def test_errors_as_the_batched_entry():
    with pytest.raises(_lib.TprError) as e:
        slice_shape(1000, 256)
    assert e.value.status == _lib.ERROR_INVALID_VALUE
    with pytest.raises(_lib.TprError) as e:
        slice_shape(65536, 256)       # 256 slices x 16 lanes: beyond one workgroup
    assert e.value.status == _lib.ERROR_NOT_SUPPORTED


def test_table_file_is_json():
    with open(A.BATCHED_PATH) as f:
        assert json.load(f)["cases"]
    assert os.path.dirname(A.BATCHED_PATH).endswith("schedules")
