"""The default slice shape (tprh_slice_shape, no device call): the
measured rule of wider shapes where they are faster without changing a
result, and the default shape everywhere else."""
import pytest

from tpr_amd import _lib
from tpr_amd.batched import slice_shape

CASES = [
    # (n, s, nsys, method, arith) -> (E, L)
    ((1024, 256, 4096, "ptpr", "fast"), (16, 16)),     # config 1: the default
    ((1024, 256, 1, "ptpr", "fast"), (16, 16)),        # fast answers would change with 4 x 64
    ((1024, 256, 4096, "tpr", "fast"), (16, 16)),
    ((1024, 256, 4096, "ptpr", "exact"), (4, 64)),     # exact PTPR, m <= 16
    ((4096, 256, 7, "ptpr", "exact"), (4, 64)),
    ((8192, 256, 7, "ptpr", "exact"), (16, 16)),       # m = 32: 4 x 64 would need 2048 lanes
    ((1024, 256, 256, "tpr", "exact"), (8, 32)),       # exact TPR, small batches
    ((1024, 256, 257, "tpr", "exact"), (16, 16)),
    ((4096, 256, 1, "tpr", "exact"), (16, 16)),        # m = 16
    ((2048, 512, 1, "ptpr", "fast"), (8, 64)),         # s = 512: PTPR up to m = 8
    ((4096, 512, 4096, "ptpr", "exact"), (8, 64)),
    ((8192, 512, 1, "ptpr", "fast"), (16, 32)),
    ((2048, 512, 1, "tpr", "fast"), (8, 64)),          # TPR up to m = 4
    ((4096, 512, 1, "tpr", "fast"), (16, 32)),
    ((512, 128, 1, "ptpr", "exact"), (8, 16)),         # other s: default_shape
    ((4096, 1024, 1, "ptpr", "fast"), (16, 64)),
    ((16, 16, 1, "tpr", "fast"), (1, 16)),
]


@pytest.mark.parametrize("args,shape", CASES)
def test_rule(args, shape):
    n, s, nsys, method, arith = args
    assert slice_shape(n, s, nsys, method, arith) == shape


def test_errors_as_the_batched_entry():
    with pytest.raises(_lib.TprError) as e:
        slice_shape(1000, 256)
    assert e.value.status == _lib.ERROR_INVALID_VALUE
    with pytest.raises(_lib.TprError) as e:
        slice_shape(65536, 256)       # 256 slices x 16 lanes: beyond one workgroup
    assert e.value.status == _lib.ERROR_NOT_SUPPORTED
