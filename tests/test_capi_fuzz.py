"""Random descriptors through the general entry point's validation, on the
host: with null arrays no descriptor can reach a launch, so every call must
return an error status (or success for an empty batch) without crashing,
and tprh_workspace_bytes must answer every descriptor."""
import ctypes

from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

from tpr_amd import _lib

STATUSES = {_lib.SUCCESS, _lib.ERROR_INVALID_VALUE, _lib.ERROR_NOT_SUPPORTED}
pow2_or_not = st.one_of(st.sampled_from([0, 1, 2, 16, 256, 512, 1024, 2048, 4096, 65536, 1 << 20]),
                        st.integers(-8, 5000))


@settings(max_examples=400, deadline=None, suppress_health_check=[HealthCheck.too_slow])
@given(method=st.integers(-1, 4), arith=st.integers(-1, 2), n=pow2_or_not, s=pow2_or_not,
       nsys=st.integers(-2, 6), coef=st.integers(-1, 2), refine=st.integers(-1, 2), shared=st.integers(-1, 2),
       E=st.sampled_from([0, 1, 4, 8, 16, 32, 3]), L=st.sampled_from([0, 2, 16, 32, 64, 5]),
       chunk=st.integers(-2, 4))
def test_random_descriptors_fail_cleanly(method, arith, n, s, nsys, coef, refine, shared, E, L, chunk):
    d = _lib.SolveDesc()
    d.method, d.arith, d.n, d.s, d.nsys, d.coef = method, arith, n, s, nsys, coef
    d.refine, d.E, d.L, d.chunk, d.coef_shared = refine, E, L, chunk, shared
    lib = _lib.lib()
    st_ = lib.tprh_solve_ex(ctypes.byref(d), None)
    assert st_ in STATUSES, st_
    if st_ == _lib.SUCCESS:
        assert nsys == 0  # nothing to solve: no launch, and no claim of one
    else:
        assert lib.tprh_last_error()  # every failure says why
    lib.tprh_workspace_bytes(ctypes.byref(d))
