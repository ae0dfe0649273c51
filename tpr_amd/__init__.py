"""tpr_amd -- MI355X-native (gfx950) TPR/PTPR tridiagonal solver engine.

The solvers are hand-written HIP kernels in libtpr_hip.so, reached through the
C ABI of include/tpr_hip.h.  This package mirrors the reference's solver
classes and CUDA entry points (``PTPR``, ``TPR``, ``PCR``, ``CR``, ``ptpr_cu``,
``tpr_cu``, ``pcr_cu``, ``cr_cu``) and adds device-resident batched solves
(``solve_batched``) and the seeded generator (``generate``).
"""
from ._lib import TprError, lib  # noqa: F401
from .batched import alloc_batch, generate, shape_supported, solve_batched, solve_ex, workspace_bytes  # noqa: F401
from .uneven import UnevenPlan  # noqa: F401
from .solvers import CR, PCR, PTPR, TPR, cr_cu, pcr_cu, ptpr_cu, tpr_cu  # noqa: F401
