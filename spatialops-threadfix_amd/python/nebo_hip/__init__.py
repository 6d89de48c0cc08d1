"""Python binding of the MI355X Nebo evaluation engine (lib/libnebo_hip.so).

The engine itself is C++/HIP behind the C-ABI of include/nebo_hip.h; this package
builds expression descriptors, owns test fields and calls that ABI.
"""
from ._abi import NeboError, lib, header_symbols, LIB_PATH  # noqa: F401
from .fields import Layout, Field, Mask, LOCATIONS, FACE_TYPES  # noqa: F401
from .expr import (E, Expr, Stencil, cond, serialize, sin, cos, tan, exp, tanh, log, log10,  # noqa: F401
                   sqrt, square, cube, erf, erfc, inv_erf, inv_erfc, abs_, max_, min_, pow_, powint, mask_shift)
from .ops import build_stencils, basic_op_types, axis_coefs, OperatorDatabase  # noqa: F401
from .api import (lower_stats, assign, interior_assign, masked_assign, assign_points,  # noqa: F401
                  assign_limits, assign_plan,
                  field_max, field_min, field_sum, field_norm, field_max_interior,
                  field_min_interior, field_sum_interior, field_norm_interior,
                  fill_ghosts_periodic, divkgrad, lapl_euler, device_count, stencil_tiles,
                  stencil_set_tile, stencil_get_tile, stencil_reset_tile, stencil_select_tile,
                  stencil_autotune, KERNEL_DIVKGRAD, KERNEL_LAPL,
                  set_fused_jit, set_report_backend, last_backend, fused_jit_stats, fused_jit_tune, fused_jit_column, fused_source)
from .bc import MaskShift, BoundaryCondition, shift_points  # noqa: F401,E402
from .batch import Batch, alloc_patches  # noqa: F401,E402
from . import adjoint  # noqa: F401,E402
from . import transport  # noqa: F401,E402
from . import particles  # noqa: F401,E402
from . import workloads  # noqa: F401,E402
