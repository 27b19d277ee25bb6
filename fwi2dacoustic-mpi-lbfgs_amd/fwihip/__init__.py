"""fwihip: host-side Python view of the MI355X wave-equation core.

The compute path is the C-ABI library libfwi_hip.so (HIP kernels for gfx950);
this package only loads it (fwihip.lib), describes workloads
(fwihip.workloads) and wraps the adjoint for torch (fwihip.autograd).
"""
from . import workloads  # noqa: F401
