"""gpupfem2_amd: the PFEM-2 hot path of andreyypopov/gpuPfem2 on MI355X (gfx950).

The compute lives in libpfem2_hip.so (hand-written HIP kernels behind the C ABI of include/pfem2.h);
this package holds its ctypes binding (capi), the mesh generators (meshgen), the case setups (cases)
and the build recipe (build).
"""
from . import capi, cases, meshgen  # noqa: F401
