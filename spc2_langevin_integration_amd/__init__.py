"""MI355X-native force-summation and stepping for the spc2 chimera and
vortex-arrays self-propelled particle models.

The compute path is the gfx950 HIP library libspc2_hip.so behind the C ABI in
include/spc2_hip.h; `_abi` moves numpy arrays across it.
"""
from ._abi import (  # noqa: F401
    CHIMERA, VORTEX, FP64, FP32_KAHAN, FP64_FACTORED, SE, SRK, RK4, STEP_NO_WRAP, NEIGH_AUTO, NEIGH_CELLS,
    NEIGH_ALL_PAIRS, NEIGH_VERLET, Batch, Params, PinnedFrame, Server, Sim, load_host,
    Spc2Error, load, params, default_params, declared_symbols, LIB_PATH,
)
from .build import build  # noqa: F401
