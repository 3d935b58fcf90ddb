"""MI355X-native EM engine for the PBM, UBM, CCM and DBN click models.

The compute path is libmclicks.so (hand-written HIP kernels for gfx950,
C ABI in include/mclicks.h); this package is its Python host interface.
"""
from .model import (  # noqa: F401
    PBM, CCM, DBN, UBM, MODEL_NAMES, MAX_SERP, REDUCE_SEGMENTED, REDUCE_ATOMIC,
    ClickModel, Dataset, SessionSet, autotune, em_train,
)
