"""MI355X-native Brownian-dynamics engine for ccp's simulation path.

The compute runs in libccpmi.so (HIP kernels for gfx950, C ABI declared in
include/ccpmi.h); this package is its Python mirror of ccp's chrsim and
chrana classes.
"""
from .ccp import (CcpError, Simulation, Analysis, Batch, batch_layout, lib, params_text, series_stats, weighted_stats,  # noqa: F401
                  HET, LAD, FF_CCP, FF_ICG, FF_LJ,
                  CCP_OK, CCP_ERR_PARAM, CCP_ERR_HIP, CCP_ERR_IO, CCP_ERR_STATE, CCP_ERR_ARG)
