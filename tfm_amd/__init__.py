"""MI355X-native Brownian-dynamics engine for ccp's simulation path.

The compute runs in libccpmi.so (HIP kernels for gfx950, C ABI declared in
include/ccpmi.h); this package is its Python mirror of ccp's chrsim class.
"""
from .ccp import (CcpError, Simulation, lib, params_text, HET, LAD, FF_CCP, FF_ICG, FF_LJ,  # noqa: F401
                  CCP_OK, CCP_ERR_PARAM, CCP_ERR_HIP, CCP_ERR_IO, CCP_ERR_STATE, CCP_ERR_ARG)
