"""The allocation-free cell list (ccp_cell_list_noalloc) against the handle's
ccp_cell_list on the same positions and geometry: bit for bit, on two calls
in a row through the C ABI, the second smaller than the first (its scratch
array keeps the first call's contents)."""
import ctypes

import numpy as np
import pytest
import torch  # before the library: torch's HIP runtime must be the one loaded first

import tfm_amd
from gpu_common import step
from helpers import radius_for, walk
from tfm_amd import ccp

pytestmark = pytest.mark.gpu


def noalloc_cell_list(xyz, csl, cps):
    n = xyz.shape[0]
    n_c = cps ** 3
    d_xyz = torch.from_numpy(np.ascontiguousarray(xyz, np.float32).reshape(-1)).cuda()
    out = {k: torch.zeros(n if k in ("sci", "soi") else n_c, dtype=torch.int32, device="cuda")
           for k in ("sci", "soi", "beg", "end")}
    torch.cuda.synchronize()
    L = ccp.lib()
    u32p = ctypes.POINTER(ctypes.c_uint32)
    ptr = lambda t: ctypes.cast(t.data_ptr(), u32p)
    rc = L.ccp_cell_list_noalloc(ctypes.cast(d_xyz.data_ptr(), ctypes.POINTER(ctypes.c_float)), n, float(csl),
                                 cps, ptr(out["sci"]), ptr(out["soi"]), ptr(out["beg"]), ptr(out["end"]), None)
    assert rc == 0, L.ccp_last_error(None).decode()
    return {k: v.cpu().numpy().view(np.uint32) for k, v in out.items()}


@pytest.mark.parametrize("N,cvf,seed", [(16384, 0.2, 4), (1000, 0.3, 3)])
def test_noalloc_cell_list_matches_cell_list(N, cvf, seed):
    step(f"allocation-free cell list N={N}")
    R_n = radius_for(N, cvf)
    r = walk(N, R_n, seed).astype(np.float32)
    sim = tfm_amd.Simulation({"number_of_particles": N, "nucleus_radius": repr(float(R_n)), "seed": 1})
    sim.positions = r
    want = sim.cell_list()
    info = sim.info
    got = noalloc_cell_list(r, info.csl, info.cps)
    for k in ("sci", "soi", "beg", "end"):
        assert np.array_equal(got[k], want[k]), k
    sim.close()


def test_noalloc_calls_in_a_row_and_limits():
    step("allocation-free cell list: two calls, the second smaller; size limits")
    out = []
    for N, cvf, seed in ((20000, 0.25, 8), (777, 0.2, 9)):
        R_n = radius_for(N, cvf)
        r = walk(N, R_n, seed).astype(np.float32)
        sim = tfm_amd.Simulation({"number_of_particles": N, "nucleus_radius": repr(float(R_n)), "seed": 1})
        sim.positions = r
        want = sim.cell_list()
        got = noalloc_cell_list(r, sim.info.csl, sim.info.cps)
        out.append(all(np.array_equal(got[k], want[k]) for k in ("sci", "soi", "beg", "end")))
        sim.close()
    assert out == [True, True]
    # more cells than the scratch array covers: refused before any launch
    L = ccp.lib()
    u32p = ctypes.POINTER(ctypes.c_uint32)
    dummy = (ctypes.c_uint32 * 4)()
    fdummy = (ctypes.c_float * 12)()
    rc = L.ccp_cell_list_noalloc(fdummy, 4, 1.0, 256, ctypes.cast(dummy, u32p), ctypes.cast(dummy, u32p),
                                 ctypes.cast(dummy, u32p), ctypes.cast(dummy, u32p), None)
    assert rc == tfm_amd.CCP_ERR_ARG
    assert "at most 255" in L.ccp_last_error(None).decode()
