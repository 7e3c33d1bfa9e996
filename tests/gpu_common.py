"""Shared set-up of GPU parity tests: the same inputs go to the HIP library
(through its C ABI) and to the CPU oracle."""
import sys

import numpy as np

import pyoracle as po
import tfm_amd


def step(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def make_pair(N, R_n, flags, r, ff=po.FF_CCP, lbs=None, chain_off=None, seed=9, dt=None, **kw):
    """A Simulation and an oracle System holding identical state."""
    params = {"number_of_particles": N, "nucleus_radius": repr(float(R_n)), "seed": seed,
              "number_of_lbs": 0 if lbs is None else len(lbs)}
    okw = {}
    if ff == po.FF_LJ:
        params.update({"force_field": "lj", "epsilon_het": kw.get("eps_het", 0.5),
                       "epsilon_lamina": kw.get("eps_lam", 1.0)})
        okw.update(bend_all=1, k_b=2.0, eps_het=kw.get("eps_het", 0.5), eps_lam=kw.get("eps_lam", 1.0))
    if dt is not None:
        params["timestep"] = repr(float(dt))
    for k in ("opening_radius", "bleb_radius", "lanes_per_bead", "fine_cell_size", "cell_reach", "integrate_pass", "group_types", "bal_rounds",
              "type_order", "cell_size"):
        if k in kw:
            params[k] = kw[k]
    if chain_off is not None:
        params["chain_lengths"] = ",".join(str(b - a) for a, b in zip(chain_off[:-1], chain_off[1:]))
    sim = tfm_amd.Simulation(params)
    r32 = np.ascontiguousarray(r, np.float32)
    sim.flags = flags
    sim.positions = r32
    if lbs is not None:
        sim.lbs = np.asarray(lbs, np.float32)
    lbs64 = None if lbs is None else np.asarray(lbs, np.float32).astype(np.float64)
    osys = po.System(N, float(np.float32(R_n)), flags=flags, lbs=lbs64, ff=ff,
                     chain_off=chain_off, seed=seed, dt=float(np.float32(dt if dt else 1 / 2048)),
                     R_o=kw.get("opening_radius", 0.0), R_b=kw.get("bleb_radius", 0.0),
                     fcap=float(sim.info.force_cap), **okw)
    # same grid as the oracle's cell list unless the test picks another cell shape
    assert osys.p.cps == sim.info.cps or sim.info.cps == 4 or "cell_reach" in kw
    assert abs(osys.p.sd - sim.info.sd) < 1e-9
    return sim, osys, r32.astype(np.float64)
