"""ctypes access to the CPU oracle (oracle/_build/libccp_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libccp_oracle.so")

FF_CCP, FF_ICG, FF_LJ = 0, 1, 2
HET, LAD = 1, 2


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int32), ("n_l", ctypes.c_int32), ("n_chains", ctypes.c_int32),
        ("ff", ctypes.c_int32), ("bend_all", ctypes.c_int32), ("bond_type", ctypes.c_int32),
        ("cps", ctypes.c_int32), ("pad0", ctypes.c_int32),
        ("R_n", ctypes.c_double), ("R_o", ctypes.c_double), ("R_b", ctypes.c_double),
        ("csl", ctypes.c_double),
        ("k_e", ctypes.c_double), ("l_0", ctypes.c_double), ("k_b", ctypes.c_double),
        ("e_p", ctypes.c_double), ("rco", ctypes.c_double), ("aco", ctypes.c_double),
        ("e_l", ctypes.c_double), ("lco", ctypes.c_double),
        ("fene_k", ctypes.c_double), ("fene_r0", ctypes.c_double),
        ("eps_wca", ctypes.c_double), ("sigma", ctypes.c_double),
        ("eps_het", ctypes.c_double), ("rc_lj", ctypes.c_double),
        ("eps_wall", ctypes.c_double), ("eps_lam", ctypes.c_double), ("rc_lam", ctypes.c_double),
        ("dt", ctypes.c_double), ("sd", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("fcap", ctypes.c_double),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        u32p, f32p, f64p = P(ctypes.c_uint32), P(ctypes.c_float), P(ctypes.c_double)
        L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.orc_noise.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, f64p]
        L.orc_noise.restype = ctypes.c_int
        L.orc_cell_list.argtypes = [f32p, ctypes.c_int32, ctypes.c_float, ctypes.c_int32] + [u32p] * 6
        L.orc_cell_list.restype = ctypes.c_int
        common = [P(OrcParams), P(ctypes.c_uint8), P(ctypes.c_int32), f64p]
        L.orc_forces.argtypes = common + [f64p, f64p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        L.orc_energy.argtypes = common + [f64p, f64p]
        L.orc_run.argtypes = common + [f64p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        L.orc_contacts.argtypes = [f32p, ctypes.c_int32, ctypes.c_float, ctypes.c_int32, ctypes.c_int32, u32p]
        L.orc_noise_batch.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int32, f64p, P(ctypes.c_int32)]
        L.orc_contacts_grid.argtypes = [f32p, ctypes.c_int32, ctypes.c_float, ctypes.c_int32, ctypes.c_int32, u32p]
        L.orc_observables.argtypes = [f64p, P(ctypes.c_uint8), ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int32, f64p, f64p]
        L.orc_sd_cp.argtypes = [f32p, ctypes.c_int32, ctypes.c_int32, f32p, f32p]
        L.orc_ccp_contact_map.argtypes = [f32p, ctypes.c_int32, f32p]
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


# ccp model constants (chrdat.cuh:15-24, chrsim.cuh:18-20) and the WCA/LJ set.
K_B = float(np.float32(0.003356))
CCP_CONST = dict(k_e=128.0, l_0=1.0, k_b=1.0, e_p=0.5, rco=float(np.float32(1.154701)), aco=2.0,
                 e_l=8.0, lco=0.5)
LJ_CONST = dict(fene_k=30.0, fene_r0=1.5, eps_wca=1.0, sigma=1.0, eps_het=0.5, rc_lj=2.5,
                eps_wall=1.0, eps_lam=1.0, rc_lam=2.5)


class System:
    """A bead system description shared by the oracle calls."""

    def __init__(self, N, R_n, chain_off=None, flags=None, lbs=None, ff=FF_CCP, csl=None,
                 R_o=0.0, R_b=0.0, T=298.0, dt=1.0 / 2048, seed=1, bend_all=0, bond_type=0, **kw):
        self.p = OrcParams()
        self.p.N = N
        self.chain_off = np.ascontiguousarray(chain_off if chain_off is not None else [0, N], dtype=np.int32)
        self.p.n_chains = len(self.chain_off) - 1
        self.flags = np.ascontiguousarray(flags if flags is not None else np.full(N, HET | LAD), dtype=np.uint8)
        self.lbs = np.ascontiguousarray(lbs if lbs is not None else np.zeros((0, 3)), dtype=np.float64).reshape(-1, 3)
        self.p.n_l = self.lbs.shape[0]
        self.p.ff = ff
        self.p.bend_all = bend_all
        self.p.bond_type = bond_type
        for k, v in {**CCP_CONST, **LJ_CONST}.items():
            setattr(self.p, k, v)
        for k, v in kw.items():
            setattr(self.p, k, v)
        self.p.R_n, self.p.R_o, self.p.R_b = R_n, R_o, R_b
        if csl is None:
            csl = 2.0 if ff != FF_LJ else 2.5
        self.p.csl = csl
        d_m = float(np.sqrt(R_n**2 - R_o**2) + np.sqrt(max(R_b**2 - R_o**2, 0.0)) + R_b)
        self.p.cps = int(2 * np.ceil(np.float32(d_m) / np.float32(csl)))
        self.p.dt = dt
        self.p.sd = float(np.float32(np.sqrt(2.0 * K_B * np.float64(np.float32(T)) * np.float64(np.float32(dt)))))
        self.p.seed = seed

    def _args(self):
        return (ctypes.byref(self.p), _p(self.flags, ctypes.c_uint8), _p(self.chain_off, ctypes.c_int32),
                _p(self.lbs, ctypes.c_double))

    def forces(self, r, method=0, prec=64, threads=0):
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 3)
        f = np.zeros_like(r)
        lib().orc_forces(*self._args(), _p(r, ctypes.c_double), _p(f, ctypes.c_double), method, prec, threads)
        return f

    def energy(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 3)
        t = np.zeros(5)
        lib().orc_energy(*self._args(), _p(r, ctypes.c_double), _p(t, ctypes.c_double))
        return t

    def run(self, r, step0, nsteps, prec=64, threads=0):
        r = np.array(r, dtype=np.float64).reshape(-1, 3)
        lib().orc_run(*self._args(), _p(r, ctypes.c_double), step0, nsteps, prec, threads)
        return r


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c, ctypes.c_uint32), _p(k, ctypes.c_uint32), _p(out, ctypes.c_uint32))
    return out


def noise(seed, bead, step):
    z = np.zeros(3)
    a = lib().orc_noise(seed, bead, step, _p(z, ctypes.c_double))
    return z, a


def noise_batch(seed, step, n):
    z = np.zeros((n, 3))
    a = np.zeros(n, np.int32)
    lib().orc_noise_batch(seed, step, n, _p(z, ctypes.c_double), _p(a, ctypes.c_int32))
    return z, a


def contacts_grid(r, rc, bin_size, n_bins):
    r = np.ascontiguousarray(r, dtype=np.float32).reshape(-1, 3)
    m = np.zeros((n_bins, n_bins), np.uint32)
    lib().orc_contacts_grid(_p(r, ctypes.c_float), r.shape[0], rc, bin_size, n_bins, _p(m, ctypes.c_uint32))
    return m


def cell_list(r, csl, cps):
    r = np.ascontiguousarray(r, dtype=np.float32).reshape(-1, 3)
    n = r.shape[0]
    n_c = cps**3
    uci, sci, uoi, soi = (np.zeros(n, np.uint32) for _ in range(4))
    beg, end = np.zeros(n_c, np.uint32), np.zeros(n_c, np.uint32)
    bad = lib().orc_cell_list(_p(r, ctypes.c_float), n, csl, cps, *(_p(a, ctypes.c_uint32) for a in
                                                                   (uci, sci, uoi, soi, beg, end)))
    return dict(uci=uci, sci=sci, uoi=uoi, soi=soi, beg=beg, end=end, bad=bad)


def contacts(r, rc, bin_size, n_bins):
    r = np.ascontiguousarray(r, dtype=np.float32).reshape(-1, 3)
    m = np.zeros((n_bins, n_bins), np.uint32)
    lib().orc_contacts(_p(r, ctypes.c_float), r.shape[0], rc, bin_size, n_bins, _p(m, ctypes.c_uint32))
    return m


def ccp_contact_map(r):
    r = np.ascontiguousarray(r, dtype=np.float32).reshape(-1, 3)
    n = r.shape[0]
    cms = (n // 2) // 4 if n == 30386 else n // 4
    cm = np.zeros(cms * (cms + 1) // 2, np.float32)
    lib().orc_ccp_contact_map(_p(r, ctypes.c_float), n, _p(cm, ctypes.c_float))
    return cm


def observables(r, flags, R_n, nod=None, n_b=64):
    """ccp's frame observables: (dcm, rg2, nop, ncf), rcd[3, n_b]."""
    r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 3)
    fl = np.ascontiguousarray(flags, dtype=np.uint8)
    out = np.zeros(4)
    rcd = np.zeros((3, n_b))
    lib().orc_observables(_p(r, ctypes.c_double), _p(fl, ctypes.c_uint8), r.shape[0], R_n,
                          R_n if nod is None else nod, n_b, _p(out, ctypes.c_double), _p(rcd, ctypes.c_double))
    return out, rcd


def sd_cp(r, chromosome=0):
    """ccp's d(s) and 1000 P(s) of one chromosome (float32 arrays of lsdcp)."""
    r = np.ascontiguousarray(r, dtype=np.float32).reshape(-1, 3)
    n = r.shape[0]
    lim = [0, 2284, 4599, 6687, 9338, 12508, 15193, 17477, 19792, 21880, 24531, 27701, 30386]
    m = (lim[chromosome + 1] - lim[chromosome]) // 2 if n == 30386 and chromosome < 6 else n // 2
    sd, cp = np.zeros(m, np.float32), np.zeros(m, np.float32)
    got = lib().orc_sd_cp(_p(r, ctypes.c_float), n, chromosome, _p(sd, ctypes.c_float), _p(cp, ctypes.c_float))
    if got != m:
        raise ValueError("no such chromosome")
    return sd, cp
