"""Python mirror of ccp's simulation class over the libccpmi C ABI.

`Simulation` stands in for mmc::chrsim (Program/src/chrsim.cuh:36-57): it is
built from the adjustable-parameters text, generates initial conditions,
saves/loads checkpoints, writes ccp's .gro/.trr files and runs the step on
the GPU. Errors raise CcpError carrying ccp's own messages (chrdat.cu:21-63).
There is no CPU path: without the built HIP library every call fails.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libccpmi.so")

CCP_OK, CCP_ERR_PARAM, CCP_ERR_HIP, CCP_ERR_IO, CCP_ERR_STATE, CCP_ERR_ARG = range(6)
HET, LAD = 1, 2
FF_CCP, FF_ICG, FF_LJ = 0, 1, 2

# Every symbol declared in include/ccpmi.h.
EXPORTS = (
    "ccp_create", "ccp_create_logged", "ccp_destroy", "ccp_last_error", "ccp_set_logger", "ccp_get_info",
    "ccp_generate_initial_condition", "ccp_save_checkpoint", "ccp_load_checkpoint", "ccp_write_frame_txt",
    "ccp_read_frame_txt", "ccp_write_lbs_txt", "ccp_read_lbs_txt", "ccp_run_simulation", "ccp_run_frames",
    "ccp_run_steps",
    "ccp_synchronize", "ccp_get_time", "ccp_set_time", "ccp_get_positions", "ccp_set_positions",
    "ccp_get_types", "ccp_set_types", "ccp_get_flags", "ccp_set_flags", "ccp_get_lbs", "ccp_set_lbs",
    "ccp_get_chains", "ccp_compute_forces", "ccp_cell_list", "ccp_noise", "ccp_count_overlaps",
    "ccp_contact_sample", "ccp_contact_reset", "ccp_contact_map", "ccp_time_kernels", "ccp_capture_graphs",
    "ccp_cell_list_noalloc", "ccp_step_grid",
    "ccp_observables",
    "ccp_distance_contact", "ccp_frame_contact_map", "ccp_analysis_create", "ccp_analysis_destroy",
    "ccp_analysis_add_frame", "ccp_analysis_add_trajectory_file", "ccp_analysis_end_simulation",
    "ccp_analysis_finish", "ccp_analysis_last_error", "ccp_series_stats", "ccp_weighted_stats",
    "ccp_batch_create", "ccp_batch_layout", "ccp_batch_destroy", "ccp_batch_last_error", "ccp_batch_info", "ccp_batch_system_params",
    "ccp_batch_generate_initial_conditions", "ccp_batch_run_steps", "ccp_batch_synchronize", "ccp_batch_get_state",
    "ccp_batch_set_state", "ccp_batch_get_step", "ccp_batch_set_step", "ccp_batch_contact_maps",
    "ccp_batch_radial_density", "ccp_batch_reset_samples", "ccp_batch_time_steps", "ccp_sequence_pattern",
)


class CcpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class Info(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "N", "n_l", "n_chains", "force_field", "cps", "n_c", "cell_capacity", "fine_cps", "spf", "fpf",
        "contact_bins", "contact_bin_size", "contact_interval")] + \
        [(n, ctypes.c_float) for n in ("csl", "fine_csl", "R_n", "R_o", "R_b", "T", "dt", "sd",
                                       "contact_cutoff", "force_cap")] + \
        [("seed", ctypes.c_uint64)]


LOG_FN = ctypes.CFUNCTYPE(None, ctypes.c_char_p, ctypes.c_void_p)
FRAME_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_uint32, ctypes.c_float, ctypes.POINTER(ctypes.c_float),
                            ctypes.c_size_t, ctypes.c_void_p)


class BatchSystem(ctypes.Structure):
    _fields_ = [("n_beads", ctypes.c_uint32), ("n_chains", ctypes.c_uint32), ("volume_fraction", ctypes.c_float),
                ("eps_het", ctypes.c_float), ("eps_lam", ctypes.c_float), ("het_fraction", ctypes.c_float),
                ("lad_fraction", ctypes.c_float), ("seed", ctypes.c_uint64), ("replica", ctypes.c_uint32),
                ("pad", ctypes.c_uint32)]

# ccp's chromosome limits at N = 30386 (chrdat.cuh:27-29)
_CHR_LIM = (0, 2284, 4599, 6687, 9338, 12508, 15193, 17477, 19792, 21880, 24531, 27701, 30386)

_lib = None


def lib():
    """Load libccpmi.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m tfm_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        f32p, u32p, i32p, u8p = P(ctypes.c_float), P(ctypes.c_uint32), P(ctypes.c_int32), P(ctypes.c_uint8)
        L.ccp_create.argtypes = [ctypes.c_char_p]
        L.ccp_create.restype = vp
        L.ccp_create_logged.argtypes = [ctypes.c_char_p, LOG_FN, vp]
        L.ccp_create_logged.restype = vp
        L.ccp_destroy.argtypes = [vp]
        L.ccp_last_error.argtypes = [vp]
        L.ccp_last_error.restype = ctypes.c_char_p
        L.ccp_set_logger.argtypes = [vp, LOG_FN, vp]
        L.ccp_get_info.argtypes = [vp, P(Info)]
        for n in ("ccp_generate_initial_condition", "ccp_synchronize", "ccp_contact_sample", "ccp_contact_reset"):
            getattr(L, n).argtypes = [vp]
        for n in ("ccp_save_checkpoint", "ccp_load_checkpoint", "ccp_write_frame_txt", "ccp_read_frame_txt",
                  "ccp_write_lbs_txt", "ccp_read_lbs_txt", "ccp_run_simulation"):
            getattr(L, n).argtypes = [vp, ctypes.c_char_p]
        L.ccp_run_steps.argtypes = [vp, ctypes.c_uint64]
        L.ccp_run_frames.argtypes = [vp, ctypes.c_uint32, FRAME_FN, vp]
        L.ccp_get_time.argtypes = [vp, u32p, f32p, P(ctypes.c_uint64)]
        L.ccp_set_time.argtypes = [vp, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint64]
        L.ccp_get_positions.argtypes = [vp, f32p, sz]
        L.ccp_set_positions.argtypes = [vp, f32p, sz]
        L.ccp_get_types.argtypes = [vp, i32p, sz]
        L.ccp_set_types.argtypes = [vp, i32p, sz]
        L.ccp_get_flags.argtypes = [vp, u8p, sz]
        L.ccp_set_flags.argtypes = [vp, u8p, sz]
        L.ccp_get_lbs.argtypes = [vp, f32p, sz]
        L.ccp_set_lbs.argtypes = [vp, f32p, sz]
        L.ccp_get_chains.argtypes = [vp, u32p, sz]
        L.ccp_compute_forces.argtypes = [vp, ctypes.c_int, f32p, sz]
        L.ccp_cell_list.argtypes = [vp, u32p, u32p, u32p, u32p, sz, sz]
        L.ccp_noise.argtypes = [vp, ctypes.c_uint64, f32p, u32p, sz]
        L.ccp_count_overlaps.argtypes = [vp, ctypes.c_float, P(ctypes.c_uint64)]
        L.ccp_contact_map.argtypes = [vp, u32p, sz, u32p]
        L.ccp_time_kernels.argtypes = [vp, ctypes.c_uint32, f32p]
        L.ccp_capture_graphs.argtypes = [vp]
        L.ccp_step_grid.argtypes = [vp, u32p, sz, u32p, sz]
        L.ccp_cell_list_noalloc.argtypes = [f32p, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint32, u32p, u32p, u32p,
                                            u32p, vp]
        f64p = P(ctypes.c_double)
        L.ccp_observables.argtypes = [vp, f64p, f64p, sz]
        L.ccp_distance_contact.argtypes = [vp, ctypes.c_uint32, f32p, f32p, sz]
        L.ccp_frame_contact_map.argtypes = [vp, f32p, sz]
        L.ccp_analysis_create.argtypes = [vp, P(vp)]
        L.ccp_analysis_destroy.argtypes = [vp]
        L.ccp_analysis_add_frame.argtypes = [vp]
        L.ccp_analysis_add_trajectory_file.argtypes = [vp, ctypes.c_char_p]
        L.ccp_analysis_end_simulation.argtypes = [vp, ctypes.c_char_p]
        L.ccp_analysis_finish.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p]
        L.ccp_analysis_last_error.argtypes = [vp]
        L.ccp_analysis_last_error.restype = ctypes.c_char_p
        L.ccp_series_stats.argtypes = [f32p, sz, f64p]
        L.ccp_weighted_stats.argtypes = [f64p, f64p, sz, f64p]
        L.ccp_batch_create.argtypes = [ctypes.c_char_p, P(BatchSystem), ctypes.c_uint32]
        L.ccp_batch_create.restype = vp
        L.ccp_batch_layout.argtypes = [ctypes.c_char_p, P(BatchSystem), ctypes.c_uint32, u32p,
                                       P(ctypes.c_uint64), P(ctypes.c_uint64), P(ctypes.c_uint64)]
        L.ccp_batch_destroy.argtypes = [vp]
        L.ccp_batch_last_error.argtypes = [vp]
        L.ccp_batch_last_error.restype = ctypes.c_char_p
        L.ccp_batch_info.argtypes = [vp, u32p, u32p, u32p, u32p]
        L.ccp_batch_system_params.argtypes = [vp, ctypes.c_uint32, ctypes.c_char_p, sz]
        for n in ("ccp_batch_generate_initial_conditions", "ccp_batch_synchronize", "ccp_batch_reset_samples"):
            getattr(L, n).argtypes = [vp]
        L.ccp_batch_run_steps.argtypes = [vp, ctypes.c_uint64]
        L.ccp_batch_get_state.argtypes = [vp, f32p, u8p, sz]
        L.ccp_batch_set_state.argtypes = [vp, f32p, u8p, sz]
        L.ccp_batch_get_step.argtypes = [vp, P(ctypes.c_uint64)]
        L.ccp_batch_set_step.argtypes = [vp, ctypes.c_uint64]
        L.ccp_batch_contact_maps.argtypes = [vp, u32p, sz, u32p]
        L.ccp_batch_radial_density.argtypes = [vp, f64p, sz]
        L.ccp_batch_time_steps.argtypes = [vp, ctypes.c_uint32, f32p]
        L.ccp_sequence_pattern.argtypes = [ctypes.c_char_p, u8p, sz]
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def params_text(params):
    if isinstance(params, str):
        return params
    return "".join(f"{k} {v}\n" for k, v in params.items())


def _quiet(line, user):
    pass


_QUIET = LOG_FN(_quiet)


class Simulation:
    """One simulation on the current GPU (ccp's chrsim)."""

    def __init__(self, params, quiet=True):
        L = lib()
        text = params_text(params).encode()
        self._log_keep = _QUIET if quiet else None
        h = L.ccp_create_logged(text, _QUIET, None) if quiet else L.ccp_create(text)
        if not h:
            msg = L.ccp_last_error(None).decode()
            code = CCP_ERR_HIP if msg.startswith("hip:") else CCP_ERR_PARAM
            raise CcpError(code, msg)
        self.h = ctypes.c_void_p(h)
        self.info = Info()
        L.ccp_get_info(self.h, ctypes.byref(self.info))
        self.N = self.info.N
        self.n_l = self.info.n_l

    def close(self):
        if getattr(self, "h", None):
            lib().ccp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != CCP_OK:
            raise CcpError(rc, lib().ccp_last_error(self.h).decode())

    # chrsim members
    def generate_initial_condition(self):
        self._check(lib().ccp_generate_initial_condition(self.h))

    def save_checkpoint(self, path):
        self._check(lib().ccp_save_checkpoint(self.h, str(path).encode()))

    def load_checkpoint(self, path):
        self._check(lib().ccp_load_checkpoint(self.h, str(path).encode()))

    def run_simulation(self, path=None):
        self._check(lib().ccp_run_simulation(self.h, None if path is None else str(path).encode()))

    def run_frames(self, n_frames, fn):
        """Frames of steps_per_frame steps; fn(i_f, t, xyz[N, 3]) after each
        (a copy of the host positions); a true return value stops the run."""
        def call(i_f, t, xyz, n, user):
            return 1 if fn(i_f, t, np.ctypeslib.as_array(xyz, shape=(n, 3)).copy()) else 0
        cb = FRAME_FN(call)
        self._check(lib().ccp_run_frames(self.h, int(n_frames), cb, None))

    def write_frame_txt(self, path):
        self._check(lib().ccp_write_frame_txt(self.h, str(path).encode()))

    def read_frame_txt(self, path):
        self._check(lib().ccp_read_frame_txt(self.h, str(path).encode()))

    def write_lbs_txt(self, path):
        self._check(lib().ccp_write_lbs_txt(self.h, str(path).encode()))

    def read_lbs_txt(self, path):
        self._check(lib().ccp_read_lbs_txt(self.h, str(path).encode()))

    # the step
    def run_steps(self, n):
        self._check(lib().ccp_run_steps(self.h, int(n)))

    def synchronize(self):
        self._check(lib().ccp_synchronize(self.h))

    # state
    @property
    def time(self):
        i_f, t, s = ctypes.c_uint32(), ctypes.c_float(), ctypes.c_uint64()
        lib().ccp_get_time(self.h, ctypes.byref(i_f), ctypes.byref(t), ctypes.byref(s))
        return i_f.value, t.value, s.value

    def set_time(self, i_f, t, step):
        self._check(lib().ccp_set_time(self.h, i_f, t, step))

    @property
    def positions(self):
        out = np.zeros((self.N, 3), np.float32)
        self._check(lib().ccp_get_positions(self.h, _ptr(out, ctypes.c_float), self.N))
        return out

    @positions.setter
    def positions(self, xyz):
        a = np.ascontiguousarray(xyz, dtype=np.float32).reshape(self.N, 3)
        self._check(lib().ccp_set_positions(self.h, _ptr(a, ctypes.c_float), self.N))

    @property
    def types(self):
        out = np.zeros(self.N, np.int32)
        self._check(lib().ccp_get_types(self.h, _ptr(out, ctypes.c_int32), self.N))
        return out

    @types.setter
    def types(self, pt):
        a = np.ascontiguousarray(pt, dtype=np.int32)
        self._check(lib().ccp_set_types(self.h, _ptr(a, ctypes.c_int32), self.N))

    @property
    def flags(self):
        out = np.zeros(self.N, np.uint8)
        self._check(lib().ccp_get_flags(self.h, _ptr(out, ctypes.c_uint8), self.N))
        return out

    @flags.setter
    def flags(self, fl):
        a = np.ascontiguousarray(fl, dtype=np.uint8)
        self._check(lib().ccp_set_flags(self.h, _ptr(a, ctypes.c_uint8), self.N))

    @property
    def lbs(self):
        out = np.zeros((self.n_l, 3), np.float32)
        self._check(lib().ccp_get_lbs(self.h, _ptr(out, ctypes.c_float), self.n_l))
        return out

    @lbs.setter
    def lbs(self, xyz):
        a = np.ascontiguousarray(xyz, dtype=np.float32).reshape(self.n_l, 3)
        self._check(lib().ccp_set_lbs(self.h, _ptr(a, ctypes.c_float), self.n_l))

    @property
    def chains(self):
        out = np.zeros(self.info.n_chains + 1, np.uint32)
        self._check(lib().ccp_get_chains(self.h, _ptr(out, ctypes.c_uint32), out.size))
        return out

    # parity / measurement access
    def compute_forces(self, ff=None):
        ff = self.info.force_field if ff is None else ff
        out = np.zeros((self.N, 3), np.float32)
        self._check(lib().ccp_compute_forces(self.h, ff, _ptr(out, ctypes.c_float), self.N))
        return out

    def cell_list(self):
        n, n_c = self.N, self.info.n_c
        sci, soi = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
        beg, end = np.zeros(n_c, np.uint32), np.zeros(n_c, np.uint32)
        self._check(lib().ccp_cell_list(self.h, *(_ptr(a, ctypes.c_uint32) for a in (sci, soi, beg, end)), n, n_c))
        return dict(sci=sci, soi=soi, beg=beg, end=end)

    def noise(self, step):
        z = np.zeros((self.N, 3), np.float32)
        a = np.zeros(self.N, np.uint32)
        self._check(lib().ccp_noise(self.h, step, _ptr(z, ctypes.c_float), _ptr(a, ctypes.c_uint32), self.N))
        return z, a

    def count_overlaps(self, dmin):
        c = ctypes.c_uint64()
        self._check(lib().ccp_count_overlaps(self.h, dmin, ctypes.byref(c)))
        return c.value

    def contact_sample(self):
        self._check(lib().ccp_contact_sample(self.h))

    def contact_reset(self):
        self._check(lib().ccp_contact_reset(self.h))

    def contact_map(self):
        b = self.info.contact_bins
        m = np.zeros((b, b), np.uint32)
        ns = ctypes.c_uint32()
        self._check(lib().ccp_contact_map(self.h, _ptr(m, ctypes.c_uint32), b * b, ctypes.byref(ns)))
        return m, ns.value

    def time_kernels(self, n_steps):
        out = np.zeros(5, np.float32)
        self._check(lib().ccp_time_kernels(self.h, n_steps, _ptr(out, ctypes.c_float)))
        return out

    def step_grid(self):
        """The step's own cell list of the current positions: (cstart, soi)."""
        fc = self.info.fine_cps
        cs = np.zeros(fc ** 3 + 1, np.uint32)
        soi = np.zeros(self.N, np.uint32)
        self._check(lib().ccp_step_grid(self.h, _ptr(cs, ctypes.c_uint32), cs.size, _ptr(soi, ctypes.c_uint32),
                                        self.N))
        return cs, soi

    def capture_graphs(self):
        """Capture every step block run_steps replays, ahead of a timed run."""
        self._check(lib().ccp_capture_graphs(self.h))

    def observables(self, n_bins=64):
        """ccp's frame observables: ((dcm, rg2, nop, ncf), rcd[3, n_bins])."""
        sc = np.zeros(4)
        rcd = np.zeros((3, n_bins))
        self._check(lib().ccp_observables(self.h, _ptr(sc, ctypes.c_double), _ptr(rcd, ctypes.c_double), n_bins))
        return sc, rcd

    def distance_contact(self, chromosome=0):
        """ccp's d(s) and 1000 P(s) of one chromosome (float32, s = 1..n)."""
        N = self.N
        if N == 30386 and chromosome < 6:
            n = (_CHR_LIM[chromosome + 1] - _CHR_LIM[chromosome]) // 2
        else:
            n = N // 2
        sd, cp = np.zeros(n, np.float32), np.zeros(n, np.float32)
        self._check(lib().ccp_distance_contact(self.h, chromosome, _ptr(sd, ctypes.c_float),
                                               _ptr(cp, ctypes.c_float), n))
        return sd, cp

    def frame_contact_map(self):
        cms = (self.N // 2) // 4 if self.N == 30386 else self.N // 4
        cm = np.zeros(cms * (cms + 1) // 2, np.float32)
        self._check(lib().ccp_frame_contact_map(self.h, _ptr(cm, ctypes.c_float), cm.size))
        return cm


class Analysis:
    """ccp's chrana on the frames of a Simulation: per-frame observables,
    d(s)/P(s) and the packed contact map, written in ccp's analysis files."""

    def __init__(self, sim):
        self.sim = sim
        h = ctypes.c_void_p()
        sim._check(lib().ccp_analysis_create(sim.h, ctypes.byref(h)))
        self.h = h.value

    def close(self):
        if self.h:
            lib().ccp_analysis_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise CcpError(rc, lib().ccp_analysis_last_error(self.h).decode())

    def add_frame(self):
        self._check(lib().ccp_analysis_add_frame(self.h))

    def add_trajectory_file(self, path):
        self._check(lib().ccp_analysis_add_trajectory_file(self.h, os.fsencode(path)))

    def end_simulation(self, path):
        self._check(lib().ccp_analysis_end_simulation(self.h, os.fsencode(path)))

    def finish(self, fin_path, cm_path):
        self._check(lib().ccp_analysis_finish(self.h, os.fsencode(fin_path), os.fsencode(cm_path)))


def series_stats(v):
    """ccp's time-series statistics: (avg, var, sem, f_n_b, i_t, thermalised)."""
    v = np.ascontiguousarray(v, np.float32)
    out = np.zeros(6)
    rc = lib().ccp_series_stats(_ptr(v, ctypes.c_float), v.size, _ptr(out, ctypes.c_double))
    if rc:
        raise CcpError(rc, "bad series")
    return out


def weighted_stats(avg, sem):
    """Inverse-variance weighted combination: (avg, var, sem)."""
    avg = np.ascontiguousarray(avg, np.float64)
    sem = np.ascontiguousarray(sem, np.float64)
    out = np.zeros(3)
    rc = lib().ccp_weighted_stats(_ptr(avg, ctypes.c_double), _ptr(sem, ctypes.c_double), avg.size,
                                  _ptr(out, ctypes.c_double))
    if rc:
        raise CcpError(rc, "bad input")
    return out


def _batch_systems(systems):
    arr = (BatchSystem * len(systems))()
    for a, d in zip(arr, systems):
        a.n_beads = int(d["n_beads"])
        a.n_chains = int(d.get("n_chains", 1))
        a.volume_fraction = float(d["volume_fraction"])
        a.eps_het = float(d.get("eps_het", 0.5))
        a.eps_lam = float(d.get("eps_lam", 1.0))
        a.het_fraction = float(d.get("het_fraction", 0.3))
        a.lad_fraction = float(d.get("lad_fraction", 0.1))
        a.seed = int(d["seed"])
        a.replica = int(d.get("replica", 0))
    return arr


def batch_layout(systems, common=None):
    """The layout a Batch of these systems would have, computed on the host
    (no GPU): dict with beads, cells, bucket_bytes and cell_offsets (first
    cell of each system, then the total). Raises CcpError as Batch would."""
    arr = _batch_systems(systems)
    nb = ctypes.c_uint32()
    nc, bb = ctypes.c_uint64(), ctypes.c_uint64()
    off = np.zeros(len(systems) + 1, np.uint64)
    rc = lib().ccp_batch_layout(params_text(common or {}).encode(), arr, len(systems), ctypes.byref(nb),
                                ctypes.byref(nc), ctypes.byref(bb), _ptr(off, ctypes.c_uint64))
    if rc != 0:
        raise CcpError(rc, lib().ccp_batch_last_error(None).decode())
    return {"beads": nb.value, "cells": nc.value, "bucket_bytes": bb.value, "cell_offsets": off}


class Batch:
    """Independent systems of the WCA/LJ set stepped together on one GPU.

    systems: list of dicts with n_beads, volume_fraction, eps_het, eps_lam,
    het_fraction, lad_fraction, seed (and optionally n_chains)."""

    def __init__(self, systems, common=None):
        self.h = None
        arr = _batch_systems(systems)
        h = lib().ccp_batch_create(params_text(common or {}).encode(), arr, len(systems))
        if not h:
            raise CcpError(1, lib().ccp_batch_last_error(None).decode())
        self.h = h
        ns, nb, nc = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        off = np.zeros(len(systems) + 1, np.uint32)
        lib().ccp_batch_info(h, ctypes.byref(ns), ctypes.byref(nb), ctypes.byref(nc), _ptr(off, ctypes.c_uint32))
        self.n_sys, self.N, self.n_cells = ns.value, nb.value, nc.value
        self.offsets = off
        common = common or {}
        self.bins = int(common.get("contact_bins", 128))
        self.rd_bins = int(common.get("radial_bins", 32))

    def close(self):
        if self.h:
            lib().ccp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise CcpError(rc, lib().ccp_batch_last_error(self.h).decode())

    def system_params(self, s):
        buf = ctypes.create_string_buffer(4096)
        self._check(lib().ccp_batch_system_params(self.h, s, buf, 4096))
        return buf.value.decode()

    def generate_initial_conditions(self):
        self._check(lib().ccp_batch_generate_initial_conditions(self.h))

    def run_steps(self, n):
        self._check(lib().ccp_batch_run_steps(self.h, n))

    def synchronize(self):
        self._check(lib().ccp_batch_synchronize(self.h))

    def state(self):
        xyz = np.zeros((self.N, 3), np.float32)
        fl = np.zeros(self.N, np.uint8)
        self._check(lib().ccp_batch_get_state(self.h, _ptr(xyz, ctypes.c_float), _ptr(fl, ctypes.c_uint8), self.N))
        return xyz, fl

    def set_state(self, xyz, flags):
        xyz = np.ascontiguousarray(xyz, np.float32)
        flags = np.ascontiguousarray(flags, np.uint8)
        self._check(lib().ccp_batch_set_state(self.h, _ptr(xyz, ctypes.c_float), _ptr(flags, ctypes.c_uint8),
                                              self.N))

    @property
    def step(self):
        v = ctypes.c_uint64()
        self._check(lib().ccp_batch_get_step(self.h, ctypes.byref(v)))
        return v.value

    @step.setter
    def step(self, value):
        self._check(lib().ccp_batch_set_step(self.h, int(value)))

    def contact_maps(self):
        m = np.zeros((self.n_sys, self.bins, self.bins), np.uint32)
        ns = ctypes.c_uint32()
        self._check(lib().ccp_batch_contact_maps(self.h, _ptr(m, ctypes.c_uint32), m.size, ctypes.byref(ns)))
        return m, ns.value

    def radial_density(self):
        rd = np.zeros((self.n_sys, self.rd_bins))
        self._check(lib().ccp_batch_radial_density(self.h, _ptr(rd, ctypes.c_double), rd.size))
        return rd

    def reset_samples(self):
        self._check(lib().ccp_batch_reset_samples(self.h))

    def time_steps(self, n):
        ms = ctypes.c_float()
        self._check(lib().ccp_batch_time_steps(self.h, n, ctypes.byref(ms)))
        return ms.value


def sequence_pattern(params, n):
    """Bead flags the initial-condition generator assigns for these parameters (host only)."""
    fl = np.zeros(n, np.uint8)
    rc = lib().ccp_sequence_pattern(params_text(params).encode(), _ptr(fl, ctypes.c_uint8), n)
    if rc:
        raise CcpError(rc, lib().ccp_last_error(None).decode())
    return fl
