"""Python mirror of the reference's TreePM gravity entry points over the C ABI (include/lagpm.h).

Names follow the reference interface they replace:
  pmmain          Cosmos/pmmain.c:390        Box.pmmain(evolfact, pmstatus)
  TSC             Cosmos/pmmain.c:112,188    Box.tsc()
  psolver         Cosmos/psolver.c:13        Box.psolver()
  fda4            Cosmos/fda4.c:31           Box.fda4(evolfact, pmstatus)
  treecorrection  Cosmos/treecorrection.c:225 Box.treecorrection(evolfact)
  i_force_spline  OST/Model4TreeForce.c:48   i_force_spline(params)
  GetEvolFactor   Cosmos/CosmosEvolFactor.c:230 get_evol_factor(params, anow, astep)

Errors: the reference exits the process on error; here every failing call raises LagpmError
carrying the library's status code and message. The shared library is loaded from this
package directory only; a missing library raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

PUSH, PULL, KICK = 0, 1, 2  # eunha.h:69-77
LAW_PMFIT, LAW_ERFC, LAW_TREEFORCE = 0, 1, 2
PREC_FP32, PREC_FP64 = 0, 1
GROUPS_TREE, GROUPS_RUNS, GROUPS_PACKED = 0, 1, 2
NSPLINE = 4096

_STATUS = {
    0: "OK", -1: "LAGPM_E_INVAL", -2: "LAGPM_E_NOMEM", -3: "LAGPM_E_HIP", -4: "LAGPM_E_FFT",
    -5: "LAGPM_E_STATE", -6: "LAGPM_E_OVERFLOW", -7: "LAGPM_E_IO",
}

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liblagpm.so")


class LagpmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32), ("law", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("deconvolve", ctypes.c_int32), ("mass", ctypes.c_double), ("efold", ctypes.c_double),
        ("epsilon", ctypes.c_double), ("rsphere", ctypes.c_double), ("theta", ctypes.c_double),
        ("amax", ctypes.c_double), ("omep", ctypes.c_double), ("omeplam", ctypes.c_double),
    ]


class EvolFact(ctypes.Structure):
    """Same fields as EvolFact, Cosmos/CosmosEvolFactor.h:1-6."""
    _fields_ = [(k, ctypes.c_float) for k in (
        "afact", "bfact", "pfact", "fact1", "fact2", "fact1_push", "fact2_push", "fact1_pull", "fact2_pull")]


class TreeStats(ctypes.Structure):
    _fields_ = [("n_pp", ctypes.c_uint64), ("n_pc", ctypes.c_uint64), ("n_nodes", ctypes.c_uint64),
                ("n_groups", ctypes.c_uint64), ("ms_build", ctypes.c_double), ("ms_walk", ctypes.c_double),
                ("deepest", ctypes.c_int32), ("depth_limit", ctypes.c_int32), ("n_pp_within", ctypes.c_uint64)]


class SnapshotInfo(ctypes.Structure):
    """The ASCII header of a GOTPM particle file (lagpm_snapshot_info)."""
    _fields_ = [("nx", ctypes.c_long), ("ny", ctypes.c_long), ("nz", ctypes.c_long),
                ("np_sum", ctypes.c_long), ("np_dm", ctypes.c_long), ("np_sph", ctypes.c_long),
                ("np_star", ctypes.c_long), ("np_agn", ctypes.c_long),
                ("ptype", ctypes.c_int), ("nspace", ctypes.c_int), ("stepcount", ctypes.c_int),
                ("xyzshift", ctypes.c_int),
                ("amax", ctypes.c_double), ("anow", ctypes.c_double), ("astep", ctypes.c_double),
                ("omep", ctypes.c_double), ("omeplam", ctypes.c_double), ("hubble", ctypes.c_double),
                ("boxsize", ctypes.c_double), ("theta", ctypes.c_double), ("epsilon", ctypes.c_double),
                ("rsphere", ctypes.c_double),
                ("header_bytes", ctypes.c_size_t), ("record_bytes", ctypes.c_size_t), ("file_bytes", ctypes.c_size_t)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# dmparticletype of a GOTPM + XYZDBL build (eunha.h:389-399): 32 bytes
DM_RECORD = np.dtype([("u4if", "<u8"), ("x", "<f4"), ("y", "<f4"), ("z", "<f4"),
                      ("vx", "<f4"), ("vy", "<f4"), ("vz", "<f4")])


class Timings(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in (
        "ms_sort", "ms_tsc", "ms_fft_r2c", "ms_green", "ms_fft_c2r", "ms_grad", "ms_interp",
        "ms_tree_sort", "ms_build", "ms_walk")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# every symbol of include/lagpm.h with its ctypes signature
_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
SIGNATURES = {
    "lagpm_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int]),
    "lagpm_destroy": (ctypes.c_int, [_P]),
    "lagpm_last_error": (ctypes.c_char_p, [_P]),
    "lagpm_abi_version": (ctypes.c_int, []),
    "lagpm_stream": (_P, [_P]),
    "lagpm_synchronize": (ctypes.c_int, [_P]),
    "lagpm_box_create": (ctypes.c_int, [_P, ctypes.POINTER(Params), ctypes.c_size_t, ctypes.POINTER(_P)]),
    "lagpm_box_destroy": (ctypes.c_int, [_P, _P]),
    "lagpm_box_params": (ctypes.c_int, [_P, ctypes.POINTER(Params)]),
    "lagpm_box_np": (ctypes.c_size_t, [_P]),
    "lagpm_box_set_cosmology": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "lagpm_box_device_bytes": (ctypes.c_size_t, [_P]),
    "lagpm_estimate_box_bytes": (ctypes.c_size_t, [ctypes.POINTER(Params), ctypes.c_size_t]),
    "lagpm_box_upload": (ctypes.c_int, [_P, _P, _D, _D, _D, _D, _D, _D]),
    "lagpm_box_download": (ctypes.c_int, [_P, _P, _D, _D, _D, _D, _D, _D]),
    "lagpm_box_upload_eunha": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t, ctypes.c_long]),
    "lagpm_box_download_eunha_vel": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    "lagpm_box_download_eunha": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    "lagpm_prefof": (ctypes.c_int, [_P, _P, ctypes.c_float, ctypes.c_int, ctypes.POINTER(ctypes.c_uint8),
                                    ctypes.POINTER(ctypes.c_size_t)]),
    "lagpm_snapshot_read_header": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.POINTER(SnapshotInfo)]),
    "lagpm_snapshot_read_dm": (ctypes.c_int, [_P, ctypes.c_char_p, _P, ctypes.c_size_t]),
    "lagpm_snapshot_write": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_char_p, _P, ctypes.c_size_t,
                                            ctypes.c_double, ctypes.c_int]),
    "lagpm_eunha_decode": (ctypes.c_int, [_P, ctypes.c_size_t, ctypes.c_long, _D, _D, _D, _D, _D, _D]),
    "lagpm_eunha_encode": (ctypes.c_int, [_P, ctypes.c_size_t, ctypes.c_long, _D, _D, _D, _D, _D, _D]),
    "lagpm_pmmain": (ctypes.c_int, [_P, _P, ctypes.POINTER(EvolFact), ctypes.c_int]),
    "lagpm_tsc": (ctypes.c_int, [_P, _P, _F]),
    "lagpm_tsc_fixed": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)]),
    "lagpm_psolver": (ctypes.c_int, [_P, _P, _F]),
    "lagpm_green_apply": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, ctypes.c_int, _F]),
    "lagpm_fda4": (ctypes.c_int, [_P, _P, ctypes.POINTER(EvolFact), ctypes.c_int]),
    "lagpm_set_potential": (ctypes.c_int, [_P, _P, _F]),
    "lagpm_get_gradient": (ctypes.c_int, [_P, _P, _F]),
    "lagpm_force_spline": (ctypes.c_int, [ctypes.POINTER(Params), _D, _D]),
    "lagpm_power_spectrum": (ctypes.c_int, [_P, _P, ctypes.c_double, _D, _D, _D, _D, _D]),
    "lagpm_power_spectrum_noalloc": (ctypes.c_int, [_P, _P, ctypes.c_double, _D, _D, _D, _D, _D]),
    "lagpm_treecorrection": (ctypes.c_int, [_P, _P, ctypes.POINTER(EvolFact)]),
    "lagpm_tree_stats_get": (ctypes.c_int, [_P, ctypes.POINTER(TreeStats)]),
    "lagpm_tree_accel_sample": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_int64), ctypes.c_size_t, _D]),
    "lagpm_tree_nodes_get": (ctypes.c_int, [_P, _P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t),
                                            ctypes.POINTER(ctypes.c_int32), _D, _D,
                                            ctypes.POINTER(ctypes.c_uint32)]),
    "lagpm_box_set_tuning": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int]),
    "lagpm_box_get_tuning": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "lagpm_box_set_group_mode": (ctypes.c_int, [_P, ctypes.c_int]),
    "lagpm_box_get_group_mode": (ctypes.c_int, [_P]),
    "lagpm_box_autotune": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "lagpm_box_timings": (ctypes.c_int, [_P, _P, ctypes.POINTER(Timings)]),
    "lagpm_evolfactor": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.c_double, ctypes.c_double,
                                        ctypes.POINTER(EvolFact)]),
    "lagpm_hubble_drag": (ctypes.c_int, [_P, _P, ctypes.c_double]),
    "lagpm_drift": (ctypes.c_int, [_P, _P, ctypes.c_double]),
    "lagpm_kdk_step": (ctypes.c_int, [_P, _P, ctypes.c_double, ctypes.c_double]),
    "lagpm_batch_step": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.c_int, _D, _D]),
    "lagpm_set_step_graphs": (ctypes.c_int, [_P, ctypes.c_int]),
    "lagpm_set_walk_census": (ctypes.c_int, [_P, ctypes.c_int]),
    "lagpm_set_batch_pipeline": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load liblagpm.so from the package directory; raise when it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LagpmError(-5, f"{path} is missing: build it with __graft_entry__.build() "
                             "(make -C lageunha_amd/csrc)")
    # rocFFT compiles some kernels at run time and caches them in a file; its default lives under
    # $HOME, which service accounts may not be able to write. Choose a per-user file in the
    # temporary directory unless the caller chose one. Done here, before the library (and with it
    # the HIP runtime's threads) is loaded: the library itself never writes the environment.
    tmp = os.environ.get("TMPDIR") or "/tmp"
    os.environ.setdefault("ROCFFT_RTC_CACHE_PATH", os.path.join(tmp, f"lagpm_rocfft_{os.getuid()}.db"))
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(_D) if a is not None else None


def _fp(a):
    return a.ctypes.data_as(_F) if a is not None else None


def make_params(n, *, law=LAW_ERFC, precision=PREC_FP32, deconvolve=1, mass=1.0, efold=2.5,
                epsilon=0.1, rsphere=5.625, theta=0.5, amax=48.0, omep=0.26, omeplam=0.74) -> Params:
    return Params(n, law, precision, deconvolve, mass, efold, epsilon, rsphere, theta, amax, omep, omeplam)


def reference_params(n, *, precision=PREC_FP32, theta=0.3, mass=1.0) -> Params:
    """The reference defaults: efold 0.9, eps 0.1, RSPHERE 4, pmcf law, no deconvolution
    (Params/header.c:501-505,651-653). efold/eps/theta/rsphere are floats there."""
    f = lambda v: float(np.float32(v))  # noqa: E731
    return Params(n, LAW_PMFIT, precision, 0, mass, f(0.9), f(0.1), f(4.0), f(theta), 48.0, 0.26, 0.74)


def evolfact(fact2=1.0, fact1=1.0, **kw) -> EvolFact:
    ef = EvolFact()
    ef.fact1, ef.fact2 = fact1, fact2
    ef.fact1_push = kw.get("fact1_push", 1.0)
    ef.fact2_push = kw.get("fact2_push", fact2)
    ef.fact1_pull = kw.get("fact1_pull", 1.0)
    ef.fact2_pull = kw.get("fact2_pull", fact2)
    ef.pfact = kw.get("pfact", 0.0)
    return ef


class Context:
    def __init__(self, device: int = 0):
        self.lib = load()
        h = _P()
        rc = self.lib.lagpm_create(ctypes.byref(h), device)
        if rc != 0:
            raise LagpmError(rc, f"lagpm_create(device={device}) failed")
        self.h = h

    def check(self, rc):
        if rc != 0:
            raise LagpmError(rc, self.lib.lagpm_last_error(self.h).decode())

    def set_step_graphs(self, enable: bool):
        self.check(self.lib.lagpm_set_step_graphs(self.h, 1 if enable else 0))

    def set_walk_census(self, enable: bool):
        """FP32 walks also count the evaluated pairs within the cutoff (TreeStats.n_pp_within)."""
        self.check(self.lib.lagpm_set_walk_census(self.h, 1 if enable else 0))

    def set_batch_pipeline(self, stages: int, reserve_cu: int = 32):
        self.check(self.lib.lagpm_set_batch_pipeline(self.h, stages, reserve_cu))

    def synchronize(self):
        self.check(self.lib.lagpm_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.lagpm_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Box:
    """One periodic box with resident particle, mesh and tree state on the device."""

    def __init__(self, ctx: Context, params: Params, np_: int):
        self.ctx, self.lib = ctx, ctx.lib
        self.params = params
        self.np = int(np_)
        self.n = params.n
        h = _P()
        ctx.check(self.lib.lagpm_box_create(ctx.h, ctypes.byref(params), self.np, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.lagpm_box_destroy(self.ctx.h, self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def device_bytes(self):
        return self.lib.lagpm_box_device_bytes(self.h)

    def set_cosmology(self, amax: float, omep: float, omeplam: float):
        """Another cosmology for the following steps (the box's memory and state are kept)."""
        rc = self.lib.lagpm_box_set_cosmology(self.h, amax, omep, omeplam)
        if rc != 0:
            raise LagpmError(rc, "bad cosmology")
        self.params.amax, self.params.omep, self.params.omeplam = amax, omep, omeplam

    def upload(self, x, y, z, vx=None, vy=None, vz=None):
        a = [np.ascontiguousarray(v, dtype=np.float64) if v is not None else None for v in (x, y, z, vx, vy, vz)]
        for v in a:
            if v is not None and v.shape != (self.np,):
                raise ValueError(f"expected arrays of shape ({self.np},)")
        self.ctx.check(self.lib.lagpm_box_upload(self.ctx.h, self.h, *[_dp(v) for v in a]))

    def upload_pos(self, pos, vel=None):
        pos = np.asarray(pos, dtype=np.float64)
        if vel is None:
            self.upload(pos[:, 0], pos[:, 1], pos[:, 2])
        else:
            vel = np.asarray(vel, dtype=np.float64)
            self.upload(pos[:, 0], pos[:, 1], pos[:, 2], vel[:, 0], vel[:, 1], vel[:, 2])

    def download(self):
        out = [np.empty(self.np) for _ in range(6)]
        self.ctx.check(self.lib.lagpm_box_download(self.ctx.h, self.h, *[_dp(v) for v in out]))
        return np.stack(out[:3], axis=1), np.stack(out[3:], axis=1)

    def velocities(self):
        return self.download()[1]

    def upload_eunha(self, records: np.ndarray):
        records = np.ascontiguousarray(records)
        self.ctx.check(self.lib.lagpm_box_upload_eunha(self.ctx.h, self.h, records.ctypes.data_as(_P),
                                                       records.shape[0], self.n))

    def download_eunha_vel(self, records: np.ndarray):
        self.ctx.check(self.lib.lagpm_box_download_eunha_vel(self.ctx.h, self.h, records.ctypes.data_as(_P),
                                                             records.shape[0]))

    def download_eunha(self, records: np.ndarray):
        """Positions (as displacements from each record's lattice site) and velocities, in place."""
        if records.dtype != DM_RECORD or not records.flags.c_contiguous:
            raise ValueError("records must be a contiguous DM_RECORD array")
        self.ctx.check(self.lib.lagpm_box_download_eunha(self.ctx.h, self.h, records.ctypes.data_as(_P),
                                                         records.shape[0]))

    def prefof(self, fof_link: float = 0.2, minnum: int = 20):
        """PreFoF halo-candidate flags (uint8 per particle, caller order) and their count."""
        flag = np.zeros(self.np, dtype=np.uint8)
        cnt = ctypes.c_size_t()
        self.ctx.check(self.lib.lagpm_prefof(self.ctx.h, self.h, fof_link, minnum,
                                             flag.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(cnt)))
        return flag, cnt.value

    # ---- the drop-in entry points ----
    def pmmain(self, ef: EvolFact, pmstatus: int = KICK):
        self.ctx.check(self.lib.lagpm_pmmain(self.ctx.h, self.h, ctypes.byref(ef), pmstatus))

    def tsc(self):
        den = np.empty(self.n ** 3, dtype=np.float32)
        self.ctx.check(self.lib.lagpm_tsc(self.ctx.h, self.h, _fp(den)))
        return den.reshape(self.n, self.n, self.n)  # [z][y][x]

    def tsc_fixed(self):
        acc = np.empty(self.n ** 3, dtype=np.int64)
        s = ctypes.c_int()
        self.ctx.check(self.lib.lagpm_tsc_fixed(self.ctx.h, self.h, acc.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                ctypes.byref(s)))
        return acc, s.value

    def psolver(self):
        phi = np.empty(self.n ** 3, dtype=np.float32)
        self.ctx.check(self.lib.lagpm_psolver(self.ctx.h, self.h, _fp(phi)))
        return phi

    def fda4(self, ef: EvolFact, pmstatus: int = KICK):
        self.ctx.check(self.lib.lagpm_fda4(self.ctx.h, self.h, ctypes.byref(ef), pmstatus))

    def set_potential(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float32).ravel()
        self.ctx.check(self.lib.lagpm_set_potential(self.ctx.h, self.h, _fp(phi)))

    def get_gradient(self):
        g = np.empty(4 * self.n ** 3, dtype=np.float32)
        self.ctx.check(self.lib.lagpm_get_gradient(self.ctx.h, self.h, _fp(g)))
        return g.reshape(-1, 4)

    def treecorrection(self, ef: EvolFact):
        self.ctx.check(self.lib.lagpm_treecorrection(self.ctx.h, self.h, ctypes.byref(ef)))

    def tree_accel_sample(self, ids) -> np.ndarray:
        """Short-range acceleration (per unit kick factor) of the particles with caller indices
        `ids`, each walked as a group of one; shape (len(ids), 3). Velocities are unchanged."""
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        acc = np.empty(3 * len(ids))
        self.ctx.check(self.lib.lagpm_tree_accel_sample(self.ctx.h, self.h, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                        len(ids), _dp(acc)))
        return acc.reshape(-1, 3)

    def tree_stats(self) -> TreeStats:
        s = TreeStats()
        self.lib.lagpm_tree_stats_get(self.h, ctypes.byref(s))
        return s

    def tree_nodes(self):
        """The octree of the last tree build: dict of info (nodes, 4: first, count, children, level),
        com (nodes, 4: x, y, z, mass), quad (nodes, 8: Qxx Qyy Qzz Qxy Qxz Qyz trQ rmax) and order
        (caller index of each particle of the box's current order)."""
        nn = ctypes.c_size_t()
        rc = self.lib.lagpm_tree_nodes_get(self.ctx.h, self.h, 0, ctypes.byref(nn), None, None, None, None)
        if rc not in (0, -1) or (rc == -1 and nn.value == 0):
            self.ctx.check(rc)
        k = nn.value
        info = np.empty((k, 4), dtype=np.int32)
        com = np.empty((k, 4))
        quad = np.empty((k, 8))
        order = np.empty(self.np, dtype=np.uint32)
        self.ctx.check(self.lib.lagpm_tree_nodes_get(
            self.ctx.h, self.h, k, ctypes.byref(nn), info.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _dp(com),
            _dp(quad), order.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        return {"info": info, "com": com, "quad": quad, "order": order}

    def timings(self) -> Timings:
        t = Timings()
        self.ctx.check(self.lib.lagpm_box_timings(self.ctx.h, self.h, ctypes.byref(t)))
        return t

    def hubble_drag(self, fact1: float):
        self.ctx.check(self.lib.lagpm_hubble_drag(self.ctx.h, self.h, fact1))

    def drift(self, pfact: float):
        self.ctx.check(self.lib.lagpm_drift(self.ctx.h, self.h, pfact))

    def set_tuning(self, nleaf: int, group: int, mode: int = None):
        """Leaf size, group size and (when given) group mode (GROUPS_TREE, GROUPS_RUNS or GROUPS_PACKED)."""
        rc = self.lib.lagpm_box_set_tuning(self.h, nleaf, group)
        if rc != 0:
            raise LagpmError(rc, f"bad tuning nleaf={nleaf} group={group}")
        if mode is not None:
            rc = self.lib.lagpm_box_set_group_mode(self.h, mode)
            if rc != 0:
                raise LagpmError(rc, f"bad group mode {mode}")

    def tuning(self):
        """(nleaf, group, group mode)."""
        a, b = ctypes.c_int(), ctypes.c_int()
        self.lib.lagpm_box_get_tuning(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value, self.lib.lagpm_box_get_group_mode(self.h)

    def autotune(self):
        """Time tree build + walk for a set of (nleaf, group, group mode), keep and return the fastest."""
        a, b = ctypes.c_int(), ctypes.c_int()
        self.ctx.check(self.lib.lagpm_box_autotune(self.ctx.h, self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value, self.lib.lagpm_box_get_group_mode(self.h)

    def power_spectrum(self, boxsize: float):
        """CorrMeasure of the current particles: (pk, npk, cor, sigma8, slope), n+1 bins."""
        pk, npk, cor = np.empty(self.n + 1), np.empty(self.n + 1), np.empty(self.n + 1)
        s8, sl = ctypes.c_double(), ctypes.c_double()
        self.ctx.check(self.lib.lagpm_power_spectrum(self.ctx.h, self.h, boxsize, _dp(pk), _dp(npk), _dp(cor),
                                                     ctypes.byref(s8), ctypes.byref(sl)))
        return pk, npk, cor, s8.value, sl.value

    def power_spectrum_noalloc(self, boxsize: float):
        """power_spectrum with the bins in the device's static scratch (no device allocation); the
        box's FFT plan must exist (one PM kick or power_spectrum on a box of this size first)."""
        pk, npk, cor = np.empty(self.n + 1), np.empty(self.n + 1), np.empty(self.n + 1)
        s8, sl = ctypes.c_double(), ctypes.c_double()
        self.ctx.check(self.lib.lagpm_power_spectrum_noalloc(self.ctx.h, self.h, boxsize, _dp(pk), _dp(npk), _dp(cor),
                                                             ctypes.byref(s8), ctypes.byref(sl)))
        return pk, npk, cor, s8.value, sl.value

    def kdk_step(self, anow: float, astep: float):
        """One GOTPM global step (drag, PM kick, tree kick, drift) at scale factor anow."""
        self.ctx.check(self.lib.lagpm_kdk_step(self.ctx.h, self.h, anow, astep))


def batch_step(ctx: Context, boxes, anow, astep):
    """One GOTPM global step of every box in `boxes` (anow/astep: scalars or one per box)."""
    nb = len(boxes)
    a = np.broadcast_to(np.asarray(anow, dtype=np.float64), (nb,)).copy()
    s = np.broadcast_to(np.asarray(astep, dtype=np.float64), (nb,)).copy()
    hs = (_P * max(nb, 1))(*[b.h for b in boxes])
    ctx.check(ctx.lib.lagpm_batch_step(ctx.h, hs, nb, _dp(a), _dp(s)))


def green_apply(ctx: Context, n: int, efold: float, deconvolve: int, spec: np.ndarray) -> np.ndarray:
    """Green multiply on a host spectrum (complex64, shape (n, n, n//2+1))."""
    s = np.ascontiguousarray(spec, dtype=np.complex64).copy()
    ctx.check(ctx.lib.lagpm_green_apply(ctx.h, n, efold, deconvolve, s.view(np.float32).ctypes.data_as(_F)))
    return s


def i_force_spline(params: Params):
    """FP64 tables diff[4096][3], slope[4096][3] (OST/Model4TreeForce.c:48-91)."""
    lib = load()
    diff = np.empty((NSPLINE, 3))
    slope = np.empty((NSPLINE, 3))
    rc = lib.lagpm_force_spline(ctypes.byref(params), _dp(diff), _dp(slope))
    if rc != 0:
        raise LagpmError(rc, "lagpm_force_spline failed")
    return diff, slope


def get_evol_factor(params: Params, anow: float, astep: float) -> EvolFact:
    lib = load()
    ef = EvolFact()
    rc = lib.lagpm_evolfactor(ctypes.byref(params), anow, astep, ctypes.byref(ef))
    if rc != 0:
        raise LagpmError(rc, "lagpm_evolfactor failed")
    return ef


def estimate_box_bytes(params: Params, np_: int) -> int:
    return int(load().lagpm_estimate_box_bytes(ctypes.byref(params), int(np_)))


@dataclass
class ForceResult:
    pm: np.ndarray
    tree: np.ndarray

    @property
    def total(self):
        return self.pm + self.tree


def treepm_accel(ctx: Context, params: Params, pos: np.ndarray) -> ForceResult:
    """Convenience: PM and tree accelerations (fact2 = 1, zero velocities) in caller order."""
    with Box(ctx, params, pos.shape[0]) as b:
        b.upload_pos(pos)
        b.pmmain(evolfact(1.0), KICK)
        v_pm = b.velocities()
        b.upload_pos(pos)
        b.treecorrection(evolfact(1.0))
        v_tree = b.velocities()
    return ForceResult(v_pm, v_tree)


# ---- GOTPM particle files (host only) ----
def _file_check(rc: int, what: str):
    if rc != 0:
        raise LagpmError(rc, what)


def read_snapshot_header(path: str) -> SnapshotInfo:
    """The header of a jread-format file (Params/header.c:72-105)."""
    info = SnapshotInfo()
    _file_check(load().lagpm_snapshot_read_header(None, os.fsencode(path), ctypes.byref(info)),
                f"cannot read the header of {path}")
    return info


def read_snapshot(path: str):
    """(header, DM records as a DM_RECORD array) of a jread-format file (kjhrw.c:51-120)."""
    info = read_snapshot_header(path)
    rec = np.zeros(info.np_dm, dtype=DM_RECORD)
    _file_check(load().lagpm_snapshot_read_dm(None, os.fsencode(path), rec.ctypes.data_as(_P), rec.shape[0]),
                f"cannot read the particles of {path}")
    return info, rec


def write_snapshot(template: str, path: str, records: np.ndarray, anow: float, stepcount: int):
    """jwrite (kjhrw.c:155-220): template's header with new Anow/Stepcount, records, template's tail."""
    records = np.ascontiguousarray(records, dtype=DM_RECORD)
    _file_check(load().lagpm_snapshot_write(None, os.fsencode(template), os.fsencode(path),
                                            records.ctypes.data_as(_P), records.shape[0], anow, stepcount),
                f"cannot write {path} from {template}")


def eunha_decode(records: np.ndarray, nx: int):
    """XofP/YofP/ZofP and velocities of DM records: (pos (N,3), vel (N,3)) in FP64."""
    records = np.ascontiguousarray(records, dtype=DM_RECORD)
    q = [np.empty(records.shape[0]) for _ in range(6)]
    _file_check(load().lagpm_eunha_decode(records.ctypes.data_as(_P), records.shape[0], nx, *[_dp(a) for a in q]),
                "bad records")
    return np.stack(q[:3], axis=1), np.stack(q[3:], axis=1)


def eunha_encode(records: np.ndarray, nx: int, pos: np.ndarray, vel: np.ndarray):
    """Write positions (as displacements from each record's lattice site) and velocities into records."""
    if records.dtype != DM_RECORD or not records.flags.c_contiguous:
        raise ValueError("records must be a contiguous DM_RECORD array")
    q = [np.ascontiguousarray(pos[:, a], dtype=np.float64) for a in range(3)]
    q += [np.ascontiguousarray(vel[:, a], dtype=np.float64) for a in range(3)]
    _file_check(load().lagpm_eunha_encode(records.ctypes.data_as(_P), records.shape[0], nx, *[_dp(a) for a in q]),
                "bad records")
