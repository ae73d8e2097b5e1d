"""Host-side mirror of DIABLO's channel driver over libdiablo_hip.so.

:class:`Channel` plays the part of the reference's Fortran driver around
``RK_CHAN_1`` (``diablo.f:59-137``): it holds the parameters of
``input.dat``/``input_chan.dat`` and the y-grid, owns one device context,
moves the COMMON-block state in and out in the reference's Fortran layout,
and advances substeps or whole time steps on the GPU.

Host arrays follow ``header:212-258`` with the index order reversed for numpy
(C order): ``cu1[j, k, i]`` is ``CU1(I,K,J)`` with shape
``(NY+2, NZ+2, NXP+1)`` complex128; physical scalars ``th[n, j, k, i]`` have
shape ``(N_TH, NY+2, NZ+2, NX+2)`` float64.
"""
import ctypes
import os

import numpy as np

from . import configs

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdiablo_hip.so")
MAX_TH = 4

DH_OK, DH_EARG, DH_EHIP, DH_ESTATE, DH_ELAYOUT = 0, -1, -2, -3, -4


class DhParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32), ("n_th", ctypes.c_int32),
        ("nu", ctypes.c_double), ("beta", ctypes.c_double), ("lx", ctypes.c_double),
        ("lz", ctypes.c_double), ("nu_v_scale", ctypes.c_double),
        ("delta_t", ctypes.c_double), ("cfl", ctypes.c_double),
        ("variable_dt", ctypes.c_int32), ("update_dt", ctypes.c_int32),
        ("f_type", ctypes.c_int32),
        ("px0", ctypes.c_double), ("omega0", ctypes.c_double), ("amp_omega0", ctypes.c_double),
        ("i_ro", ctypes.c_double), ("grav_x", ctypes.c_double), ("grav_y", ctypes.c_double),
        ("grav_z", ctypes.c_double),
        ("u_bc_ymin", ctypes.c_int32), ("v_bc_ymin", ctypes.c_int32), ("w_bc_ymin", ctypes.c_int32),
        ("u_bc_ymax", ctypes.c_int32), ("v_bc_ymax", ctypes.c_int32), ("w_bc_ymax", ctypes.c_int32),
        ("u_bc_ymin_c1", ctypes.c_double), ("v_bc_ymin_c1", ctypes.c_double),
        ("w_bc_ymin_c1", ctypes.c_double),
        ("u_bc_ymax_c1", ctypes.c_double), ("v_bc_ymax_c1", ctypes.c_double),
        ("w_bc_ymax_c1", ctypes.c_double),
        ("th_bc_ymin", ctypes.c_int32 * MAX_TH), ("th_bc_ymax", ctypes.c_int32 * MAX_TH),
        ("th_bc_ymin_c1", ctypes.c_double * MAX_TH), ("th_bc_ymax_c1", ctypes.c_double * MAX_TH),
        ("ri", ctypes.c_double * MAX_TH), ("pr", ctypes.c_double * MAX_TH),
        ("drhodx", ctypes.c_double * MAX_TH), ("drhodz", ctypes.c_double * MAX_TH),
        ("create_new_flow", ctypes.c_int32),
    ]


_DP = ctypes.POINTER(ctypes.c_double)


class DhHostState(ctypes.Structure):
    _fields_ = [(n, _DP) for n in ("cu1", "cu2", "cu3", "cp", "cf1", "cf2", "cf3", "cth", "cfth", "th")]


# Every entry point declared in include/diablo_hip.h: (restype, argtypes)
EXPORTS = {
    "dh_create": (ctypes.c_int, [ctypes.POINTER(DhParams), ctypes.POINTER(ctypes.c_void_p)]),
    "dh_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "dh_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "dh_set_grid": (ctypes.c_int, [ctypes.c_void_p, _DP, _DP, _DP, _DP]),
    "dh_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DhHostState)]),
    "dh_download": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DhHostState)]),
    "dh_set_clock": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int32]),
    "dh_get_clock": (ctypes.c_int, [ctypes.c_void_p, _DP, _DP, _DP, ctypes.POINTER(ctypes.c_int32)]),
    "dh_rk_substep": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "dh_time_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "dh_set_mixed_precision": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "dh_set_plane_threads": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "dh_plane_threads": (ctypes.c_int, [ctypes.c_void_p]),
    "dh_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "dh_ensemble_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "dh_mixed_precision": (ctypes.c_int, [ctypes.c_void_p]),
    "dh_set_les": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "dh_les": (ctypes.c_int, [ctypes.c_void_p]),
    "dh_set_les_closure": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]),
    "dh_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "dh_profile_substep": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _DP, ctypes.c_int32]),
    "dh_phase_name": (ctypes.c_char_p, [ctypes.c_int32]),
    "dh_rem_div": (ctypes.c_int, [ctypes.c_void_p]),
    "dh_poisson_p_chan": (ctypes.c_int, [ctypes.c_void_p]),
    "dh_h5_read_f64": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.POINTER(ctypes.c_int64), ctypes.c_char_p, ctypes.c_int32]),
    "dh_set_grid_h5": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "dh_ensemble_create": (ctypes.c_int, [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]),
    "dh_ensemble_time_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "dh_ensemble_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "dh_ensemble_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "dh_save_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]),
    "dh_create_flow_chan": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_uint64,
                                           ctypes.c_double, ctypes.c_double]),
    "dh_create_th_chan": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "dh_apply_bc_vel": (ctypes.c_int, [ctypes.c_void_p]),
    "dh_fft_xz_to_fourier_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]),
    "dh_fft_xz_to_physical_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]),
    "dh_field_dev": (ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_int32]),
    "dh_batched_last_error": (ctypes.c_char_p, []),
    "dh_les_features": (ctypes.c_int, [ctypes.c_void_p]),
    "dh_mlp_last_error": (ctypes.c_char_p, []),
    "dh_mlp_nparams": (ctypes.c_int32, []),
    "dh_mlp_packed_bytes": (ctypes.c_int64, []),
    "dh_mlp_work_bytes": (ctypes.c_int64, [ctypes.c_int64]),
    "dh_mlp_pack_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "dh_mlp_infer_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.c_int64, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]),
    "dh_mlp_grad_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p]),
    "dh_mlp_adam_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_float, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]),
    "dh_cn_solve_batched_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                               ctypes.c_void_p]),
    "dh_fft512_xz_to_fourier_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                                   ctypes.c_void_p]),
    "dh_fft512_xz_to_physical_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                                    ctypes.c_void_p]),
    "dh_fft512_xz_to_physical_noalloc_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                            ctypes.c_int32, ctypes.c_void_p]),
    "dh_stream": (ctypes.c_void_p, [ctypes.c_void_p]),
}

_lib = None


# profiles of dh_save_stats, in the order of the DH_ST_* enum
STATS = ("ume", "vme", "wme", "urms", "vrms", "wrms", "uv", "uw", "wv", "dudy", "dwdy", "shear")


def load_library(path=LIB_PATH):
    """Load libdiablo_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with `python -m diablo_amd.build` "
                               "(the channel step has no CPU fallback)")
        # torch carries HIP and HSA runtimes of its own.  In a process that
        # uses both, torch's must be loaded first: loaded after the library
        # had started the system runtime, torch found no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(path)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def make_params(d):
    p = DhParams()
    for k, v in d.items():
        if isinstance(v, (list, tuple)):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(p, k, v)
    return p


def grid_metrics(gy):
    """GY(1..NY+1) -> (GY, GYF, DY, DYF) on 0..NY+1 as ReadGridHDF5
    (``hdf5.f:690-705``) and CREATE_GRID_CHAN (``channel.f:1646-1652``)."""
    ny = len(gy) - 1
    GY = np.zeros(ny + 2)
    GYF = np.zeros(ny + 2)
    DY = np.zeros(ny + 2)
    DYF = np.zeros(ny + 2)
    GY[1:ny + 2] = gy
    for j in range(1, ny + 1):
        GYF[j] = 0.5 * (GY[j] + GY[j + 1])
    GYF[0] = 2.0 * GYF[1] - GYF[2]
    GYF[ny + 1] = 2.0 * GYF[ny] - GYF[ny - 1]
    for j in range(1, ny + 2):
        DY[j] = GYF[j] - GYF[j - 1]
    for j in range(1, ny + 1):
        DYF[j] = GY[j + 1] - GY[j]
    DYF[ny + 1] = DYF[ny]
    return GY, GYF, DY, DYF


def read_h5_f64(path, dataset):
    """A float64 dataset of an HDF5 grid or flow file through the library's
    reader (dh_h5_read_f64; ReadGridHDF5's subset, hdf5.f:548-710)."""
    lib = load_library()
    n = ctypes.c_int64(0)
    err = ctypes.create_string_buffer(256)
    rc = lib.dh_h5_read_f64(str(path).encode(), dataset.encode(), None, 0, ctypes.byref(n), err, 256)
    if rc != 0 and b"too small" not in err.value:
        raise DiabloError(f"dh_h5_read_f64({path}, {dataset}): {err.value.decode()} ({rc})")
    out = np.zeros(max(n.value, 1))
    rc = lib.dh_h5_read_f64(str(path).encode(), dataset.encode(), out.ctypes.data, out.size, ctypes.byref(n),
                            err, 256)
    if rc != 0:
        raise DiabloError(f"dh_h5_read_f64({path}, {dataset}): {err.value.decode()} ({rc})")
    return out[: n.value].copy()


def _ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"] and a.dtype in (np.float64, np.complex128)
    return a.ctypes.data_as(_DP)


class DiabloError(RuntimeError):
    pass


class Channel:
    """One device context running the channel timestep."""

    def __init__(self, params: dict, gy=None):
        self.lib = load_library()
        self.params = dict(params)
        self._p = make_params(params)
        h = ctypes.c_void_p()
        rc = self.lib.dh_create(ctypes.byref(self._p), ctypes.byref(h))
        if rc != DH_OK:
            raise DiabloError(f"dh_create failed ({rc})")
        self.h = h
        self.nx, self.ny, self.nz = params["nx"], params["ny"], params["nz"]
        self.n_th = params.get("n_th", 0)
        self.nxp = self.nx // 2
        self.nkx, self.nkz = self.nx // 3, self.nz // 3
        self.tnkz = 2 * self.nkz
        if gy is not None:
            self.set_grid(gy)
        if params.get("les"):
            self.set_les(True)

    def _check(self, rc, what):
        if rc != DH_OK:
            msg = self.lib.dh_last_error(self.h).decode()
            raise DiabloError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.dh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- set-up ----------------------------------------------------------
    def set_grid(self, gy):
        self.GY, self.GYF, self.DY, self.DYF = grid_metrics(np.asarray(gy, dtype=np.float64))
        self._check(self.lib.dh_set_grid(self.h, _ptr(self.GY), _ptr(self.GYF), _ptr(self.DY), _ptr(self.DYF)),
                    "dh_set_grid")

    def set_grid_h5(self, path):
        """CREATE_GRID_CHAN from a grid file (/grids/y, ReadGridHDF5 hdf5.f:548-710)."""
        self._check(self.lib.dh_set_grid_h5(self.h, str(path).encode()), "dh_set_grid_h5")
        self.GY, self.GYF, self.DY, self.DYF = grid_metrics(read_h5_f64(path, "/grids/y"))

    def set_clock(self, time, delta_t, time_step):
        self._check(self.lib.dh_set_clock(self.h, time, delta_t, time_step), "dh_set_clock")

    def clock(self):
        t, dt = ctypes.c_double(), ctypes.c_double()
        hb = (ctypes.c_double * 3)()
        ts = ctypes.c_int32()
        self._check(self.lib.dh_get_clock(self.h, ctypes.byref(t), ctypes.byref(dt), hb, ctypes.byref(ts)),
                    "dh_get_clock")
        return t.value, dt.value, list(hb), ts.value

    # -- state -------------------------------------------------------------
    def empty_state(self):
        cshape = (self.ny + 2, self.nz + 2, self.nxp + 1)
        nth = max(self.n_th, 1)
        st = {k: np.zeros(cshape, np.complex128) for k in ("cu1", "cu2", "cu3", "cp", "cf1", "cf2", "cf3")}
        st["cth"] = np.zeros((nth,) + cshape, np.complex128)
        st["cfth"] = np.zeros((nth,) + cshape, np.complex128)
        st["th"] = np.zeros((nth, self.ny + 2, self.nz + 2, self.nx + 2))
        return st

    def _hs(self, st):
        hs = DhHostState()
        for k in ("cu1", "cu2", "cu3", "cp", "cf1", "cf2", "cf3", "cth", "cfth", "th"):
            a = st.get(k)
            setattr(hs, k, _ptr(a) if a is not None else None)
        return hs

    def upload(self, st):
        self._check(self.lib.dh_upload(self.h, ctypes.byref(self._hs(st))), "dh_upload")

    def download(self, st=None):
        if st is None:
            st = self.empty_state()
        self._check(self.lib.dh_download(self.h, ctypes.byref(self._hs(st))), "dh_download")
        return st

    # -- time stepping -------------------------------------------------------
    def rk_substep(self, rk):
        self._check(self.lib.dh_rk_substep(self.h, rk), "dh_rk_substep")

    def time_step(self, n=1):
        self._check(self.lib.dh_time_step(self.h, n), "dh_time_step")

    def set_mixed_precision(self, on):
        """Switch the FP32/FP64 mixed form of the substep on or off."""
        self._check(self.lib.dh_set_mixed_precision(self.h, 1 if on else 0), "dh_set_mixed_precision")

    @property
    def mixed_precision(self):
        return bool(self.lib.dh_mixed_precision(self.h))

    def set_les(self, on):
        """Constant-Smagorinsky LES (the reference's LES = .TRUE., LES_MODEL_TYPE = 1)."""
        self._check(self.lib.dh_set_les(self.h, 1 if on else 0), "dh_set_les")

    @property
    def les(self):
        return bool(self.lib.dh_les(self.h))

    def set_les_closure(self, params=None, ts=1.0, ls=1.0, nus=1.0):
        """Learned eddy viscosity NU_T = nus softplus(y) (the network of
        dh_mlp_*: 6 -> 64 -> 64 -> 64 -> 1, GELU, BF16 MFMA): params the flat
        float32 parameter vector, ts and ls the time and length scales of its
        inputs; None returns to the Smagorinsky model."""
        if params is None:
            self._check(self.lib.dh_set_les_closure(self.h, None, 0), "dh_set_les_closure")
            return
        w = np.concatenate([np.asarray(params, dtype=np.float32).ravel(), np.float32([ts, ls, nus])])
        self._check(self.lib.dh_set_les_closure(self.h, w.ctypes.data, w.size), "dh_set_les_closure")

    def les_features(self):
        """The closure's inputs from the current state (dh_field_dev 301)."""
        self._check(self.lib.dh_les_features(self.h), "dh_les_features")

    @property
    def plane_threads(self):
        return int(self.lib.dh_plane_threads(self.h))

    def set_plane_threads(self, threads):
        self._check(self.lib.dh_set_plane_threads(self.h, threads), "dh_set_plane_threads")

    def autotune(self, nsteps=2):
        """Time the valid block sizes on this context's shape and keep the fastest."""
        best = ctypes.c_int32(0)
        self._check(self.lib.dh_autotune(self.h, nsteps, ctypes.byref(best)), "dh_autotune")
        return best.value

    def synchronize(self):
        self._check(self.lib.dh_synchronize(self.h), "dh_synchronize")

    def profile_substep(self, rk, nrep):
        """Mean milliseconds per phase of substep rk over nrep repetitions."""
        ms = (ctypes.c_double * 8)()
        self._check(self.lib.dh_profile_substep(self.h, rk, nrep, ms, 8), "dh_profile_substep")
        return {self.lib.dh_phase_name(i).decode(): ms[i] for i in range(6)}

    def rem_div(self):
        self._check(self.lib.dh_rem_div(self.h), "dh_rem_div")

    def poisson_p(self):
        """POISSON_P_CHAN (channel.f:1217-1453): CP on planes 1..NY from CU and CTH."""
        self._check(self.lib.dh_poisson_p_chan(self.h), "dh_poisson_p_chan")

    def save_stats(self, ly):
        """SAVE_STATS_CHAN's plane statistics on the device (save_stats.f:136-262):
        dict of (NY+2,) arrays in the order of STATS, plus 'ubulk'."""
        out = np.zeros((len(STATS), self.ny + 2))
        ub = ctypes.c_double(0.0)
        self._check(self.lib.dh_save_stats(self.h, ly, out.ctypes.data, ctypes.byref(ub)), "dh_save_stats")
        d = {k: out[i].copy() for i, k in enumerate(STATS)}
        d["ubulk"] = ub.value
        return d

    def create_flow(self, ic_type, seed, amp, ymask=0.0, ubulk0=0.0):
        self._check(self.lib.dh_create_flow_chan(self.h, ic_type, ubulk0, seed, amp, ymask), "dh_create_flow_chan")

    def create_th(self, ic_type):
        self._check(self.lib.dh_create_th_chan(self.h, ic_type), "dh_create_th_chan")

    def apply_bc_vel(self):
        self._check(self.lib.dh_apply_bc_vel(self.h), "dh_apply_bc_vel")

    def initial_condition(self, ic):
        """CREATE_TH_CHAN then CREATE_FLOW_CHAN (diablo_io.f init order)."""
        if self.n_th > 0 and ic.get("th_ic_type") is not None:
            self.create_th(ic["th_ic_type"])
        self.create_flow(ic["ic_type"], ic["seed"], ic["amp"], ic["ymask"], ic["ubulk0"])

    def field_ptr(self, which):
        return self.lib.dh_field_dev(self.h, which)

    def stream(self):
        return self.lib.dh_stream(self.h)

    def fft_to_fourier_dev(self, phys_ptr, four_ptr, nplanes):
        self._check(self.lib.dh_fft_xz_to_fourier_dev(self.h, phys_ptr, four_ptr, nplanes), "dh_fft_xz_to_fourier_dev")

    def fft_to_physical_dev(self, four_ptr, phys_ptr, nplanes):
        self._check(self.lib.dh_fft_xz_to_physical_dev(self.h, four_ptr, phys_ptr, nplanes),
                    "dh_fft_xz_to_physical_dev")


def make_channel(name, **kw):
    p, gy, ic = configs.CONFIGS[name](**kw)
    return Channel(p, gy), p, gy, ic


class Ensemble:
    """Independent Channel instances of any sizes stepped together
    (dh_ensemble_*): members of one shape run each launch of their step as
    one batch, shapes on streams of their own, all in one captured graph.
    The members stay usable on their own between ensemble steps and must
    outlive it."""

    def __init__(self, members):
        self.lib = load_library()
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = ctypes.c_void_p()
        rc = self.lib.dh_ensemble_create(len(self.members), arr, ctypes.byref(h))
        if rc != DH_OK:
            raise DiabloError(f"dh_ensemble_create failed ({rc})")
        self.h = h

    def time_step(self, n=1):
        rc = self.lib.dh_ensemble_time_step(self.h, n)
        if rc != DH_OK:
            raise DiabloError(f"dh_ensemble_time_step failed ({rc})")

    def autotune(self, nsteps=2):
        """Tune the block size per member shape (dh_ensemble_autotune)."""
        rc = self.lib.dh_ensemble_autotune(self.h, nsteps)
        if rc != DH_OK:
            raise DiabloError(f"dh_ensemble_autotune failed ({rc})")
        return [m.plane_threads for m in self.members]

    def synchronize(self):
        rc = self.lib.dh_ensemble_synchronize(self.h)
        if rc != DH_OK:
            raise DiabloError(f"dh_ensemble_synchronize failed ({rc})")

    def close(self):
        if getattr(self, "h", None):
            self.lib.dh_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

