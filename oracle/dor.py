"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of
``bench.py`` may import this module, and only to check or to time the CPU
restatement; the product never calls it.

Arrays are exposed as numpy views of the oracle's Fortran-layout storage with
the index order reversed, so ``cu1[j, k, i]`` is the reference's
``CU1(I,K,J)`` (``header:212-243``).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
MAX_TH = 4


class DorParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("n_th", ctypes.c_int),
        ("nu", ctypes.c_double), ("beta", ctypes.c_double), ("lx", ctypes.c_double),
        ("lz", ctypes.c_double), ("nu_v_scale", ctypes.c_double),
        ("delta_t", ctypes.c_double), ("cfl", ctypes.c_double),
        ("variable_dt", ctypes.c_int), ("update_dt", ctypes.c_int),
        ("f_type", ctypes.c_int),
        ("px0", ctypes.c_double), ("omega0", ctypes.c_double), ("amp_omega0", ctypes.c_double),
        ("i_ro", ctypes.c_double), ("grav_x", ctypes.c_double), ("grav_y", ctypes.c_double),
        ("grav_z", ctypes.c_double),
        ("u_bc_ymin", ctypes.c_int), ("v_bc_ymin", ctypes.c_int), ("w_bc_ymin", ctypes.c_int),
        ("u_bc_ymax", ctypes.c_int), ("v_bc_ymax", ctypes.c_int), ("w_bc_ymax", ctypes.c_int),
        ("u_bc_ymin_c1", ctypes.c_double), ("v_bc_ymin_c1", ctypes.c_double),
        ("w_bc_ymin_c1", ctypes.c_double),
        ("u_bc_ymax_c1", ctypes.c_double), ("v_bc_ymax_c1", ctypes.c_double),
        ("w_bc_ymax_c1", ctypes.c_double),
        ("th_bc_ymin", ctypes.c_int * MAX_TH), ("th_bc_ymax", ctypes.c_int * MAX_TH),
        ("th_bc_ymin_c1", ctypes.c_double * MAX_TH), ("th_bc_ymax_c1", ctypes.c_double * MAX_TH),
        ("ri", ctypes.c_double * MAX_TH), ("pr", ctypes.c_double * MAX_TH),
        ("drhodx", ctypes.c_double * MAX_TH), ("drhodz", ctypes.c_double * MAX_TH),
        ("create_new_flow", ctypes.c_int),
        ("les", ctypes.c_int),
    ]


(CU1, CU2, CU3, CP, CR1, CR2, CR3, CF1, CF2, CF3, CS1, CTH, CFTH, CRTH, TH, RTH,
 GY, GYF, DY, DYF, KX, KX2, KZ, KZ2, HBAR, BETABAR, ZETABAR, NUT) = range(28)

_lib = None


# plane statistics of dor_save_stats, in its order
STATS = ("ume", "vme", "wme", "urms", "vrms", "wrms", "uv", "uw", "wv", "dudy", "dwdy", "shear")


def load():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libdor.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle` (or __graft_entry__.build())")
        lib = ctypes.CDLL(path)
        vp, dp, ip, i64p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.POINTER(ctypes.c_int64)
        lib.dor_create.restype = vp
        lib.dor_create.argtypes = [ctypes.POINTER(DorParams)]
        lib.dor_destroy.argtypes = [vp]
        lib.dor_array.restype = dp
        lib.dor_array.argtypes = [vp, ip, i64p]
        lib.dor_set_grid_gy.argtypes = [vp, dp]
        lib.dor_make_gy.argtypes = [ip, ctypes.c_double, ctypes.c_double, ip, dp]
        lib.dor_init.argtypes = [vp]
        lib.dor_get_clock.argtypes = [vp, dp, dp, ctypes.POINTER(ctypes.c_int)]
        lib.dor_set_clock.argtypes = [vp, ctypes.c_double, ctypes.c_double, ip]
        for f in ("dor_fft_xz_to_fourier", "dor_fft_xz_to_physical"):
            getattr(lib, f).argtypes = [vp, dp, ip, ip]
        lib.dor_rk_chan_1.argtypes = [vp, ip]
        lib.dor_rem_div_chan.argtypes = [vp]
        lib.dor_poisson_p_chan.argtypes = [vp]
        lib.dor_courant.argtypes = [vp]
        lib.dor_time_step.argtypes = [vp]
        lib.dor_thomas_real.argtypes = [dp, dp, dp, dp, ip, ip]
        lib.dor_thomas_complex.argtypes = [dp, dp, dp, dp, ip, ip]
        lib.dor_fft_c2c.argtypes = [ip, ip, dp, dp]
        lib.dor_create_flow_chan.argtypes = [vp, ip, ctypes.c_double, ctypes.c_uint64, ctypes.c_double, ctypes.c_double]
        lib.dor_create_th_chan.argtypes = [vp, ip]
        lib.dor_apply_bc_vel.argtypes = [vp]
        lib.dor_les_chan.argtypes = [vp]
        lib.dor_save_stats.argtypes = [vp, ctypes.c_double, dp, dp]
        lib.dor_gauss.restype = ctypes.c_double
        lib.dor_gauss.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        _lib = lib
    return _lib


def _dptr(a):
    assert a.dtype in (np.float64, np.complex128) and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_params(d):
    p = DorParams()
    for k, v in d.items():
        if k in ("th_bc_ymin", "th_bc_ymax", "th_bc_ymin_c1", "th_bc_ymax_c1", "ri", "pr", "drhodx", "drhodz"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        elif k in DorParams.__dict__:
            setattr(p, k, v)
        else:
            raise KeyError(k)
    return p


def make_gy(ny, L, cs, grid_type=1):
    out = np.zeros(ny + 1)
    load().dor_make_gy(ny, L, cs, grid_type, _dptr(out))
    return out


def fft_c2c(x, sign):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(x)
    load().dor_fft_c2c(len(x), sign, _dptr(x), _dptr(out))
    return out


def thomas_real(a, b, c, g):
    """Arrays shaped [NY+2, INX] (reference (0:INX-1, 0:INY+1)); in place."""
    iny = a.shape[0] - 2
    load().dor_thomas_real(_dptr(a), _dptr(b), _dptr(c), _dptr(g), iny, a.shape[1])


def thomas_complex(a, b, c, g):
    """a,b,c real [NY+2, INX+1]; g complex [NY+2, INX+1]; in place."""
    iny = a.shape[0] - 2
    load().dor_thomas_complex(_dptr(a), _dptr(b), _dptr(c), _dptr(g), iny, a.shape[1] - 1)


def gauss(seed, counter):
    return load().dor_gauss(seed, counter)


class Oracle:
    """One-rank restatement of DIABLO's channel state and RK_CHAN_1."""

    def __init__(self, params: dict):
        self.lib = load()
        self.pd = dict(params)
        self._p = make_params(params)
        self.h = self.lib.dor_create(ctypes.byref(self._p))
        if not self.h:
            raise ValueError("dor_create rejected the parameters")
        self.nx, self.ny, self.nz = params["nx"], params["ny"], params["nz"]
        self.n_th = params.get("n_th", 0)
        self.nxp = self.nx // 2
        self.nkx = self.nx // 3
        self.nkz = self.nz // 3
        self.tnkz = 2 * self.nkz

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.dor_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def _raw(self, which):
        n = ctypes.c_int64()
        ptr = self.lib.dor_array(self.h, which, ctypes.byref(n))
        return np.ctypeslib.as_array(ptr, shape=(n.value,))

    def carr(self, which):
        """Complex field view [(n_th,) NY+2, NZ+2, NXP+1]."""
        a = self._raw(which).view(np.complex128)
        per = (self.ny + 2) * (self.nz + 2) * (self.nxp + 1)
        if a.size > per:
            return a.reshape(-1, self.ny + 2, self.nz + 2, self.nxp + 1)
        return a.reshape(self.ny + 2, self.nz + 2, self.nxp + 1)

    def rarr(self, which):
        """Real physical view [(n_th,) NY+2, NZ+2, NX+2] (aliases the complex one for CU*, CR*, ...)."""
        a = self._raw(which)
        per = (self.ny + 2) * (self.nz + 2) * (self.nx + 2)
        if a.size > per:
            return a.reshape(-1, self.ny + 2, self.nz + 2, self.nx + 2)
        return a.reshape(self.ny + 2, self.nz + 2, self.nx + 2)

    def vec(self, which):
        return self._raw(which)

    def set_grid_gy(self, gy):
        gy = np.ascontiguousarray(gy, dtype=np.float64)
        assert gy.shape == (self.ny + 1,)
        self.lib.dor_set_grid_gy(self.h, _dptr(gy))

    def clock(self):
        t, dt, ts = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        self.lib.dor_get_clock(self.h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(ts))
        return t.value, dt.value, ts.value

    def set_clock(self, time, dt, time_step):
        self.lib.dor_set_clock(self.h, time, dt, time_step)

    def fft_to_fourier(self, which, jmin, jmax):
        self.lib.dor_fft_xz_to_fourier(self.h, _dptr(self._raw(which)), jmin, jmax)

    def fft_to_physical(self, which, jmin, jmax):
        self.lib.dor_fft_xz_to_physical(self.h, _dptr(self._raw(which)), jmin, jmax)

    def rk_chan_1(self, rk):
        self.lib.dor_rk_chan_1(self.h, rk)

    def rem_div(self):
        self.lib.dor_rem_div_chan(self.h)

    def poisson_p(self):
        """POISSON_P_CHAN: CP from CU (and CTH); CF1..3 and CS1 become scratch."""
        self.lib.dor_poisson_p_chan(self.h)

    def courant(self):
        self.lib.dor_courant(self.h)

    def time_step(self, n=1):
        for _ in range(n):
            self.lib.dor_time_step(self.h)

    def create_flow(self, ic_type, seed, amp, ymask=0.0, ubulk0=0.0):
        self.lib.dor_create_flow_chan(self.h, ic_type, ubulk0, seed, amp, ymask)

    def create_th(self, ic_type):
        self.lib.dor_create_th_chan(self.h, ic_type)

    def apply_bc_vel(self):
        self.lib.dor_apply_bc_vel(self.h)

    def les_chan(self):
        """les_chan: NU_T and the SGS divergence added to CF1..3; U1..3 left physical."""
        self.lib.dor_les_chan(self.h)

    def save_stats(self, ly):
        """SAVE_STATS_CHAN's plane statistics (dict of (NY+2,) arrays) and UBULK."""
        out = np.zeros((len(STATS), self.ny + 2))
        ub = ctypes.c_double(0.0)
        self.lib.dor_save_stats(self.h, ly, _dptr(out), ctypes.byref(ub))
        d = {k: out[i].copy() for i, k in enumerate(STATS)}
        d["ubulk"] = ub.value
        return d
