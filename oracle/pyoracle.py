"""ctypes wrapper of oracle/libggx_oracle.so (TEST INFRASTRUCTURE ONLY; see oracle/__init__.py)."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libggx_oracle.so")
c_int, c_dbl = ctypes.c_int32, ctypes.c_double
P = ctypes.POINTER


class Params(ctypes.Structure):
    _fields_ = [("real_domain_size", c_int), ("grid_blowup", c_int), ("convolution_mode", c_int),
                ("dt", c_dbl), ("analysis_timestep", c_int), ("start_time", c_int), ("isotropic", c_int),
                ("nthreads", c_int)]


class Info(ctypes.Structure):
    _fields_ = [(n, c_int) for n in (
        "present", "exists", "motion_regular", "intersects_boundary", "xmin", "xmax", "ymin", "ymax",
        "id_xmin", "id_xmax", "id_ymin", "id_ymax", "contour_len", "n_neighbours", "n_compare", "error")] + [
        (n, c_dbl) for n in ("volume", "perimeter", "energy", "mean_da")]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        sig = {
            "ggxo_create": (c_int, [P(Params), c_int, P(vp)]),
            "ggxo_destroy": (None, [vp]),
            "ggxo_add_grain": (c_int, [vp, c_int, c_int, P(c_dbl), P(c_dbl)]),
            "ggxo_init": (c_int, [vp]),
            "ggxo_step": (c_int, [vp, c_int]),
            "ggxo_get_info": (c_int, [vp, c_int, P(Info)]),
            "ggxo_grain_phi": (c_int, [vp, c_int, P(c_dbl)]),
            "ggxo_grain_idlocal": (c_int, [vp, c_int, P(ctypes.c_uint32)]),
            "ggxo_grain_contour": (c_int, [vp, c_int, P(c_dbl)]),
            "ggxo_grain_neighbours": (c_int, [vp, c_int, P(c_int), P(c_dbl)]),
            "ggxo_grain_compare_list": (c_int, [vp, c_int, P(c_int)]),
            "ggxo_current_grains": (c_int, [vp]),
            "ggxo_convolve": (c_int, [c_int, c_int, c_dbl, c_int, P(c_dbl), P(c_dbl)]),
            "ggxo_dft": (c_int, [c_int, c_int, P(c_dbl), P(c_dbl)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(P(c_dbl))


def convolve(phi, mode, dt, R):
    L = load()
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    out = np.empty_like(phi)
    assert L.ggxo_convolve(phi.shape[0], mode, dt, R, _dp(phi), _dp(out)) == 0
    return out


def dft(x, sign=-1):
    L = load()
    x = np.ascontiguousarray(x, dtype=np.complex128)
    inp = np.ascontiguousarray(np.stack([x.real, x.imag], -1).reshape(-1))
    out = np.empty_like(inp)
    assert L.ggxo_dft(len(x), sign, _dp(inp), _dp(out)) == 0
    o = out.reshape(-1, 2)
    return o[:, 0] + 1j * o[:, 1]


class Oracle:
    """One oracle simulation: grains from a network, then step(loop)."""

    def __init__(self, R, gb, dt, max_id, analysis_timestep=1, nthreads=0, mode=2, start_time=0):
        self.L = load()
        p = Params(R, gb, mode, dt, analysis_timestep, start_time, 1, nthreads)
        self._c = ctypes.c_void_p()
        assert self.L.ggxo_create(ctypes.byref(p), max_id, ctypes.byref(self._c)) == 0
        self.max_id = max_id
        self.loop = start_time

    def add_grain(self, gid, pts, quat):
        xy = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1)
        q = np.ascontiguousarray(quat, dtype=np.float64)
        assert self.L.ggxo_add_grain(self._c, gid, len(xy) // 2, _dp(xy), _dp(q)) == 0

    def init(self):
        assert self.L.ggxo_init(self._c) == 0

    def step(self):
        rc = self.L.ggxo_step(self._c, self.loop)
        self.loop += 1
        return rc

    def info(self, gid):
        i = Info()
        self.L.ggxo_get_info(self._c, gid, ctypes.byref(i))
        return i

    def phi(self, gid):
        i = self.info(gid)
        n = i.xmax - i.xmin
        out = np.zeros(max(n * n, 1))
        cnt = self.L.ggxo_grain_phi(self._c, gid, _dp(out))
        return out[:cnt].reshape(n, n)

    def idlocal(self, gid):
        i = self.info(gid)
        n = i.id_xmax - i.id_xmin
        out = np.zeros(max(n * n, 1), dtype=np.uint32)
        cnt = self.L.ggxo_grain_idlocal(self._c, gid, out.ctypes.data_as(P(ctypes.c_uint32)))
        return out[:cnt]

    def contour(self, gid):
        i = self.info(gid)
        out = np.zeros(max(2 * i.contour_len, 2))
        cnt = self.L.ggxo_grain_contour(self._c, gid, _dp(out))
        return out[:2 * cnt].reshape(-1, 2)

    def neighbours(self, gid):
        i = self.info(gid)
        ids = np.zeros(max(i.n_neighbours, 1), dtype=np.int32)
        lens = np.zeros(max(i.n_neighbours, 1))
        cnt = self.L.ggxo_grain_neighbours(self._c, gid, ids.ctypes.data_as(P(c_int)), _dp(lens))
        return ids[:cnt], lens[:cnt]

    def compare_list(self, gid):
        i = self.info(gid)
        ids = np.zeros(max(i.n_compare, 1), dtype=np.int32)
        cnt = self.L.ggxo_grain_compare_list(self._c, gid, ids.ctypes.data_as(P(c_int)))
        return ids[:cnt]

    def current_grains(self):
        return self.L.ggxo_current_grains(self._c)

    def __del__(self):
        if getattr(self, "_c", None):
            self.L.ggxo_destroy(self._c)
            self._c = None
