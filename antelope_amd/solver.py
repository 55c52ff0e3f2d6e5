"""Python binding of the C ABI (libantelope_sbe.so) for tests and bench.

`SBESolver` mirrors the reference's `SBEs` object as driven by
antelope.cpp:159-338: construct for a k block (SBEs::SBEs(cfg, gauge,
idx_start, idx_end)), run the time loop (the GenInterIntraCurrent +
RunSBEs sequence of antelope.cpp:297-338), then read the k-summed series
(inter_rad / intra_rad / density_matrix_integrated, antelope.cpp:196-205).

There is no fallback: if the HIP library is not built, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Optional

import numpy as np

from . import _abi as A

_LIB_PATH = Path(__file__).resolve().parent / "lib" / "libantelope_sbe.so"
_lib: Optional[C.CDLL] = None


class SBEError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{A.STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def library_path() -> Path:
    return Path(os.environ.get("ANTELOPE_SBE_LIB", _LIB_PATH))


def load_library() -> C.CDLL:
    """Load the HIP library; raises if it is not built (no silent fallback)."""
    global _lib
    if _lib is None:
        p = library_path()
        if not p.exists():
            raise RuntimeError(
                f"HIP library {p} is not built; run `python -m antelope_amd.build` first")
        _lib = A.bind(C.CDLL(str(p)))
    return _lib


def _check(lib, ctx, rc: int) -> None:
    if rc != A.SBE_OK:
        msg = lib.sbe_last_error(ctx)
        raise SBEError(rc, msg.decode() if msg else "")


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(A.c_double_p)


def laser_eval(config, t: np.ndarray):
    """A(t), E(t) [n][3] and (atmin, Nt) of a configuration (laser.h:301-322)."""
    lib = load_library()
    cfg, keep = config.to_c()
    t = np.ascontiguousarray(t, dtype=np.float64)
    Av = np.zeros((t.size, 3))
    Ev = np.zeros((t.size, 3))
    atmin = C.c_double()
    nt = C.c_int64()
    _check(lib, None, lib.sbe_laser_eval(C.byref(cfg), t.size, _ptr(t), _ptr(Av), _ptr(Ev),
                                         C.byref(atmin), C.byref(nt)))
    return Av, Ev, atmin.value, nt.value


def query_footprint(config, model, k_begin: int = -1, k_end: int = -1) -> int:
    lib = load_library()
    cfg, k1 = config.to_c()
    mdl, k2 = model.to_c()
    out = C.c_size_t()
    _check(lib, None, lib.sbe_query_footprint(C.byref(cfg), C.byref(mdl), k_begin, k_end, C.byref(out)))
    return out.value


def autotune(config, model, trial_steps: int = 64) -> dict:
    """Time the plan candidates of this problem on the GPU and cache the fastest
    for its class (sbe_autotune); returns the chosen plan."""
    lib = load_library()
    cfg, k1 = config.to_c()
    mdl, k2 = model.to_c()
    t = A.sbe_tuning()
    _check(lib, None, lib.sbe_autotune(C.byref(cfg), C.byref(mdl), trial_steps, C.byref(t)))
    return t.as_dict()


def batch_autotune(configs, models, trial_steps: int = 2048) -> dict:
    """sbe_batch_autotune: time the batch plan candidates and cache the fastest."""
    lib = load_library()
    n = len(configs)
    cs, ms, keep = [], [], []
    for c, m in zip(configs, models):
        cc, k1 = c.to_c()
        mm, k2 = m.to_c()
        cs.append(cc)
        ms.append(mm)
        keep += [k1, k2]
    t = A.sbe_tuning()
    _check(lib, None, lib.sbe_batch_autotune(n, (A.sbe_config * n)(*cs), (A.sbe_model * n)(*ms), trial_steps,
                                             C.byref(t)))
    return t.as_dict()


def tuning_get(config, model) -> dict:
    lib = load_library()
    cfg, k1 = config.to_c()
    mdl, k2 = model.to_c()
    t = A.sbe_tuning()
    _check(lib, None, lib.sbe_tuning_get(C.byref(cfg), C.byref(mdl), C.byref(t)))
    return t.as_dict()


def tuning_set(config, model, **fields) -> None:
    lib = load_library()
    cfg, k1 = config.to_c()
    mdl, k2 = model.to_c()
    t = A.sbe_tuning(**fields)
    _check(lib, None, lib.sbe_tuning_set(C.byref(cfg), C.byref(mdl), C.byref(t)))


def tuning_clear() -> None:
    load_library().sbe_tuning_clear()


class SBESolver:
    """One k block of one SBE solve, resident on the GPU."""

    def __init__(self, config, model, k_begin: int = -1, k_end: int = -1):
        self._lib = load_library()
        self.config, self.model = config, model
        cfg, k1 = config.to_c()
        mdl, k2 = model.to_c()
        ctx = C.c_void_p()
        rc = self._lib.sbe_create(C.byref(cfg), C.byref(mdl), k_begin, k_end, C.byref(ctx))
        _check(self._lib, None, rc)
        self._ctx = ctx
        self.nband = model.nband

    # -- info -------------------------------------------------------------
    @property
    def info(self) -> A.sbe_info:
        inf = A.sbe_info()
        _check(self._lib, self._ctx, self._lib.sbe_get_info(self._ctx, C.byref(inf)))
        return inf

    @property
    def nt(self) -> int:
        return self.info.nt

    # -- time loop ----------------------------------------------------------
    def run(self, step_begin: Optional[int] = None, step_end: Optional[int] = None) -> None:
        inf = self.info
        b = inf.steps_done if step_begin is None else step_begin
        e = inf.nt if step_end is None else step_end
        _check(self._lib, self._ctx, self._lib.sbe_run(self._ctx, b, e))

    def synchronize(self) -> None:
        _check(self._lib, self._ctx, self._lib.sbe_synchronize(self._ctx))

    # -- results ------------------------------------------------------------
    def series(self) -> dict:
        """k sums per time sample (before x dV), as sbe_get_series."""
        nt, N = self.nt, self.nband
        out = {
            "j_inter": np.zeros((nt, 3)),
            "j_intra": np.zeros((nt, 3)),
            "rho_bar": np.zeros((nt, N, N, 2)),
            "band_pop": np.zeros((nt, N)),
        }
        rc = self._lib.sbe_get_series(self._ctx, _ptr(out["j_inter"]), _ptr(out["j_intra"]),
                                      _ptr(out["rho_bar"]), _ptr(out["band_pop"]))
        _check(self._lib, self._ctx, rc)
        rb = out.pop("rho_bar")
        out["rho_bar"] = rb[..., 0] + 1j * rb[..., 1]
        return out

    def state(self, basis: str = "orbital") -> np.ndarray:
        inf = self.info
        nk = inf.k_end - inf.k_begin
        buf = np.zeros((nk, self.nband, self.nband, 2))
        b = {"orbital": A.SBE_BASIS_ORBITAL, "band": A.SBE_BASIS_BAND}[basis]
        _check(self._lib, self._ctx, self._lib.sbe_get_state(self._ctx, b, _ptr(buf)))
        return buf[..., 0] + 1j * buf[..., 1]

    def set_timing(self, enable: bool = True) -> None:
        _check(self._lib, self._ctx, self._lib.sbe_set_timing(self._ctx, int(enable)))

    def step_stats(self) -> A.sbe_step_stats:
        """Accepted / rejected steps per block of k points (sbe_get_step_stats)."""
        st = A.sbe_step_stats()
        _check(self._lib, self._ctx, self._lib.sbe_get_step_stats(self._ctx, C.byref(st)))
        return st

    def timing(self) -> dict:
        """Summed HIP-event durations (ms) of the kernels since the last call."""
        ms = (C.c_double * 2)()
        n = (C.c_int64 * 2)()
        _check(self._lib, self._ctx, self._lib.sbe_get_timing(self._ctx, ms, n))
        return {"step_ms": ms[0], "step_launches": n[0], "reduce_ms": ms[1], "reduce_launches": n[1]}

    def close(self) -> None:
        if getattr(self, "_ctx", None):
            self._lib.sbe_destroy(self._ctx)
            self._ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SBEBatch:
    """Independent 2-band solves advanced together (sbe_batch_*): one launch of
    the step kernel per chunk covers every unfinished solve."""

    def __init__(self, configs, models):
        self._lib = load_library()
        n = len(configs)
        cs, ms, self._keep = [], [], []
        for c, m in zip(configs, models):
            cc, k1 = c.to_c()
            mm, k2 = m.to_c()
            cs.append(cc)
            ms.append(mm)
            self._keep += [k1, k2]
        carr = (A.sbe_config * n)(*cs)
        marr = (A.sbe_model * n)(*ms)
        h = C.c_void_p()
        rc = self._lib.sbe_batch_create(n, carr, marr, C.byref(h))
        if rc != A.SBE_OK:
            raise SBEError(rc, self._lib.sbe_batch_last_error(None).decode())
        self._h = h
        self.n = n
        self.nband = [m.nband for m in models]

    def _chk(self, rc):
        if rc != A.SBE_OK:
            raise SBEError(rc, self._lib.sbe_batch_last_error(self._h).decode())

    def run(self, max_steps: int = 0) -> None:
        self._chk(self._lib.sbe_batch_run(self._h, max_steps))

    def synchronize(self) -> None:
        self._chk(self._lib.sbe_batch_synchronize(self._h))

    def reset(self) -> None:
        """Every solve back to its initial state (sbe_batch_reset)."""
        self._chk(self._lib.sbe_batch_reset(self._h))

    def info(self, i: int) -> A.sbe_info:
        inf = A.sbe_info()
        self._chk(self._lib.sbe_batch_get_info(self._h, i, C.byref(inf)))
        return inf

    def series(self, i: int) -> dict:
        nt, N = self.info(i).nt, self.nband[i]
        out = {"j_inter": np.zeros((nt, 3)), "j_intra": np.zeros((nt, 3)),
               "rho_bar": np.zeros((nt, N, N, 2)), "band_pop": np.zeros((nt, N))}
        self._chk(self._lib.sbe_batch_get_series(self._h, i, _ptr(out["j_inter"]), _ptr(out["j_intra"]),
                                                 _ptr(out["rho_bar"]), _ptr(out["band_pop"])))
        rb = out.pop("rho_bar")
        out["rho_bar"] = rb[..., 0] + 1j * rb[..., 1]
        return out

    def step_stats(self, i: int) -> A.sbe_step_stats:
        st = A.sbe_step_stats()
        self._chk(self._lib.sbe_batch_get_step_stats(self._h, i, C.byref(st)))
        return st

    def timing(self) -> dict:
        ms = C.c_double()
        n = C.c_int64()
        self._chk(self._lib.sbe_batch_get_timing(self._h, C.byref(ms), C.byref(n)))
        return {"step_ms": ms.value, "launches": n.value}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sbe_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
