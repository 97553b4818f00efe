"""Batched matte solves (gm_batch_*): many images of uneven size, several
constraint sets (right-hand sides) per image, one launch per phase for all."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import SolveInfo, call, lib
from .matting import _dev, _p, _stream, solve_opts


class _DevView:
    """__cuda_array_interface__ over library-owned device memory."""

    def __init__(self, ptr: int, n: int):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "<f4", "data": (ptr, False), "version": 3,
                                         "strides": None}


class Batch:
    def __init__(self, sizes, nrhs: int, precision: int = _lib.PREC_FP32):
        self.sizes = [(int(w), int(h)) for (w, h) in sizes]
        self.nrhs = nrhs
        W = np.array([w for w, _ in self.sizes], np.int32)
        H = np.array([h for _, h in self.sizes], np.int32)
        self.h = C.c_void_p()
        call("gm_batch_create", C.byref(self.h), len(self.sizes), W.ctypes.data, H.ctypes.data, nrhs, precision)

    def assemble(self, i: int, image, eps: float = 1e-7):
        im = _dev(image)
        Hh, Ww, nch = im.shape
        assert (Ww, Hh) == self.sizes[i]
        self._keep = im
        call("gm_batch_assemble", self.h, i, _p(im), nch, Ww, eps, _stream())

    def set_rhs(self, i: int, j: int, lam, tgt, x0=None):
        la, tg, x = _dev(lam), _dev(tgt), _dev(x0)
        call("gm_batch_set_rhs", self.h, i, j, _p(la), _p(tg), _p(x), _stream())
        torch.cuda.current_stream().synchronize()

    def set_grid(self, per_cu: int):
        call("gm_batch_set_grid", self.h, per_cu)

    def set_tiles(self, tile_width: int):
        call("gm_batch_set_tiles", self.h, tile_width)

    def autotune(self, reps: int = 3, solve: dict | None = None) -> float:
        """Pick the fused iteration's tile width and grid by timing them on this
        batch: full-load launches (solve None; returns microseconds per launch)
        or whole solves with the given solve() options (returns microseconds
        per solve)."""
        us = C.c_double(0)
        o = None
        if solve is not None:
            kw = dict(tol=1e-6, max_iter=20000, check_every=10, refine_every=0, verify=True)
            kw.update(solve)
            o = C.byref(solve_opts(method=_lib.METHOD_PCG, **kw))
        call("gm_batch_autotune", self.h, o, reps, C.byref(us), _stream())
        return us.value

    def config(self):
        """(tile width, workgroups per CU, tuned microseconds per launch)."""
        tw, pc, us = C.c_int(0), C.c_int(0), C.c_double(0)
        call("gm_batch_get_config", self.h, C.byref(tw), C.byref(pc), C.byref(us))
        return tw.value, pc.value, us.value

    def solve(self, tol=1e-6, max_iter=20000, check_every=10, refine_every=0, verify=True, raise_on_error=True):
        o = solve_opts(method=_lib.METHOD_PCG, tol=tol, max_iter=max_iter, check_every=check_every,
                       refine_every=refine_every, verify=verify)
        st = lib().gm_batch_solve(self.h, C.byref(o), _stream())
        if raise_on_error and st not in (0, 6):
            _lib.check(st)
        return st

    def alpha(self, i: int, j: int) -> torch.Tensor:
        ptr = C.c_void_p()
        call("gm_batch_alpha", self.h, i, j, C.byref(ptr))
        w, h = self.sizes[i]
        return torch.as_tensor(_DevView(ptr.value, w * h), device="cuda")

    def info(self, i: int, j: int) -> SolveInfo:
        inf = SolveInfo()
        call("gm_batch_info", self.h, i, j, C.byref(inf))
        return inf

    def close(self):
        if self.h:
            lib().gm_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
