"""Serving loop over the HIP library: continuous batching of CI solves (pqg_serve_*)."""
from __future__ import annotations

import ctypes as ct
import time
from typing import Dict, List, Optional

import numpy as np

from . import _lib
from ._lib import DavParams, Result, check, load
from .ci import Context, default_context, make_desc


class Server:
    """Admits CI requests into `max_slots` resident slots by the context's HBM budget."""

    def __init__(self, ctx: Optional[Context] = None, max_slots: int = 8):
        self.ctx = ctx or default_context()
        self.h = ct.c_void_p()
        check(load().pqg_serve_create(self.ctx.h, max_slots, ct.byref(self.h)))
        self._buf = (Result * 256)()

    def submit(self, prob, nev: Optional[int] = None, nvec_init: int = 5, nvec_max: int = 0, maxit: int = 500,
               tol: float = 1e-6, precision: int = _lib.PQG_FP64, perturb: float = 0.0) -> int:
        keep: list = []
        d = make_desc(prob, keep)
        p = DavParams(nev or prob.nroots, nvec_init, nvec_max, maxit, tol, precision, perturb)
        rid = ct.c_int64()
        check(load().pqg_serve_submit(self.h, ct.byref(d), ct.byref(p), ct.byref(rid)))
        return rid.value

    def poll(self) -> List[dict]:
        n = ct.c_int()
        check(load().pqg_serve_step(self.h, self._buf, len(self._buf), ct.byref(n)))
        out = []
        for i in range(n.value):
            r = self._buf[i]
            out.append(dict(id=r.request_id, status=r.status, energies=np.array(r.energies[: r.nev]),
                            t_submit=r.t_submit, t_admit=r.t_admit, t_done=r.t_done, iterations=r.iterations,
                            message=r.message.decode(errors="replace")))
        return out

    def stats(self):
        q, r, b = ct.c_int(), ct.c_int(), ct.c_uint64()
        check(load().pqg_serve_stats(self.h, ct.byref(q), ct.byref(r), ct.byref(b)))
        return q.value, r.value, b.value

    def clock(self) -> float:
        return load().pqg_serve_clock(self.h)

    def drain(self, expected: int, timeout_s: float = 600.0) -> Dict[int, dict]:
        """Poll until `expected` results arrived."""
        got: Dict[int, dict] = {}
        t0 = time.time()
        while len(got) < expected:
            for r in self.poll():
                got[r["id"]] = r
            if len(got) < expected:
                if time.time() - t0 > timeout_s:
                    raise TimeoutError(f"{len(got)}/{expected} results after {timeout_s}s")
                time.sleep(0.0005)
        return got

    def close(self):
        if self.h:
            load().pqg_serve_destroy(self.h)
            self.h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
