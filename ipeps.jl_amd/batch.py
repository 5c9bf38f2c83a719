"""Many independent CTMRG environments advanced together on one GPU.

The reference advances independent environments one after another (the
per-basis-vector loop of src/optimize/optim_es.jl:127-153, split into jobs by
es_resume / es_num / part_save); here ``BatchedCTM`` hands every instance to
the engine at once (``ipeps_batch_*`` in include/ipeps_hip.h): a pool of
worker contexts, each with its own HIP stream and host thread, runs
``CTMTensors`` (src/ctm/ctm_tensor.jl:63-69) + ``rg_step``
(src/ctm/ctmrg.jl:57-64) + ``get_gs_energy`` (src/graph/evaluation.jl:15-29)
for the instances assigned to it.

``c4_instances`` builds the uneven workload: 64 instances with D in {2, 3, 4},
chi in {16, 32, 48, 64, 96} and the model (ising(2) or heisenberg(1)) drawn
from numpy default_rng(3); instance i has A = default_rng(3 + i)
standard_normal in Fortran order, rounded to float32.
"""
import ctypes as C

import numpy as np

from . import models
from ._lib import C128, F64, MIXED, check, lib
from .ctm import _hpair, _i64

_DT = {"f64": F64, "c128": C128, "mixed": MIXED}


def c4_instances(n=64, seed=3):
    """[(A float32 values as float64, chi, H, model name)] of the uneven batch."""
    rng = np.random.default_rng(seed)
    Ds = rng.choice([2, 3, 4], size=n)
    chis = rng.choice([16, 32, 48, 64, 96], size=n)
    kinds = rng.choice(["ising", "heisenberg"], size=n)
    out = []
    for i in range(n):
        D = int(Ds[i])
        A = models.seeded_tensor((D, D, D, D, 2), seed + i).astype(np.float32).astype(np.float64)
        H = models.ising(2.0) if kinds[i] == "ising" else models.heisenberg(1.0)
        out.append((np.asfortranarray(A), int(chis[i]), H, str(kinds[i])))
    return out


class BatchedCTM:
    """Independent environments on one device.

    ``As``: list of (D,D,D,D,d) arrays; ``chis``: per-instance chi; ``Hs``:
    per-instance [H1, H2]; ``dtype``: "f64", "c128" or "mixed"; ``workers``:
    contexts / streams / host threads; ``options``: engine options applied to
    every instance (see ``ENV_OPTIONS``).
    """

    def __init__(self, As, chis, Hs, dtype="mixed", workers=8, options=None, device=0):
        n = len(As)
        if not (len(chis) == n and len(Hs) == n and n > 0):
            raise ValueError("As, chis and Hs must have the same nonzero length")
        self.code = _DT[dtype] if isinstance(dtype, str) else int(dtype)
        hnp = np.complex128 if self.code == C128 else np.float64
        self._A = [np.asfortranarray(np.asarray(a), dtype=hnp) for a in As]
        self._H = [_hpair(H)[0] for H in Hs]
        dims = []
        for a in self._A:
            if a.ndim != 5:
                raise ValueError("A must be (D, D, D, D, d)")
            dims.extend(a.shape)
        hdim = [h[0].shape[0] for h in self._H]
        Aps = (C.c_void_p * n)(*[a.ctypes.data for a in self._A])
        H1s = (C.POINTER(C.c_double) * n)(*[h[0].ctypes.data_as(C.POINTER(C.c_double)) for h in self._H])
        H2s = (C.POINTER(C.c_double) * n)(*[h[1].ctypes.data_as(C.POINTER(C.c_double)) for h in self._H])
        opts = ";".join(f"{k}={float(v)}" for k, v in (options or {}).items()).encode()
        h = C.c_void_p()
        check(lib().ipeps_batch_create(int(device), n, self.code, _i64(dims), Aps, _i64(chis), H1s, H2s,
                                       _i64(hdim), int(workers), opts, C.byref(h)))
        self.h = h
        self.n = n
        self.shapes = [a.shape for a in self._A]

    def run(self, steps):
        """Advance every instance by ``steps`` rg_steps; returns (E complex per
        instance, wall seconds of the call)."""
        E = np.zeros(2 * self.n)
        sec = np.zeros(1)
        check(lib().ipeps_batch_run(self.h, int(steps), E.ctypes.data_as(C.POINTER(C.c_double)),
                                    sec.ctypes.data_as(C.POINTER(C.c_double))))
        return E[0::2] + 1j * E[1::2], float(sec[0])

    def stats(self):
        """Graph-mode counters: dict(captured, replayed, undone, eager) steps."""
        out = (C.c_int64 * 4)()
        check(lib().ipeps_batch_stats(self.h, out))
        return dict(captured=out[0], replayed=out[1], undone=out[2], eager=out[3])

    def tensor(self, inst, which):
        dims = (C.c_int64 * 5)()
        nd = C.c_int()
        check(lib().ipeps_batch_get_tensor(self.h, int(inst), int(which), dims, C.byref(nd), None))
        shape = tuple(dims[i] for i in range(nd.value))
        out = np.empty(shape, dtype=np.complex128 if self.code == C128 else np.float64, order="F")
        check(lib().ipeps_batch_get_tensor(self.h, int(inst), int(which), dims, C.byref(nd),
                                           out.ctypes.data_as(C.c_void_p)))
        return out

    def close(self):
        if self.h:
            lib().ipeps_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
