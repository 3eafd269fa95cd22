"""Continuous-batching simulation server (simuse_server_*)."""
from __future__ import annotations

import ctypes as C

from ._lib import P, check, lib, u64p
from .batch import GenDesc, _L


class KuraGenDesc(C.Structure):
    _fields_ = [("n", C.c_uint32), ("modules", C.c_uint32), ("k", C.c_uint32), ("beta", C.c_float),
                ("inter", C.c_float), ("coupling", C.c_float), ("dt", C.c_float), ("seed", C.c_uint64)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_int32), ("net", GenDesc), ("kura", KuraGenDesc), ("horizon", C.c_uint64)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("steps", C.c_uint64), ("spikes", C.c_uint64), ("events", C.c_uint64),
                ("order_parameter", C.c_double), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("bytes", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("live", C.c_uint64), ("queued", C.c_uint64),
                ("live_bytes", C.c_uint64), ("peak_bytes", C.c_uint64), ("peak_live", C.c_uint32),
                ("now", C.c_double)]


REQ_NETWORK, REQ_KURAMOTO = 0, 1
QUEUED, RUNNING, DONE, REJECTED, FAILED = 0, 1, 2, -1, -2

_SIGS = {
    "simuse_server_create": ([C.c_uint64, C.c_uint32, C.POINTER(P)], C.c_int),
    "simuse_server_submit": ([P, C.POINTER(Request), u64p], C.c_int),
    "simuse_server_result": ([P, C.c_uint64, C.POINTER(Result)], C.c_int),
    "simuse_server_drain": ([P, C.c_double], C.c_int),
    "simuse_server_stats": ([P, C.POINTER(Stats)], C.c_int),
    "simuse_server_destroy": ([P], None),
    "simuse_kura_create_generated": ([C.POINTER(KuraGenDesc), C.c_int32, C.POINTER(P)], C.c_int),
    "simuse_kura_download_graph": ([P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_float),
                                    C.POINTER(C.c_float)], C.c_int),
    "simuse_kura_info": ([P, C.POINTER(C.c_uint32), u64p], C.c_int),
}
_bound = False


def _S():
    global _bound
    L = _L()
    if not _bound:
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _bound = True
    return L


def network_request(g: GenDesc, horizon: int) -> Request:
    r = Request()
    r.kind = REQ_NETWORK
    r.net = g
    r.horizon = int(horizon)
    return r


def kuramoto_request(n, modules=64, k=20, beta=0.1, inter=0.05, coupling=2.0, dt=0.001, seed=1,
                     horizon=1000) -> Request:
    r = Request()
    r.kind = REQ_KURAMOTO
    r.kura = KuraGenDesc(int(n), int(modules), int(k), float(beta), float(inter), float(coupling), float(dt),
                         int(seed) & ((1 << 64) - 1))
    r.horizon = int(horizon)
    return r


class Server:
    def __init__(self, hbm_budget_bytes: int, chunk_steps: int = 16):
        h = C.c_void_p()
        check(_S().simuse_server_create(int(hbm_budget_bytes), int(chunk_steps), C.byref(h)))
        self._h = h

    def submit(self, req: Request) -> int:
        rid = C.c_uint64()
        check(_S().simuse_server_submit(self._h, C.byref(req), C.byref(rid)))
        return int(rid.value)

    def result(self, rid: int) -> Result:
        r = Result()
        check(_S().simuse_server_result(self._h, int(rid), C.byref(r)))
        return r

    def drain(self, timeout_s: float = 600.0):
        check(_S().simuse_server_drain(self._h, float(timeout_s)))

    def stats(self) -> Stats:
        s = Stats()
        check(_S().simuse_server_stats(self._h, C.byref(s)))
        return s

    def close(self):
        if getattr(self, "_h", None) is not None:
            _S().simuse_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
