"""Client of the continuous-batching MIS request scheduler (tbh_submit /
tbh_poll / tbh_wait in include/tbh.h)."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .api import _dtype_code, _Marshal


class Server:
    """Requests (lists of SlicedBranch) are served in pipelined rounds: a
    round admits slices while their summed arena footprint fits the pool's
    region, their estimated work stays within 2^round_log2_macs MACs (0: the
    library default) and at most max_resident_slices (0: no cap) are in it.
    A request with a slice too large for the pool is refused alone.
    Results are the MIS values."""

    def __init__(self, backend, max_resident_slices=0, round_log2_macs=0.0):
        self.be = backend
        cfg = _lib.TbhServeConfig(int(max_resident_slices), float(round_log2_macs))
        check(_lib.tbh().tbh_serve_start_ex(backend.handle, C.byref(cfg)))
        self._open = True
        self._dt = {}

    def stats(self):
        """{'rounds', 'retried_rounds', 'failed_requests'} since start."""
        v = [C.c_uint64() for _ in range(3)]
        check(_lib.tbh().tbh_serve_stats(self.be.handle, *(C.byref(x) for x in v)))
        return dict(zip(("rounds", "retried_rounds", "failed_requests"), (x.value for x in v)))

    def submit(self, branches, element_type=np.float32, priority=0):
        """priority=1: the latency lane (scheduled ahead of normal requests)."""
        code, npdt = _dtype_code(element_type)
        m = _Marshal(branches, npdt)
        req = _lib.TbhRequest(code, m.ptr, m.n, int(priority))
        t = C.c_uint64()
        check(_lib.tbh().tbh_submit(self.be.handle, C.byref(req), C.byref(t)))
        self._dt[t.value] = npdt
        return t.value

    def poll(self, ticket):
        """(status, value, t_submit, t_done); a finished request is reported
        once, after which its ticket is unknown."""
        npdt = self._dt.get(ticket, np.float32)
        out = np.zeros(1, npdt)
        st = C.c_int()
        t0, t1 = C.c_double(), C.c_double()
        check(_lib.tbh().tbh_poll(self.be.handle, ticket, out.ctypes.data, C.byref(st),
                                  C.byref(t0), C.byref(t1)))
        status = _lib.STATUS[st.value]
        if status in ("done", "failed"):
            self._dt.pop(ticket, None)
        return status, out[0], t0.value, t1.value

    def wait(self, ticket):
        """Blocks until the request finishes; returns its value or raises."""
        npdt = self._dt.pop(ticket, np.float32)
        out = np.zeros(1, npdt)
        st = C.c_int()
        check(_lib.tbh().tbh_wait(self.be.handle, ticket, out.ctypes.data, C.byref(st)))
        if _lib.STATUS[st.value] != "done":
            raise _lib.TbhError(-1, f"request {ticket} {_lib.STATUS[st.value]}: "
                                    f"{_lib.tbh().tbh_last_error().decode()}")
        return out[0]

    def cancel(self, ticket):
        check(_lib.tbh().tbh_cancel(self.be.handle, ticket))

    def close(self):
        if self._open:
            _lib.tbh().tbh_serve_stop(self.be.handle)
            self._open = False

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
