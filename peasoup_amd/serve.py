"""Python client of the continuous-batching search server (psx_server_*).

A Server owns one scheduler thread in libpsx; requests are submitted
without blocking, admitted against the HBM budget, their trials batched with
those of other requests, and their candidates returned when complete.
"""
import ctypes as C

import numpy as np

import peasoup_amd as P
from peasoup_amd.workloads import CAND_DTYPE, SearchArgs, top_level


class Schedule(C.Structure):
    _fields_ = [("size", C.c_uint64), ("batch", C.c_uint32), ("fft_npass", C.c_int32), ("fft_len", C.c_int32 * 3),
                ("fft_cols", C.c_int32 * 3)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("batch", C.c_uint32), ("max_active", C.c_uint32),
                ("use_graph", C.c_int32), ("nschedules", C.c_uint32), ("schedules", C.POINTER(Schedule)),
                ("dispatch", C.c_int32)]


DISPATCH_STREAMS, DISPATCH_MIXED_GRAPH = 0, 1


class Request(C.Structure):
    _fields_ = [("args", SearchArgs), ("series", C.c_void_p), ("series_on_device", C.c_int32),
                ("nsamps", C.c_uint64), ("dm_list", C.c_void_p), ("ndm", C.c_uint32), ("accs", C.c_void_p),
                ("naccs", C.c_uint32)]


class Status(C.Structure):
    _fields_ = [("state", C.c_int32), ("error", C.c_int32), ("ncand", C.c_uint32), ("ntrials", C.c_uint64),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]


class Counters(C.Structure):
    _fields_ = [("used_bytes", C.c_uint64), ("budget_bytes", C.c_uint64), ("waiting", C.c_uint32),
                ("active", C.c_uint32), ("engines", C.c_uint32), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("trials_done", C.c_uint64)]


P.register_signatures({
    "psx_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "psx_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "psx_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(Status)]),
    "psx_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double, C.POINTER(Status)]),
    "psx_server_result": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "psx_server_stats": (C.c_int, [C.c_void_p, C.POINTER(Counters)]),
    "psx_server_destroy": (C.c_int, [C.c_void_p]),
})

STATES = {0: "waiting", 1: "active", 2: "done", 3: "failed"}


def _status(st):
    return {"state": STATES[st.state], "error": st.error, "ncand": st.ncand, "ntrials": st.ntrials,
            "t_submit": st.t_submit, "t_admit": st.t_admit, "t_done": st.t_done,
            "latency": st.t_done - st.t_submit if st.state >= 2 else None}


def schedules_from(table):
    """psx_schedule array from {size: {"batch": B, "decomp": [(L, C), ...] or None}}."""
    arr = (Schedule * max(len(table), 1))()
    for i, (n, e) in enumerate(sorted(table.items())):
        arr[i].size = int(n)
        arr[i].batch = int(e.get("batch") or 0)
        dec = e.get("decomp") or []
        arr[i].fft_npass = len(dec)
        for q, (L, Cc) in enumerate(dec):
            arr[i].fft_len[q] = int(L)
            arr[i].fft_cols[q] = int(Cc)
    return arr, len(table)


class Server:
    def __init__(self, budget_bytes, batch=16, max_active=0, use_graph=True, schedules=None,
                 dispatch=DISPATCH_STREAMS):
        """schedules: optional {fft size: {"batch", "decomp"}} (the tuned table).
        dispatch: DISPATCH_STREAMS (one stream per FFT length) or
        DISPATCH_MIXED_GRAPH (one graph launch per mixed-length batch)."""
        arr, n = schedules_from(schedules or {})
        self.cfg = ServerConfig(int(budget_bytes), batch, max_active, 1 if use_graph else 0, n,
                                C.cast(arr, C.POINTER(Schedule)), dispatch)
        h = C.c_void_p()
        P.check(P.lib().psx_server_create(C.byref(self.cfg), C.byref(h)))
        self.h = h.value
        self._keep = {}

    def submit(self, args, series, dm_list, accs=None, nsamps=None):
        dm_list = np.ascontiguousarray(dm_list, np.float32)
        if isinstance(series, P.DeviceArray):
            ptr, on_dev, ns = series.ptr, 1, nsamps
        else:
            series = np.ascontiguousarray(series, np.float32)
            ptr, on_dev, ns = series.ctypes.data, 0, series.shape[-1]
        a = None if accs is None else np.ascontiguousarray(accs, np.float32)
        req = Request(args, ptr, on_dev, ns, dm_list.ctypes.data, dm_list.size,
                      None if a is None else a.ctypes.data, 0 if a is None else a.size)
        rid = C.c_uint64()
        P.check(P.lib().psx_server_submit(self.h, C.byref(req), C.byref(rid)))
        self._keep[rid.value] = (series, dm_list, a)  # buffers stay valid until the request completes
        return rid.value

    def poll(self, rid):
        st = Status()
        P.check(P.lib().psx_server_poll(self.h, rid, C.byref(st)))
        return _status(st)

    def wait(self, rid, timeout=600.0):
        st = Status()
        rc = P.lib().psx_server_wait(self.h, rid, timeout, C.byref(st))
        s = _status(st)
        if rc:
            s["error_text"] = P.lib().psx_last_error_string().decode()
        return s

    def result(self, rid, cap=1 << 20, forest=False):
        """Top-level candidates of a finished request (the whole forest with
        forest=True); releases the request."""
        out = np.zeros(cap, CAND_DTYPE)
        cnt = C.c_uint32()
        P.check(P.lib().psx_server_result(self.h, rid, out.ctypes.data, cap, C.byref(cnt)))
        self._keep.pop(rid, None)
        out = out[:cnt.value]
        return out if forest else top_level(out)

    def stats(self):
        c = Counters()
        P.check(P.lib().psx_server_stats(self.h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in Counters._fields_}

    def close(self):
        if self.h:
            P.check(P.lib().psx_server_destroy(self.h))
            self.h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
