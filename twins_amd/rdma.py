"""Python access to the RDMA model (`twins-rdma`) in libtwins_core.so: a ctypes layer over
include/twins_rdma.h. Worlds come from an HPCC config file, KEY=VALUE overrides, a TwiNS topology
file and a flow file; a batch of them runs on the GPU; results are per-flow completion times and
the reference's counters. It simulates nothing itself: without the built library or a GPU every
call raises.
"""
import ctypes

from . import lib as _core_lib, TwError, _track, _release_ok

EV_NAMES = {0: "drops", 1: "marks", 2: "pauses", 3: "resumes", 4: "tx-ovf", 5: "queue-ovf", 6: "ackq-ovf", 7: "qps-ovf",
            8: "undecided-ties", 9: "lost-tails", 10: "acks-for-done", 11: "nacks", 12: "due-ovf", 32: "late"}
BAD = (4, 5, 6, 7, 12, 32)  # counters that make a run inexact (the reference's exit status 3)
EV_LOST = 9  # ties decided on an overwritten tail node: the run is redone with larger tail rings
INEXACT = BAD + (EV_LOST,)  # a result that is not known to be exact


class PackCfg(ctypes.Structure):
    _fields_ = [("q3cap", ctypes.c_int32), ("tail_log2", ctypes.c_int32), ("team_blocks", ctypes.c_int32),
                ("hosts_per_wave", ctypes.c_int32), ("tail_exact", ctypes.c_int32), ("serial", ctypes.c_int32)]


class RunStats(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int64), ("sim_ns", ctypes.c_int64), ("wall_s", ctypes.c_double),
                ("kernel_s", ctypes.c_double), ("flows_done", ctypes.c_uint32), ("launches", ctypes.c_int32)]


class ServerCfg(ctypes.Structure):
    _fields_ = [("q3cap", ctypes.c_int32), ("tail_log2", ctypes.c_int32), ("hosts_per_wave", ctypes.c_int32),
                ("max_blocks", ctypes.c_int32), ("arena_bytes", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("waiting", ctypes.c_int32), ("resident", ctypes.c_int32), ("done", ctypes.c_int32), ("failed", ctypes.c_int32),
                ("arena_used", ctypes.c_uint64), ("arena_bytes", ctypes.c_uint64), ("launches", ctypes.c_uint64)]


_bound = False


def lib():
    global _bound
    L = _core_lib()
    if not _bound:
        vp, i32, i64, u64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64
        L.trd_last_error.restype = ctypes.c_char_p
        L.trd_world_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), i32, ctypes.c_char_p, ctypes.c_char_p,
                                       u64, i64, ctypes.POINTER(vp)]
        L.trd_world_info.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32),
                                     ctypes.POINTER(u64), ctypes.POINTER(ctypes.c_uint32)]
        L.trd_world_destroy.argtypes = [vp]
        L.trd_batch_create.argtypes = [ctypes.POINTER(vp), i32, ctypes.POINTER(PackCfg), ctypes.POINTER(vp)]
        L.trd_batch_run.argtypes = [vp, ctypes.POINTER(RunStats)]
        L.trd_fct_get.argtypes = [vp, i32, ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.trd_counters_get.argtypes = [vp, ctypes.POINTER(u64)]
        L.trd_world_stats_get.argtypes = [vp, i32, ctypes.POINTER(i64), ctypes.POINTER(ctypes.c_double)]
        L.trd_batch_tail_log2.argtypes = [vp, ctypes.POINTER(i32)]
        L.trd_batch_destroy.argtypes = [vp]
        L.trd_txtrace_enable.argtypes = [vp, i64, i64, u64]
        L.trd_txtrace_write.argtypes = [vp, i32, ctypes.c_char_p, ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.trd_server_create.argtypes = [ctypes.POINTER(ServerCfg), ctypes.POINTER(vp)]
        L.trd_server_submit.argtypes = [vp, vp, ctypes.POINTER(i32)]
        L.trd_server_poll.argtypes = [vp, ctypes.POINTER(ServerStats)]
        L.trd_server_state.argtypes = [vp, i32, ctypes.POINTER(i32)]
        L.trd_server_result.argtypes = [vp, i32, ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(u64)]
        L.trd_server_destroy.argtypes = [vp]
        _bound = True
    return L


def _check(rc):
    if rc != 0:
        raise TwError(f"twins rdma error {rc}: {lib().trd_last_error().decode()}")


class World:
    """One RDMA world: config file, overrides (["CC_MODE=3", ...]), topology and flow files."""

    def __init__(self, config, topo, flows, overrides=(), rate_bps=0, delay_ns=0):
        L = lib()
        ov = (ctypes.c_char_p * max(1, len(overrides)))(*[o.encode() for o in overrides])
        self._h = ctypes.c_void_p()
        _check(L.trd_world_create(config.encode(), ov, len(overrides), topo.encode(), flows.encode(), rate_bps, delay_ns,
                                  ctypes.byref(self._h)))
        h, s, f = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        rtt, bdp = ctypes.c_uint64(), ctypes.c_uint32()
        _check(L.trd_world_info(self._h, ctypes.byref(h), ctypes.byref(s), ctypes.byref(f), ctypes.byref(rtt), ctypes.byref(bdp)))
        self.hosts, self.switches, self.flows = h.value, s.value, f.value
        self.max_rtt, self.max_bdp = rtt.value, bdp.value
        _track(self, 2)

    def close(self):
        if getattr(self, "_h", None) and _release_ok():
            lib().trd_world_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Worlds side by side on the GPU (the reference's --batch). tail_log2 with tail_exact: every
    world's ring of chain tails holds 2^tail_log2 nodes (the reference's --tail); without, it caps
    the automatic size."""

    TAIL_MAX = 27  # the largest ring run_exact() asks for (the reference's rerun limit)

    def __init__(self, worlds, q3cap=0, tail_log2=0, team_blocks=0, hosts_per_wave=0, tail_exact=False, serial=False):
        self.worlds = list(worlds)
        self._cfg = (q3cap, tail_log2, team_blocks, hosts_per_wave, 1 if tail_exact else 0, 1 if serial else 0)
        self._h = ctypes.c_void_p()
        self._create()
        _track(self, 1)

    def _create(self):
        self._h = ctypes.c_void_p()  # (a fresh handle: close() left None)
        arr = (ctypes.c_void_p * len(self.worlds))(*[w._h for w in self.worlds])
        _check(lib().trd_batch_create(arr, len(self.worlds), ctypes.byref(PackCfg(*self._cfg)), ctypes.byref(self._h)))

    def run(self):
        st = RunStats()
        _check(lib().trd_batch_run(self._h, ctypes.byref(st)))
        return st

    def tail_log2(self):
        v = ctypes.c_int32()
        _check(lib().trd_batch_tail_log2(self._h, ctypes.byref(v)))
        return v.value

    def run_exact(self):
        """run(); if a tie met an overwritten tail node, the worlds again from the start on larger
        rings (2^26, then 2^27 nodes), as the reference's twins-rdma reruns itself
        (src/rdma/rmain.cu:2332-2363). Returns (stats, reruns)."""
        reruns = 0
        while True:
            st = self.run()
            lg = self.tail_log2()
            if not self.counters()[EV_LOST] or lg >= self.TAIL_MAX:
                return st, reruns
            self.close()
            q3cap, _, team_blocks, hpw, _, serial = self._cfg
            self._cfg = (q3cap, 26 if lg < 26 else self.TAIL_MAX, team_blocks, hpw, 1, serial)
            self._create()
            reruns += 1

    def fct(self, w):
        """(start_ns, fct_ns) lists of world w's flows by id; fct -1: not finished."""
        n = self.worlds[w].flows
        a, b = (ctypes.c_int64 * max(1, n))(), (ctypes.c_int64 * max(1, n))()
        _check(lib().trd_fct_get(self._h, w, a, b))
        return list(a)[:n], list(b)[:n]

    def fct_text(self, w):
        """The reference's output file: "id start_ns fct_ns" per finished flow, by id."""
        st, f = self.fct(w)
        return "".join(f"{i} {st[i]} {f[i]}\n" for i in range(len(f)) if f[i] >= 0)

    def counters(self):
        ev = (ctypes.c_uint64 * 40)()
        _check(lib().trd_counters_get(self._h, ev))
        return list(ev)

    def summary_line(self):
        """The reference's counter line (src/rdma/rmain.cu:2257-2258)."""
        e = self.counters()
        return (f"drops {e[0]} marks {e[1]} pauses {e[2]} resumes {e[3]} nacks {e[11]} | undecided ties {e[8]}, on lost tails {e[9]} | "
                f"overflows: tx {e[4]} queue {e[5]} ackq {e[6]} qps {e[7]} due {e[12]} | acks for done QPs {e[10]}")

    def txtrace_enable(self, from_ns=0, until_ns=(1 << 63) - 1, max_records=0):
        """Trace every transmission start at from_ns <= t <= until_ns (the reference's --txtrace); before run()."""
        _check(lib().trd_txtrace_enable(self._h, from_ns, until_ns, max_records))

    def txtrace_write(self, w, path):
        """World w's trace in the reference's text format; returns (lines written, records made over all worlds)."""
        n, seen = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().trd_txtrace_write(self._h, w, str(path).encode(), ctypes.byref(n), ctypes.byref(seen)))
        return n.value, seen.value

    def world_stats(self, w):
        slots, dev = ctypes.c_int64(), ctypes.c_double()
        _check(lib().trd_world_stats_get(self._h, w, ctypes.byref(slots), ctypes.byref(dev)))
        return slots.value, dev.value

    def close(self):
        if getattr(self, "_h", None) and _release_ok():
            lib().trd_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving loop for RDMA requests: submit worlds, poll, take results (include/twins_rdma.h)."""

    def __init__(self, q3cap=0, tail_log2=0, hosts_per_wave=0, max_blocks=0, arena_bytes=0):
        self._h = ctypes.c_void_p()
        _check(lib().trd_server_create(ctypes.byref(ServerCfg(q3cap, tail_log2, hosts_per_wave, max_blocks, arena_bytes)),
                                       ctypes.byref(self._h)))
        self.flows = []
        _track(self, 0)

    def submit(self, world):
        i = ctypes.c_int32()
        _check(lib().trd_server_submit(self._h, world._h, ctypes.byref(i)))
        self.flows.append(world.flows)
        return i.value

    def poll(self):
        st = ServerStats()
        _check(lib().trd_server_poll(self._h, ctypes.byref(st)))
        return st

    def state(self, i):
        v = ctypes.c_int32()
        _check(lib().trd_server_state(self._h, i, ctypes.byref(v)))
        return v.value

    def result(self, i):
        """(start_ns, fct_ns, counters) of a done request; its device memory is released."""
        n = self.flows[i]
        a, b, ev = (ctypes.c_int64 * max(1, n))(), (ctypes.c_int64 * max(1, n))(), (ctypes.c_uint64 * 40)()
        _check(lib().trd_server_result(self._h, i, a, b, ev))
        return list(a)[:n], list(b)[:n], list(ev)

    def close(self):
        if getattr(self, "_h", None) and _release_ok():
            lib().trd_server_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
