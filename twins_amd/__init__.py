"""twins_amd: Python access to libtwins_core.so, the MI355X (gfx950) TCP packet-simulation core.

This is a thin ctypes layer over the C ABI in include/twins.h (the same calls the `twins` command
line makes): build worlds from `twins`-style arguments, pack them into one device batch, run it on
the GPU, read per-flow establishment / completion times and the counters. It never simulates
anything itself: without the built library, or without a GPU, every call raises.
"""
import atexit
import ctypes
import os
import weakref

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtwins_core.so")
CLI_PATH = os.path.join(HERE, "bin", "twins")

EV_NAMES = {0: "drops", 1: "marks", 2: "retx", 3: "rto", 4: "tx-cap", 5: "active-ovf", 6: "ooo-ovf", 7: "due-ovf",
            13: "max-inflight", 14: "rtt-ovf", 19: "ring-ovf", 25: "undecided-ties", 26: "packet-events", 31: "slots"}


class TwError(RuntimeError):
    pass


class Plan(ctypes.Structure):
    _fields_ = [("nteams", ctypes.c_int32), ("nqueues", ctypes.c_int32), ("team_blocks", ctypes.POINTER(ctypes.c_int32)),
                ("team_queue", ctypes.POINTER(ctypes.c_int32)), ("world_queue", ctypes.POINTER(ctypes.c_int32)),
                ("world_order", ctypes.POINTER(ctypes.c_int32))]


class PackCfg(ctypes.Structure):
    _fields_ = [("chain_runs", ctypes.c_int32), ("block_threads", ctypes.c_int32), ("blocks", ctypes.c_int32),
                ("team_blocks", ctypes.c_int32), ("hosts_per_wave", ctypes.c_int32), ("no_graph", ctypes.c_int32),
                ("slot_budget", ctypes.c_int64), ("plan", ctypes.POINTER(Plan))]


class RunStats(ctypes.Structure):
    _fields_ = [("wall_s", ctypes.c_double), ("kernel_s", ctypes.c_double), ("launches", ctypes.c_int64),
                ("slots_max", ctypes.c_int64), ("packet_events", ctypes.c_uint64), ("flows_done", ctypes.c_uint64),
                ("worlds_done", ctypes.c_int32), ("graph", ctypes.c_int32)]


class WorldStats(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int64), ("device_s", ctypes.c_double), ("team_blocks", ctypes.c_int32),
                ("finished", ctypes.c_int32), ("packet_events", ctypes.c_uint64)]


class WorldInfo(ctypes.Structure):
    _fields_ = [("hosts", ctypes.c_int32), ("switches", ctypes.c_int32), ("ports_per_switch", ctypes.c_int32),
                ("links", ctypes.c_int32), ("conns", ctypes.c_int32), ("flows", ctypes.c_int32),
                ("device_bytes", ctypes.c_uint64)]


class ServerCfg(ctypes.Structure):
    _fields_ = [("chain_runs", ctypes.c_int32), ("block_threads", ctypes.c_int32), ("small_teams", ctypes.c_int32),
                ("mid_teams", ctypes.c_int32), ("large_teams", ctypes.c_int32), ("max_requests", ctypes.c_int32),
                ("arena_bytes", ctypes.c_uint64), ("idle_s", ctypes.c_double), ("mid_size", ctypes.c_int32),
                ("large_size", ctypes.c_int32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("waiting", ctypes.c_int32), ("resident", ctypes.c_int32), ("done", ctypes.c_int32), ("failed", ctypes.c_int32),
                ("arena_used", ctypes.c_uint64), ("arena_bytes", ctypes.c_uint64), ("relaunches", ctypes.c_uint64)]


REQ_WAITING, REQ_RESIDENT, REQ_DONE, REQ_FAILED, REQ_TAKEN = range(5)

_lib = None

# Live handles, released in order (servers, batches, worlds) by an exit hook while the interpreter
# and the HIP runtime are still whole; after it, finalizers release nothing (a hipFree from a
# finalizer run during interpreter shutdown can meet a runtime already torn down).
_live = {rank: weakref.WeakSet() for rank in range(3)}
_exiting = False


def _track(obj, rank):
    _live[rank].add(obj)


@atexit.register
def _close_all():
    global _exiting
    for rank in range(3):
        for obj in list(_live[rank]):
            try:
                obj.close()
            except Exception:
                pass
    _exiting = True


def _release_ok():
    return _lib is not None and not _exiting


def lib():
    """The loaded core library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TwError(f"{LIB_PATH} is not built: run __graft_entry__.build() or make -C twins_amd")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
        L.tw_last_error.restype = ctypes.c_char_p
        L.tw_world_create_args.argtypes = [i32, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(vp)]
        L.tw_world_info_get.argtypes = [vp, ctypes.POINTER(WorldInfo)]
        L.tw_world_destroy.argtypes = [vp]
        L.tw_batch_create.argtypes = [ctypes.POINTER(vp), i32, ctypes.POINTER(PackCfg), ctypes.POINTER(vp)]
        L.tw_batch_run.argtypes = [vp, ctypes.POINTER(RunStats)]
        L.tw_batch_run_noalloc.argtypes = [vp, ctypes.POINTER(RunStats)]
        L.tw_fct_get.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        L.tw_counters_get.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_uint64)]
        L.tw_world_stats_get.argtypes = [vp, i32, ctypes.POINTER(WorldStats)]
        L.tw_launch_capacity.argtypes = [i32, i32, ctypes.POINTER(ctypes.c_int32)]
        L.tw_server_create.argtypes = [ctypes.POINTER(ServerCfg), ctypes.POINTER(vp)]
        L.tw_server_submit.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_int32)]
        L.tw_server_poll.argtypes = [vp, ctypes.POINTER(ServerStats)]
        L.tw_server_state.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_int32)]
        L.tw_server_result.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                       ctypes.POINTER(ctypes.c_uint64)]
        L.tw_server_destroy.argtypes = [vp]
        L.tw_batch_destroy.argtypes = [vp]
        L.tw_qtrace_enable.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64]
        L.tw_qtrace_write.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise TwError(f"twins core error {rc}: {lib().tw_last_error().decode()}")


class World:
    """One simulation, from `twins` arguments (flow and topology files)."""

    def __init__(self, args):
        L = lib()
        arr = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
        self._h = ctypes.c_void_p()
        _check(L.tw_world_create_args(len(args), arr, ctypes.byref(self._h)))
        self.info = WorldInfo()
        _check(L.tw_world_info_get(self._h, ctypes.byref(self.info)))
        _track(self, 2)

    def close(self):
        if getattr(self, "_h", None) and _release_ok():
            _lib.tw_world_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()


def launch_capacity(chain_runs=9, block_threads=0):
    """Workgroups the GPU holds at once for the slot kernel (a plan's limit)."""
    n = ctypes.c_int32()
    _check(lib().tw_launch_capacity(chain_runs, block_threads, ctypes.byref(n)))
    return n.value


def _i32(xs):
    return (ctypes.c_int32 * len(xs))(*xs)


class Batch:
    """Worlds packed into one device batch and run by the persistent world-queue kernel.

    plan: None (one queue, teams of `team_blocks`, automatic when 0) or a dict with "team_blocks"
    (workgroups per team), "team_queue", "world_queue" and optionally "world_order" (see tw_plan in
    include/twins.h); the batched sweep's autotuner (twins_amd.autotune) makes them."""

    def __init__(self, worlds, chain_runs=9, block_threads=0, blocks=0, team_blocks=0, hosts_per_wave=0, slot_budget=0,
                 plan=None, graph=True):
        L = lib()
        self.worlds = list(worlds)
        hs = (ctypes.c_void_p * len(self.worlds))(*[w._h for w in self.worlds])
        p = None
        if plan is not None:
            keep = [_i32(plan["team_blocks"]), _i32(plan["team_queue"]), _i32(plan["world_queue"])]
            order = plan.get("world_order")
            keep.append(_i32(order) if order is not None else None)
            nq = max(max(plan["team_queue"]), max(plan["world_queue"])) + 1
            p = Plan(len(plan["team_blocks"]), nq, keep[0], keep[1], keep[2],
                     keep[3] if keep[3] is not None else ctypes.POINTER(ctypes.c_int32)())
        cfg = PackCfg(chain_runs, block_threads, blocks, team_blocks, hosts_per_wave, 0 if graph else 1, slot_budget,
                      ctypes.pointer(p) if p is not None else ctypes.POINTER(Plan)())
        self._h = ctypes.c_void_p()
        _check(L.tw_batch_create(hs, len(self.worlds), ctypes.byref(cfg), ctypes.byref(self._h)))
        _track(self, 1)

    def run(self):
        st = RunStats()
        _check(lib().tw_batch_run(self._h, ctypes.byref(st)))
        return st

    def run_noalloc(self):
        """run() through tw_batch_run_noalloc: the launch's scratch in the library's static device
        array, nothing allocated (the batch's launch must fit it: at most about 33 workgroups)."""
        st = RunStats()
        _check(lib().tw_batch_run_noalloc(self._h, ctypes.byref(st)))
        return st

    def fct(self, w):
        n = self.worlds[w].info.flows
        est, done = (ctypes.c_int64 * n)(), (ctypes.c_int64 * n)()
        _check(lib().tw_fct_get(self._h, w, est, done))
        return list(est), list(done)

    def qtrace_enable(self, from_ps=0, until_ps=0, max_records=0):
        """Trace every enqueue at from_ps <= t < until_ps (the reference's --qtrace); before run()."""
        _check(lib().tw_qtrace_enable(self._h, from_ps, until_ps, max_records))

    def qtrace_write(self, path):
        """The trace in the reference's text format; returns the records written."""
        n = ctypes.c_uint64()
        _check(lib().tw_qtrace_write(self._h, str(path).encode(), ctypes.byref(n)))
        return n.value

    def world_stats(self, w):
        st = WorldStats()
        _check(lib().tw_world_stats_get(self._h, w, ctypes.byref(st)))
        return st

    def counters(self, w):
        ev = (ctypes.c_uint64 * 32)()
        _check(lib().tw_counters_get(self._h, w, ev))
        return list(ev)

    def close(self):
        if getattr(self, "_h", None) and _release_ok():
            _lib.tw_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()


class Server:
    """The serving loop (tw_server_* in include/twins.h): one persistent kernel runs requests (one
    world each) as they are submitted; poll() collects completions and admits waiting requests
    (in arrival order within each size class, while their device footprint fits and their class's
    teams have at most 4 resident requests each); result() returns a finished request's flow times
    and counters."""

    def __init__(self, chain_runs=9, block_threads=0, small_teams=0, mid_teams=0, large_teams=0, max_requests=0,
                 arena_bytes=0, idle_s=0.0, mid_size=0, large_size=0):
        L = lib()
        cfg = ServerCfg(chain_runs, block_threads, small_teams, mid_teams, large_teams, max_requests, arena_bytes, idle_s,
                        mid_size, large_size)
        self._h = ctypes.c_void_p()
        self._worlds = {}
        _check(L.tw_server_create(ctypes.byref(cfg), ctypes.byref(self._h)))
        _track(self, 0)

    def submit(self, world):
        i = ctypes.c_int32()
        _check(lib().tw_server_submit(self._h, world._h, ctypes.byref(i)))
        self._worlds[i.value] = world.info.flows
        return i.value

    def poll(self):
        st = ServerStats()
        _check(lib().tw_server_poll(self._h, ctypes.byref(st)))
        return st

    def state(self, rid):
        x = ctypes.c_int32()
        _check(lib().tw_server_state(self._h, rid, ctypes.byref(x)))
        return x.value

    def result(self, rid):
        n = self._worlds[rid]
        est, done, ev = (ctypes.c_int64 * n)(), (ctypes.c_int64 * n)(), (ctypes.c_uint64 * 32)()
        _check(lib().tw_server_result(self._h, rid, est, done, ev))
        return list(est), list(done), list(ev)

    def close(self):
        if getattr(self, "_h", None) and _release_ok():
            _lib.tw_server_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()
