"""Autotuner of the batched sweep's launch plan: which worlds run on teams of how many workgroups,
and how many teams of each size one persistent launch holds.

The reference runs a sweep (`--worlds`) of one topology in one fixed launch shape
(src/main.cu:807-1015); here the worlds of a batch differ in size and length by orders of
magnitude, so the plan matters: a world's run time is its slot count times its per-slot time,
and a larger team shortens the slot (its entities are dealt over more wavefronts) at the cost of
a device-wide barrier per slot and of workgroups that other worlds could use. The makespan of the
launch is the longest queue's.

The tuner learns from the batch's own runs: every run reports each world's slots and the device
time its team spent on it (tw_world_stats_get), and the next plan is solved from a per-world model
fitted to those measurements:

    per-slot time s_w(T) = c_w / min(T, sat_w) + o(T)

with c_w the world's work per slot on one workgroup, sat_w the team size beyond which its work
items (about one per wavefront, with as few hosts or switches each as the team's waves allow)
run out, and o(T) the barrier and ticket overhead (one workgroup's
barrier is cheap; a team's is an agent-scope barrier). c_w is refitted from each new measurement.
The plan is then solved for the least makespan M: each world takes the smallest team size whose
predicted time fits M, worlds of one size share a queue in longest-first order, and each queue
gets the fewest teams whose longest-first list schedule fits M; M is feasible when the teams fit
in the workgroups the GPU holds at once.
"""
import math

TEAM_SIZES = (1, 2, 4, 8, 16, 32, 64)
OVERHEAD_ONE = 1.0e-6   # s per slot: a workgroup's barrier and ticket
OVERHEAD_TEAM = 4.0e-6  # s per slot: a team's device-wide barrier


class WorldShape:
    """What the model needs of a world: entities and work items per slot."""

    def __init__(self, info):
        self.hosts = info.hosts
        self.switches = info.switches
        self.k = max(1, info.ports_per_switch)
        self.flows = info.flows

    def items_at(self, per):
        """Work items of one slot with `per` hosts per host item (and min(per, 64 / k) switches per
        switch item)."""
        sp = min(per, max(1, 64 // self.k))
        return -(-self.hosts // per) + -(-self.switches // sp)

    def items(self, T, waves):
        """Work items of one slot for a team of T workgroups of `waves` wavefronts, as the kernel
        deals them (slot_items in tcp_state.h): the fewest hosts per item whose items do not
        outnumber the team's waves, at most 64."""
        per = 1
        while per < 64 and self.items_at(per) > T * waves:
            per += 1
        return self.items_at(per)

    def saturation(self, waves):
        """Team size beyond which more workgroups add nothing: one host or switch per item."""
        for T in TEAM_SIZES:
            if T * waves >= self.hosts + self.switches:
                return T
        return TEAM_SIZES[-1]


def overhead(T):
    return OVERHEAD_ONE if T == 1 else OVERHEAD_TEAM


class Autotuner:
    """Plans for one batch of worlds, improved by observe() after each run."""

    def __init__(self, infos, resident_blocks, block_threads=256, max_team=64):
        self.shapes = [WorldShape(i) for i in infos]
        self.resident = resident_blocks
        self.waves = max(1, block_threads // 64)
        self.sizes = tuple(T for T in TEAM_SIZES if T <= max_team)
        self.slots = [None] * len(infos)  # measured slot counts
        self.work = [None] * len(infos)   # c_w: per-slot work on one workgroup, s
        self.history = []                 # (plan, measured makespan)

    # ------------------------------------------------------------------ the model
    def predict(self, w, T):
        """Predicted device time of world w on a team of T workgroups (None: not measured yet)."""
        if self.slots[w] is None:
            return None
        sat = self.shapes[w].saturation(self.waves)
        return self.slots[w] * (self.work[w] / min(T, sat) + overhead(T))

    def observe(self, batch):
        """Reads each world's slots and device time from a finished run."""
        makespan = 0.0
        for w in range(len(self.shapes)):
            st = batch.world_stats(w)
            if not st.finished or st.slots <= 0:
                continue
            T = max(1, st.team_blocks)
            sat = self.shapes[w].saturation(self.waves)
            per_slot = st.device_s / st.slots
            self.slots[w] = st.slots
            self.work[w] = max(1e-8, (per_slot - overhead(T)) * min(T, sat))
            makespan = max(makespan, st.device_s)
        return makespan

    # ------------------------------------------------------------------ plans
    def initial_plan(self):
        """Before any measurement: a workgroup per world, largest worlds first (the library's own
        choice for a batch of at least as many worlds as CUs), as an explicit plan."""
        W = len(self.shapes)
        order = sorted(range(W), key=lambda w: -(self.shapes[w].hosts + self.shapes[w].switches))
        teams = min(W, self.resident)
        return {"team_blocks": [1] * teams, "team_queue": [0] * teams, "world_queue": [0] * W, "world_order": order}

    def _schedule(self, times, n):
        """Longest-first list schedule of `times` on n teams: the makespan."""
        loads = [0.0] * n
        for t in sorted(times, reverse=True):
            i = min(range(n), key=loads.__getitem__)
            loads[i] += t
        return max(loads) if loads else 0.0

    def _fit(self, M):
        """The plan for makespan M, or None."""
        W = len(self.shapes)
        pick = {}
        for w in range(W):
            for T in self.sizes:
                p = self.predict(w, T)
                if p <= M:
                    pick[w] = (T, p)
                    break
            else:
                return None
        blocks, classes = 0, {}
        for w, (T, p) in pick.items():
            classes.setdefault(T, []).append((p, w))
        teams = {}
        for T, ws in classes.items():
            times = [p for p, _ in ws]
            n = max(1, min(len(ws), math.ceil(sum(times) / M)))
            while n < len(ws) and self._schedule(times, n) > M:
                n += 1
            teams[T] = n
            blocks += n * T
            if blocks > self.resident:
                return None
        return classes, teams

    def plan(self):
        """The next plan: solved from the model once every world has been measured."""
        if any(s is None for s in self.slots):
            return self.initial_plan()
        W = len(self.shapes)
        lo = max(min(self.predict(w, T) for T in self.sizes) for w in range(W))
        hi = max(lo, sum(self.predict(w, 1) for w in range(W)) / max(1, self.resident)) * 4 + lo
        while self._fit(hi) is None:
            hi *= 2
        for _ in range(40):
            mid = (lo + hi) / 2
            if self._fit(mid) is None:
                lo = mid
            else:
                hi = mid
        classes, teams = self._fit(hi)
        team_blocks, team_queue, world_queue, order = [], [], [0] * W, []
        for q, T in enumerate(sorted(classes)):
            team_blocks += [T] * teams[T]
            team_queue += [q] * teams[T]
            for p, w in sorted(classes[T], reverse=True):
                world_queue[w] = q
                order.append(w)
        return {"team_blocks": team_blocks, "team_queue": team_queue, "world_queue": world_queue, "world_order": order,
                "predicted_makespan_s": hi}

    def best(self):
        """The measured best plan so far."""
        return min(self.history, key=lambda x: x[1])[0] if self.history else None

    def record(self, plan, makespan):
        self.history.append((plan, makespan))


def describe(plan):
    """Team sizes and counts of a plan, e.g. {1: 380, 4: 20}."""
    out = {}
    for T in plan["team_blocks"]:
        out[T] = out.get(T, 0) + 1
    return dict(sorted(out.items()))
