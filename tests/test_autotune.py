"""The batched sweep's plan solver (twins_amd.autotune), on synthetic measurements: every plan it
returns is valid for the library (teams fit in the resident workgroups, every world sits in a
queue that has teams, the order is a permutation), and its makespan model behaves as the
scheduling arguments say it must."""
import random
import types

import pytest

from twins_amd import autotune


def info(hosts, switches, k, flows=1000):
    return types.SimpleNamespace(hosts=hosts, switches=switches, ports_per_switch=k, flows=flows)


class FakeBatch:
    """world_stats() of a finished run: slots and device time per world, from a hidden truth."""

    def __init__(self, truth, plan):
        self.truth, self.team = truth, {}
        sizes = {}
        for T, q in zip(plan["team_blocks"], plan["team_queue"]):
            sizes[q] = T
        for w, q in enumerate(plan["world_queue"]):
            self.team[w] = sizes[q]

    def world_stats(self, w):
        slots, c, sat = self.truth[w]
        T = self.team[w]
        t = slots * (c / min(T, sat) + autotune.overhead(T))
        return types.SimpleNamespace(slots=slots, device_s=t, team_blocks=T, finished=1, packet_events=slots * 10)


def check_plan(plan, W, resident):
    assert sorted(plan["world_order"]) == list(range(W))
    queues_with_teams = set(plan["team_queue"])
    assert all(q in queues_with_teams for q in plan["world_queue"])
    if any(T > 1 for T in plan["team_blocks"]):
        assert sum(plan["team_blocks"]) <= resident
    # a queue's teams all have one size (the library picks hosts per wave per team)
    size = {}
    for T, q in zip(plan["team_blocks"], plan["team_queue"]):
        assert size.setdefault(q, T) == T


def world_set(n, seed):
    rng = random.Random(seed)
    infos, truth = [], []
    for _ in range(n):
        kind = rng.choice(["small", "mid", "big"])
        h, s, k = {"small": (16, 20, 4), "mid": (128, 80, 8), "big": (1024, 320, 16)}[kind]
        infos.append(info(h, s, k))
        slots = rng.randint(500, 45000)
        truth.append((slots, {"small": 2e-6, "mid": 8e-6, "big": 60e-6}[kind], autotune.WorldShape(infos[-1]).saturation(4)))
    return infos, truth


def test_initial_plan_is_one_workgroup_per_world():
    infos, _ = world_set(40, 1)
    t = autotune.Autotuner(infos, resident_blocks=512)
    p = t.initial_plan()
    check_plan(p, 40, 512)
    assert p["team_blocks"] == [1] * 40
    # largest worlds first
    sizes = [infos[w].hosts + infos[w].switches for w in p["world_order"]]
    assert sizes == sorted(sizes, reverse=True)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_solved_plan_is_valid_and_not_worse_than_one_workgroup_each(seed):
    W, R = 256, 512
    infos, truth = world_set(W, seed)
    t = autotune.Autotuner(infos, resident_blocks=R)
    p0 = t.initial_plan()
    b0 = FakeBatch(truth, p0)
    m0 = t.observe(b0)
    p1 = t.plan()
    check_plan(p1, W, R)
    # the model is exact for this synthetic truth, so the measured makespan of the new plan is the
    # longest queue's longest-first schedule, which the solver bounded by its predicted makespan
    b1 = FakeBatch(truth, p1)
    loads = {}
    per_team = {}
    for T, q in zip(p1["team_blocks"], p1["team_queue"]):
        per_team.setdefault(q, 0)
        per_team[q] += 1
    for q, n in per_team.items():
        times = sorted((b1.world_stats(w).device_s for w in range(W) if p1["world_queue"][w] == q), reverse=True)
        lanes = [0.0] * n
        for x in times:
            i = lanes.index(min(lanes))
            lanes[i] += x
        loads[q] = max(lanes)
    makespan = max(loads.values())
    assert makespan <= p1["predicted_makespan_s"] * (1 + 1e-9)
    # one workgroup per world on min(W, R) teams: its makespan is at least the longest world's
    assert makespan <= m0 * (1 + 1e-9)
    # the longest world alone bounds any plan from below
    lower = max(min(t.predict(w, T) for T in t.sizes) for w in range(W))
    assert makespan >= lower * (1 - 1e-9)


def test_saturation_counts_work_items():
    s = autotune.WorldShape(info(16, 20, 4))
    # k=4 FatTree on one workgroup of 4 waves: the fewest hosts per item whose items fit the 4
    # waves is 10 (2 host items, 2 switch items of 10 switches)
    assert s.items(1, 4) == 2 + 2
    assert s.items_at(9) == 2 + 3
    # 36 entities: one per wave from 9 workgroups of 4 waves (team sizes are powers of two: 16)
    assert s.saturation(4) == 16
    big = autotune.WorldShape(info(1024, 320, 16))
    # one workgroup: 64 hosts per item (16 items), 4 switches of 16 ports per item (80 items)
    assert big.items(1, 4) == 16 + 80
    assert big.saturation(4) == 64
