"""The ambient-occlusion chain (scene/rendering/gpu-ray-tracing/src/feature/
ao.rs:192-304) on two schedules of the task-graph model:

* the executor's rounds (shader/task-graph/src/runtime/mod.rs:396-460): every
  group in the order trace, ray gen, closest hit, occlusion closest hit, miss,
  each prepared, polled and compacted in turn;
* the schedule of the device AO path: the four stage groups polled together
  as one phase, the trace group as the other, no prepare or compaction
  between them (a pending task sleeps at once), no trace poll in the first
  round, the trace queue filled at spawn;
* that schedule with rounds 4 and 5 run as one pass over the chains still
  open (the device path's k_finish_chains).

For random camera / occlusion hit patterns both must give every pixel the
same occlusion value after the same number of rounds, and the same task
states (awake, asleep or finished, free) after every round count.  Runs on
the CPU; the device path is checked against the oracle image in test_ao.py.
"""

import random

import pytest

from task_graph_model import FINISHED, GO_TO_SLEEP, NOT_EXIST, SLEEP, WAKEN

TRACE, RAYGEN, CLOSEST, CLOSEST_TEST, MISS = range(5)


class AoChain:
    """Task pools and polls of the AO chain.  camera_hit[p] / occ_hit[p]:
    whether pixel p's camera ray and occlusion ray hit."""

    def __init__(self, camera_hit, occ_hit):
        n = len(camera_hit)
        self.camera_hit, self.occ_hit = camera_hit, occ_hit
        sizes = [2 * n, n, n, n, n]
        self.state = [[NOT_EXIST] * (s + 1) for s in sizes]
        self.parent = [[None] * (s + 1) for s in sizes]
        self.stage = [[0] * (s + 1) for s in sizes]
        self.child = [[None] * (s + 1) for s in sizes]
        self.pay = [[None] * (s + 1) for s in sizes]
        self.size = sizes
        self.empty = [list(range(s, 0, -1)) for s in sizes]
        self.removed = [[] for _ in sizes]
        self.active = [[] for _ in sizes]
        self.snapshot = [0] * 5
        self.value = [None] * n  # the pixel's payload when its ray gen resolves
        for p in range(n):
            self.spawn(RAYGEN, None, ("px", p))

    # task_group.rs
    def spawn(self, g, parent, payload):
        idx = self.empty[g].pop()
        self.state[g][idx], self.parent[g][idx], self.stage[g][idx], self.child[g][idx] = WAKEN, parent, 0, None
        self.pay[g][idx] = payload
        self.active[g].append(idx)
        return idx

    def wake(self, g, idx):
        was = self.state[g][idx]
        self.state[g][idx] = WAKEN
        if was != GO_TO_SLEEP:
            self.active[g].append(idx)

    def free(self, g, idx):
        self.removed[g].append(idx)
        self.state[g][idx] = NOT_EXIST

    def end(self, g, idx, resolved, phased):
        if not resolved:
            self.state[g][idx] = SLEEP if phased else GO_TO_SLEEP
            return
        self.state[g][idx] = FINISHED
        if self.parent[g][idx] is None:
            self.free(g, idx)
        else:
            self.wake(*self.parent[g][idx])

    # one poll of task idx of group g (ao.rs semantics)
    def poll_one(self, g, idx, phased):
        if g == TRACE:
            kind, px = self.pay[g][idx][0], self.pay[g][idx][1]
            if self.stage[g][idx] == 0:  # traced now: the hit group's task or the miss task
                hit = self.camera_hit[px] if kind == "camera" else self.occ_hit[px]
                cg = (CLOSEST if kind == "camera" else CLOSEST_TEST) if hit else MISS
                c = self.spawn(cg, (TRACE, idx), ("value", px))
                self.stage[g][idx], self.child[g][idx] = 1, (cg, c)
                self.end(g, idx, False, phased)
                return
            cg, c = self.child[g][idx]
            done = self.state[cg][c] == FINISHED
            if done:
                self.pay[g][idx] = (kind, px, self.pay[cg][c][2])
                self.free(cg, c)
            self.end(g, idx, done, phased)
            return
        px = self.pay[g][idx][1]
        if g in (CLOSEST_TEST, MISS):  # store 0 / 1 and resolve
            self.pay[g][idx] = ("value", px, 0 if g == CLOSEST_TEST else 1)
            self.end(g, idx, True, phased)
            return
        if self.stage[g][idx] == 0:  # ray gen: camera ray; closest hit: occlusion ray
            kind = "camera" if g == RAYGEN else "occlusion"
            c = self.spawn(TRACE, (g, idx), (kind, px, 0 if g == RAYGEN else 1))
            self.stage[g][idx], self.child[g][idx] = 1, c
            self.end(g, idx, False, phased)
            return
        c = self.child[g][idx]
        done = self.state[TRACE][c] == FINISHED
        if done:
            v = self.pay[TRACE][c][2]
            self.free(TRACE, c)
            if g == RAYGEN:
                self.value[px] = v
            else:
                self.pay[g][idx] = ("value", px, v)
        self.end(g, idx, done, phased)

    # the executor's turn of one group
    def prepare(self, g):
        self.empty[g].extend(self.removed[g])
        self.removed[g] = []
        self.snapshot[g] = len(self.active[g])

    def compact(self, g):
        keep = []
        for idx in self.active[g]:
            s = self.state[g][idx]
            if s == GO_TO_SLEEP:
                self.state[g][idx] = SLEEP
            if s == WAKEN:
                keep.append(idx)
        self.active[g] = keep

    def round_sequential(self):
        for g in range(5):
            self.prepare(g)
            for i in range(self.snapshot[g]):
                self.poll_one(g, self.active[g][i], False)
            self.compact(g)

    def phase(self, groups):
        """Polls `groups` as one phase: the lists are fixed while it runs
        (no task of these groups is spawned or woken by their polls), the
        groups are interleaved task by task, and each list ends empty."""
        work = [(g, idx) for g in groups for idx in list(self.active[g])]
        random.Random(len(work)).shuffle(work)  # any order among them
        for g in groups:
            self.active[g] = []
        for g, idx in work:
            self.poll_one(g, idx, True)

    def round_phased(self, first):
        if not first:
            self.phase([TRACE])
        self.phase([RAYGEN, CLOSEST, CLOSEST_TEST, MISS])

    def finish_chains(self):
        """Rounds 4 and 5 in one pass (the device path's k_finish_chains):
        every awake task is an occlusion trace whose closest-hit-test / miss
        child has finished; its pixel's chain takes the four remaining polls
        in turn (occlusion trace, closest hit, camera trace, ray gen), each
        woken task polled at once by the same pass."""
        traces = list(self.active[TRACE])
        assert all(not self.active[g] for g in range(1, 5))
        self.active[TRACE] = []
        for idx in traces:
            g, t = TRACE, idx
            for expect in (TRACE, CLOSEST, TRACE, RAYGEN):
                assert g == expect
                self.poll_one(g, t, True)
                if g == RAYGEN:
                    assert self.state[g][t] == NOT_EXIST
                    break
                assert self.state[g][t] == FINISHED
                g, t = self.parent[g][t]
                assert self.active[g] == [t]  # the one task it woke
                self.active[g] = []

    def states(self):
        out = []
        for g in range(5):
            self.prepare(g)
            self.compact(g)
            w, e = len(self.active[g]), len(self.empty[g])
            out.append((w, self.size[g] - min(self.size[g], w + e), e))
        return out


@pytest.mark.parametrize("seed", range(6))
def test_phases_equal_executor_rounds(seed):
    rng = random.Random(seed)
    n = rng.randrange(1, 300)
    camera_hit = [rng.random() < 0.4 for _ in range(n)]
    occ_hit = [rng.random() < 0.5 for _ in range(n)]
    for rounds in range(1, 7):
        a, b = AoChain(camera_hit, occ_hit), AoChain(camera_hit, occ_hit)
        for r in range(rounds):
            a.round_sequential()
            b.round_phased(r == 0)
        assert a.value == b.value, rounds
        assert a.states() == b.states(), rounds
    # five rounds finish every pixel: a miss is 1, a hit the occlusion test's
    # complement; and the graph is drained
    c = AoChain(camera_hit, occ_hit)
    for r in range(5):
        c.round_phased(r == 0)
    assert c.value == [0 if (ch and oh) else 1 for ch, oh in zip(camera_hit, occ_hit)]
    assert all(w == 0 and s == 0 for w, s, _ in c.states())


@pytest.mark.parametrize("seed", range(6))
def test_fused_last_two_rounds_equal_executor_rounds(seed):
    rng = random.Random(100 + seed)
    n = rng.randrange(1, 300)
    camera_hit = [rng.random() < 0.6 for _ in range(n)]
    occ_hit = [rng.random() < 0.5 for _ in range(n)]
    a, b = AoChain(camera_hit, occ_hit), AoChain(camera_hit, occ_hit)
    for r in range(5):
        a.round_sequential()
    for r in range(3):
        b.round_phased(r == 0)
    b.finish_chains()
    assert a.value == b.value
    assert a.states() == b.states()
    assert all(w == 0 and s == 0 for w, s, _ in b.states())


def test_four_rounds_leave_hit_pixels_unfinished():
    camera_hit = [True, False, True, False]
    occ_hit = [True, True, False, False]
    c = AoChain(camera_hit, occ_hit)
    for r in range(4):
        c.round_phased(r == 0)
    # camera misses resolve in round 3; camera hits need round 5
    assert c.value == [None, 1, None, 1]
