"""The device task graph (shader/task-graph/src/runtime) against the
reference's own task-graph tests (shader/task-graph/src/test.rs), whose
asserted counts are the golden data here, and against a sequential model of
the same semantics (tests/task_graph_model.py) on further graphs.

CPU tests pin the model on the reference's counts; GPU tests run the device
runtime through the C-ABI on the same sequences."""

import numpy as np
import pytest

from tests.task_graph_model import READY, SPAWN_THEN, Model


# The three tests of shader/task-graph/src/test.rs as scripts: graph
# (kind, child, max_in_flight) per task, capacity, and steps with the
# counts the reference asserts after them.
def simple_map():
    # test_simple_map (test.rs:1-29)
    tasks = [(READY, 0, 2), (READY, 0, 2)]
    steps = [("spawn", 3, 0), ("spawn", 4, 1), ("check", {"wake": [3, 4]}),
             ("execute", 1), ("check", {"wake": [0, 0]})]
    return tasks, 12, steps


def then_task_spawn():
    # test_task_graph_then_task_spawn (test.rs:31-128); the round runs twice
    tasks = [(READY, 0, 2), (SPAWN_THEN, 0, 2)]
    rnd = [("spawn", 3, 1), ("check", {"wake": [0, 3]}),
           ("execute", 1), ("check", {"wake": [3, 0], "sof": [0, 3]}),
           ("execute", 1), ("check", {"wake": [0, 0], "sof": [0, 0]})]
    return tasks, 4, rnd + rnd


def self_spawn_recursive():
    # test_task_graph_then_task_self_spawn_recursive (test.rs:130-198)
    tasks = [(SPAWN_THEN, 0, 4)]
    steps = [("spawn", 3, 0), ("check", {"wake": [3]}),
             ("execute", 1), ("check", {"wake": [3], "sof": [3]}),
             ("execute", 1), ("check", {"wake": [3], "sof": [6]}),
             ("execute", 1), ("check", {"wake": [3], "sof": [9]})]
    return tasks, 4, steps


REFERENCE_SCRIPTS = {"simple_map": simple_map, "then_task_spawn": then_task_spawn,
                     "self_spawn_recursive": self_spawn_recursive}


def run_model(tasks, capacity, steps):
    m = Model(tasks, capacity)
    out = []
    for st in steps:
        if st[0] == "spawn":
            m.dispatch_allocate_init_task(st[1], st[2])
        elif st[0] == "execute":
            m.execute(st[1])
        else:
            out.append(m.states())
    return out


@pytest.mark.parametrize("name", sorted(REFERENCE_SCRIPTS))
def test_model_matches_reference_counts(name):
    tasks, cap, steps = REFERENCE_SCRIPTS[name]()
    got = iter(run_model(tasks, cap, steps))
    for st in steps:
        if st[0] != "check":
            continue
        wake, sof, empty, errors = next(got)
        if "wake" in st[1]:
            assert wake == st[1]["wake"], (name, st)
        if "sof" in st[1]:
            assert sof == st[1]["sof"], (name, st)
        assert all(e == 0 for e in errors)


def test_model_chain_depth_and_conservation():
    """A chain of k SPAWN_THEN groups ending in a READY group, each group
    spawning into the next: the way down takes one round (each group polls
    after its parent has spawned into it), the way up one round per level,
    so the graph drains in k + 1 rounds with every slot free again; slots
    are conserved throughout."""
    for k in range(1, 5):
        tasks = [(SPAWN_THEN, i + 1, 1) for i in range(k)] + [(READY, 0, 1)]
        m = Model(tasks, 8)
        m.dispatch_allocate_init_task(5, 0)
        rounds = 0
        while True:
            wake, sof, empty, errors = m.states()
            assert all(w + s + e == g.size for w, s, e, g in zip(wake, sof, empty, m.g))
            if all(w == 0 for w in wake) and all(s == 0 for s in sof):
                break
            m.execute(1)
            rounds += 1
            assert rounds < 50
        assert rounds == k + 1, (k, rounds)


def test_model_pool_exhaustion_sets_error():
    m = Model([(SPAWN_THEN, 0, 1)], 4)  # 4 slots, recursive self spawn
    m.dispatch_allocate_init_task(3, 0)
    m.execute(1)
    wake, sof, empty, errors = m.states()
    assert errors == [1] and wake == [1] and sof == [3]


# ------------------------------------------------------------------ GPU --
def _device_system():
    from rendiation_amd.acceleration import AccelerationStructureSystem

    return AccelerationStructureSystem(device=0)


def run_device(system, tasks, capacity, steps):
    from rendiation_amd.task_graph import DeviceTaskGraphBuildSource

    src = DeviceTaskGraphBuildSource()
    for kind, child, mif in tasks:
        src.define_task(kind, mif, child)
    src.capacity = capacity
    exe = src.build(system)
    out = []
    for st in steps:
        if st[0] == "spawn":
            exe.dispatch_allocate_init_task(st[1], st[2])
        elif st[0] == "execute":
            exe.execute(st[1])
        else:
            s = exe.read_back_execution_states()
            out.append((s.wake_counts, s.sleep_or_finished_counts, s.empty_counts, s.errors))
    exe.close()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(REFERENCE_SCRIPTS))
def test_device_matches_reference_counts(name, device):
    tasks, cap, steps = REFERENCE_SCRIPTS[name]()
    sysm = _device_system()
    got = run_device(sysm, tasks, cap, steps)
    want = run_model(tasks, cap, steps)
    assert [tuple(map(list, g)) for g in got] == [tuple(map(list, w)) for w in want]
    sysm.close()


def _random_script(rng):
    n = int(rng.integers(1, 6))
    tasks = []
    for i in range(n):
        kind = int(rng.integers(0, 2))
        tasks.append((kind, int(rng.integers(0, n)), int(rng.integers(1, 4))))
    cap = int(rng.integers(1, 300))
    steps = []
    for _ in range(int(rng.integers(2, 7))):
        if rng.random() < 0.5:
            steps.append(("spawn", int(rng.integers(0, cap + 1)), int(rng.integers(0, n))))
        steps.append(("execute", int(rng.integers(1, 4))))
        steps.append(("check", {}))
    return tasks, cap, steps


@pytest.mark.gpu
def test_device_matches_model_random_graphs(device):
    """Random graphs (chains, self recursion, pool exhaustion included): the
    device counts and error bits equal the model's after every step."""
    rng = np.random.default_rng(17)
    sysm = _device_system()
    for case in range(60):
        tasks, cap, steps = _random_script(rng)
        print(f"case {case}: tasks={tasks} capacity={cap}", flush=True)
        got = run_device(sysm, tasks, cap, steps)
        want = run_model(tasks, cap, steps)
        assert [tuple(map(list, g)) for g in got] == [tuple(map(list, w)) for w in want], (case, tasks, cap, steps)
    sysm.close()


@pytest.mark.gpu
def test_device_large_chain(device):
    """200k root tasks through a three-level chain: the graph drains in the
    model's round count with no errors."""
    tasks = [(SPAWN_THEN, 1, 1), (SPAWN_THEN, 2, 1), (READY, 0, 1)]
    steps = [("spawn", 200_000, 0)] + [x for _ in range(6) for x in (("execute", 1), ("check", {}))]
    sysm = _device_system()
    got = run_device(sysm, tasks, 200_000, steps)
    want = run_model(tasks, 200_000, steps)
    assert [tuple(map(list, g)) for g in got] == [tuple(map(list, w)) for w in want]
    assert got[-1][0] == [0, 0, 0] and got[-1][1] == [0, 0, 0]
    sysm.close()
