"""A sequential model of the device task graph's observable behaviour, for
the two generic task kinds (shader/task-graph/src/runtime: task_pool.rs
states, task_group.rs spawn / wake / poll / cleanup / compaction, mod.rs
execute and read_back_execution_states).

It is pinned by the counts the reference's own task-graph tests assert
(shader/task-graph/src/test.rs) and then serves as the expectation for the
device runtime on other graphs.  Only counts are compared: which slot a
task gets, and the order of the active list, are not observable there.
"""

NOT_EXIST, WAKEN, GO_TO_SLEEP, SLEEP, FINISHED = 0, 1, 2, 3, 4
NONE = None
READY, SPAWN_THEN = 0, 1


class Group:
    def __init__(self, size, kind, child):
        self.size, self.kind, self.child_group = size, kind, child
        self.state = [NOT_EXIST] * (size + 1)
        self.parent = [None] * (size + 1)
        self.stage = [0] * (size + 1)
        self.child = [None] * (size + 1)
        self.payload = [0] * (size + 1)
        self.active = []
        self.empty = list(range(size, 0, -1))  # popped from the end: slot 1 first
        self.removed = []
        self.snapshot = 0
        self.errors = 0


class Model:
    def __init__(self, tasks, capacity):
        """tasks: list of (kind, child_group, max_in_flight)."""
        self.g = [Group(mif * capacity, kind, child) for kind, child, mif in tasks]

    # -- task_group.rs operations
    def _push(self, G, idx):
        if len(G.active) < G.size:
            G.active.append(idx)
        else:
            G.errors |= 2

    def spawn(self, gid, parent, payload):
        G = self.g[gid]
        if not G.empty:
            G.errors |= 1
            return None
        idx = G.empty.pop()
        G.state[idx], G.parent[idx], G.stage[idx], G.child[idx] = WAKEN, parent, 0, None
        G.payload[idx] = payload
        self._push(G, idx)
        return idx

    def wake(self, gid, idx):
        G = self.g[gid]
        was = G.state[idx]
        G.state[idx] = WAKEN
        if was != GO_TO_SLEEP:
            self._push(G, idx)

    def cleanup(self, gid, idx):
        G = self.g[gid]
        if len(G.removed) < G.size:
            G.removed.append(idx)
        else:
            G.errors |= 4
        G.state[idx] = NOT_EXIST

    def end_poll(self, gid, idx, resolved):
        G = self.g[gid]
        if not resolved:
            G.state[idx] = GO_TO_SLEEP
            return
        G.state[idx] = FINISHED
        if G.parent[idx] is None:
            self.cleanup(gid, idx)
        else:
            pg, pi = G.parent[idx]
            self.wake(pg, pi)

    # -- one group's turn
    def prepare(self, gid):
        G = self.g[gid]
        G.empty.extend(G.removed)
        G.removed = []
        G.snapshot = len(G.active)

    def poll(self, gid):
        G = self.g[gid]
        for i in range(G.snapshot):
            idx = G.active[i]
            if G.kind == READY:
                self.end_poll(gid, idx, True)
                continue
            C = self.g[G.child_group]
            if G.stage[idx] == 0:
                c = self.spawn(G.child_group, (gid, idx), G.payload[idx] + 1)
                G.child[idx], G.stage[idx] = c, 1
                self.end_poll(gid, idx, False)
                continue
            c = G.child[idx]
            if c is not None and C.state[c] == FINISHED:
                G.payload[idx] = C.payload[c]
                self.cleanup(G.child_group, c)
                self.end_poll(gid, idx, True)
            else:
                self.end_poll(gid, idx, False)

    def compact(self, gid):
        G = self.g[gid]
        keep = []
        for idx in G.active:
            s = G.state[idx]
            if s == GO_TO_SLEEP:
                G.state[idx] = SLEEP
            if s == WAKEN:
                keep.append(idx)
        G.active = keep

    # -- mod.rs
    def dispatch_allocate_init_task(self, count, gid):
        self.prepare(gid)
        for i in range(count):
            self.spawn(gid, None, i)

    def execute(self, rounds):
        for _ in range(rounds):
            for gid in range(len(self.g)):
                self.prepare(gid)
                self.poll(gid)
                self.compact(gid)

    def states(self):
        wake, sof, empty, errors = [], [], [], []
        for gid, G in enumerate(self.g):
            self.prepare(gid)
            self.compact(gid)
            w, e = len(G.active), len(G.empty)
            wake.append(w)
            empty.append(e)
            sof.append(G.size - min(G.size, w + e))
            errors.append(G.errors)
        return wake, sof, empty, errors
