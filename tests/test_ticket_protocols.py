"""Host-side model of the GPU ticket protocols, run under many random
interleavings (no GPU).

The reference synchronises its tree with one spinning grid barrier on a
module-global counter (kernelQR.h:297, :324-327, :360, :367-375), which needs
every block co-resident and is not reentrant.  This library replaces it with
last-arriver tickets: a workgroup writes its result, issues an agent-scope
release fence and a relaxed fetch_add on a ticket, and only the workgroup
that draws the last ticket issues an acquire fence and goes on to read the
others' results.  No workgroup ever waits.  Three protocols use this:

  - the tree walk (tree_up_body in tsqr_kernels.hpp, and the fused walk of
    the superleaf R kernel): a child stores its R into its parent's block,
    the last child factors the parent and climbs;
  - the streaming Gram passes (stream_finish / last_group in gram.hpp): every
    workgroup stores its partial sum, the last of each group of G adds the
    group's partials in index order;
  - the final sum of the group partials in the next launch (tsqr_gram_flat),
    in index order.

The model runs each workgroup as a sequence of atomic steps, picks the next
step from a random runnable workgroup, and tracks happens-before with vector
clocks under the C++ / HIP rules the kernels rely on: a release fence
followed by a read-modify-write on the ticket, and an acquire fence after the
RMW that read it, order the writer's stores before the reader's loads (RMWs
on one location always read the latest value, so every earlier ticket holder
is covered).  The checks: exactly one workgroup continues from each node or
group, it reads only stores that happen before its reads, the root is reached
once, and the sums do not depend on the interleaving.  A model with the
acquire fence left out must fail the visibility check (a planted mistake).
"""
import random

import pytest

# ---------------------------------------------------------------------------
# vector clocks
# ---------------------------------------------------------------------------


def vc_join(a, b):
    return {k: max(a.get(k, 0), b.get(k, 0)) for k in set(a) | set(b)}


def vc_leq(a, b):
    return all(v <= b.get(k, 0) for k, v in a.items())


class Memory:
    """Stores carry the writer's clock; a load checks that the store
    happens-before it."""

    def __init__(self):
        self.cells = {}
        self.tickets = {}  # location -> (value, clock of the release sequence)
        self.violations = []

    def store(self, actor, key, value):
        actor.tick()
        self.cells[key] = (value, dict(actor.clock))

    def load(self, actor, key):
        if key not in self.cells:
            self.violations.append(f"{actor.name} read {key} before any store")
            return None
        value, wclock = self.cells[key]
        if not vc_leq(wclock, actor.clock):
            self.violations.append(f"{actor.name} read {key} without happens-before")
        return value

    def fetch_add(self, actor, key, released):
        """Relaxed RMW; `released` when a release fence precedes it."""
        actor.tick()
        value, clock = self.tickets.get(key, (0, {}))
        new_clock = vc_join(clock, actor.clock) if released else clock
        self.tickets[key] = (value + 1, new_clock)
        actor.last_rmw_clock = clock  # what an acquire fence after this RMW synchronises with
        return value


class Actor:
    """One workgroup: a generator of steps."""

    def __init__(self, name, body):
        self.name = name
        self.clock = {name: 0}
        self.body = body(self)
        self.last_rmw_clock = {}
        self.done = False

    def tick(self):
        self.clock[self.name] = self.clock.get(self.name, 0) + 1

    def acquire(self):
        self.clock = vc_join(self.clock, self.last_rmw_clock)

    def step(self):
        try:
            next(self.body)
        except StopIteration:
            self.done = True


def run_random(actors, rng):
    live = list(actors)
    while live:
        a = rng.choice(live)
        a.step()
        if a.done:
            live.remove(a)


# ---------------------------------------------------------------------------
# the tree (plan.hpp make_tree restated) and the walk
# ---------------------------------------------------------------------------


def make_tree(m, n, H, leaf_rows, arity_cap=0):
    """Levels of the reduction tree as plan.hpp builds them: level 0 has
    tiles of leaf_rows rows, level l > 0 blocks of arity * n rows (arity =
    H / n, or a tuned cap below it), each child contributing min(rows, n)
    rows; stops at the first level with one tile.  Returns [(rows, tiles,
    tile_rows)] per level and the arity."""
    arity = H // n
    if 2 <= arity_cap < arity:
        arity = arity_cap
    assert arity >= 2
    levels = []
    rows, tile_rows = m, leaf_rows
    while True:
        tiles = -(-rows // tile_rows)
        levels.append((rows, tiles, tile_rows))
        if tiles == 1:
            break
        last = rows - (tiles - 1) * tile_rows
        rows = (tiles - 1) * n + min(last, n)
        tile_rows = arity * n
    return levels, arity


def tree_walk(levels, arity, start_level, mem, acquire=True, release=True):
    """Actors for the walk that starts at `start_level` (1: tsqr_tree_up, one
    workgroup per level-1 block; 0: the fused walk of the superleaf R kernel,
    one per superleaf).  Node (l, t) reads its children's R from its block,
    'factors' (its value is the tuple of the children's values), stores its
    R into the parent's slot, releases, takes the parent's ticket; the last
    child acquires and becomes the parent."""
    nodes_done = []
    root = []

    def body(tile):
        def gen(actor):
            level, t = start_level, tile
            while True:
                lv_rows, lv_tiles, _ = levels[level]
                if level == start_level:
                    # the workgroup's own data: a superleaf, or a level-1
                    # block the previous launch completed
                    value = mem.load(actor, ("start", t))
                    yield
                else:
                    # children of node t on the level below
                    below_tiles = levels[level - 1][1]
                    kids = range(t * arity, min((t + 1) * arity, below_tiles))
                    value = tuple(mem.load(actor, ("slot", level, t, k - t * arity)) for k in kids)
                    yield
                nodes_done.append((level, t))
                if lv_tiles == 1:
                    root.append(value)
                    return
                parent, slot = t // arity, t % arity
                mem.store(actor, ("slot", level + 1, parent, slot), value)
                yield
                old = mem.fetch_add(actor, ("ticket", level + 1, parent), released=release)
                yield
                nchild = min(arity, lv_tiles - parent * arity)
                if old != nchild - 1:
                    return
                if acquire:
                    actor.acquire()
                level, t = level + 1, parent

        return gen

    for t in range(levels[start_level][1]):  # written before the launch: visible to every workgroup
        mem.cells[("start", t)] = (("start", start_level, t), {})
    actors = [Actor(f"wg{t}", body(t)) for t in range(levels[start_level][1])]
    return actors, nodes_done, root


def expected_root(levels, arity, start_level):
    """The root's value: nested tuples of the start-level data, in child
    order."""

    def value(level, t):
        if level == start_level:
            return ("start", start_level, t)
        below = levels[level - 1][1]
        return tuple(value(level - 1, k) for k in range(t * arity, min((t + 1) * arity, below)))

    return value(len(levels) - 1, 0)


# (m, n, H, leaf_rows, arity cap): ragged last tiles, arity 2 (n = H / 2),
# wide trees, the level shapes of real calls (FP32 n = 32: H = 256, arity
# 8), and the tuned narrower trees (cap 2, 3, 4 below H / n)
TREES = [
    (27648, 32, 256, 256, 0),
    (100000, 32, 256, 512, 0),
    (600000, 32, 256, 1024, 0),
    (5000, 7, 512, 512, 0),
    (70001, 20, 256, 256, 0),
    (9000, 128, 256, 256, 0),
    (1 << 20, 16, 512, 4096, 0),
    (100000, 32, 256, 512, 2),
    (70001, 8, 256, 256, 3),
    (600000, 32, 256, 1024, 4),
]


@pytest.mark.parametrize("start_level", [0, 1])
@pytest.mark.parametrize("m,n,H,leaf_rows,cap", TREES)
def test_tree_walk_under_random_interleavings(m, n, H, leaf_rows, cap, start_level):
    levels, arity = make_tree(m, n, H, leaf_rows, cap)
    if len(levels) <= start_level:
        pytest.skip("single-level tree")
    want = expected_root(levels, arity, start_level)
    for seed in range(12):
        mem = Memory()
        actors, nodes_done, root = tree_walk(levels, arity, start_level, mem)
        run_random(actors, random.Random(seed))
        assert not mem.violations, mem.violations[:3]
        # every node at or above the start level is factored exactly once
        want_nodes = [(l, t) for l in range(start_level, len(levels)) for t in range(levels[l][1])]
        assert sorted(nodes_done) == sorted(want_nodes)
        assert len(root) == 1 and root[0] == want


def test_tree_walk_without_acquire_is_caught():
    """Planted mistake: the last child reads its siblings' R without the
    acquire fence; some interleaving must expose a read that is not ordered
    after the sibling's store."""
    levels, arity = make_tree(100000, 32, 256, 512)
    caught = 0
    for seed in range(20):
        mem = Memory()
        actors, _, _ = tree_walk(levels, arity, 1, mem, acquire=False)
        run_random(actors, random.Random(seed))
        caught += bool(mem.violations)
    assert caught > 0


def test_tree_walk_without_release_is_caught():
    levels, arity = make_tree(100000, 32, 256, 512)
    caught = 0
    for seed in range(20):
        mem = Memory()
        actors, _, _ = tree_walk(levels, arity, 1, mem, release=False)
        run_random(actors, random.Random(seed))
        caught += bool(mem.violations)
    assert caught > 0


# ---------------------------------------------------------------------------
# streaming Gram partial sums: group tickets, then the ordered final sum
# ---------------------------------------------------------------------------


def stream_group_size(P):  # plan.hpp
    return 8 if P <= 64 else (16 if P <= 256 else 32)


def stream_groups(m, cap=256):  # plan.hpp: workgroups of the streaming Gram for m rows
    chunks = (m + 63) // 64
    p = (chunks + 15) // 16
    if p < 256:
        p = min((chunks + 3) // 4, 256)
    return max(1, min(p, cap))  # cap: the tuned largest grid, 256 .. 1024


def stream_pass(P, partial_values, mem, acquire=True):
    G = stream_group_size(P)
    group_sums = {}

    def body(b):
        def gen(actor):
            mem.store(actor, ("partial", b), partial_values[b])
            yield
            g, g0 = b // G, (b // G) * G
            members = min(P - g0, G)
            old = mem.fetch_add(actor, ("tk", g), released=True)
            yield
            if old != members - 1:
                return
            if acquire:
                actor.acquire()
            s = 0.0
            for k in range(g0, g0 + members):  # index order
                s += mem.load(actor, ("partial", k))
                yield
            group_sums[g] = s

        return gen

    return [Actor(f"wg{b}", body(b)) for b in range(P)], group_sums


@pytest.mark.parametrize("cap", [256, 1000, 1024])
@pytest.mark.parametrize("m", [27648, 262144, 1 << 20, 23887872])
def test_stream_group_sums_deterministic(m, cap):
    P = stream_groups(m, cap)
    G = stream_group_size(P)
    rng0 = random.Random(m)
    # values with wide exponent spread: a different summation order would
    # change the floating-point result
    vals = [rng0.uniform(-1, 1) * 10 ** rng0.randint(-8, 8) for _ in range(P)]
    reference = None
    for seed in range(6):
        mem = Memory()
        actors, group_sums = stream_pass(P, vals, mem)
        run_random(actors, random.Random(seed))
        assert not mem.violations, mem.violations[:3]
        ngroups = -(-P // G)
        assert sorted(group_sums) == list(range(ngroups))  # one last arriver per group
        # the next launch adds the group sums in index order (tsqr_gram_flat)
        total = 0.0
        for g in range(ngroups):
            total += group_sums[g]
        if reference is None:
            reference = total
        assert total == reference  # bit-identical whatever the interleaving


def test_stream_group_without_acquire_is_caught():
    P = stream_groups(1 << 20, 1024)
    vals = [float(i) for i in range(P)]
    caught = 0
    for seed in range(10):
        mem = Memory()
        actors, _ = stream_pass(P, vals, mem, acquire=False)
        run_random(actors, random.Random(seed))
        caught += bool(mem.violations)
    assert caught > 0
