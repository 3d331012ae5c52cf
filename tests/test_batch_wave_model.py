"""Host model of the batched kernel's wave loop (skwarka_amd/csrc/kernels/batch.hip).

A wave is 64 lanes; `__ballot(p)` is the bitmask of the lanes for which p holds; a walk is a
sequence of advance steps with leaves held in between (the Walker's advance() / leaf() contract:
advance() either moves one node, stops at a leaf (at_leaf), or ends the walk; leaf() clears
at_leaf). The model runs the kernel's control flow over a chunk list shaped like the GPU tests'
(scenes of various ray counts, chunk sizes 64..4096, optional gather chunks) and checks:

* every loop iteration moves the wave (claims a chunk, starts rays, tests leaves or advances a
  lane): the invariant the kernel's stall guard relies on;
* every ray is started exactly once and finished exactly once, and every chunk is accounted once;
* the loop exit is taken by all lanes in the same iteration.

It models both the current single loop and the earlier nested form (a per-chunk inner loop inside
a lane-0 atomicAdd + shuffle claim loop) that was removed after a GPU run hung. Both terminate
here for every chunk list and walk shape drawn when all 64 lanes run every iteration together.

The hang came from the compiled form of the nested loop, not from its source-level control flow:
in the gfx950 code of that kernel the outer loop's back edge was threaded past the loop-invariant
branch `lane == 0` (the claim's atomicAdd), so after the first chunk lanes 1..63 went straight
back to the shuffle, without a claim, and ran a chunk's ray loop by themselves while lane 0 waited
at the loop's exit. `test_compiled_nested_form_spins_without_lane_0` models that structure: the
63-lane subset never sees lane 0's bit in `__ballot(!busy)`, so `idle == ~0ull` never holds and
the loop never ends. The current kernel tests idle lanes against the active mask (`__ballot(1)`)
and takes its leader from it, which the model shows ends the same subset's loop.
"""
import random

import pytest

FULL = (1 << 64) - 1
REFILL_MIN, LEAF_MIN = 8, 16


def ballot(pred):
    m = 0
    for i, p in enumerate(pred):
        if p:
            m |= 1 << i
    return m


def popc(x):
    return bin(x).count("1")


class Walk:
    """One ray's walk: segments of advance steps, a held leaf between consecutive segments."""

    def __init__(self, rng):
        nseg = rng.randint(1, 6)
        self.segs = [rng.randint(0, 12) for _ in range(nseg)]
        self.seg = 0
        self.left = self.segs[0]
        self.at_leaf = False

    def advance(self):
        assert not self.at_leaf
        if self.left > 0:
            self.left -= 1
            return False
        if self.seg + 1 < len(self.segs):
            self.seg += 1
            self.left = self.segs[self.seg]
            self.at_leaf = True
            return False
        return True

    def leaf(self):
        assert self.at_leaf
        self.at_leaf = False


def make_queue(rng, n_items, chunk, gather_items=0):
    items = []
    for i in range(n_items):
        items.append(("ray", rng.choice([1, 7, 63, 64, 65, 200, 1000, 4096 + 3])))
    for j in range(gather_items):
        items.append(("gather", rng.choice([1, 64, 100, 3000])))
    order = sorted(range(len(items)), key=lambda i: items[i][1])
    chunks = []
    for it in order:
        n = items[it][1]
        for b in range(0, n, chunk):
            chunks.append((it, b, min(b + chunk, n)))
    return items, chunks


class Shared:
    def __init__(self, items, chunks):
        self.items, self.chunks = items, chunks
        self.head = 0
        self.done = [0] * len(items)
        self.started = {}
        self.finished = {}

    def claim(self):
        q = self.head
        self.head += 1
        return q


def run_single_loop(sh, rng, walk_rng, max_iter=10 ** 6):
    """batch.hip k_batch_trace (current): one loop, chunk claim when the wave is idle and its
    chunk has no rays left."""
    busy = [False] * 64
    walk = [None] * 64
    idx = [0] * 64
    cursor = end = chunk_len = 0
    sid = None
    for _ in range(max_iter):
        moved = False
        idle = ballot([not b for b in busy])
        if idle == FULL and cursor >= end:
            moved = True
            if sid is not None:
                sh.done[sid] += chunk_len
            sid = None
            q = sh.claim()
            if q >= len(sh.chunks):
                return True
            it, b, e = sh.chunks[q]
            if sh.items[it][0] == "gather":
                for qi in range(b, e):
                    sh.finished[(it, qi)] = sh.finished.get((it, qi), 0) + 1
                sh.done[it] += e - b
                cursor = end = 0
                continue
            sid, cursor, end, chunk_len = it, b, e, e - b
        if idle != 0 and cursor < end and (idle == FULL or popc(idle) >= REFILL_MIN):
            moved = True
            avail = end - cursor
            for lane in range(64):
                rank = popc(idle & ((1 << lane) - 1))
                if not busy[lane] and rank < avail:
                    idx[lane] = cursor + rank
                    key = (sid, idx[lane])
                    sh.started[key] = sh.started.get(key, 0) + 1
                    walk[lane] = Walk(walk_rng)
                    busy[lane] = True
            took = popc(idle)
            cursor += took if took < avail else avail
        held = ballot([busy[l] and walk[l].at_leaf for l in range(64)])
        if held:
            moving = ballot([busy[l] and not walk[l].at_leaf for l in range(64)])
            if popc(held) >= LEAF_MIN or moving == 0 or cursor >= end:
                moved = True
                for l in range(64):
                    if busy[l] and walk[l].at_leaf:
                        walk[l].leaf()
        adv = ballot([busy[l] and not walk[l].at_leaf for l in range(64)])
        moved = moved or adv != 0
        assert moved, "an iteration moved nothing"
        for l in range(64):
            if busy[l] and not walk[l].at_leaf:
                if walk[l].advance():
                    key = (sid, idx[l])
                    sh.finished[key] = sh.finished.get(key, 0) + 1
                    busy[l] = False
    return False


def run_nested(sh, rng, walk_rng, max_iter=10 ** 6):
    """The removed form: outer claim loop (lane 0 atomicAdd + shuffle), per-chunk inner loop whose
    walk takes one step per iteration (advance or leaf)."""
    iters = 0
    while True:
        q = sh.claim()
        if q >= len(sh.chunks):
            return True
        it, b, e = sh.chunks[q]
        if sh.items[it][0] == "gather":
            for qi in range(b, e):
                sh.finished[(it, qi)] = sh.finished.get((it, qi), 0) + 1
            sh.done[it] += e - b
            continue
        busy = [False] * 64
        walk = [None] * 64
        idx = [0] * 64
        cursor = b
        while True:
            iters += 1
            if iters > max_iter:
                return False
            idle = ballot([not x for x in busy])
            exits = [idle == FULL and cursor >= e for _ in range(64)]
            assert all(exits) or not any(exits), "divergent exit"
            if exits[0]:
                break
            moved = False
            if idle != 0 and cursor < e:
                moved = True
                avail = e - cursor
                for lane in range(64):
                    rank = popc(idle & ((1 << lane) - 1))
                    if not busy[lane] and rank < avail:
                        idx[lane] = cursor + rank
                        key = (it, idx[lane])
                        sh.started[key] = sh.started.get(key, 0) + 1
                        walk[lane] = Walk(walk_rng)
                        busy[lane] = True
                took = popc(idle)
                cursor += took if took < avail else avail
            for l in range(64):
                if busy[l]:
                    moved = True
                    w = walk[l]
                    if w.at_leaf:
                        w.leaf()
                    elif w.advance():
                        key = (it, idx[l])
                        sh.finished[key] = sh.finished.get((it, idx[l]), 0) + 1
                        busy[l] = False
            assert moved
        sh.done[it] += e - b


@pytest.mark.parametrize("form", ["single", "nested"])
@pytest.mark.parametrize("seed", range(6))
def test_wave_loop_terminates_and_covers_every_ray(form, seed):
    rng = random.Random(seed)
    walk_rng = random.Random(1000 + seed)
    chunk = rng.choice([64, 256, 1024, 4096])
    items, chunks = make_queue(rng, rng.randint(1, 7), chunk, gather_items=rng.randint(0, 2))
    sh = Shared(items, chunks)
    # several waves share the queue; they run one after another in this model (each wave's loop
    # touches only the queue head and its own lanes)
    for _ in range(3):
        ok = (run_single_loop if form == "single" else run_nested)(sh, rng, walk_rng)
        assert ok, "a wave did not leave its loop"
    for it, (kind, n) in enumerate(items):
        assert sh.done[it] == n
        for i in range(n):
            assert sh.finished.get((it, i), 0) == 1, (kind, it, i)
            if kind == "ray":
                assert sh.started.get((it, i), 0) == 1


def test_missing_lane_would_spin_and_the_guard_catches_it():
    """A lane absent from the ballots (the failure a hang of this loop would need) never lets the
    idle mask become full; the single loop's guard turns that into an exit."""
    sh = Shared(*make_queue(random.Random(0), 2, 64))
    busy = [False] * 64
    idle = ballot([not b for b in busy]) & ~1  # lane 0 missing from every ballot
    assert idle != FULL
    # with no lane busy and the chunk empty, the real loop claims only when idle == FULL, takes no
    # refill (cursor >= end) and has no lane to move: the iteration moves nothing
    cursor = end = 0
    claimed = idle == FULL and cursor >= end
    refill = idle != 0 and cursor < end
    assert not claimed and not refill


def run_chunk_subset(lanes, b, e, walk_rng, exit_mask, max_iter=20000):
    """The ray loop of one chunk [b, e) run by `lanes` only (the others are masked off: they are
    missing from every ballot). exit_mask: the mask `idle` is compared with to leave the loop.
    Returns the iterations taken, or None when the loop did not end within max_iter."""
    busy = {l: False for l in lanes}
    walk = {}
    cursor = b
    for it in range(max_iter):
        idle = ballot([l in busy and not busy[l] for l in range(64)])
        if idle == exit_mask and cursor >= e:
            return it
        if idle != 0 and cursor < e:
            avail = e - cursor
            for l in lanes:
                rank = popc(idle & ((1 << l) - 1))
                if not busy[l] and rank < avail:
                    walk[l] = Walk(walk_rng)
                    busy[l] = True
            cursor += min(popc(idle), avail)
        for l in lanes:
            if busy[l]:
                w = walk[l]
                if w.at_leaf:
                    w.leaf()
                elif w.advance():
                    busy[l] = False
    return None


@pytest.mark.parametrize("stale_q", [0, 1])
def test_compiled_nested_form_spins_without_lane_0(stale_q):
    """The removed kernel as compiled: all 64 lanes run the first chunk; then lanes 1..63 re-enter
    the ray loop on a chunk index they did not claim (0, or the stale value of lane 0's register)
    while lane 0 waits at the loop exit. With the `idle == ~0ull` exit the subset spins forever;
    with the exit taken against the subset's active mask it ends."""
    walk_rng = random.Random(7)
    chunks = [(0, 0, 256), (0, 256, 512)]
    assert run_chunk_subset(list(range(64)), 0, 256, walk_rng, FULL) is not None  # first chunk: whole wave
    b, e = chunks[stale_q][1:]
    subset = list(range(1, 64))
    assert run_chunk_subset(subset, b, e, walk_rng, FULL) is None  # the hang
    active = ballot([l in subset for l in range(64)])
    assert run_chunk_subset(subset, b, e, walk_rng, active) is not None


def test_subset_with_its_own_leader_claims_fresh_chunks():
    """In the current loop the claim's leader is the first lane of `__ballot(1)`: a subset of the
    wave that goes round the loop alone still claims its own chunks and finishes, so every chunk is
    claimed and the queue drains (duplicated rays would be traced twice with identical results)."""
    sh = Shared(*make_queue(random.Random(3), 4, 64))
    walk_rng = random.Random(9)
    for subset in (list(range(64)), list(range(1, 64)), [0]):
        active = ballot([l in subset for l in range(64)])
        leader = (active & -active).bit_length() - 1
        assert leader == subset[0]
        while True:
            q = sh.claim()  # done by `leader`, broadcast from it
            if q >= len(sh.chunks):
                break
            it, b, e = sh.chunks[q]
            assert run_chunk_subset(subset, b, e, walk_rng, active) is not None
    assert sh.head >= len(sh.chunks)
