"""Host model of the persistent sweep's document queues, run under adversarial interleavings.

The persistent kernels (k_sweep_persistent, k_sweep_persistent_long, k_sweep_block, the fold-in
worker pool's claims aside) share one protocol: a fixed grid of wavefronts, each looping
    item = atomicAdd(head, 1); if item >= n_items: leave; sweep(order[item])
with the head cleared by a fill kernel on the same stream before the launch (also inside a
captured iteration, so every replay starts from 0).  The split and block schedules run two such
launches on two streams, each with its own head.

Termination argument: every loop iteration of every wavefront performs one atomicAdd on its
head, which only grows; after at most n_items + (grid size) additions every wavefront has read
an index >= n_items and left.  No wavefront waits on another, so no interleaving can stall the
grid.  Exactly-once: atomicAdd hands out each index once, and indices below n_items are exactly
the queue.

Per-XCD queues with stealing (one head per XCD, a wavefront tries its own queue first, then the
others in order) keep both properties when a wavefront leaves only after an atomicAdd on every
queue returned an index past that queue's end: each attempt grows one head, so the bound is
8 * (n_items + grid size) additions.  The model below checks these claims on random schedules,
including graph replays, and checks that a replay whose heads are not cleared processes
nothing.
"""
import random

import pytest


class Atomic:
    def __init__(self):
        self.v = 0

    def add(self, x):
        old = self.v
        self.v += x
        return old


def wave_single(head, n_items, done):
    while True:
        item = head.add(1)
        yield
        if item >= n_items:
            return
        done.append(item)
        yield


def wave_stealing(heads, sizes, home, done):
    """Own queue first, then the others; leave once every queue returned an index past its end."""
    q = len(heads)
    while True:
        progressed = False
        for off in range(q):
            j = (home + off) % q
            item = heads[j].add(1)
            yield
            if item < sizes[j]:
                done.append((j, item))
                progressed = True
                yield
                break
        if not progressed:
            return


def run(waves, rng, max_steps):
    live = list(waves)
    steps = 0
    while live:
        i = rng.randrange(len(live))
        try:
            next(live[i])
        except StopIteration:
            live.pop(i)
        steps += 1
        assert steps <= max_steps, "a wavefront did not leave"
    return steps


@pytest.mark.parametrize("seed", range(20))
def test_single_queue_exactly_once_and_terminates(seed):
    rng = random.Random(seed)
    n_items, grid = rng.randrange(0, 300), rng.randrange(1, 64)
    head, done = Atomic(), []
    run([wave_single(head, n_items, done) for _ in range(grid)], rng, 4 * (n_items + grid) + 10)
    assert sorted(done) == list(range(n_items))


@pytest.mark.parametrize("seed", range(20))
def test_two_launches_split_or_block(seed):
    """Split / block schedules: two grids on two streams, each with its own head, interleaved."""
    rng = random.Random(100 + seed)
    n_long, n_short = rng.randrange(0, 50), rng.randrange(0, 400)
    heads, done_l, done_s = [Atomic(), Atomic()], [], []
    waves = [wave_single(heads[1], n_long, done_l) for _ in range(rng.randrange(1, 16))]
    waves += [wave_single(heads[0], n_short, done_s) for _ in range(rng.randrange(1, 64))]
    run(waves, rng, 10 * (n_long + n_short + len(waves)) + 10)
    assert sorted(done_l) == list(range(n_long)) and sorted(done_s) == list(range(n_short))


def test_graph_replays_clear_the_head():
    rng = random.Random(7)
    head = Atomic()
    for replay in range(5):
        head.v = 0  # the fill kernel captured before the launch
        done = []
        run([wave_single(head, 120, done) for _ in range(17)], rng, 10_000)
        assert sorted(done) == list(range(120)), replay


def test_replay_without_clearing_processes_nothing():
    """What a head left at its end value does on the next replay: every wavefront leaves at once
    (no hang, but no document is swept) -- why the head is cleared inside the captured graph."""
    rng = random.Random(8)
    head, done = Atomic(), []
    run([wave_single(head, 50, done) for _ in range(9)], rng, 10_000)
    again = []
    run([wave_single(head, 50, again) for _ in range(9)], rng, 10_000)
    assert sorted(done) == list(range(50)) and again == []


@pytest.mark.parametrize("seed", range(20))
def test_per_xcd_queues_with_stealing(seed):
    rng = random.Random(200 + seed)
    sizes = [rng.randrange(0, 60) for _ in range(8)]
    heads, done = [Atomic() for _ in range(8)], []
    grid = rng.randrange(1, 48)
    waves = [wave_stealing(heads, sizes, w % 8, done) for w in range(grid)]
    run(waves, rng, 8 * 4 * (sum(sizes) + grid) + 10)
    assert sorted(done) == sorted((j, i) for j in range(8) for i in range(sizes[j]))
