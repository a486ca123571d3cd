"""CPU model check of the scheduling protocol of the XCD-cooperative two-step
kernel (minfft_amd/csrc/coop.hpp): work-groups of one queue claim positions
of an interleaved queue (step-1 chunks of transform r, then step-2 chunks of
transform r - LA), a claimed step-2 chunk is run by its claimant only if an
atomic OR on the transform's word shows step 1 complete, and the work-group
whose atomic ADD completes step 1 runs the chunks deferred before it.

Simulated work-groups advance one atomic operation at a time in a random
order (every interleaving the hardware could produce is a possible schedule
here), and each run checks:
- every step-2 chunk runs exactly once and every step-1 chunk exactly once;
- no step-2 chunk starts before all step-1 chunks of its transform are done;
- the run ends (no work-group waits on another, so every schedule drains).
The model mirrors the kernel's decision code step by step; the kernel's
arithmetic is checked on the GPU (tests/test_gpu_batch.py, schedule tests)."""
import random

import pytest


def run_model(G, CH1, CH2, LA, nwg, seed, slow_prob=0.0):
    """One queue of the kernel's loop (coop.hpp), one atomic step at a time:
    claim a position; a step-2 claim ORs its bit and runs the chunk only if
    step 1 is complete; a step-1 chunk stores, then ADDs, and the ADD that
    completes step 1 runs the chunks deferred before it."""
    rng = random.Random(seed)
    per_round = CH1 + CH2
    total = (G + LA) * per_round
    head = [0]
    word = [0] * G
    step1_done = [0] * G          # chunks whose slot stores are acknowledged
    runs1, runs2, violations = {}, {}, []

    def run2(e, c):
        if step1_done[e] != CH1:
            violations.append(("step 2 before step 1", e, c))
        runs2[(e, c)] = runs2.get((e, c), 0) + 1

    def wg():
        while True:
            s = head[0]                 # claim (one atomic)
            head[0] += 1
            yield
            if s >= total:
                return
            r, j = divmod(s, per_round)
            if j < CH1:
                if r >= G:
                    continue
                e = r
                while rng.random() < slow_prob:   # loads, FFT, stores
                    yield
                runs1[(e, j)] = runs1.get((e, j), 0) + 1
                step1_done[e] += 1      # stores acknowledged, work-group synchronised
                yield
                old = word[e]           # atomic ADD
                word[e] = old + 1
                yield
                mask = (old >> 8) if (old & 0xFF) + 1 == CH1 else 0
                while mask:
                    c = (mask & -mask).bit_length() - 1
                    mask &= mask - 1
                    run2(e, c)
                    yield
            else:
                if r < LA or r - LA >= G:
                    continue
                e, c = r - LA, j - CH1
                old = word[e]           # atomic OR
                word[e] = old | (1 << (8 + c))
                yield
                if (old & 0xFF) == CH1:
                    run2(e, c)
                    yield

    live = [wg() for _ in range(nwg)]
    steps = 0
    while live:
        k = rng.randrange(len(live))
        try:
            next(live[k])
        except StopIteration:
            live.pop(k)
        steps += 1
        assert steps < 10_000_000, "did not drain"
    assert not violations, violations[:5]
    for e in range(G):
        for j in range(CH1):
            assert runs1.get((e, j)) == 1, ("step 1", e, j, runs1.get((e, j)))
        for c in range(CH2):
            assert runs2.get((e, c)) == 1, ("step 2", e, c, runs2.get((e, c)))


@pytest.mark.parametrize("CH1,CH2", [(8, 8), (4, 4), (16, 16), (8, 24)])
@pytest.mark.parametrize("LA", [0, 1, 3, 5, 8])
@pytest.mark.parametrize("nwg", [1, 3, 64])
def test_protocol_random_schedules(CH1, CH2, LA, nwg):
    for seed in range(4):
        run_model(G=12, CH1=CH1, CH2=CH2, LA=LA, nwg=nwg, seed=seed, slow_prob=0.5)


@pytest.mark.parametrize("G", [0, 1, 2, 94])
def test_protocol_queue_sizes(G):
    for seed in range(3):
        run_model(G=G, CH1=8, CH2=8, LA=5, nwg=64, seed=100 + seed, slow_prob=0.3)


def test_model_rejects_a_planted_mistake():
    """A protocol that lets a claimant run a step-2 chunk whenever any step-1
    chunk is done (the wrong readiness test) must be caught by the checks."""
    def broken(G, CH1, CH2, LA, nwg, seed):
        rng = random.Random(seed)
        per_round = CH1 + CH2
        total = (G + LA) * per_round
        head = [0]
        done = [0] * G
        bad = []

        def wg():
            while True:
                s = head[0]
                head[0] += 1
                yield
                if s >= total:
                    return
                r, j = divmod(s, per_round)
                if j < CH1:
                    if r < G:
                        while rng.random() < 0.5:
                            yield
                        done[r] += 1
                        yield
                elif LA <= r < G + LA:
                    e = r - LA
                    if done[e] >= 1:      # wrong: should require CH1
                        if done[e] != CH1:
                            bad.append(e)
                    yield

        live = [wg() for _ in range(nwg)]
        while live:
            k = rng.randrange(len(live))
            try:
                next(live[k])
            except StopIteration:
                live.pop(k)
        return bad

    assert any(broken(12, 8, 8, 0, 16, s) for s in range(8))
