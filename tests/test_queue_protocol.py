"""Host-side interleaving model of the batched kernel's task hand-over (csrc/uw_batched.hip,
k_batched's main loop): one lane (pf, in the last wave) claims the next task from the queue
(atomic head, then the task record) and stages it in the shared word s_next; every wave reads
s_next after a barrier, and exits when it reads the end of the queue.

Waves run independently between barriers. The model follows the hardware barrier rule of AMD GPUs:
s_barrier releases when every wave of the workgroup that has not ended waits at a barrier -- any
barrier instruction, not a particular one -- and ended waves are no longer waited for. Random
schedules of the waves' steps are explored, and each run checks that
  (1) every wave reads the same task sequence (all waves work on one chunk at a time), and
  (2) every wave ends (the launch finishes).

The shipped protocol claims either after the loop-top barrier (followed by a second barrier before
the reads) or, for one-pass tasks with lookahead, during the density phase, after the table
barrier that follows every wave's read of the current task. Both writes are therefore ordered
after all reads of the previous value and before all reads of the new one; the model finds no
violation in any schedule.

The dropped variant issued the claim just before the loop-top barrier. A task whose resolution is
out of range ends its iteration with `continue` before any barrier, so pf's wave can come back to
the loop top and overwrite s_next while a slower wave has not yet read the current task. That
wave then reads the next task and runs one barrier phase apart from the others (barriers pair up
across different barrier instructions). The model reproduces this split for that variant, and
shows it needs the barrier-free exit: without out-of-range tasks the early claim is ordered after
every read. In the control flow alone the split waves still end (the end-of-queue value stays in
s_next); in the kernel, waves one phase apart read the block-scan partials, chunk counts and
active-list lengths (uw_batched.hip wtot, scan_lds, LIST/RING) written for another phase, and
those values bound the emission loops (up to 2^32 entries per chunk), which is how one run can
exceed any time limit.
"""
import random

import pytest

BARRIER, READ, CLAIM, EXIT = "barrier", "read", "claim", "exit"


def _body(kind, lookahead):
    """Steps of one task after its read of s_next (and whether the claim was staged)."""
    if kind == "invalid":
        return [], False               # continue before any barrier
    steps = [BARRIER]                  # x / y tables
    staged = False
    if kind in ("blank", "surface"):  # one-pass chunk: densities, then the s_bits barrier
        if lookahead:
            steps.append(CLAIM)        # pf stages the next task during the density phase (pf's wave only)
            staged = True
        steps.append(BARRIER)
        if kind == "surface":
            steps += [BARRIER, BARRIER, BARRIER]  # wave totals, reservation, index map
    else:                              # x-slab chunk: slab barriers, scan, reservation, flushes
        steps += [BARRIER] * 6
    return steps, staged


class Wave:
    def __init__(self, wid, pf):
        self.wid, self.pf = wid, pf
        self.prog = []
        self.staged = True   # the first task is staged before the loop
        self.seen = []
        self.done = False
        self.at_barrier = False


def run(tasks, variant, waves=4, grid=1, seed=0, max_steps=20000):
    """Returns (sequences read by each wave, finished). tasks: kinds in queue order."""
    rng = random.Random(seed)
    state = {"head": grid, "s_next": 0}   # workgroup 0 starts with task 0
    ntask = len(tasks)
    W = [Wave(i, i == waves - 1) for i in range(waves)]

    def loop_top(w):
        if variant == "shipped":
            steps = [BARRIER]
            if not w.staged:
                steps += ([CLAIM] if w.pf else []) + [BARRIER]
        else:  # claim before the loop-top barrier
            steps = ([CLAIM] if (not w.staged and w.pf) else []) + [BARRIER]
        return steps + [READ]

    for w in W:
        w.prog = loop_top(w)
    for _ in range(max_steps):
        live = [w for w in W if not w.done]
        if not live:
            return [w.seen for w in W], True
        if all(w.at_barrier for w in live):  # hardware barrier: every live wave waits
            for w in live:
                w.at_barrier = False
                w.prog.pop(0)
                if not w.prog:
                    w.prog = loop_top(w)
            continue
        runnable = [w for w in live if not w.at_barrier]
        w = rng.choice(runnable)
        op = w.prog[0]
        if op == BARRIER:
            w.at_barrier = True
            continue
        w.prog.pop(0)
        if op == CLAIM and w.pf:  # only the claiming lane's wave writes
            t = state["head"]
            state["head"] += 1
            state["s_next"] = t
        elif op == READ:
            t = state["s_next"]
            w.seen.append(t)
            if t >= ntask:
                w.done = True
                continue
            lookahead = t + 2 * grid < ntask
            body, staged = _body(tasks[t], lookahead)
            w.staged = staged
            w.prog = body
        if not w.prog and not w.done:
            w.prog = loop_top(w)
    return [w.seen for w in W], False


def _mixes():
    g = random.Random(11)
    kinds = ["blank", "surface", "xslab", "invalid"]
    out = [["surface"] * 12, ["invalid", "surface", "invalid", "blank", "xslab", "invalid"] * 3]
    for _ in range(30):
        out.append([g.choice(kinds) for _ in range(g.randint(1, 24))])
    return out


def _consistent(seqs):
    return all(s == seqs[0] for s in seqs)


def test_shipped_hand_over_is_race_free():
    for mix in _mixes():
        for seed in range(40):
            seqs, finished = run(mix, "shipped", seed=seed)
            assert finished, (mix, seed)
            assert _consistent(seqs), (mix, seed, seqs)
            assert seqs[0] == list(range(len(mix) + 1))  # every task once, in queue order, then the end


def test_claim_before_loop_top_barrier_splits_waves():
    split = 0
    for mix in _mixes():
        for seed in range(40):
            seqs, finished = run(mix, "claim_before_barrier", seed=seed)
            split += not _consistent(seqs)
    assert split > 0


@pytest.mark.parametrize("mix", [["surface"] * 10, ["blank", "surface", "xslab"] * 4])
def test_claim_before_barrier_is_safe_without_barrier_free_exits(mix):
    # without a task that leaves the iteration before any barrier, the early claim is ordered
    # after every read (the race needs the invalid-resolution exit)
    for seed in range(60):
        seqs, finished = run(mix, "claim_before_barrier", seed=seed)
        assert finished and _consistent(seqs)
