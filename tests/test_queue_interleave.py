"""Queue slots of the self-placed batched set-up (uw_batched.hip, estimate_block): without the bulk
path every estimate workgroup of 256 chunks places its chunks in its own bucket order, and round q
of the queue holds every workgroup's q-th chunk. The slots must cover 0 .. nchunks - 1 exactly once
(a gap would end a persistent workgroup's walk early, a collision would drop a chunk), and each
workgroup's chunks must keep their bucket order (rank q before rank q + 1)."""
import numpy as np
import pytest

K = 256  # chunks per estimate workgroup (kPrepThreads)


def slot(q, j, nchunks):
    """Mirror of the device mapping: local rank q of estimate workgroup j."""
    m = (nchunks + K - 1) // K
    if m == 1:
        return q
    r = nchunks - K * (m - 1)
    return q * m + j if q < r else r * m + (q - r) * (m - 1) + j


@pytest.mark.parametrize("nchunks", [1, 2, 255, 256, 257, 300, 511, 512, 513, 1000, 4095, 4096, 4097, 8192, 12345])
def test_interleaved_slots_are_a_bijection(nchunks):
    m = (nchunks + K - 1) // K
    seen = np.zeros(nchunks, dtype=np.int64)
    for j in range(m):
        size = K if j < m - 1 else nchunks - K * (m - 1)
        slots = [slot(q, j, nchunks) for q in range(size)]
        assert all(0 <= s < nchunks for s in slots)
        assert slots == sorted(slots), "a workgroup's bucket order is kept"
        for s in slots:
            seen[s] += 1
    assert np.all(seen == 1)


def test_rounds_hold_one_chunk_of_every_workgroup():
    # 4096 chunks, 16 workgroups: round q is slots [16 q, 16 q + 16), one per workgroup
    n = 4096
    for q in (0, 1, 100, 255):
        assert sorted(slot(q, j, n) for j in range(16)) == list(range(16 * q, 16 * q + 16))
