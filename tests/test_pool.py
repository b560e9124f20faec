"""Runtime pool semantics (tenet/src/runtime.rs:484-519, :561-576, :795-818):
mint on empty, reuse, bounded idle count, quarantine on failure, distinct
runtimes for concurrent leases.  Host logic, checked with stand-in runtimes."""
import threading

import pytest

from tenet_amd.pool import RuntimePool


class FakeRuntime:
    count = 0

    def __init__(self):
        FakeRuntime.count += 1
        self.id = FakeRuntime.count
        self.closed = False

    def close(self):
        self.closed = True


def test_mint_reuse_and_idle_cap():
    pool = RuntimePool(max_idle=1, factory=FakeRuntime)
    with pool.lease() as a:
        pass
    with pool.lease() as b:
        assert b is a and not a.closed  # reused
        with pool.lease() as c:          # the only idle one is leased: mint
            assert c is not a
    assert pool.minted == 2 and pool.idle == 1
    assert c.closed != a.closed          # one returned, the other over the cap and closed
    pool.close()
    assert a.closed and c.closed and pool.idle == 0


def test_failed_lease_is_quarantined():
    pool = RuntimePool(max_idle=4, factory=FakeRuntime)
    with pytest.raises(RuntimeError):
        with pool.lease() as bad:
            raise RuntimeError("op failed")
    assert bad.closed and pool.quarantined == 1 and pool.idle == 0
    with pool.lease() as good:
        assert good is not bad
    assert pool.idle == 1


def test_concurrent_leases_get_distinct_runtimes():
    pool = RuntimePool(max_idle=8, factory=FakeRuntime)
    held, barrier, lock = [], threading.Barrier(4), threading.Lock()

    def worker():
        with pool.lease() as rt:
            barrier.wait(timeout=10)  # all four leases are live at once
            with lock:
                held.append(rt)

    ts = [threading.Thread(target=worker) for _ in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert len({id(r) for r in held}) == 4 and pool.minted == 4 and pool.idle == 4
    pool.close()
