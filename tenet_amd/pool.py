"""Runtime pool: one backend context leased per operation.

Mirrors the reference's context pool (``Runtime::lease_context`` and
``ContextLease``'s drop, tenet/src/runtime.rs:484-519, :561-576, :795-818):
``lease()`` pops an idle runtime or mints one; a lease that ends normally
returns its runtime while fewer than ``max_idle`` are idle (otherwise the
runtime is closed); a lease that ends with an exception quarantines its
runtime -- it is closed and never reused, since a failed operation may have
left queued work or half-built caches behind.  Each runtime is one backend
context (one HIP stream, its own pools and artifact tables), so operations
leased from one pool by several host threads run on separate streams.
Results a caller reads back are synchronised by that caller's own
``to_numpy`` / ``synchronize``; work still queued on a returned runtime stays
ordered before the next lease's work on the same stream.
"""
import os
import threading
from contextlib import contextmanager


class RuntimePool:
    def __init__(self, device=0, precision="f64", max_idle=None, factory=None):
        """max_idle: idle runtimes kept warm (default: host cores, as the
        reference's idle cap); factory: builds a runtime (default
        Runtime(device, precision))."""
        if factory is None:
            from .tensor import Runtime

            def factory():
                return Runtime(device, precision)
        self._factory = factory
        self.max_idle = max_idle if max_idle is not None else (os.cpu_count() or 1)
        self._idle = []
        self._lock = threading.Lock()
        self.minted = 0
        self.quarantined = 0

    @contextmanager
    def lease(self):
        with self._lock:
            rt = self._idle.pop() if self._idle else None
        if rt is None:
            rt = self._factory()
            with self._lock:
                self.minted += 1
        try:
            yield rt
        except BaseException:
            with self._lock:
                self.quarantined += 1
            rt.close()
            raise
        with self._lock:
            if len(self._idle) < self.max_idle:
                self._idle.append(rt)
                rt = None
        if rt is not None:
            rt.close()

    @property
    def idle(self):
        with self._lock:
            return len(self._idle)

    def close(self):
        with self._lock:
            idle, self._idle = self._idle, []
        for rt in idle:
            rt.close()
