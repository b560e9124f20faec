"""Tile-shape autotuner (tenet_amd.autotune) on the device.

The tuner picks one tile configuration per contraction artifact by timing
the workload step (eagerly, or on hipGraph replays of it).  Whatever it
picks, the step's result stays the FP64 result up to rounding, the graphs it
records for timing are released, and a second call reuses the table.
"""
import numpy as np
import pytest

from tenet_amd import _native as N
from tenet_amd.autotune import CONFIGS, autotune, signature
from tenet_amd.tensor import Runtime
from tenet_amd.workloads import MPSEnvUpdate

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rule,graph", [("U1", True), ("SU2", True), ("U1", False)])
def test_autotune_keeps_the_result_and_releases_its_graphs(monkeypatch, rule, graph):
    monkeypatch.delenv("TENET_AMD_TILE_CACHE", raising=False)
    rt = Runtime(0)
    try:
        wl = MPSEnvUpdate(rt, rule, 256, "f64")
        want = wl.run().to_numpy().copy()
        table = autotune(rt, wl.contractions, wl.run, reps=2, graph_step=wl if graph else None)
        assert set(table) == {signature(c) for c in wl.contractions}
        assert all(v in (-1,) + CONFIGS for v in table.values())
        assert getattr(wl, "graph", None) is None  # timing graphs destroyed
        got = wl.run().to_numpy()
        assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want)
        # tuned: a second call applies the stored choices without timing
        calls = []
        again = autotune(rt, wl.contractions, lambda: calls.append(1) or wl.run(), reps=2,
                         graph_step=wl if graph else None)
        assert again == table and not calls
        # the tuned step still records and replays as a graph
        wl.capture()
        wl.replay()
        assert np.linalg.norm(wl.out.to_numpy() - want) <= 1e-13 * np.linalg.norm(want)
        rt._check(N.hip().tnt_graph_destroy(rt._ctx, wl.graph))
        wl.graph = None
    finally:
        rt.close()
