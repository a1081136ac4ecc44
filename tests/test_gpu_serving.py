"""GPU: the continuous-batching server (one persistent traversal launch)
against bst_predict_device on the same rows.

Every request's output must equal the device prediction of its rows bit for
bit (the server runs the prediction kernel's own tile code): requests of one
row to many tiles, batches small enough to be cut into leaf and sum units
and batches traversed tile by tile, more requests than slots (slots reused),
submissions from several threads at once, a pause longer than the kernel's
idle time (the kernel leaves and is launched again), softmax outputs, and
requests still queued when the server is destroyed.
"""
import threading
import time

import numpy as np
import pytest

from boosters_amd import _abi as A
from boosters_amd.model import GBDTModel, Server
from tests.test_zz_no_alloc import _random_forest

pytestmark = pytest.mark.gpu

F32 = np.float32


def _alone(gm, X):
    return gm.predict_feature_major(X, raw=True)


@pytest.mark.parametrize("K", [1, 3])
def test_server_outputs_equal_device_prediction(ctx, K):
    info, arr = _random_forest(120, 7, 30, seed=11 + K, K=K)
    gm = GBDTModel.from_arrays(info, arr, ctx)
    rng = np.random.default_rng(K)
    sizes = [1, 2, 255, 256, 257, 1000, 70000, 3, 511, 4096] + list(rng.integers(1, 5000, 30))
    reqs = [rng.standard_normal((30, int(n)), dtype=F32) for n in sizes]
    reqs[3][5, :] = np.nan
    want = [_alone(gm, X) for X in reqs]
    srv = Server(gm, hbm_budget_bytes=1 << 26, max_inflight=3)
    tickets = [srv.submit(X) for X in reqs]
    outs = [srv.wait(t)[0] for t in tickets]
    for o, w in zip(outs, want):
        np.testing.assert_array_equal(o, w)
    # A pause past the kernel's idle time: it leaves, and the next batch
    # launches it again.
    time.sleep(0.2)
    t = srv.submit(reqs[5])
    np.testing.assert_array_equal(srv.wait(t)[0], want[5])
    srv.close()


def test_server_concurrent_submitters(ctx):
    info, arr = _random_forest(200, 8, 40, seed=21)
    gm = GBDTModel.from_arrays(info, arr, ctx)
    # Expected outputs first, on this thread: a bst_ctx is not re-entrant, so
    # the clients below only talk to the (thread-safe) server.
    work = {}
    for seed in range(6):
        rng = np.random.default_rng(seed)
        xs = [rng.standard_normal((40, int(rng.integers(1, 20000))), dtype=F32) for _ in range(12)]
        work[seed] = [(X, _alone(gm, X)) for X in xs]
    srv = Server(gm, hbm_budget_bytes=1 << 28, max_inflight=4)
    errors = []

    def client(seed):
        try:
            for X, want in work[seed]:
                t = srv.submit(X)
                out, lat = srv.wait(t)
                np.testing.assert_array_equal(out, want)
                assert lat > 0
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=client, args=(s,)) for s in range(6)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    srv.close()
    assert not errors, errors


def test_server_budget_and_shape_errors(ctx):
    info, arr = _random_forest(20, 4, 8, seed=31)
    gm = GBDTModel.from_arrays(info, arr, ctx)
    # The 20-tree forest fits one LDS chunk, so batches are never cut into
    # leaf units and no leaf buffer is taken from the budget: 1 MiB over 2
    # slots of 9 floats per row, 14,563 rows per slot, 14,336 after rounding
    # to 256-row tiles; a larger request is refused.
    srv = Server(gm, hbm_budget_bytes=1 << 20, max_inflight=2)
    with pytest.raises(A.BoostersError) as e:
        srv.submit(np.zeros((8, 14337), F32))
    assert e.value.code == A.BST_E_OOM
    with pytest.raises(A.BoostersError) as e:
        srv.submit(np.zeros((7, 10), F32))
    assert e.value.code == A.BST_E_SHAPE
    X = np.random.default_rng(1).standard_normal((8, 14336), dtype=F32)
    t = srv.submit(X)
    np.testing.assert_array_equal(srv.wait(t)[0], _alone(gm, X))
    # Requests still queued at destroy are finished first.
    ts = [srv.submit(X[:, : 1000 + i]) for i in range(8)]
    srv.close()
    del ts
