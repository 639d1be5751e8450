"""The learned direction selector on the GPU: forward/backward kernels
against the float64 numpy oracle (itself pinned by finite differences),
training, persistence, and what the trained network gains over the
heuristic switch it replaces -- with levels and parents unchanged."""
import os

import numpy as np
import pytest

import albba_amd as A
from oracle import mlp_oracle as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu
SELECTOR = os.path.join(os.path.dirname(A.__file__), "selector.txt")


def batch(ns, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((ns, 12)).astype(np.float32)
    t = rng.normal(0, 1, (ns, 4)).astype(np.float32)
    return x, t


@pytest.mark.parametrize("ns", [1, 7, 64, 300, 2049, 40000])
def test_gpu_loss_and_gradient_match_oracle(ctx, ns):
    """Sizes below, at and across the 64-sample chunks of the gradient's
    partial sums (ragged last chunks of 44 and 1 samples)."""
    sel = A.Selector(ctx, seed=ns)
    x, t = batch(ns, ns)
    loss, grad = sel.loss_grad(x, t)
    p = sel.params.astype(np.float64)
    rl, rg = M.loss_grad(p, x, t)
    # float32 kernels vs float64 oracle
    assert abs(loss - rl) <= 1e-5 * abs(rl) + 1e-7
    np.testing.assert_allclose(grad, rg, rtol=2e-4, atol=2e-6 * np.abs(rg).max())
    # no atomics: a second call gives the same bits
    loss2, grad2 = sel.loss_grad(x, t)
    assert loss2 == loss
    np.testing.assert_array_equal(grad2, grad)
    y = sel.predict(x)
    np.testing.assert_allclose(y, M.forward(p, x)[0], rtol=1e-5, atol=1e-5)
    sel.close()


def test_training_reduces_loss_and_round_trips(ctx, tmp_path):
    rng = np.random.default_rng(0)
    x = rng.random((512, 12)).astype(np.float32)
    # a learnable target: smooth functions of the features
    t = np.stack([x[:, 0] - x[:, 3], x[:, 1] * x[:, 2], -x[:, 4], np.sin(3 * x[:, 5])], 1)
    t = (t - t.mean(1, keepdims=True)).astype(np.float32)
    sel = A.Selector(ctx, seed=5)
    before = sel.loss_grad(x, t)[0]
    after = sel.train(x, t, steps=800, lr=3e-3)
    assert after < 0.1 * before
    # the reported loss is the loss of the final parameters
    assert abs(after - M.loss(sel.params.astype(np.float64), x, t)) < 1e-4
    path = str(tmp_path / "sel.txt")
    sel.save(path)
    other = A.Selector(ctx, seed=9)
    other.load(path)
    np.testing.assert_array_equal(other.params, sel.params)
    np.testing.assert_array_equal(other.predict(x), sel.predict(x))
    assert other.queue_ratio == 0.0
    sel.queue_ratio = 2.5e-4
    sel.save(path)
    other.load(path)
    assert other.queue_ratio == np.float32(2.5e-4)
    # a version-1 file (parameters only) loads with the ratio unset
    lines = open(path).read().split("\n")
    v1 = ["albba-selector 1 1604"] + lines[1:1605]
    (tmp_path / "v1.txt").write_text("\n".join(v1) + "\n")
    other.load(str(tmp_path / "v1.txt"))
    assert other.queue_ratio == 0.0
    np.testing.assert_array_equal(other.params, sel.params)
    with pytest.raises(A.AlbbaError):
        other.queue_ratio = -1.0
    (tmp_path / "bad.txt").write_text("not a selector\n")
    with pytest.raises(A.AlbbaError):
        other.load(str(tmp_path / "bad.txt"))


def test_trained_selector_gains_over_heuristic(ctx):
    """Held-out RMAT graph (seed outside the training set): the committed
    trained network gives the same levels and parents as the heuristic and
    needs less kernel time per BFS."""
    assert os.path.exists(SELECTOR), "albba_amd/selector.txt (trained weights) missing"
    sel = A.Selector(ctx)
    sel.load(SELECTOR)
    g = A.Graph.rmat(21, 16, 0.57, 0.19, 0.19, seed=424242, perm_seed=77)
    rp, col = g.row_ptr, g.col
    dg = A.DeviceGraph(ctx, g.sellcs(64, 4096, 1))
    t_heur, t_learn = 0.0, 0.0
    for root in g.pick_roots(6, 99):
        h = dg.bfs(int(root), policy=A.SWITCH_HEURISTIC, time_levels=True)
        lrn = dg.bfs(int(root), policy=A.SWITCH_LEARNED, selector=sel, time_levels=True)
        np.testing.assert_array_equal(h["level"], lrn["level"])
        np.testing.assert_array_equal(h["parent"], lrn["parent"])
        t_heur += float(h["level_ms"].sum())
        t_learn += float(lrn["level_ms"].sum())
    assert O.validate(rp, col, int(root), lrn["level"], lrn["parent"]) == 0
    print(f"kernel time per 6 BFS: heuristic {t_heur:.3f} ms, learned {t_learn:.3f} ms")
    assert t_learn < 0.8 * t_heur
    dg.close()
