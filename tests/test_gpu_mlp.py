"""GPU parity of the learned score's kernels against the numpy restatement
(oracle/mlp_ref.py), and of its inputs and its use inside Analysis."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_ref as M
from oracle.oracle import OracleTree
from remobench_amd import lib as L
from remobench_amd.score_model import ScoreMLP

from .cases import lidar_scene
from .helpers import keyed
from .test_gpu_construct import build_hip, build_oracle


def _theta(seed, scale=0.2):
    return np.random.default_rng(seed).normal(0, scale, M.PARAMS).astype(np.float32)


@pytest.mark.gpu
def test_forward_matches_numpy(gpu):
    theta = _theta(1)
    x = np.random.default_rng(2).normal(0, 1, (1000, M.IN)).astype(np.float32)
    m = ScoreMLP(seed=3, max_batch=1024)
    m.set_params(theta)
    got = m.forward(x)
    # BF16 weights and BF16 layer inputs on the matrix cores, FP32 accumulation
    want = M.forward_infer(theta, x)
    err = np.abs(got - want)
    # summation order and erf differ in the last bits; where such a difference
    # flips an activation's BF16 rounding the output moves by ~1e-4
    assert np.mean(err <= 1e-5 + 1e-5 * np.abs(want)) > 0.9, np.quantile(err, [0.5, 0.9, 0.99])
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-3)
    # and within BF16's precision of the FP32-activation network the training kernels run
    np.testing.assert_allclose(got, M.forward(M.bf16(theta), x), rtol=0, atol=3e-2 * np.abs(want).max())
    np.testing.assert_array_equal(m.params(), theta)  # the FP32 master is kept as given


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 31, 100, 8192])
def test_gradient_matches_numpy(gpu, n):
    theta = _theta(4)
    rng = np.random.default_rng(5)
    x = rng.normal(0, 1, (n, M.IN)).astype(np.float32)
    t = rng.normal(0, 1, (n, M.OUT)).astype(np.float32)
    m = ScoreMLP(seed=6, max_batch=8192)
    m.set_params(theta)
    loss, g = m.gradient(x, t)
    want_loss, want_g = M.loss_and_grad(M.bf16(theta), x, t, dtype=np.float64)
    assert loss == pytest.approx(want_loss, rel=1e-5)
    np.testing.assert_allclose(g, want_g, rtol=1e-3, atol=1e-5 * np.abs(want_g).max())


@pytest.mark.gpu
def test_adam_steps_match_numpy(gpu):
    theta = _theta(7)
    rng = np.random.default_rng(8)
    x = rng.normal(0, 1, (500, M.IN)).astype(np.float32)
    t = rng.normal(0, 1, (500, M.OUT)).astype(np.float32)
    m = ScoreMLP(seed=9, max_batch=512)
    m.set_params(theta)
    bx, bt = L.DeviceBuffer.of(x), L.DeviceBuffer.of(t)
    ref, mm, vv = theta.astype(np.float64), np.zeros(M.PARAMS), np.zeros(M.PARAMS)
    for step in (1, 2, 3):
        _, g = M.loss_and_grad(M.bf16(ref.astype(np.float32)), x, t, dtype=np.float64)
        ref, mm, vv = M.adam(ref, g, mm, vv, step, 1e-3)
        m.train_step_device(bx.ptr, bt.ptr, len(x), 1e-3)
        # each step moves a parameter by <= ~lr; FP32 vs FP64 gradients differ in low bits
        np.testing.assert_allclose(m.params(), ref, rtol=0, atol=3e-6 * step)
    bx.free()
    bt.free()


@pytest.mark.gpu
def test_training_fits_a_smooth_target(gpu):
    rng = np.random.default_rng(10)
    x = rng.uniform(-1, 1, (8192, M.IN)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) * 0.5, x[:, 1] * x[:, 2]], axis=1).astype(np.float32)
    m = ScoreMLP(seed=11, max_batch=8192)
    bx, bt = L.DeviceBuffer.of(x), L.DeviceBuffer.of(t)
    first = m.train_step_device(bx.ptr, bt.ptr, len(x), 3e-3)
    for _ in range(400):
        last = m.train_step_device(bx.ptr, bt.ptr, len(x), 3e-3)
    bx.free()
    bt.free()
    assert last < 0.1 * first, (first, last)


@pytest.fixture(scope="module")
def built(gpu):
    points, meta = lidar_scene()
    tree, ring = build_hip(points, meta, 262_144, closure_batches=3)
    o = build_oracle(points, meta, 262_144)
    return points, meta, tree, ring, o


@pytest.mark.gpu
def test_features_match_the_oracle_roll_up(built):
    points, meta, tree, ring, o = built
    u = L.SharedUniforms()
    tree.analysis(u, L.REMO_ANALYSIS_ANALYTIC)
    f = tree.features()
    want = o.analyse(u, closure_batches=3)
    nodes_g, nodes_o = tree.nodes(), o.nodes()
    kg, ko = keyed(nodes_g), keyed(nodes_o)
    assert len(f) == len(nodes_g)
    for k, ig in kg.items():
        w, rec = want[ko[k]], nodes_o[ko[k]]
        n = int(w["n"])
        exp = np.zeros(13)
        exp[0] = np.log1p(n) / np.log1p(50_000)
        exp[10] = rec["level"] / 20
        exp[11] = rec["num_voxels"] / 2 ** 21 if rec["first_child"] >= 0 else min(1.0, n / 50_000)
        exp[12] = 1.0  # every batch ingested
        if n >= 1:
            exp[7:10] = 2 * w["mean"]
        if n >= 2:
            tr = w["cov"][0] + w["cov"][3] + w["cov"][5]
            if tr > 0:
                exp[1:7] = w["cov"] / tr
        np.testing.assert_allclose(f[ig], exp, rtol=1e-5, atol=1e-6, err_msg=str(k))


@pytest.mark.gpu
def test_learned_analysis_uses_the_model_for_open_nodes_only(built):
    """The model predicts a change to the analytic values: an open node takes it
    where it exceeds 0.05 in surface variation or 0.1 in planarity, clamped to the
    quantities' ranges; closed nodes and smaller changes keep the analytic values."""
    points, meta, tree, ring, o = built
    u = L.SharedUniforms()
    tree.analysis(u, L.REMO_ANALYSIS_ANALYTIC)
    analytic = tree.analysis_results()
    m = ScoreMLP(seed=1, max_batch=64)
    theta = np.zeros(M.PARAMS, dtype=np.float32)
    # a constant prediction: surface variation +0.1 (> 0.05: taken), planarity -0.05
    theta[-2:] = [0.1, -0.05]
    m.set_params(theta)
    tree.set_score_model(m)
    tree.analysis(u, L.REMO_ANALYSIS_LEARNED)
    r = tree.analysis_results()
    open_ = (r["state"] == 0) & (r["n"] >= 3)
    closed = r["state"] == 2
    assert open_.sum() > 0 and closed.sum() > 0
    assert np.all(r["flags"][open_] & 2)
    d = M.bf16(np.float32([0.1, -0.05]))
    np.testing.assert_array_equal(r["surface_variation"][open_],
                                  np.clip(analytic["surface_variation"][open_] + d[0], 0, np.float32(1 / 3)))
    np.testing.assert_array_equal(r["planarity"][open_], np.clip(analytic["planarity"][open_] + d[1], 0, 1))
    for k in ("linearity", "lambda"):
        np.testing.assert_array_equal(r[k], analytic[k])
    for k in ("surface_variation", "planarity", "score"):
        np.testing.assert_array_equal(r[k][closed], analytic[k][closed])
    assert not np.any(r["flags"][closed] & 2)
    # changes under both thresholds leave every node analytic
    theta[-2:] = [0.04, -0.09]
    m.set_params(theta)
    tree.set_score_model(m)
    tree.analysis(u, L.REMO_ANALYSIS_LEARNED)
    r2 = tree.analysis_results()
    for k in ("surface_variation", "planarity", "linearity", "score", "lambda", "flags"):
        np.testing.assert_array_equal(r2[k], analytic[k])
    # a planarity change alone is enough
    theta[-2:] = [0.0, 0.2]
    m.set_params(theta)
    tree.set_score_model(m)
    tree.analysis(u, L.REMO_ANALYSIS_LEARNED)
    r3 = tree.analysis_results()
    assert np.all(r3["flags"][open_] & 2)
    np.testing.assert_array_equal(r3["planarity"][open_],
                                  np.clip(analytic["planarity"][open_] + M.bf16(np.float32(0.2)), 0, 1))


@pytest.mark.gpu
def test_learned_analysis_predictions_equal_inference_on_the_features(built):
    """Analysis computes each node's inputs and prediction in one kernel: the
    changes it applies are those of batched inference on the features entry
    point's rows, and exactly the open nodes with a large predicted change take
    them."""
    points, meta, tree, ring, o = built
    u = L.SharedUniforms()
    tree.analysis(u, L.REMO_ANALYSIS_ANALYTIC)
    analytic = tree.analysis_results()
    f = tree.features()
    m = ScoreMLP(seed=1, max_batch=len(f))
    theta = _theta(7, 0.1)
    theta[-2:] = 0.0
    m.set_params(theta)
    p0 = m.forward(f)
    # surface variation centred on its threshold (about half the nodes over it), planarity on 0
    theta[-2:] = [0.05 - np.median(p0[:, 0]), -np.median(p0[:, 1])]
    m.set_params(theta)
    pred = m.forward(f)
    tree.set_score_model(m)
    tree.analysis(u, L.REMO_ANALYSIS_LEARNED)
    r = tree.analysis_results()
    open_ = (r["state"] == 0) & (r["n"] >= 3)
    take = open_ & ((np.abs(pred[:, 0]) > np.float32(0.05)) | (np.abs(pred[:, 1]) > np.float32(0.1)))
    assert take.sum() > 10 and (open_ & ~take).sum() > 10
    np.testing.assert_array_equal((r["flags"] & 2) != 0, take)
    np.testing.assert_array_equal(r["surface_variation"][take],
                                  np.clip(analytic["surface_variation"][take] + pred[take, 0], 0, np.float32(1 / 3)))
    np.testing.assert_array_equal(r["planarity"][take], np.clip(analytic["planarity"][take] + pred[take, 1], 0, 1))
    np.testing.assert_array_equal(r["surface_variation"][~take], analytic["surface_variation"][~take])


@pytest.mark.gpu
def test_learned_mode_without_a_model_is_refused(gpu):
    points, meta = lidar_scene(300_000, seed=3)
    tree, ring = build_hip(points, meta, 100_000)
    with pytest.raises(L.RemoError) as e:
        tree.analysis(L.SharedUniforms(), L.REMO_ANALYSIS_LEARNED)
    assert e.value.code == -5
