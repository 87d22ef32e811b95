"""Learned densification inside the trainer (needs a gfx950 GPU).

- mode 1 (learn) applies the reference's decisions unchanged: the model after training is
  bitwise the plain run's, while decisions are logged, labelled and fit;
- the scorer features of the resident model equal the oracle's restatement (binary32; the
  GPU's log2 may differ from glibc's by an ulp, hence 1e-6);
- the lineage planes: rows created at an event are born at that iteration and descend from
  logged decisions;
- mode 2 (learned) trains with the scorer's decisions and reaches a PSNR well above the
  starting point; the gain over the plain method at matched Gaussian count is pinned at
  the densifying configuration's size in test_gpu_learned_gain.py.
"""
import math
import sys

import numpy as np
import pytest

import oracle_lib as O
from helpers import ground_truth, init_points, ring_cameras
from splatrs_amd import abi, scenes
from splatrs_amd import records as rec

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = abi.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def scene(ctx):
    rng = np.random.default_rng(300)
    gt = ground_truth(rng, 60)
    cams = ring_cameras(8, 64, 48, 2.6)
    views = [abi.to_rgb8(ctx.render(gt, c, rec.proj_opts(1), (0, 0, 0))) for c in cams]
    xyz, rgb = init_points(gt, rng)
    return cams, views, xyz[:20], rgb[:20]


def train_cfg(**kw):
    base = dict(iterations=700, densify=1, densify_from_iter=40, densify_until_iter=560,
                densification_interval=40, opacity_reset_interval=300,
                densify_grad_threshold=0.0005, sh_degree=1, sh_degree_interval=150,
                position_lr_max_steps=700, log_interval=0, seed=21, ordered_backward=1)
    base.update(kw)
    return abi.TrainConfig.default(**base)


def make_trainer(ctx, scene):
    cams, views, xyz, rgb = scene
    tr = abi.Trainer(ctx, ctx.from_points(xyz, rgb))
    for c, v in zip(cams, views):
        tr.add_view(c, v)
    return tr


def learn_cfg(mode):
    return rec.learn_config(mode=mode, samples_per_event=2048, outcome_delay=120, eval_views=4,
                            fit=rec.scorer_fit_config(epochs=3, batch=256, lr=2e-3, seed=4))


def test_learn_mode_keeps_the_reference_decisions(ctx, scene):
    extent = scenes.scene_extent(scene[0])
    plain = make_trainer(ctx, scene)
    say("plain run")
    plain.train(extent, train_cfg())
    learner = make_trainer(ctx, scene)
    scorer = abi.Scorer(ctx, seed=7)
    learner.attach_scorer(scorer, learn_cfg(1))
    say("learn run")
    learner.train(extent, train_cfg())
    a, b = plain.download(moments=True), learner.download(moments=True)
    for k in a:
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    st = learner.learn_stats()
    say(f"logged {st['logged']} resolved {st['resolved']} positives {st['positives']} "
        f"updates {st['updates']} loss {st['last_loss']:.4f} events {st['events']}")
    assert st["logged"] > 0 and st["resolved"] > 0 and st["updates"] >= 1
    assert 0 < st["positives"] <= st["resolved"]
    f, y, w = learner.learn_samples()
    assert np.isfinite(f).all() and (w > 0).all() and (y < 8).all()
    lin = learner.lineage()
    n = len(b["params"])
    born = lin["birth"] > 0
    assert born.any() and (lin["birth"] <= 560).all()
    tagged = lin["origin"] != 0xFFFFFFFF
    assert tagged.any() and (lin["origin"][tagged] < st["logged"]).all()
    assert (lin["birth"][~born] == 0).all() and len(lin["birth"]) == n


def test_features_match_oracle(ctx, scene):
    cams, views, xyz, rgb = scene
    extent = scenes.scene_extent(cams)
    tr = make_trainer(ctx, scene)
    tr.attach_scorer(abi.Scorer(ctx, seed=1), learn_cfg(1))
    for it in range(1, 31):
        lr = rec.learning_rates(rec.position_lr(it, extent))
        tr.step(it % len(cams), rec.step_params(1, lr, accumulate_stats=True))
    # strong steps use the threshold of the last splat_train (default 2e-4 before any)
    d = tr.download(stats=True)
    lin = tr.lineage()
    got = tr.features(extent, 0.0002, 0.01, 30)
    want = O.scorer_features(d["params"], d["grad_accum"], d["denom"], d["max_radii"],
                             lin["max_grad"], lin["strong"], lin["birth"], 0.0002,
                             np.float32(math.log(np.float32(0.01) * np.float32(extent))), 30)
    assert np.abs(got - want).max() <= 1e-6, np.abs(got - want).max()
    assert (lin["strong"] <= d["denom"]).all()
    assert (lin["max_grad"] * d["denom"] >= d["grad_accum"] * (1 - 1e-6)).all()


def test_learned_mode_trains(ctx, scene):
    cams, views, xyz, rgb = scene
    extent = scenes.scene_extent(cams)
    teacher = make_trainer(ctx, scene)
    scorer = abi.Scorer(ctx, seed=3)
    teacher.attach_scorer(scorer, learn_cfg(1))
    teacher.train(extent, train_cfg())
    assert teacher.learn_stats()["updates"] >= 1
    student = make_trainer(ctx, scene)
    student.attach_scorer(scorer, learn_cfg(2))
    say("learned run")
    student.train(extent, train_cfg())

    def psnr(tr):
        tot = 0.0
        for i, (c, v) in enumerate(zip(cams, views)):
            tot += abi.psnr(tr.render(i, 1), v.astype(np.float32) / np.float32(255.0))
        return tot / len(cams)
    p_plain, p_learned = psnr(teacher), psnr(student)
    n_plain, n_learned = teacher.info()[0], student.info()[0]
    say(f"plain {p_plain:.2f} dB / {n_plain} Gaussians, learned {p_learned:.2f} dB / "
        f"{n_learned} Gaussians")
    assert p_learned > 20.0 and n_learned > 20
