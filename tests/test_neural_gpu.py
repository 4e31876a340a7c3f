"""The learned PBR shader on the GPU: MFMA forward and gradient against the numpy network, FP64 targets
against the numpy FP64 shader, training gain, and the shader in the pipeline."""
import math

import numpy as np
import pytest

from oracle import nn_ref as nn
from softglrender_amd import neural, scenes
from softglrender_amd import stream as S

pytestmark = pytest.mark.gpu


def _synthetic_records(n, seed):
    rng = np.random.default_rng(seed)
    rec = np.zeros((n, nn.REC_FLOATS), np.float32)
    N = rng.normal(size=(n, 3)); N /= np.linalg.norm(N, axis=1, keepdims=True)
    V = rng.normal(size=(n, 3)); V /= np.linalg.norm(V, axis=1, keepdims=True)
    rec[:, nn.REC_N:nn.REC_N + 3], rec[:, nn.REC_V:nn.REC_V + 3] = N, V
    rec[:, nn.REC_ALBEDO:nn.REC_ALBEDO + 4] = rng.uniform(0, 1, (n, 4))
    rec[:, nn.REC_ROUGH], rec[:, nn.REC_METAL], rec[:, nn.REC_AO] = rng.uniform(0, 1, n), rng.uniform(0, 1, n), 1.0
    rec[:, nn.REC_MATERIAL] = rng.integers(0, 3, n)
    rec[:, nn.REC_TARGET:nn.REC_TARGET + 3] = rng.uniform(0, 1, (n, 3))
    rec[:, nn.REC_PLAIN:nn.REC_PLAIN + 3] = rec[:, nn.REC_TARGET:nn.REC_TARGET + 3]
    return nn.exact_from_inputs(rec)


def _random_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(nn.N_PARAMS)
    p[nn.OFF_W1:nn.OFF_B1] = rng.uniform(-1, 1, nn.N_HID * nn.N_IN) * math.sqrt(6 / 24)
    p[nn.OFF_B1:nn.OFF_W2] = rng.uniform(-0.2, 0.2, nn.N_HID)
    p[nn.OFF_W2:nn.OFF_B2] = rng.uniform(-1, 1, nn.N_HID * nn.N_HID) * math.sqrt(6 / 64)
    p[nn.OFF_B2:nn.OFF_W3] = rng.uniform(-0.2, 0.2, nn.N_HID)
    p[nn.OFF_W3:nn.OFF_B3] = rng.uniform(-1, 1, nn.N_OUT * nn.N_HID) * math.sqrt(6 / 64)
    p[nn.OFF_B3:] = rng.uniform(-0.2, 0.2, nn.N_OUT)
    return p.astype(np.float32)


def test_mfma_forward_matches_numpy_network(renderer):
    rec = _synthetic_records(1000, 1)
    p = _random_params(2)
    renderer.neural_records(len(rec))
    renderer.neural_write_records(0, rec)
    renderer.neural_set_weights(p)
    _, _, out = renderer.neural_evaluate(0, len(rec), outputs=True)
    want, _ = nn.forward(p.astype(np.float64), nn.features(rec), bf=True)
    err = np.abs(out - want)
    print(f"[gpu] forward: max |gpu - numpy| {err.max():.2e}, mean {err.mean():.2e}", flush=True)
    # bf16 operands, FP32 accumulation on both sides: the summation order differs, and where a hidden FP32
    # activation lies within an ulp of a bf16 rounding boundary the two sides' bf16 operands for the next layer
    # differ by one bf16 ulp (2^-8 relative), which moves that record's output by up to about 1e-2
    rows = err.max(1)
    assert err.mean() < 1e-5 and np.mean(rows > 1e-4) < 0.01 and err.max() < 1e-2


def test_mfma_gradient_matches_numpy_network(renderer):
    rec = _synthetic_records(1000, 3)
    p = _random_params(4)
    renderer.neural_records(len(rec))
    renderer.neural_write_records(0, rec)
    renderer.neural_set_weights(p)
    g, loss = renderer.neural_gradient(0, len(rec))
    gw, lw = nn.gradient(p.astype(np.float64), nn.features(rec), rec[:, nn.REC_TARGET:nn.REC_TARGET + 3], bf=True)
    rel = np.linalg.norm(g - gw) / np.linalg.norm(gw)
    print(f"[gpu] gradient: relative L2 error {rel:.2e}, loss {loss:.6f} vs {lw:.6f}", flush=True)
    # as for the forward pass: summation order, plus the odd bf16 operand one ulp apart
    assert abs(loss - lw) <= 1e-4 * lw
    assert rel < 1e-4
    for a, b in ((nn.OFF_W1, nn.OFF_B1), (nn.OFF_B1, nn.OFF_W2), (nn.OFF_W2, nn.OFF_B2), (nn.OFF_B2, nn.OFF_W3),
                 (nn.OFF_W3, nn.OFF_B3), (nn.OFF_B3, nn.N_PARAMS)):
        assert np.linalg.norm(g[a:b] - gw[a:b]) <= 1e-3 * np.linalg.norm(gw[a:b]) + 1e-9, (a, b)


def _small_views(n_cams, seed, materials, normal, w=320, h=180):
    cams = neural.random_cameras(n_cams, seed)
    return neural.scene_views(cams, materials, normal, width=w, height=h, freq=40)


@pytest.fixture(scope="module")
def small_setup():
    albedo, normal, mr = scenes.pbr_textures(256, 3)
    mats = neural.neural_materials(2, 256, 11, base=(albedo, mr))
    return mats, normal


def test_capture_targets_match_numpy_fp64_shader(renderer, small_setup):
    mats, normal = small_setup
    renderer.neural_records(200_000)
    ns = _small_views(2, 5, mats, normal)
    first, count = neural.capture(renderer, ns, 0)
    assert count > 1000
    rec = renderer.neural_read_records(first, count)
    irr, pref = scenes.ibl_maps(3 + 4)
    want = np.clip(nn.pbr_ibl_f64(rec, irr, pref), 0, 1)
    d = np.abs(rec[:, nn.REC_TARGET:nn.REC_TARGET + 3] - want)
    plain = np.abs(rec[:, nn.REC_PLAIN:nn.REC_PLAIN + 3] - rec[:, nn.REC_TARGET:nn.REC_TARGET + 3]) * 255
    print(f"[gpu] capture: {count} records, FP64 target vs numpy max {d.max():.2e}; plain GPU vs FP64 mean "
          f"{plain.mean():.3f}/255 max {plain.max():.2f}/255", flush=True)
    assert d.max() < 1e-6
    # the plain shader truncates to bytes (texels and output): within about a byte of the FP64 value
    assert plain.mean() < 1.0 and np.percentile(plain, 99) < 2.0 and plain.max() < 8.0


def test_training_gains_over_untrained_network(renderer, small_setup):
    mats, normal = small_setup
    renderer.neural_records(2_000_000)
    tr_first, tr_count = neural.capture(renderer, _small_views(8, 21, mats, normal), 0)
    ho_first, ho_count = neural.capture(renderer, _small_views(2, 99, mats, normal), 0)
    assert tr_count > 100_000 and ho_count > 20_000, (tr_count, ho_count)
    renderer.neural_init(7)
    before, plain, _ = renderer.neural_evaluate(ho_first, ho_count)
    losses = renderer.neural_train(tr_first, tr_count, steps=6000, batch=8192, lr_start=1e-2, lr_end=2e-5, seed=3)
    after, _, _ = renderer.neural_evaluate(ho_first, ho_count)
    print(f"[gpu] training: loss {losses[0]:.4f} -> {losses[-1]:.6f}; held-out mean error {before['mean']:.2f} -> "
          f"{after['mean']:.3f}/255, p99 {after['p99']:.2f}/255 (plain shader: mean {plain['mean']:.3f})", flush=True)
    assert losses[-1] < losses[0] / 100
    assert after["mean"] < before["mean"] / 10
    assert after["mean"] < 4.0


def test_neural_draw_writes_the_network_output(renderer, small_setup):
    """A learned-shader draw writes, at every sample of every fragment, exactly the network's output for that
    fragment's inputs (captured from the plain draw of the same view); after training its image is close to
    the plain shader's."""
    mats, normal = small_setup
    renderer.neural_records(2_000_000)
    tr_first, tr_count = neural.capture(renderer, _small_views(8, 21, mats, normal), 0)
    assert tr_count > 100_000
    renderer.neural_init(7)
    renderer.neural_train(tr_first, tr_count, steps=6000, batch=8192, lr_start=1e-2, lr_end=2e-5, seed=3)
    from softglrender_amd.replay import Replayer, parse
    cams = neural.random_cameras(1, 77)
    w, h = 320, 180
    shots = {}
    for sh in (S.SHADING_PBR, S.SHADING_NEURAL_PBR):
        ns = neural.scene_views(cams, mats[:1], normal, width=w, height=h, freq=40, shading=sh)
        ns.scene.stream.read(6, which=1)
        if sh == S.SHADING_PBR:
            first, count = neural.capture(renderer, ns, 0)
        rp = Replayer(renderer, layouts=ns.scene.meta.get("layouts"))
        reads = []
        for op, p in parse(ns.scene.stream.bytes()):
            if op != S.OP_FRAME:
                rp.exec(op, p, reads)
        shots[sh] = reads[-1].reshape(h, w, 4, 4)
    rec = renderer.neural_read_records(first, count)
    _, _, out = renderer.neural_evaluate(first, count, outputs=True)
    rgb = np.trunc(np.clip(out, 0, 1) * np.float32(255)).astype(np.int64)
    a = np.trunc(np.clip(rec[:, nn.REC_ALBEDO + 3], 0, 1) * np.float32(255)).astype(np.int64)
    pix = rec[:, nn.REC_PIXEL].astype(np.int64)
    mask = rec[:, nn.REC_MASK].astype(np.int64)
    samples = shots[S.SHADING_NEURAL_PBR].reshape(-1, 4, 4).astype(np.int64)
    bad = 0
    for s_ in range(4):
        sel = (mask >> s_) & 1 == 1
        got = samples[pix[sel], s_]
        bad += int(np.sum(np.any(got[:, :3] != rgb[sel], axis=1) | (got[:, 3] != a[sel])))
    plain = shots[S.SHADING_PBR].astype(np.float64).mean(2)
    learned = shots[S.SHADING_NEURAL_PBR].astype(np.float64).mean(2)
    cov = np.zeros(h * w, bool)
    cov[pix] = True
    diff = np.abs(plain.reshape(-1, 4)[cov, :3] - learned.reshape(-1, 4)[cov, :3]).mean()
    print(f"[gpu] neural draw: {count} fragments, {bad} samples differ from the network output; "
          f"mean |learned - plain| {diff:.3f}/255", flush=True)
    assert count > 5000 and bad == 0
    assert diff < 10.0


def test_training_is_reproducible(renderer):
    """Each block writes its gradient sums to its own row and the update kernel adds the rows in a fixed
    order (no float atomics), so training from the same weights on the same records repeats bit for bit."""
    rec = _synthetic_records(20_000, 5)
    renderer.neural_records(len(rec))
    renderer.neural_write_records(0, rec)
    runs = []
    for _ in range(2):
        renderer.neural_init(3)
        losses = renderer.neural_train(0, len(rec), steps=50, batch=4096, lr_start=1e-2, lr_end=1e-3, seed=9)
        runs.append((renderer.neural_get_weights(), np.asarray(losses)))
    print(f"[gpu] training twice: loss {runs[0][1][0]:.5f} -> {runs[0][1][-1]:.5f}", flush=True)
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    assert runs[0][1][-1] < runs[0][1][0]


def _shot(renderer, mats, normal, m, shading, cams, w=320, h=180):
    """Material m of `mats` drawn alone (texture ids of list position 0), tagged as material m."""
    from softglrender_amd.replay import Replayer, parse
    ns = neural.scene_views(cams, mats[m:m + 1], normal, width=w, height=h, freq=40, shading=shading)
    ns.scene.stream.read(6, which=1)
    renderer.neural_material(m)
    rp = Replayer(renderer, layouts=ns.scene.meta.get("layouts"))
    reads = []
    for op, p in parse(ns.scene.stream.bytes()):
        if op != S.OP_FRAME:
            rp.exec(op, p, reads)
    return reads[-1]


def test_fallback_materials_draw_with_the_plain_shader(renderer, small_setup):
    """The per-material fallback acts in program selection: an untrained network misses the p99 bound for
    every material; a flagged material's learned-shader draws then write exactly the plain shader's samples,
    while an unflagged material's draws still run the network."""
    mats, normal = small_setup
    renderer.neural_records(400_000)
    first, count = neural.capture(renderer, _small_views(2, 31, mats, normal), 0)
    renderer.neural_init(5)
    table = neural.fallback_table(renderer, first, count, len(mats))
    print(f"[gpu] untrained network p99 per material: {[round(t['network']['p99'], 1) for t in table]}/255",
          flush=True)
    assert all(t["fallback"] for t in table)
    cams = neural.random_cameras(1, 78)
    plain = [_shot(renderer, mats, normal, m, S.SHADING_PBR, cams) for m in range(2)]
    try:
        renderer.neural_fallback([1])
        learned = [_shot(renderer, mats, normal, m, S.SHADING_NEURAL_PBR, cams) for m in range(2)]
        assert np.array_equal(learned[1], plain[1])
        assert not np.array_equal(learned[0], plain[0])
        assert neural.apply_fallback(renderer, table) == [0, 1]
        assert np.array_equal(_shot(renderer, mats, normal, 0, S.SHADING_NEURAL_PBR, cams), plain[0])
    finally:
        renderer.neural_fallback([])
        renderer.neural_material(0)
