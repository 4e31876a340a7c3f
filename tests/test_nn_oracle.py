"""Pins the numpy restatement of the learned shader's network and FP64 shader (oracle/nn_ref.py)."""
import numpy as np

from oracle import nn_ref as nn


def _random_params(rng, scale=0.3):
    return rng.normal(0, scale, nn.N_PARAMS)


def test_float64_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = _random_params(rng)
    x = rng.normal(0, 1, (12, nn.N_IN))
    t = rng.normal(0, 1, (12, nn.N_OUT))
    g, loss = nn.gradient(p, x, t, bf=False)
    idx = rng.choice(nn.N_PARAMS, 60, replace=False)
    idx = np.concatenate([idx, [nn.OFF_B1 + 3, nn.OFF_B2 + 5, nn.OFF_B3 + 1]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = np.sum((nn.forward(pp, x, bf=False)[0] - t) ** 2)
        lm = np.sum((nn.forward(pm, x, bf=False)[0] - t) ** 2)
        num = (lp - lm) / (2 * h)
        assert abs(num - g[i]) <= 1e-5 * max(1.0, abs(num)), (i, num, g[i])
    assert abs(loss - np.sum((nn.forward(p, x, bf=False)[0] - t) ** 2)) < 1e-9


def test_forward_single_path_closed_form():
    """One active path: unit 0 of layer 1 reads input 0 (+bias), unit 0 of layer 2 copies it, output 1 = 2 h + 0.25."""
    w1, b1 = np.zeros((64, 24)), np.zeros(64)
    w2, b2 = np.zeros((64, 64)), np.zeros(64)
    w3, b3 = np.zeros((3, 64)), np.zeros(3)
    w1[0, 0], b1[0] = 1.5, -0.5
    w2[0, 0] = 1.0
    w3[1, 0], b3[1] = 2.0, 0.25
    p = nn.pack(w1, b1, w2, b2, w3, b3)
    x = np.zeros((4, 24))
    x[:, 0] = [-1.0, 0.0, 0.5, 2.0]
    y, _ = nn.forward(p, x, bf=False)
    want = 2.0 * np.maximum(1.5 * x[:, 0] - 0.5, 0) + 0.25
    assert np.allclose(y[:, 1], want) and np.all(y[:, 0] == 0) and np.all(y[:, 2] == 0)


def test_bf16_rounding_ties_to_even():
    one = 1.0
    step = 2.0 ** -7  # bf16 spacing at 1.0
    vals = np.array([one, one + step / 2, one + 3 * step / 2, one + step * 0.49, -2.0 - step / 2], np.float32)
    got = nn.bf16(vals)
    assert got.tolist() == [1.0, 1.0, one + 2 * step, 1.0, -2.0]


def test_features_of_axis_vectors():
    rec = np.zeros((2, nn.REC_FLOATS), np.float32)
    rec[0, nn.REC_N:nn.REC_N + 3] = (0, 0, 1)
    rec[0, nn.REC_V:nn.REC_V + 3] = (0, 0, 1)
    rec[1, nn.REC_N:nn.REC_N + 3] = (0, 1, 0)
    rec[1, nn.REC_V:nn.REC_V + 3] = (1, 0, 0)
    rec[:, nn.REC_ROUGH], rec[:, nn.REC_METAL] = 0.25, 0.75
    rec[:, nn.REC_ALBEDO:nn.REC_ALBEDO + 3] = (0.1, 0.2, 0.3)
    f = nn.features(rec)
    assert f[0, 6] == 1.0 and np.allclose(f[0, 12:15], (0, 0, 1))   # N.V = 1, R = N
    assert f[1, 6] == 0.0 and np.allclose(f[1, 12:15], (-1, 0, 0))  # N.V = 0, R = -V
    assert np.allclose(f[:, 7:12], [[0.25, 0.75, 0.1, 0.2, 0.3]] * 2)
    assert np.allclose(f[1, 15:18], np.sin(np.pi * np.array([-1, 0, 0])), atol=1e-6)
    assert np.allclose(f[0, 21:24], np.sin(4 * np.pi * np.array([0, 0, 1])), atol=1e-6)


def test_fp64_shader_closed_forms():
    """Uniform cubes make the cube lookups constants: the colour is the split-sum formula itself."""
    irr = [np.full((4, 4, 4), (100, 150, 200, 255), np.uint8) for _ in range(6)]
    pref = [[np.full((8 >> k, 8 >> k, 4), (60, 90, 120, 255), np.uint8) for _ in range(6)] for k in range(4)]
    rng = np.random.default_rng(3)
    n = 50
    rec = np.zeros((n, nn.REC_FLOATS))
    N = rng.normal(size=(n, 3)); N /= np.linalg.norm(N, axis=1, keepdims=True)
    V = rng.normal(size=(n, 3)); V /= np.linalg.norm(V, axis=1, keepdims=True)
    # the shader reads the exact-input fields (and V); the network-input fields get other values
    rec[:, nn.REC_XN:nn.REC_XN + 3], rec[:, nn.REC_V:nn.REC_V + 3] = N, V
    rec[:, nn.REC_XALBEDO:nn.REC_XALBEDO + 3] = rng.uniform(0, 1, (n, 3))
    rec[:, nn.REC_XROUGH], rec[:, nn.REC_XMETAL], rec[:, nn.REC_AO] = rng.uniform(0, 1, n), rng.uniform(0, 1, n), 1.0
    rec[:, nn.REC_N:nn.REC_N + 3] = -N
    rec[:, nn.REC_ALBEDO:nn.REC_ALBEDO + 3] = 1 - rec[:, nn.REC_XALBEDO:nn.REC_XALBEDO + 3]
    rec[:, nn.REC_ROUGH], rec[:, nn.REC_METAL] = 1 - rec[:, nn.REC_XROUGH], 1 - rec[:, nn.REC_XMETAL]
    got = nn.pbr_ibl_f64(rec, irr, pref)
    alb = rec[:, nn.REC_XALBEDO:nn.REC_XALBEDO + 3] ** 2.2
    m, r = rec[:, nn.REC_XMETAL][:, None], rec[:, nn.REC_XROUGH][:, None]
    nv = np.maximum(np.sum(N * V, 1), 0)[:, None]
    F0 = 0.04 * (1 - m) + alb * m
    F = F0 + (np.maximum(1 - r, F0) - F0) * (1 - nv) ** 5
    c = np.array([1.0, -0.0275, -0.572, 0.022]) * r + np.array([0.0, 0.0425, 1.04, -0.04])
    c[:, 0] = 1 - r[:, 0]
    a004 = np.minimum(c[:, 0] ** 2, 2 ** (-9.28 * nv[:, 0])) * c[:, 0] + c[:, 1]
    A = -1.04 * a004 + c[:, 2]
    B = (1.04 * a004 + c[:, 3]) * np.clip(50 * F[:, 1], 0, 1)
    want = ((1 - F) * (1 - m) * np.array([100, 150, 200]) / 255 * alb
            + np.array([60, 90, 120]) / 255 * (F * A[:, None] + B[:, None])) ** (1 / 2.2)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
