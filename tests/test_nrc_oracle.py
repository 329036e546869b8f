"""The float64 radiance-cache model (oracle/nrc_oracle.py) pinned by
closed forms and central finite differences."""
import numpy as np

from oracle import nrc_oracle as N


def _batch(rng, n):
    p = rng.uniform(-1, 1, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    nn /= np.linalg.norm(nn, axis=1, keepdims=True)
    return np.concatenate([p, d, nn, rng.random((n, 1)), rng.random((n, 3))], 1)


def test_encoding_closed_forms():
    f = np.zeros((1, 13))
    f[0, 0:3] = [0.25, 0.5, 0.0]   # normalised by bounds [0,1]^3
    f[0, 3:6] = [0.0, 0.0, 1.0]
    f[0, 6:9] = [1.0, 0.0, 0.0]
    f[0, 9] = 0.3
    f[0, 10:13] = [0.1, 0.2, 0.3]
    e = N.encode(f, [0, 0, 0], [1, 1, 1])[0]
    assert np.isclose(e[0], np.sin(np.pi * 0.25)) and np.isclose(e[1], np.cos(np.pi * 0.25))
    assert np.isclose(e[2], np.sin(2 * np.pi * 0.25)) and np.isclose(e[24], np.sin(np.pi * 0.5))
    assert np.isclose(e[72], 0.5 / np.sqrt(np.pi))            # Y00
    assert np.isclose(e[74], np.sqrt(3 / (4 * np.pi)))         # Y10 at z = 1
    assert np.isclose(e[78], np.sqrt(5 / (16 * np.pi)) * 2)    # Y20 at z = 1: (3z^2-1) * 0.3154
    assert np.isclose(e[91], -np.sqrt(3 / (4 * np.pi)))        # normal x = 1 -> -0.4886 x
    assert e[104] == 0.3 and list(e[105:108]) == [0.1, 0.2, 0.3] and e[108] == 1 and not e[109:].any()
    # SH basis is orthonormal: Monte Carlo of Y_i Y_j over the sphere ~ delta
    rng = np.random.default_rng(0)
    d = rng.normal(size=(400000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    Y = N.sh4(d)
    G = Y.T @ Y * (4 * np.pi / len(d))
    assert np.allclose(G, np.eye(16), atol=0.02)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    ws = N.init_weights(3)
    x = N.encode(_batch(rng, 24), [-1, -1, -1], [1, 1, 1])
    t = rng.random((24, 3)) * 2
    loss, grads = N.gradients(ws, x, t)
    flat = N.flatten(ws)
    gflat = N.flatten(grads)
    idx = rng.choice(np.nonzero(gflat)[0], 60, replace=False)
    h = 1e-6
    for i in idx:
        wp = flat.copy()
        wp[i] += h
        wm = flat.copy()
        wm[i] -= h
        # relative L2 treats the denominator's pred as a constant: evaluate with
        # the stop-gradient pred fixed at the unperturbed forward
        pred0, _ = N.forward(ws, x)
        den = pred0[:, :3] ** 2 + N.REL_EPS

        def f(v):
            p, _ = N.forward(N.unflatten(v), x)
            return np.mean((p[:, :3] - t) ** 2 / den)

        fd = (f(wp) - f(wm)) / (2 * h)
        assert abs(fd - gflat[i]) <= 1e-5 * max(1e-3, abs(fd)), (i, fd, gflat[i])


def test_adam_first_step_is_signed_lr():
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    p, m, v = N.adam(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, lr=1e-2)
    assert np.allclose(p, [-1e-2, 1e-2, -1e-2, 0.0])


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159265, 1e-40], np.float32)
    r = N.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0078125
    assert abs(r[3] + 3.140625) < 1e-7
