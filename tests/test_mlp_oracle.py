"""Pins for the oracle of the neural shadow-visibility MLP (oracle/oracle_mlp.hpp):
finite differences of the exact FP64 function, closed forms, and a numpy forward pass
written from the parameter layout and the encoding definition."""
import numpy as np

import oracle_lib as O

CENTER = np.array([0.0, 40.0, 0.0], np.float32)
INV_HALF = np.array([1 / 150.0, 1 / 150.0, 1 / 150.0], np.float32)


def samples(n, seed):
    rng = np.random.default_rng(seed)
    s = np.zeros((n, 8), np.float32)
    s[:, :3] = rng.uniform(-140, 140, (n, 3)) + CENTER
    d = rng.normal(size=(n, 3))
    s[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    return s


def numpy_forward(P, smp):
    P = P.astype(np.float64)
    x = np.concatenate([(smp[:, :3].astype(np.float64) - CENTER) * INV_HALF, smp[:, 4:7].astype(np.float64)], 1)
    k = 2.0 ** np.arange(8)
    ang = np.pi * x[:, :, None] * k[None, None, :]                 # (n, 6, 8)
    enc = np.stack([np.sin(ang), np.cos(ang)], -1).reshape(len(smp), 96)  # f = (dim*8+k)*2+s
    h = enc
    dims = [96, 64, 64, 64]
    for L, (w, b) in enumerate([("W1", "b1"), ("W2", "b2"), ("W3", "b3"), ("W4", "b4")]):
        W = P[O.MLP_OFF[w]:O.MLP_OFF[w] + 64 * dims[L]].reshape(64, dims[L])
        h = np.maximum(h @ W.T + P[O.MLP_OFF[b]:O.MLP_OFF[b] + 64], 0.0)
    z = h @ P[O.MLP_OFF["w5"]:O.MLP_OFF["w5"] + 64] + P[O.MLP_OFF["b5"]]
    return 1.0 / (1.0 + np.exp(-z))


def test_forward_matches_layout_definition():
    P = O.mlp_init(1)
    P[O.MLP_OFF["b1"]:O.MLP_OFF["b1"] + 64] = 0.05
    smp = samples(300, 2)
    ref = numpy_forward(P, smp)
    assert np.allclose(O.mlp_forward(P, CENTER, INV_HALF, smp, fp64=True), ref, atol=1e-6)
    # the bf16 form stays close to the exact function
    assert np.abs(O.mlp_forward(P, CENTER, INV_HALF, smp) - ref).max() < 0.03


def test_closed_forms():
    smp = samples(16, 3)
    P = np.zeros(O.MLP_PARAMS, np.float32)
    P[O.MLP_OFF["b5"]] = 0.7
    assert np.allclose(O.mlp_forward(P, CENTER, INV_HALF, smp, fp64=True), 1 / (1 + np.exp(-0.7)))
    # one path: h1[0] = relu(cos(pi x0) + 1) = cos(pi x0) + 1 >= 0 -> ... -> z = 2 * h4[0]
    P[:] = 0
    P[O.MLP_OFF["W1"] + 0 * 96 + 1] = 1.0      # feature 1 = cos(pi * x0)
    P[O.MLP_OFF["b1"] + 0] = 1.0
    for w in ("W2", "W3", "W4"):
        P[O.MLP_OFF[w] + 0] = 1.0              # h_L[0] = h_{L-1}[0]
    P[O.MLP_OFF["w5"]] = 2.0
    x0 = (smp[:, 0].astype(np.float64) - CENTER[0]) * INV_HALF[0]
    z = 2.0 * (np.cos(np.pi * x0) + 1.0)
    assert np.allclose(O.mlp_forward(P, CENTER, INV_HALF, smp, fp64=True), 1 / (1 + np.exp(-z)), atol=1e-6)


def test_gradient_finite_differences():
    P = O.mlp_init(4).astype(np.float64) * 0.7
    P[O.MLP_OFF["b1"]:O.MLP_OFF["b1"] + 64] = 0.02
    smp = samples(24, 5)
    tgt = (np.random.default_rng(6).random(24) < 0.5).astype(np.float32)
    g, loss = O.mlp_grad(P, CENTER, INV_HALF, smp, tgt, fp64=True)
    rng = np.random.default_rng(7)
    idx = np.concatenate([rng.choice(O.MLP_PARAMS, 60, replace=False),
                          [O.MLP_OFF["b5"], O.MLP_OFF["w5"] + 3, O.MLP_OFF["b1"] + 5, O.MLP_OFF["W1"] + 17]])
    h = 1e-6
    bad = 0
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (O.mlp_grad(Pp, CENTER, INV_HALF, smp, tgt, fp64=True)[1] -
              O.mlp_grad(Pm, CENTER, INV_HALF, smp, tgt, fp64=True)[1]) / (2 * h)
        if abs(fd - g[i]) > 1e-6 + 1e-4 * abs(fd):
            bad += 1
    assert bad <= 1, bad  # at most one parameter sitting on a ReLU kink


def test_bf16_gradient_close_to_exact():
    P = O.mlp_init(8)
    smp = samples(256, 9)
    tgt = (np.random.default_rng(10).random(256) < 0.4).astype(np.float32)
    g64, l64 = O.mlp_grad(P.astype(np.float64), CENTER, INV_HALF, smp, tgt, fp64=True)
    g32, l32 = O.mlp_grad(P, CENTER, INV_HALF, smp, tgt)
    assert abs(l64 - l32) < 0.02
    assert np.linalg.norm(g32 - g64) / np.linalg.norm(g64) < 0.05


def test_feature_sincospi_accuracy_and_closed_forms():
    """The FP32 sin(pi v) / cos(pi v) that the kernels and the oracle's bf16 form evaluate:
    within 1.5e-7 of the float64 functions over the feature range |v| <= 64, exact at the
    multiples of 1/2, and odd / even in v."""
    L = O.lib()
    rng = np.random.default_rng(5)
    v = np.concatenate([np.linspace(-64, 64, 400001), rng.uniform(-64, 64, 100000),
                        np.arange(-256, 257) / 4.0,                     # quadrant boundaries and 1/4 ties
                        np.arange(-256, 257) / 4.0 + 1e-6, [0.0, -0.0]]).astype(np.float32)
    s = np.zeros_like(v)
    c = np.zeros_like(v)
    L.orc_sincospi_f32(O.ptr(v), len(v), O.ptr(s), O.ptr(c))
    vd = v.astype(np.float64)
    es = np.abs(s - np.sin(np.pi * vd)).max()
    ec = np.abs(c - np.cos(np.pi * vd)).max()
    assert es < 1.5e-7 and ec < 1.5e-7, (es, ec)
    half = np.arange(-128, 129) / 2.0
    hv = half.astype(np.float32)
    hs = np.zeros_like(hv)
    hc = np.zeros_like(hv)
    L.orc_sincospi_f32(O.ptr(hv), len(hv), O.ptr(hs), O.ptr(hc))
    assert np.array_equal(hs, np.round(np.sin(np.pi * half)).astype(np.float32))
    assert np.array_equal(hc, np.round(np.cos(np.pi * half)).astype(np.float32))
    w = rng.uniform(0, 64, 20000).astype(np.float32)
    a, b, a2, b2 = (np.zeros_like(w) for _ in range(4))
    nw = -w
    L.orc_sincospi_f32(O.ptr(w), len(w), O.ptr(a), O.ptr(b))
    L.orc_sincospi_f32(O.ptr(nw), len(w), O.ptr(a2), O.ptr(b2))
    assert np.array_equal(a2, -a) and np.array_equal(b2, b)
