"""Pins of the closure oracle (oracle/closure_oracle.c): bf16 rounding, the
subgrid targets against closed forms, the network's feature indexing against
a hand-built network, and its gradient against finite differences."""
import numpy as np
import pytest

from oracle import pyoracle as po

F = po.CLOSURE_FIELDS


def _fine(shape, seed=0, amp=1.0):
    rng = np.random.default_rng(seed)
    return {n: amp * rng.standard_normal(shape) + (2.0 if n in ("r", "t") else 0.0) for n in F}


def test_bf16_round_to_nearest_even():
    one = np.float32(1.0)
    cases = {
        1.0 / 3.0: 0.333984375,
        float(one + np.float32(2.0 ** -8)): 1.0,                          # tie -> even (down)
        float(one + np.float32(3 * 2.0 ** -8)): 1.0 + 2.0 ** -6,          # tie -> even (up)
        float(one + np.float32(2.0 ** -8 + 2.0 ** -20)): 1.0 + 2.0 ** -7,  # above the tie
        -2.5: -2.5,
    }
    for x, want in cases.items():
        assert float(po.lib().orc_bf16(x)) == want, x
        assert float(po.bf16(np.array([x], np.float32))[0]) == want, x


def test_targets_vanish_for_block_constant_fields():
    # constant over each f x f column block and uniform in z: nothing is left
    # below the filter
    f = 2
    rng = np.random.default_rng(1)
    plane = {n: rng.standard_normal((4, 4)) + 3.0 for n in F}
    fine = {n: np.broadcast_to(np.kron(plane[n], np.ones((f, f)))[None], (6, 8, 8)).copy() for n in F}
    c, t = po.closure_targets(fine, f, fmag=0.7)
    assert np.abs(t).max() < 1e-12
    for n in F:
        assert np.allclose(c[n], np.broadcast_to(plane[n], (3, 4, 4)), rtol=0, atol=1e-13)


@pytest.mark.parametrize("f", [2, 4])
def test_targets_closed_forms_for_linear_fields(f):
    nz, ny, nx = 2 * f, f, 2 * f     # two coarse planes: the walls, no z averaging
    shape = (nz, ny, nx)
    i = np.broadcast_to(np.arange(nx, dtype=float), shape).copy()   # fine x index
    var = (f * f - 1) / 12.0                                        # variance of f consecutive integers
    zero = np.zeros(shape)
    fmag = 0.3
    # Reynolds part: u = i, r = 2
    fine = {n: zero.copy() for n in F}
    fine["r"] = 2.0 + zero
    fine["u"] = i
    _, t = po.closure_targets(fine, f, fmag)
    assert np.allclose(t[0], 2.0 * var) and np.abs(t[1:]).max() < 1e-12
    # Maxwell part: Bx = i -> tau_xx = -fmag var / 2, tau_yy = tau_zz = +fmag var / 2
    fine = {n: zero.copy() for n in F}
    fine["bx"] = i
    _, t = po.closure_targets(fine, f, fmag)
    assert np.allclose(t[0], -0.5 * fmag * var)
    assert np.allclose(t[1], 0.5 * fmag * var) and np.allclose(t[2], 0.5 * fmag * var)
    assert np.abs(t[3:]).max() < 1e-12
    # EMF: w = i, Bx = i -> eps_y = F(w Bx) - F(w) F(Bx) = var; Maxwell of Bx as above
    fine["w"] = i
    _, t = po.closure_targets(fine, f, fmag)
    assert np.allclose(t[7], var) and np.abs(t[6]).max() < 1e-12 and np.abs(t[8]).max() < 1e-12
    assert np.allclose(t[2], 0.5 * fmag * var)   # R_zz = var enters with F(r) = 0
    # heat flux: u = i, T = i -> q_x = var
    fine = {n: zero.copy() for n in F}
    fine["u"] = i
    fine["t"] = i
    _, t = po.closure_targets(fine, f, 0.0)
    assert np.allclose(t[9], var) and np.abs(t[10:]).max() < 1e-12


def test_coarse_fields_horizontal_block_mean_and_z_vertices():
    # horizontally the coarse fields are the f x f block means of each fine
    # plane; in z the wall planes are the fine wall planes, and a linear
    # profile a + b z is reproduced exactly at the coarse planes K / (cz - 1)
    nz, ny, nx, f = 24, 8, 12, 4
    cz = nz // f
    fine = _fine((nz, ny, nx), seed=2)
    c, _ = po.closure_targets(fine, f, 1.0)
    for n in F:
        for k, K in ((0, 0), (nz - 1, cz - 1)):
            want = fine[n][k].reshape(ny // f, f, nx // f, f).mean(axis=(1, 3))
            assert np.allclose(c[n][K], want, rtol=0, atol=1e-13)
    z = np.arange(nz) / (nz - 1.0)
    Z = np.arange(cz) / (cz - 1.0)
    x = np.arange(nx) / nx
    lin = {n: (1.5 + 0.7 * z)[:, None, None] + 0 * x[None, None, :] + np.zeros((nz, ny, nx)) for n in F}
    c, _ = po.closure_targets(lin, f, 1.0)
    assert np.abs(c["t"] - (1.5 + 0.7 * Z)[:, None, None]).max() < 1e-14


@pytest.mark.parametrize("nz,cz", [(24, 6), (384, 96), (40, 7), (13, 13)])
def test_z_weights_are_a_centred_window(nz, cz):
    """Each coarse plane's weights sum to 1; an interior plane's weights
    centre on Z_K (first moment) and give the window mean of the fine
    linear interpolant of z^2 (second moment, against dense sampling)."""
    z = np.arange(nz) / (nz - 1.0)
    H = 1.0 / (cz - 1)
    for K in range(cz):
        w = np.array([po.lib().orc_zweight(nz, cz, K, k) for k in range(nz)])
        assert abs(w.sum() - 1.0) < 1e-14
        if K in (0, cz - 1):
            assert w[0 if K == 0 else -1] == 1.0
            continue
        Zk = K * H
        assert abs((w * z).sum() - Zk) < 1e-14
        # the interpolant of z^2 integrated over the window, by fine sub-sampling
        a, b = Zk - H / 2, Zk + H / 2
        zz = np.linspace(a, b, 200001)
        want = np.trapezoid(np.interp(zz, z, z * z), zz) / H
        assert abs((w * z * z).sum() - want) < 1e-9


def _norm(seed=3):
    rng = np.random.default_rng(seed)
    mu = rng.standard_normal(8) * 0.1
    sd = 0.5 + rng.random(8)
    tmu = rng.standard_normal(12) * 0.1
    tsd = 0.5 + rng.random(12)
    return np.concatenate([mu, sd, tmu, tsd])


def _gelu(z):
    # tanh-form GELU
    return 0.5 * z * (1.0 + np.tanh(np.sqrt(2.0 / np.pi) * (z + 0.044715 * z ** 3)))


@pytest.mark.parametrize("nb,fld", [(0, 0), (13, 2), (26, 7), (4, 5), (22, 1)])
def test_forward_feature_indexing(nb, fld):
    """A network that passes one feature through: y_0 = w3 gelu(w2 gelu(x_f))."""
    nz, ny, nx = 4, 3, 5
    rng = np.random.default_rng(nb)
    coarse = {n: rng.standard_normal((nz, ny, nx)) for n in F}
    norm = _norm()
    p = np.zeros(po.N_PARAM)
    p[(nb * 8 + fld) * 64 + 0] = 1.0              # W1[feature][0]
    p[216 * 64 + 64 + 0] = 0.8                    # W2[0][0]
    p[216 * 64 + 64 + 4096 + 64 + 0] = 1.5        # W3[0][0]
    y = po.mlp_forward(coarse, p, norm, round_bf16=False)
    dz, dy, dx = nb // 9 - 1, (nb // 3) % 3 - 1, nb % 3 - 1
    k = np.clip(np.arange(nz) + dz, 0, nz - 1)
    src = coarse[F[fld]][k][:, (np.arange(ny) + dy) % ny][:, :, (np.arange(nx) + dx) % nx]
    x = ((src - norm[fld]) / norm[8 + fld]).astype(np.float32).astype(np.float64)
    want = 1.5 * _gelu(0.8 * _gelu(x))
    assert np.allclose((y[0] - norm[16]) / norm[28], want, rtol=1e-12, atol=1e-14)
    for o in range(1, 12):
        assert np.allclose(y[o], norm[16 + o])


def test_gradient_matches_finite_differences():
    nz, ny, nx = 4, 4, 4
    fine = _fine((2 * nz, 2 * ny, 2 * nx), seed=5)
    snaps = [po.closure_targets(fine, 2, 0.5)]
    norm = _norm()
    rng = np.random.default_rng(7)
    p = rng.standard_normal(po.N_PARAM) * 0.2
    cells = rng.integers(0, nz * ny * nx, size=20)
    L0, g = po.mlp_grad(snaps, p, norm, cells, round_bf16=False)
    idx = np.concatenate([rng.integers(0, 13824, 12), 13824 + rng.integers(0, 64, 3),
                          13888 + rng.integers(0, 4096, 8), 17984 + rng.integers(0, 64, 3),
                          18048 + rng.integers(0, 768, 6), 18816 + np.arange(12)])
    h = 1e-5
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        Lp, _ = po.mlp_grad(snaps, pp, norm, cells, round_bf16=False)
        Lm, _ = po.mlp_grad(snaps, pm, norm, cells, round_bf16=False)
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * abs(g[k]) + 1e-8, (k, fd, g[k])


def test_loss_matches_forward():
    nz, ny, nx = 4, 4, 6
    fine = _fine((2 * nz, 2 * ny, 2 * nx), seed=9)
    coarse, tgt = po.closure_targets(fine, 2, 0.5)
    norm = _norm()
    rng = np.random.default_rng(11)
    p = rng.standard_normal(po.N_PARAM) * 0.1
    cells = np.arange(nz * ny * nx)
    for rb in (False, True):
        L, _ = po.mlp_grad([(coarse, tgt)], p, norm, cells, round_bf16=rb)
        y = po.mlp_forward(coarse, p, norm, round_bf16=rb)
        yn = (y - norm[16:28, None, None, None]) / norm[28:40, None, None, None]
        tn = ((tgt - norm[16:28, None, None, None]) / norm[28:40, None, None, None]).astype(np.float32)
        assert np.isclose(L, np.mean((yn - tn) ** 2), rtol=1e-10)
