"""CPU tests that pin the FP64 oracle before anything is checked against it.

Each check is independent of the oracle's own formula: known answers, the
reference's Random.h compiled into oracle/_ref, numpy/torch golden vectors
(tests/golden/make_golden.py), closed forms and invariants.
"""
import numpy as np
import pytest

import oracle as O
from golden.make_golden import KERNELS, WORLDS

SHAPES = {"shell": O.SHAPE_SHELL, "ring": O.SHAPE_RING}
CORES = {"gaussian": O.CORE_GAUSSIAN, "exp": O.CORE_EXP}


def as_oracle_kernel(k):
    d = dict(k)
    d["shape"] = SHAPES[k["shape"]]
    d["core"] = CORES[k["core"]]
    return d


def as_oracle_world(w):
    d = dict(w)
    d["kernels"] = [as_oracle_kernel(k) for k in w["kernels"]]
    return d


# ------------------------------------------------------------------ hashes
def test_hash_known_answers():
    # independent transcription of the integer ops (values quoted with the
    # hot-path survey of the reference)
    assert O.pcg_hash(0) == 129708002
    assert O.pcg_hash(1) == 2831084092
    s = O.derive_seed(7777, 0x4C454E31)
    assert s == 3628426083
    assert O.lib().lo_hash3(0, 0, 503, s) == 3478061327
    assert abs(O.rand01(0, 0, 503, s) - 0.8097992539405823) < 1e-15
    assert O.module_seed(7777) == s


def test_hash_golden(golden):
    for v, h in zip(golden["hash_in"], golden["hash_out"]):
        assert O.pcg_hash(int(v)) == int(h)
    xy = golden["rand_xy"]
    for i in range(xy.shape[1]):
        assert O.rand01(int(xy[0, i]), int(xy[1, i]), 503, 3628426083) == golden["rand_out"][i]


def test_hash_matches_reference_random_h():
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(5)
    for v in list(range(64)) + [int(x) for x in rng.integers(0, 2**32, 200)]:
        assert O.pcg_hash(v) == R.ref_pcg_hash(v)
        assert O.derive_seed(v, 0x4C454E31) == R.ref_derive_seed(v, 0x4C454E31)
    import ctypes as C

    for seed in (0, 1, 7777, 2**63 + 5):
        buf = (C.c_uint64 * 16)()
        R.ref_splitmix_stream(seed, 16, buf)
        assert O.splitmix_stream(seed, 16) == list(buf)


def test_step_seed_semantics():
    m = O.module_seed(7777)
    # pcg(seed_ ^ (frame*197 + s)) with uint32 wrap
    assert O.step_seed(m, 3, 1) == O.pcg_hash(m ^ (3 * 197 + 1))
    big = 30000000
    assert O.step_seed(m, big, 0) == O.pcg_hash(m ^ ((big * 197) & 0xFFFFFFFF))


# ------------------------------------------------------------------ kernels
@pytest.mark.parametrize("name", sorted(KERNELS))
def test_kernel_window_golden(golden, name):
    k = as_oracle_kernel(KERNELS[name])
    win = O.kernel_window(k)
    np.testing.assert_allclose(win, golden["win_" + name], rtol=0, atol=1e-15)


@pytest.mark.parametrize("name", [n for n in sorted(KERNELS) if KERNELS[n]["shape"] == "ring"])
def test_kernel_image_golden(golden, name):
    k = as_oracle_kernel(KERNELS[name])
    img = O.kernel_image(k, 24, 20)
    np.testing.assert_allclose(img, golden["img_" + name], rtol=0, atol=1e-15)


@pytest.mark.parametrize("name", sorted(KERNELS))
def test_kernel_invariants(name):
    k = as_oracle_kernel(KERNELS[name])
    win = O.kernel_window(k)
    h = (win.shape[0] - 1) // 2
    assert abs(win.sum() - 1.0) < 1e-14
    assert win[h, h] == 0.0  # zero at r = 0
    # 8-fold dihedral symmetry
    for t in (win.T, win[::-1], win[:, ::-1], win[::-1, ::-1]):
        np.testing.assert_array_equal(win, t)
    # zero outside r > R
    d = np.arange(-h, h + 1)
    r = np.hypot(*np.meshgrid(d, d, indexing="ij"))
    assert np.all(win[r > k["radius"] + 1e-12] == 0.0)
    assert np.all(win >= 0)


def test_shell_peak_and_ratio():
    # Gaussian shell: K(r1)/K(r2) = exp(-(t1^2 - t2^2)/2), t = (r/R - mu)/sigma
    k = as_oracle_kernel(KERNELS["shell_gauss_R7_mu04"])
    win = O.kernel_window(k)
    h = 7
    r1, r2 = 3.0, 5.0
    t1 = (r1 / 7 - 0.4) / 0.2
    t2 = (r2 / 7 - 0.4) / 0.2
    assert abs(win[h, h + 3] / win[h, h + 5] - np.exp(-(t1**2 - t2**2) / 2)) < 1e-13
    # exp core peaks at r = R/2 with raw value exp(0) = 1
    ke = as_oracle_kernel(KERNELS["shell_exp_R6"])
    assert abs(O.lib().lo_kernel_raw(O.kernel_struct(ke), 3, 0) - 1.0) < 1e-15


def test_ring_equals_shell_window_for_single_beta():
    # the two reference builders agree when betas=[1] and shellMu=0.5
    shell = as_oracle_kernel(KERNELS["shell_gauss_R5"])
    ring = as_oracle_kernel(KERNELS["ring_b1_R5"])
    np.testing.assert_allclose(O.kernel_window(shell), O.kernel_window(ring), atol=1e-15)


def test_kernel_image_is_wrapped_window():
    k = as_oracle_kernel(KERNELS["ring_b2_R6p5"])
    win = O.kernel_window(k)
    img = O.kernel_image(k, 24, 20)
    h = (win.shape[0] - 1) // 2
    placed = np.zeros((20, 24))
    for iy in range(win.shape[0]):
        for ix in range(win.shape[1]):
            placed[(iy - h) % 20, (ix - h) % 24] += win[iy, ix]
    np.testing.assert_allclose(img, placed, atol=1e-15)


# ------------------------------------------------------------------ FFT
@pytest.mark.parametrize("n", [15, 60, 64, 96, 120])
def test_fft_golden(golden, n):
    x = golden[f"fft_in_{n}"]
    np.testing.assert_allclose(O.fft(x), golden[f"fft_out_{n}"], atol=1e-12)
    np.testing.assert_allclose(O.fft(O.fft(x), inverse=True) / n, x, atol=1e-14)


def test_fft2_golden(golden):
    np.testing.assert_allclose(O.fft2(golden["fft2_in"]), golden["fft2_out"], atol=1e-12)


@pytest.mark.parametrize("n", [8, 30, 45, 1080])
def test_fft_closed_forms(n):
    d = np.zeros(n, complex)
    d[0] = 1
    np.testing.assert_allclose(O.fft(d), np.ones(n), atol=1e-13)  # delta -> constant
    c = np.ones(n, complex)
    e = np.zeros(n)
    e[0] = n
    np.testing.assert_allclose(O.fft(c), e, atol=1e-9)  # constant -> N delta
    m = 3
    x = np.cos(2 * np.pi * m * np.arange(n) / n)
    X = O.fft(x)
    want = np.zeros(n, complex)
    want[m] += n / 2
    want[n - m] += n / 2
    np.testing.assert_allclose(X, want, atol=1e-9)  # cosine -> two bins
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    Z = O.fft(z)
    assert abs(np.sum(abs(Z) ** 2) / n - np.sum(abs(z) ** 2)) < 1e-9 * n  # Parseval


# ------------------------------------------------------------------ potentials
def test_delta_response_is_kernel():
    k = as_oracle_kernel(KERNELS["ring_b3_R7_exp"])
    win = O.kernel_window(k)
    h = (win.shape[0] - 1) // 2
    A = np.zeros((20, 24))
    A[10, 12] = 1.0
    u = O.potential_direct(A, win)
    # u(x) = sum K(d) A(x+d) -> u(10-dy, 12-dx) = K(dy, dx); K even so = K
    np.testing.assert_allclose(u[10 - h:10 + h + 1, 12 - h:12 + h + 1], win[::-1, ::-1],
                               atol=1e-16)
    img = O.kernel_image(k, 24, 20)
    u2 = O.potential_fft(A, img)
    np.testing.assert_allclose(u2, u, atol=1e-15)


def test_uniform_potential_and_direct_equals_fft():
    rng = np.random.default_rng(3)
    k = as_oracle_kernel(KERNELS["ring_b2_R6p5"])
    win = O.kernel_window(k)
    img = O.kernel_image(k, 30, 22)
    A = np.full((22, 30), 0.37)
    np.testing.assert_allclose(O.potential_direct(A, win), 0.37, atol=1e-15)
    A = rng.random((22, 30))
    np.testing.assert_allclose(O.potential_direct(A, win), O.potential_fft(A, img), atol=1e-14)


# ------------------------------------------------------------------ steps
@pytest.mark.parametrize("name", sorted(WORLDS))
@pytest.mark.parametrize("use_fft", [False, True])
def test_step_golden(golden, name, use_fft):
    w = as_oracle_world(WORLDS[name])
    A = golden[f"step_{name}_A"]
    got = O.step(w, A, use_fft=use_fft, seed=2831084092, frame=29)
    want = golden[f"step_{name}_{'fft' if use_fft else 'direct'}"]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)


def test_uniform_field_closed_form():
    w = as_oracle_world(WORLDS["w1_direct"])
    for c in (0.0, 0.05, 0.15, 0.16, 0.5, 1.0):
        A = np.full((24, 32), c)
        bell = np.exp(-0.5 * ((c - 0.15) / 0.015) ** 2)
        want = np.clip(c + 0.1 * (2 * bell - 1), 0, 1)
        np.testing.assert_allclose(O.step(w, A), want, atol=1e-15)
        np.testing.assert_allclose(O.step(w, A, use_fft=True), want, atol=1e-14)


def _mode1_world(weights):
    ks = [dict(shape=O.SHAPE_RING, core=O.CORE_GAUSSIAN, radius=6.0, shellSigma=0.15,
               betas=[1.0], mu=0.15, sigma=0.017, weight=weights[0]),
          dict(shape=O.SHAPE_RING, core=O.CORE_GAUSSIAN, radius=4.0, shellSigma=0.15,
               betas=[1.0, 0.6], mu=0.30, sigma=0.035, weight=weights[1]),
          dict(shape=O.SHAPE_SHELL, core=O.CORE_EXP, radius=5.0, shellMu=0.5, shellSigma=0.15,
               betas=[1.0], mu=0.22, sigma=0.028, weight=weights[2])]
    return dict(width=32, height=24, growthMode=1, dt=0.35, radius=6, kernels=ks)


def test_uniform_field_closed_form_mode1():
    """growth mode 1 (Lenia.metal:95-100, 188-203): a uniform field c has
    potential c under every kernel (each sums to 1), so
    A' = clip(c + dt (sum_k w_k bell_k(c) / sum_k w_k - c)), and A' = c when
    sum_k w_k <= 1e-6; both convolution forms"""
    mus, sigmas = (0.15, 0.30, 0.22), (0.017, 0.035, 0.028)
    for weights in ((1.0, 0.6, 0.4), (0.0, 0.0, 2.5), (0.0, 0.0, 0.0)):
        w = _mode1_world(weights)
        for c in (0.0, 0.1, 0.15, 0.22, 0.3, 0.5, 0.97, 1.0):
            A = np.full((24, 32), c)
            sw = sum(weights)
            if sw > 1e-6:
                tgt = sum(wk * np.exp(-0.5 * ((c - m) / sg) ** 2)
                          for wk, m, sg in zip(weights, mus, sigmas)) / sw
            else:
                tgt = c
            want = np.clip(c + 0.35 * (tgt - c), 0, 1)
            np.testing.assert_allclose(O.step(w, A), want, atol=1e-15)
            np.testing.assert_allclose(O.step(w, A, use_fft=True), want, atol=1e-14)
    # the single-kernel leniaStep form A' = A + dt (bell - A)
    w = dict(_mode1_world((1.0, 0.0, 0.0)), kernels=_mode1_world((1.0, 0, 0))["kernels"][:1])
    for c in (0.12, 0.15, 0.4):
        want = np.clip(c + 0.35 * (np.exp(-0.5 * ((c - 0.15) / 0.017) ** 2) - c), 0, 1)
        np.testing.assert_allclose(O.step(w, np.full((24, 32), c)), want, atol=1e-15)


def test_zero_field_stays_zero_mode0():
    w = as_oracle_world(WORLDS["w3_mode0"])
    A = np.zeros((28, 36))
    assert np.all(O.step(w, A) == 0.0)


def test_planted_transposed_kernel_is_caught(golden):
    # a plausible mistake (kernel window transposed / dy-dx swapped) must be
    # detectable by the checks above: use an asymmetric window on purpose
    rng = np.random.default_rng(11)
    win = rng.random((5, 5))
    A = golden["step_w1_direct_A"]
    u = O.potential_direct(A, win)
    u_t = O.potential_direct(A, win.T.copy())
    assert np.abs(u - u_t).max() > 1e-3


# ------------------------------------------------------------------ gradient
def _grad_world():
    return dict(width=24, height=20, growthMode=0, dt=0.1, radius=4, kernels=[
        dict(shape=O.SHAPE_RING, core=O.CORE_GAUSSIAN, radius=4.0, shellSigma=0.15,
             betas=[1.0], mu=0.2, sigma=0.05, weight=0.6),
        dict(shape=O.SHAPE_RING, core=O.CORE_GAUSSIAN, radius=5.0, shellSigma=0.15,
             betas=[1.0, 0.6], mu=0.25, sigma=0.06, weight=0.4)])


def test_rollout_and_gradient_match_torch(golden):
    w = _grad_world()
    A0, gT = golden["grad_A0"], golden["grad_gT"]
    np.testing.assert_allclose(O.rollout(w, A0, 6), golden["grad_AT"], atol=1e-13)
    g = O.rollout_grad(w, A0, 6, gT)
    np.testing.assert_allclose(g["A0"], golden["grad_dA0"], rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(g["mu"], golden["grad_mu"], rtol=1e-9)
    np.testing.assert_allclose(g["sigma"], golden["grad_sigma"], rtol=1e-9)
    np.testing.assert_allclose(g["weight"], golden["grad_weight"], rtol=1e-9)
    np.testing.assert_allclose(g["dt"], golden["grad_dt"][0], rtol=1e-9)
    np.testing.assert_allclose(g["betas"][0][:1], golden["grad_betas0"], rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(g["betas"][1][:2], golden["grad_betas1"], rtol=1e-8)


def test_gradient_central_differences():
    w = _grad_world()
    rng = np.random.default_rng(9)
    A0 = 0.2 + 0.6 * rng.random((20, 24))
    gT = rng.standard_normal((20, 24))
    T = 4
    g = O.rollout_grad(w, A0, T, gT)
    h = 1e-6

    def loss(wd, A):
        return float(np.sum(O.rollout(wd, A, T) * gT))

    for (k, key) in [(0, "mu"), (1, "sigma"), (0, "weight")]:
        wp = {**w, "kernels": [dict(kk) for kk in w["kernels"]]}
        wm = {**w, "kernels": [dict(kk) for kk in w["kernels"]]}
        wp["kernels"][k][key] += h
        wm["kernels"][k][key] -= h
        fd = (loss(wp, A0) - loss(wm, A0)) / (2 * h)
        assert abs(fd - g[key][k]) <= 1e-5 * max(1.0, abs(fd)), (key, fd, g[key][k])
    wp, wm = dict(w), dict(w)
    wp["dt"] += h
    wm["dt"] -= h
    fd = (loss(wp, A0) - loss(wm, A0)) / (2 * h)
    assert abs(fd - g["dt"]) <= 1e-5 * max(1.0, abs(fd))
    for j in range(2):
        wp = {**w, "kernels": [dict(kk) for kk in w["kernels"]]}
        wm = {**w, "kernels": [dict(kk) for kk in w["kernels"]]}
        wp["kernels"][1]["betas"] = list(wp["kernels"][1]["betas"])
        wm["kernels"][1]["betas"] = list(wm["kernels"][1]["betas"])
        wp["kernels"][1]["betas"][j] += h
        wm["kernels"][1]["betas"][j] -= h
        fd = (loss(wp, A0) - loss(wm, A0)) / (2 * h)
        assert abs(fd - g["betas"][1][j]) <= 1e-5 * max(1.0, abs(fd))
    for (y, x) in [(3, 4), (10, 11), (17, 20)]:
        Ap, Am = A0.copy(), A0.copy()
        Ap[y, x] += h
        Am[y, x] -= h
        fd = (loss(w, Ap) - loss(w, Am)) / (2 * h)
        assert abs(fd - g["A0"][y, x]) <= 1e-5 * max(1.0, abs(fd))


def test_rollout_piece_replays_the_recorded_trajectory():
    """the clip-frozen rollout: recording gives lo_rollout's trajectory, and
    replaying the record from the same start reproduces it bit for bit"""
    w = _grad_world()
    rng = np.random.default_rng(2)
    A0 = rng.random((20, 24))
    A0[rng.random((20, 24)) < 0.3] = 0.0
    AT, mask = O.rollout_piece(w, A0, 7)
    assert np.array_equal(AT, O.rollout(w, A0, 7))
    assert mask.shape == (7, 20, 24) and set(np.unique(mask)) <= {0, 1, 2}
    assert np.count_nonzero(mask == 1) > 0  # the trajectory does clip
    AT2, _ = O.rollout_piece(w, A0, 7, mask)
    assert np.array_equal(AT2, AT)


def test_gradient_central_differences_config4_size():
    """The oracle gradient at the differentiable-rollout size (256x256, two
    ring kernels R=16, T=64) against FP64 central differences, h = 1e-6.
    A parameter moves every cell, so over 64 x 65536 cell-steps some pre-clip
    value sits within reach of 0 or 1 and a plain difference straddles the
    kink (mu: 43 flips, 0.8% off); the parameter probes therefore difference
    lo_rollout_piece with the base trajectory's clip states held, which is the
    smooth branch the adjoint differentiates.  State probes are local and
    difference the plain rollout, after checking that no clip state flips."""
    from lenia_cases import INIT_NONE, config4

    w = dict(config4(), initMode=INIT_NONE)
    A0 = O.step(w, O.init_patch(256, 256, O.module_seed(1234), 0, 0, 256, 256), use_fft=True)
    gT = np.random.default_rng(5).standard_normal((256, 256))
    T, h = 64, 1e-6
    g = O.rollout_grad(w, A0, T, gT)
    _, mask = O.rollout_piece(w, A0, T)

    def piece_loss(wd):
        return float(np.sum(O.rollout_piece(wd, A0, T, mask)[0] * gT))

    def with_kernel(k, key, v):
        wd = {**w, "kernels": [dict(kk) for kk in w["kernels"]]}
        wd["kernels"][k][key] = v
        return wd

    probes = []
    for k in range(2):
        for key in ("mu", "sigma", "weight"):
            v = w["kernels"][k][key]
            probes.append((f"{key}[{k}]", with_kernel(k, key, v + h), with_kernel(k, key, v - h),
                           g[key][k]))
    probes.append(("dt", dict(w, dt=w["dt"] + h), dict(w, dt=w["dt"] - h), g["dt"]))
    for j in range(2):
        bp, bm = list(w["kernels"][1]["betas"]), list(w["kernels"][1]["betas"])
        bp[j] += h
        bm[j] -= h
        probes.append((f"betas[1][{j}]", with_kernel(1, "betas", bp), with_kernel(1, "betas", bm),
                       g["betas"][1][j]))
    for name, wp, wm, ana in probes:
        fd = (piece_loss(wp) - piece_loss(wm)) / (2 * h)
        rel = abs(fd - ana) / max(1.0, abs(fd))
        print(f"{name}: FD {fd:.9g} adjoint {ana:.9g} rel {rel:.2e}")
        assert rel <= 1e-5, (name, fd, ana)
    for (y, x) in [(40, 50), (128, 128), (200, 17)]:
        Ap, Am = A0.copy(), A0.copy()
        Ap[y, x] += h
        Am[y, x] -= h
        (ATp, mp), (ATm, mm) = O.rollout_piece(w, Ap, T), O.rollout_piece(w, Am, T)
        assert np.array_equal(mp, mm), "state probe straddles a clip kink"
        fd = float(np.sum((ATp - ATm) * gT)) / (2 * h)
        rel = abs(fd - g["A0"][y, x]) / max(1.0, abs(fd))
        print(f"A0[{y},{x}]: FD {fd:.9g} adjoint {g['A0'][y, x]:.9g} rel {rel:.2e}")
        assert rel <= 1e-5


def test_value_noise_init_invariants():
    A = O.init_value_noise(64, 48, 1234, 0.42, 26.0)
    assert A.min() >= 0.0 and A.max() <= 0.9
    assert 0.0 < A.mean() < 0.9
    # continuity across lattice cells: neighbouring cells differ by a bounded
    # amount (value noise is C1 between lattice points)
    assert np.abs(np.diff(A, axis=1)).max() < 0.5


def test_patch_init():
    A = O.init_patch(40, 30, 99, 10, 5, 20, 16)
    assert np.all(A[:5] == 0) and np.all(A[21:] == 0)
    assert np.all(A[:, :10] == 0) and np.all(A[:, 30:] == 0)
    assert A[5, 10] == O.rand01(10, 5, O.TAG_PATCH, 99)
    assert 0.3 < A[5:21, 10:30].mean() < 0.7


def test_liveness_guard_rejects_dead_uniform_and_static_fields():
    """the guard that the multi-step GPU comparisons call before they compare
    anything: an empty field, a uniform fixed point and an unchanged field
    fail it; a structured evolving field passes"""
    from lenia_cases import assert_alive

    with pytest.raises(AssertionError, match="dead"):
        assert_alive(np.zeros((64, 64)), min_live=10)
    with pytest.raises(AssertionError, match="dead"):
        assert_alive(np.ones((64, 64)), min_live=10)
    with pytest.raises(AssertionError, match="uniform"):
        assert_alive(np.full((64, 64), 0.275), min_live=10, min_std=0.05)
    A = O.init_value_noise(64, 64, 1234, 0.42, 26.0)
    with pytest.raises(AssertionError, match="static"):
        assert_alive(A, A, min_live=10, min_change=1e-3)
    assert A.std() > 0.01
    assert assert_alive(A, 0.9 * A, min_live=10, min_std=0.01, min_change=1e-3) >= 10
