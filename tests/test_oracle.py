"""Pins of the FP64 oracle (CPU only).

The reference ships no golden vectors (its tests print norms), so each oracle
function is pinned by something other than its own formula: known answers,
closed forms, brute force, textbook identities, and the executable CPU spec
of the stencil in test/tvd.cpp:14-75.
"""

import math

import numpy as np
import pytest

from oracle import tomocam_oracle as O
from conftest import rel_l2


# ---------------------------------------------------------------- NPRandom


def test_nprandom_known_answers():
    # std::mt19937 with the default seed 5489: first output 3499211612
    raw = np.random.RandomState(5489)._bit_generator.random_raw(8)
    assert int(raw[0]) == 3499211612
    # test/timer.h:50-55 construction on raw words == random_sample
    got = O.nprandom_from_words(raw)
    assert np.array_equal(got, np.random.RandomState(5489).random_sample(4))
    # the widely published first values of this generator/construction
    np.testing.assert_allclose(
        O.NPRandom().rand(4),
        [0.81472368639317894, 0.90579193707561922, 0.12698681629350606,
         0.91337585613901939], rtol=0, atol=1e-16)


# ---------------------------------------------------------------- grid


@pytest.mark.parametrize("n", [15, 16, 362])
def test_polar_grid_closed_form(n):
    ang = np.array([0.0, np.pi / 2, 1.0])
    x, y = O.polar_grid(ang, n)
    c = n // 2
    r = 2 * np.pi * (np.arange(n) - c) / (n - 1)
    np.testing.assert_allclose(x[0], r)
    np.testing.assert_allclose(y[2], r * np.sin(1.0))
    assert abs(x[1]).max() < 1e-12 * n
    if n % 2:
        np.testing.assert_allclose(r, -r[::-1], atol=1e-14)
        assert abs(r[-1] - np.pi) < 1e-14
    else:
        assert abs(r[0] + np.pi * n / (n - 1)) < 1e-12


# ---------------------------------------------------------------- NUDFT


def test_nudft_known_answers():
    n = 9
    # one unit point at the origin: F(k) = 1 for all k; type 2 of f gives sum
    F = O.nudft_type1(np.ones((1, 1)), np.zeros(1), np.zeros(1), n)
    np.testing.assert_allclose(F, np.ones((1, n, n)))
    f = np.random.default_rng(0).standard_normal((1, n, n))
    c = O.nudft_type2(f, np.zeros(1), np.zeros(1))
    np.testing.assert_allclose(c[0, 0], f.sum())
    # points on the uniform grid 2 pi m / n reduce to numpy's FFT
    m = np.arange(n)
    xx, yy = np.meshgrid(2 * np.pi * m / n, 2 * np.pi * m / n)
    vals = np.random.default_rng(1).standard_normal(n * n)
    F = O.nudft_type1(vals[None], xx.ravel(), yy.ravel(), n)[0]
    k = O.modes(n)
    # F[k2,k1] = sum_{m2,m1} v[m2,m1] e^{+2pi i (k1 m1 + k2 m2)/n}
    ref = np.fft.ifft2(vals.reshape(n, n)) * n * n
    np.testing.assert_allclose(F, ref[np.ix_(k % n, k % n)], atol=1e-10)


def test_nudft_adjoint():
    n, P = 12, 7
    x, y = O.polar_grid(np.arange(P) * np.pi / P, n)
    rng = np.random.default_rng(2)
    c = rng.standard_normal((1, P * n)) + 1j * rng.standard_normal((1, P * n))
    f = rng.standard_normal((1, n, n)) + 1j * rng.standard_normal((1, n, n))
    lhs = np.vdot(O.nudft_type2(f, x, y)[0], c[0])
    # <A f, c> = <f, A^H c>; A^H is type 1 with the conjugate sign
    rhs = np.vdot(f[0], O.nudft_type1(c, x, y, n)[0])
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("n,P", [(15, 7), (16, 9), (33, 20)])
def test_es_nufft_vs_nudft(n, P):
    ang = np.arange(P) * np.pi / P
    x, y = O.polar_grid(ang, n)
    rng = np.random.default_rng(n)
    c = rng.standard_normal((2, P * n)) + 1j * rng.standard_normal((2, P * n))
    assert rel_l2(O.nufft_type1(c, x, y, n), O.nudft_type1(c, x, y, n)) < 1e-11
    f = rng.standard_normal((2, n, n))
    assert rel_l2(O.nufft_type2(f, x, y), O.nudft_type2(f, x, y)) < 1e-11
    # PSF mode count 2n - 1 (odd)
    assert rel_l2(O.nufft_type1(c[:1], x, y, 2 * n - 1),
                  O.nudft_type1(c[:1], x, y, 2 * n - 1)) < 1e-11


# ---------------------------------------------------------------- projectors


@pytest.mark.parametrize("n", [15, 16])
def test_centered_ray_dft_definition(n):
    v = np.random.default_rng(3).standard_normal((2, n)) + 0j
    c = n // 2
    j = np.arange(n)
    for sign in (-1, 1):
        E = np.exp(sign * 2j * np.pi * np.outer(j - c, j - c) / n)
        np.testing.assert_allclose(O.centered_ray_dft(v, sign), v @ E.T,
                                   atol=1e-12)


@pytest.mark.parametrize("n,P", [(15, 7), (16, 9), (20, 11)])
def test_projector_adjoint(n, P):
    ang = np.arange(P) * np.pi / P
    rng = np.random.default_rng(n)
    img = rng.standard_normal((2, n, n))
    s = rng.standard_normal((2, P, n))
    lhs = np.sum(O.project(img, ang, exact=True) * s)
    rhs = np.sum(img * O.backproject(s, ang, exact=True))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_odd_n_projection_is_real():
    n, P = 15, 7
    ang = np.arange(P) * np.pi / P
    img = np.random.default_rng(4).standard_normal((1, n, n))
    x, y = O.polar_grid(ang, n)
    c = O.nudft_type2(img, x, y).reshape(1, P, n)
    out = O.centered_ray_dft(c, +1)
    assert np.abs(out.imag).max() < 1e-12 * np.abs(out.real).max()


def test_ramp_filter():
    np.testing.assert_allclose(O.ramlak(5), [0.4, 0.2, 0.0, 0.2, 0.4])
    np.testing.assert_allclose(O.ramlak(4), [0.5, 0.25, 0.0, 0.25])


# ---------------------------------------------------------------- Toeplitz


@pytest.mark.parametrize("n,P", [(15, 7), (16, 9)])
def test_toeplitz_identities(n, P):
    ang = np.arange(P) * np.pi / P
    ps = O.psf(ang, n, exact=True)
    # PSF is real and even: PSF(m) == PSF(-m)
    np.testing.assert_allclose(ps, ps[::-1, ::-1], atol=1e-10)
    x = np.random.default_rng(n).standard_normal((2, n, n))
    brute = O.toeplitz_bruteforce(x, ps)
    assert rel_l2(O.toeplitz(x, ps), brute) < 1e-13
    for m in (2 * n - 1, 2 * n + 5, O.smooth_size(2 * n - 1)):
        assert rel_l2(O.toeplitz_circulant(x, ps, m), brute) < 1e-13
    # Toeplitz == Re(A^H A x)/n^3 (src/gradient.cpp vs src/gradient2.cpp)
    g = O.gradient_nufft(x, 0 * x, ang, exact=True)
    assert rel_l2(g, brute) < 1e-13
    # function_value2 == ||P x - y||^2 for odd n
    if n % 2:
        y = np.random.default_rng(1).standard_normal((2, P, n))
        b = O.backproject(y, ang, exact=True)
        f1 = O.function_value(x, y, ang, exact=True)
        f2 = O.function_value2(x, b, ps, float(np.sum(y * y)))
        assert abs(f1 - f2) <= 1e-11 * f1


# ---------------------------------------------------------------- qGGMRF


def test_qggmrf_weights_and_hessian():
    assert abs(O.WEIGHTS.sum() - 0.9994) < 1e-12
    assert O.WEIGHTS[1, 1, 1] == 0.0
    # sum_d w_d q / (sigma^q c)
    for sigma in (0.1, 10.0, 100.0):
        assert abs(O.tv_hessian_zero(sigma) - 0.9994 * 2 / (sigma ** 2 * 0.001)) \
            <= 1e-5 * O.tv_hessian_zero(sigma)


def test_potential_derivative_is_derivative():
    for sigma, p in ((10.0, 1.2), (0.1, 1.2), (3.0, 1.5)):
        d = np.linspace(-3 * sigma, 3 * sigma, 41)
        d = d[np.abs(d) > 1e-3 * sigma]
        h = 1e-6 * sigma
        fd = (O.potential(d + h, sigma, p) - O.potential(d - h, sigma, p)) / (2 * h)
        np.testing.assert_allclose(O.d_potential(d, sigma, p), fd, rtol=1e-6)
    assert O.d_potential(np.array([0.0]), 10.0, 1.2)[0] == 0.0


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 3, 4), (3, 9, 11)])
def test_qggmrf_matches_tvd_cpp_loop_spec(shape):
    x = O.NPRandom().rand(int(np.prod(shape))).reshape(shape).astype(np.float32)
    loops = O.qggmrf_grad_loops(x, 10.0, 1.2)
    vec = O.qggmrf_grad(x, 10.0, 1.2, dtype=np.float32)
    np.testing.assert_allclose(vec, loops, rtol=2e-5, atol=1e-7)
    vec64 = O.qggmrf_grad(x, 10.0, 1.2)
    np.testing.assert_allclose(vec64, loops, rtol=1e-4, atol=1e-6)


def test_qggmrf_is_gradient_in_interior():
    rng = np.random.default_rng(9)
    x = rng.standard_normal((5, 6, 7))
    sigma, p = 1.5, 1.2
    g = O.qggmrf_grad(x, sigma, p)
    h = 1e-6
    for v in [(2, 2, 3), (1, 3, 2), (3, 2, 4)]:
        xp = x.copy()
        xm = x.copy()
        xp[v] += h
        xm[v] -= h
        fd = (O.qggmrf_value(xp, sigma, p) - O.qggmrf_value(xm, sigma, p)) / (2 * h)
        assert abs(fd - g[v]) <= 1e-6 * max(1.0, abs(g[v]))


def test_qggmrf_clamp_boundary():
    # a constant volume has zero gradient (all deltas are zero, even at the
    # clamped boundary); a single-slice volume uses itself as neighbours
    assert np.abs(O.qggmrf_grad(np.full((2, 4, 4), 3.0), 1.0, 1.2)).max() == 0
    x = np.random.default_rng(1).random((1, 4, 5))
    g1 = O.qggmrf_grad(x, 1.0, 1.2)
    # with one slice, z-neighbours clamp to the slice itself: same as a
    # 3-slice stack of identical slices at its middle slice
    g3 = O.qggmrf_grad(np.repeat(x, 3, axis=0), 1.0, 1.2)
    np.testing.assert_allclose(g1[0], g3[1], rtol=1e-12)


# ---------------------------------------------------------------- run2


def _small_setup():
    P, N, S = 12, 16, 2
    ang = np.arange(P) * np.pi / P
    rng = np.random.default_rng(5)
    img = rng.random((S, N, N))
    sino = O.project(img, ang, exact=True)
    return O.mbir2_setup(sino, ang, N / 2, exact=True)


def test_run2_invariants_and_linear_form():
    st = _small_setup()
    sigma = 10.0
    log = O.Run2Log()

    def grad_fn(x):
        return O.gradient2(x, st.sinoT, st.psf) + O.qggmrf_grad(x, sigma, 1.2)

    def err_fn(x):
        return math.sqrt(O.function_value2(x, st.sinoT, st.psf, st.sino_norm))

    ref = O.run2(st.x0, grad_fn, err_fn, 15, st.step, log)
    assert len(log.errors) == 15
    l2 = O.Run2Log()
    lin = O.run2_linear(st.x0, lambda v: O.toeplitz(v, st.psf), st.sinoT,
                        lambda v: O.qggmrf_grad(v, sigma, 1.2), st.sino_norm,
                        15, st.step, l2)
    assert rel_l2(lin, ref) < 1e-12
    np.testing.assert_allclose(l2.errors, log.errors, rtol=1e-10)
    np.testing.assert_allclose(l2.xerrs, log.xerrs, rtol=1e-8)


def test_run2_retry_semantics():
    """Rejected candidates become `sol`; the linear form follows (4 rejects
    in two iterations from a perturbed converged point, step 6)."""
    st = _small_setup()
    sigma = 1e4
    T = lambda v: O.toeplitz(v, st.psf)
    reg = lambda v: O.qggmrf_grad(v, sigma, 1.2)
    xs = O.run2_linear(st.x0, T, st.sinoT, reg, st.sino_norm, 150, 3.0)
    x0 = xs + 1e-3

    def grad_fn(x):
        return O.gradient2(x, st.sinoT, st.psf) + reg(x)

    def err_fn(x):
        return math.sqrt(O.function_value2(x, st.sinoT, st.psf, st.sino_norm))

    log = O.Run2Log()
    ref = O.run2(x0, grad_fn, err_fn, 2, 6.0, log)
    assert log.backtracks > 0
    # every accepted step satisfies the reference's test
    l2 = O.Run2Log()
    lin = O.run2_linear(x0, T, st.sinoT, reg, st.sino_norm, 2, 6.0, l2)
    assert l2.backtracks == log.backtracks
    assert rel_l2(lin, ref) < 1e-10
    assert all(s < 6.0 for s in log.steps[:1])


def _odd_setup():
    """Odd widths (15 -> padded 21), where the projector-path error
    ||P v - y||^2 equals the Toeplitz expansion exactly, so the accept test
    can be judged without the oracle's own function_value2."""
    P, N, S = 12, 15, 2
    ang = np.arange(P) * np.pi / P
    rng = np.random.default_rng(5)
    img = rng.random((S, N, N))
    sino = O.project(img, ang, exact=True)
    return O.mbir2_setup(sino, ang, N / 2, exact=True), ang


def _run2_traced(st, x0, sigma, step0, iters, err_sq=False):
    reg = lambda v: O.qggmrf_grad(v, sigma, 1.2)

    def grad_fn(x):
        return O.gradient2(x, st.sinoT, st.psf) + reg(x)

    def err_fn(x):
        f = O.function_value2(x, st.sinoT, st.psf, st.sino_norm)
        return f if err_sq else math.sqrt(f)

    log = O.Run2Log()
    log.attempts = []
    out = O.run2(x0, grad_fn, err_fn, iters, step0, log)
    return out, log, reg


def _run2_violations(st, ang, x0, step0, iters, log, reg):
    """Checks the trace of O.run2 against src/optimize.h:96-154 with
    quantities computed on other paths than the oracle's run2 uses:
      - f(v) = ||P v - y||^2 by the exact projector (function_value), and
        the accept test sqrt f(cand) <= sqrt f(y) + s/2 sum g^2
        (src/optimize.h:123-127: a norm against a squared norm);
      - g = P^T (P y - y) + R'(y) by projector and backprojector;
      - tnew^2 = t^2 (1 - tnew) (the root the t-update takes, :110-112);
      - beta = tnew_prev (1/t - 1), y = sol + beta (sol - x) (:109-116);
      - the step is step0 on an iteration's first attempt and shrinks by
        0.9 per rejection (:128-132); the momentum term vanishes on the
        first attempt after an acceptance (x = sol there, :134);
      - exactly `iters` acceptances (tol/xtol unused).
    Returns a list of messages (empty when every check holds)."""
    bad = []
    fp = lambda v: O.function_value(v, st.sino2, ang, exact=True)
    x_acc = np.asarray(x0, dtype=np.float64)
    sol = x_acc
    t_exp, tnew_prev = 1.0, 1.0
    step_exp, first = step0, True
    for k, a in enumerate(log.attempts):
        t, tn, beta, s = a["t"], a["tnew"], a["beta"], a["step"]
        if abs(t - t_exp) > 1e-15:
            bad.append(f"{k}: t {t} != last accepted tnew {t_exp}")
        if abs(tn * tn - t * t * (1.0 - tn)) > 1e-14:
            bad.append(f"{k}: tnew^2 != t^2 (1 - tnew)")
        if abs(beta - tnew_prev * (1.0 / t - 1.0)) > 1e-15:
            bad.append(f"{k}: beta != tnew_prev (1/t - 1)")
        if s != step_exp:
            bad.append(f"{k}: step {s} != {step_exp}")
        if not np.array_equal(a["x"], x_acc) or not np.array_equal(
                a["sol_prev"], sol):
            bad.append(f"{k}: x / sol not the last accepted / candidate")
        y_exp = sol + beta * (sol - x_acc)
        if np.abs(a["y"] - y_exp).max() > 1e-15 * np.abs(y_exp).max():
            bad.append(f"{k}: y != sol + beta (sol - x)")
        if first and not np.array_equal(a["y"], x_acc):
            bad.append(f"{k}: momentum term non-zero after an acceptance")
        g_ind = O.backproject(O.project(a["y"], ang, exact=True) - st.sino2,
                              ang, exact=True) + reg(a["y"])
        if np.abs(a["g"] - g_ind).max() > 1e-10 * np.abs(g_ind).max():
            bad.append(f"{k}: g != P^T (P y - y) + R'(y)")
        cand = a["y"] - s * g_ind
        fx, fy = math.sqrt(fp(cand)), math.sqrt(fp(a["y"]))
        gy = 0.5 * s * float(np.sum(g_ind * g_ind))
        margin = (fx - fy - gy) / (fy + gy)
        if abs(margin) < 1e-9:
            bad.append(f"{k}: accept test too close to call ({margin:.1e})")
        elif a["accepted"] != (fx <= fy + gy):
            bad.append(f"{k}: accepted={a['accepted']} but sqrt f(cand) - "
                       f"sqrt f(y) - s/2 |g|^2 = {margin:.3e} (relative)")
        tnew_prev = tn
        sol = a["cand"]
        if a["accepted"]:
            x_acc, t_exp, step_exp, first = sol, tn, step0, True
        else:
            step_exp, first = s * 0.9, False
    n_acc = sum(a["accepted"] for a in log.attempts)
    if n_acc != iters:
        bad.append(f"{n_acc} acceptances, expected {iters}")
    return bad


def test_run2_control_flow_pinned_by_projector_path():
    """sigma = 10 from the FBP guess: 15 accepted attempts, every one
    checked against projector-path errors and gradients."""
    st, ang = _odd_setup()
    _, log, reg = _run2_traced(st, st.x0, 10.0, st.step, 15)
    assert len(log.attempts) == 15
    assert _run2_violations(st, ang, st.x0, st.step, 15, log, reg) == []


def _backtracking_start(st):
    T = lambda v: O.toeplitz(v, st.psf)
    reg = lambda v: O.qggmrf_grad(v, 1e4, 1.2)
    xs = O.run2_linear(st.x0, T, st.sinoT, reg, st.sino_norm, 150, 3.0)
    return xs + 1e-3


def test_run2_backtracking_pinned_by_projector_path():
    """A perturbed converged point with step 6 (sigma = 1e4): rejections
    with and without momentum, each rejected attempt violating the
    projector-path accept test and each accepted one satisfying it."""
    st, ang = _odd_setup()
    x0 = _backtracking_start(st)
    _, log, reg = _run2_traced(st, x0, 1e4, 6.0, 2)
    acc = [a["accepted"] for a in log.attempts]
    assert acc.count(False) >= 4 and acc.count(True) == 2
    # a retry with non-zero momentum (sol != x after a rejection)
    assert any(a["beta"] != 0.0 and not np.array_equal(a["sol_prev"], a["x"])
               for a in log.attempts)
    assert _run2_violations(st, ang, x0, 6.0, 2, log, reg) == []


def test_run2_pin_detects_squared_norm_accept_test():
    """The checks above fail for a plausible mistake planted in the
    oracle: the accept test on squared norms f(cand) > f(y) + s/2 |g|^2
    (what a 'fixed' reference would do) takes decisions the reference's
    norm-against-squared-norm test does not."""
    st, ang = _odd_setup()
    x0 = _backtracking_start(st)
    _, log, reg = _run2_traced(st, x0, 1e4, 6.0, 2, err_sq=True)
    bad = _run2_violations(st, ang, x0, 6.0, 2, log, reg)
    assert any("accepted=" in m for m in bad), bad


# ---------------------------------------------------------------- pre/post


@pytest.mark.parametrize("n,npad,n2", [(256, 53, 362), (512, 107, 726),
                                       (1024, 215, 1454), (2048, 430, 2908),
                                       (768, 161, 1090), (1536, 322, 2180)])
def test_pad_amount_workload_sizes(n, npad, n2):
    p, off = O.pad_amount(n, n / 2)
    assert (p, off, n + 2 * p) == (npad, 0, n2)


def test_preproc_postproc_bruteforce():
    rng = np.random.default_rng(0)
    for N, center in [(16, 8.0), (17, 5.4), (20, 14.5), (9, -1.0)]:
        s = rng.standard_normal((2, 3, N))
        out = O.preproc(s, center)
        npad, off = O.pad_amount(N, center)
        N2 = N + 2 * npad
        assert out.shape == (2, 3, N2)
        for z in range(N2):
            src = (z + off) % N2 - npad
            exp = s[..., src] if 0 <= src < N else 0.0
            np.testing.assert_array_equal(out[..., z], exp)
    img = rng.standard_normal((2, 5, 5))
    np.testing.assert_array_equal(O.postproc(O.pad2d(img, 6), 5), img)


def test_std_round_halves_away_from_zero():
    assert O.pad_amount(10, 4.5)[1] == (11 // 2) - 5
    assert O.pad_amount(10, 5.5)[1] == (11 // 2) - 6
    assert O.pad_amount(10, -0.5)[1] == (11 // 2) + 1


def test_qggmrf_grad_free_is_gradient_of_value():
    """The free-boundary gradient is the exact derivative of qggmrf_value at
    every voxel (faces included): FP64 central differences."""
    rng = np.random.default_rng(3)
    x = rng.random((3, 5, 6))
    for sigma in (0.1, 10.0):
        g = O.qggmrf_grad_free(x, sigma, 1.2)
        eps = 1e-6
        fd = np.zeros_like(x)
        for idx in np.ndindex(x.shape):
            xp = x.copy()
            xp[idx] += eps
            xm = x.copy()
            xm[idx] -= eps
            fd[idx] = (O.qggmrf_value(xp, sigma, 1.2) -
                       O.qggmrf_value(xm, sigma, 1.2)) / (2 * eps)
        assert np.abs(fd - g).max() <= 1e-7 * np.abs(g).max()
        # interior voxels: the reference (clamped) stencil
        gi = O.qggmrf_grad(x, sigma, 1.2)
        assert np.array_equal(gi[1:-1, 1:-1, 1:-1], g[1:-1, 1:-1, 1:-1])


def test_power_lipschitz_is_largest_eigenvalue():
    """Power iteration on T: its Rayleigh quotient never exceeds the largest
    eigenvalue of the dense (symmetric) operator and converges to it."""
    n, P = 12, 7
    ang = np.arange(P) * np.pi / P
    ps = O.psf(ang, n, exact=True)
    eye = np.eye(n * n).reshape(n * n, n, n)
    M = O.toeplitz(eye, ps).reshape(n * n, n * n)
    assert np.abs(M - M.T).max() <= 1e-13 * np.abs(M).max()
    lam = np.linalg.eigvalsh(0.5 * (M + M.T))
    assert lam.min() >= -1e-12 * lam.max()   # A^H A / n^3: PSD
    l20 = O.power_lipschitz(ps, n, 20)
    assert 0.5 * lam.max() < l20 <= lam.max() * (1 + 1e-12)
    assert abs(O.power_lipschitz(ps, n, 3000) - lam.max()) <= 1e-8 * lam.max()


def test_cpu_baseline_port_matches_oracle():
    """The benchmark's CPU baseline (FP32 restatement with w = 8 ES NUFFT
    and C/OpenMP kernels) reconstructs what the FP64 oracle does, to the
    FP32 bound (sigma = 100, 2 x 40^2, 20 iterations)."""
    from oracle import cpu_port
    P, N, S = 30, 40, 2
    ang = np.arange(P) * np.pi / P
    img = np.random.default_rng(9).random((S, N, N))
    sino = O.project(img, ang, exact=True)
    ref = O.mbir2(sino, ang, N / 2, 20, 100.0, exact=True)
    got = cpu_port.mbir2(sino.astype(np.float32), ang, N / 2, 20, 100.0)
    assert rel_l2(got, ref) < 1e-4
    # split set-up / resumed iterations give the same result
    m = cpu_port.Mbir2CPU(sino.astype(np.float32), ang, N / 2, 100.0)
    m.iterate(7)
    m.iterate(13)
    assert np.array_equal(m.result(), got)
