"""CPU check of the GPU kernels' per-cell arithmetic (csrc/rr_cell.hpp).

tests/hostcheck/rr_cell_host.cpp compiles the same templates the gfx950
kernels instantiate with g++; these tests compare them with the oracle's
per-direction restatement of collision.cu / BC.cu.  FP64: 1e-15 absolute on
populations; FP32 arithmetic: 1e-7 absolute.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import PKG, REPO

HC_SRC = os.path.join(REPO, "tests", "hostcheck", "rr_cell_host.cpp")
HC_LIB = os.path.join(REPO, "tests", "hostcheck", "librr_cell_host.so")
_D = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def hc():
    if not os.path.exists(HC_LIB) or os.path.getmtime(HC_LIB) < max(
            os.path.getmtime(HC_SRC), os.path.getmtime(os.path.join(PKG, "csrc", "rr_cell.hpp"))):
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                        "-I", os.path.join(PKG, "csrc"), HC_SRC, "-o", HC_LIB], check=True)
    lib = ctypes.CDLL(HC_LIB)
    lib.hc_wale.restype = ctypes.c_double
    return lib


def p(a):
    return a.ctypes.data_as(_D)


def hc_cell(hc, q, single, fast, f, rho, u, F, tau, mode):
    out = np.zeros(q)
    F = np.asarray(F, dtype=np.float64)
    d = ctypes.c_double
    hc.hc_rr_cell(q, single, fast, p(np.ascontiguousarray(f)), d(rho), d(u[0]), d(u[1]), d(u[2]), p(F), d(tau),
                  mode, p(out))
    return out


def random_cell(kind, rng, consistent):
    c, w, _ = oracle.lattice(kind)
    rho = 1 + 0.03 * rng.standard_normal()
    u = 0.1 * rng.uniform(-1, 1, 3)
    f = oracle.feq(kind, rho, *u) + 3e-3 * w * rng.standard_normal(len(w))
    F = 1e-4 * rng.standard_normal(3)
    if consistent:
        rho = f.sum()
        u = (f @ c + 0.5 * F) / rho
    return f, rho, u, F


@pytest.mark.parametrize("q,kind", [(27, oracle.D3Q27_CUDA), (19, oracle.D3Q19)])
@pytest.mark.parametrize("fast", [0, 1])
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("consistent", [True, False])
def test_cell_fp64_matches_oracle(hc, q, kind, fast, mode, consistent):
    if q == 19 and fast:
        pytest.skip("moment-space path is D3Q27 only")
    rng = np.random.default_rng(10 + mode)
    for _ in range(50):
        f, rho, u, F = random_cell(kind, rng, consistent)
        tau = rng.uniform(0.5005, 1.5)
        got = hc_cell(hc, q, 0, fast, f, rho, u, F, tau, mode)
        ref = oracle.rr_cell(kind, f, rho, *u, F, tau, mode)
        assert np.abs(got - ref).max() < 1e-15


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("fast,force", [(2, "x"), (3, "none"), (4, "all"), (5, "x"), (6, "none")])
def test_cell_force_modes(hc, fast, force, mode):
    """The compile-time force modes (channel: force along x only; Taylor-Green
    and cavity: none) equal the general formulation with those forces; modes
    4-6 are the FP16-storage form of the polynomial (output scaled by S/w_i)."""
    rng = np.random.default_rng(20 + fast)
    for _ in range(50):
        f, rho, u, F = random_cell(oracle.D3Q27_CUDA, rng, True)
        if force == "x":
            F = np.array([F[0], 0.0, 0.0])
        elif force == "none":
            F = np.zeros(3)
        got = hc_cell(hc, 27, 0, fast, f, rho, u, F, 0.57, mode)
        ref = oracle.rr_cell(oracle.D3Q27_CUDA, f, rho, *u, F, 0.57, mode)
        assert np.abs(got - ref).max() < 1e-15


@pytest.mark.parametrize("q,kind", [(27, oracle.D3Q27_CUDA), (19, oracle.D3Q19)])
def test_cell_fp32_arithmetic(hc, q, kind):
    rng = np.random.default_rng(3)
    for _ in range(50):
        f, rho, u, F = random_cell(kind, rng, True)
        got = hc_cell(hc, q, 1, 1, f, rho, u, F, 0.53, 0)
        ref = oracle.rr_cell(kind, f, rho, *u, F, 0.53, 0)
        assert np.abs(got - ref).max() < 1e-7


@pytest.mark.parametrize("q,kind", [(27, oracle.D3Q27_CUDA), (19, oracle.D3Q19)])
def test_moments_and_feq(hc, q, kind):
    rng = np.random.default_rng(4)
    c, w, _ = oracle.lattice(kind)
    for _ in range(20):
        f = w * (1 + 0.1 * rng.standard_normal(q))
        m = np.zeros(10)
        hc.hc_moments(q, p(f), p(m))
        h = f - w
        assert abs(m[0] - h.sum()) < 1e-16
        assert np.allclose(m[1:4], c.T @ h, atol=1e-16)
        assert np.allclose(m[4:7], (c ** 2).T @ h, atol=1e-16)
        assert np.allclose(m[7:10], [h @ (c[:, 0] * c[:, 1]), h @ (c[:, 0] * c[:, 2]), h @ (c[:, 1] * c[:, 2])],
                           atol=1e-16)
        rho, u = 1 + 0.02 * rng.standard_normal(), 0.1 * rng.uniform(-1, 1, 3)
        out = np.zeros(q)
        hc.hc_feq(q, ctypes.c_double(rho), ctypes.c_double(u[0]), ctypes.c_double(u[1]), ctypes.c_double(u[2]),
                  p(out))
        assert np.abs(out - oracle.feq(kind, rho, *u)).max() < 1e-16


@pytest.mark.parametrize("single", [0, 1])
def test_weight_scaled_moments(hc, single):
    """moments_q27_wscaled -- the FP16 kernels' moment sum, which reads the
    stored values g_i = h_i S / w_i and takes the D3Q27 weights as the
    per-axis factors (1, 4, 1)/6 inside its line sums -- against the raw
    moments of h_i = g_i w_i / S.  A wrong axis factor, sign or index moves
    a moment by O(|h|)."""
    rng = np.random.default_rng(11)
    c, w, _ = oracle.lattice(oracle.D3Q27_CUDA)
    S = 4096.0
    for _ in range(50):
        # FP16-representable stored values
        g = (rng.standard_normal(27) * 64).astype(np.float16).astype(np.float64)
        m = np.zeros(10)
        hc.hc_moments_wscaled(single, p(g), ctypes.c_double(S), p(m))
        h = g * w / S
        ref = np.concatenate([[h.sum()], c.T @ h, (c ** 2).T @ h,
                              [h @ (c[:, 0] * c[:, 1]), h @ (c[:, 0] * c[:, 2]), h @ (c[:, 1] * c[:, 2])]])
        scale = np.abs(h).sum()
        assert np.abs(m - ref).max() <= (2e-7 if single else 1e-15) * scale


def test_wale_kernel_formula(hc):
    """The kernels' WALE (sqrt forms of the powers, 0/0 -> 0) against the
    long-double matrix definition used to pin the oracle."""
    from test_oracle import wale_matrix
    rng = np.random.default_rng(5)
    for _ in range(50):
        g = 1e-3 * rng.standard_normal((3, 3))
        flat = np.array([g[0, 0], g[0, 1], g[0, 2], g[1, 0], g[1, 1], g[1, 2], g[2, 0], g[2, 1], g[2, 2]])
        got = hc.hc_wale(p(flat), ctypes.c_double(0.5))
        assert abs(got - wale_matrix(g)) <= 1e-13 * abs(wale_matrix(g)) + 1e-30
    assert hc.hc_wale(p(np.zeros(9)), ctypes.c_double(0.5)) == 0.0


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("fast,force", [(7, "all"), (8, "x"), (9, "none"), (10, "all")])
@pytest.mark.parametrize("consistent", [True, False])
def test_sixth_order_cell_matches_oracle(hc, fast, force, mode, consistent):
    """The kernels' factorised sixth-order form (rr_post_q27_o6) against the
    oracle's per-direction Hermite sums with the fourth- to sixth-order terms."""
    rng = np.random.default_rng(30 + fast + 4 * mode)
    with oracle.rr_order(6):
        for _ in range(50):
            f, rho, u, F = random_cell(oracle.D3Q27_CUDA, rng, consistent)
            if force == "x":
                F = np.array([F[0], 0.0, 0.0])
            elif force == "none":
                F = np.zeros(3)
            tau = rng.uniform(0.5005, 1.5)
            got = hc_cell(hc, 27, 0, fast, f, rho, u, F, tau, mode)
            ref = oracle.rr_cell(oracle.D3Q27_CUDA, f, rho, *u, F, tau, mode)
            assert np.abs(got - ref).max() < 1e-15
        # the third-order oracle differs (the extra orders are seen)
    ref3 = oracle.rr_cell(oracle.D3Q27_CUDA, f, rho, *u, F, tau, mode)
    assert np.abs(got - ref3).max() > 1e-10


def test_sixth_order_cell_fp32(hc):
    rng = np.random.default_rng(40)
    with oracle.rr_order(6):
        for _ in range(50):
            f, rho, u, F = random_cell(oracle.D3Q27_CUDA, rng, True)
            got = hc_cell(hc, 27, 1, 7, f, rho, u, F, 0.53, 0)
            ref = oracle.rr_cell(oracle.D3Q27_CUDA, f, rho, *u, F, 0.53, 0)
            assert np.abs(got - ref).max() < 1e-7


@pytest.mark.parametrize("fast", [11, 12])
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("consistent", [True, False])
def test_d3q19_moment_space_matches_oracle(hc, fast, mode, consistent):
    """The kernels' D3Q19 polynomial form (rr_post_q19, Malaspinas third-order
    pairs, no body force) against the oracle's per-direction restatement."""
    rng = np.random.default_rng(50 + fast + 2 * mode)
    for _ in range(50):
        f, rho, u, _ = random_cell(oracle.D3Q19, rng, consistent)
        F = np.zeros(3)
        if consistent:
            c, _, _ = oracle.lattice(oracle.D3Q19)
            u = (f @ c) / rho
        tau = rng.uniform(0.5005, 1.5)
        got = hc_cell(hc, 19, 0, fast, f, rho, u, F, tau, mode)
        ref = oracle.rr_cell(oracle.D3Q19, f, rho, *u, F, tau, mode)
        assert np.abs(got - ref).max() < 1e-15


def test_d3q19_moment_space_fp32(hc):
    rng = np.random.default_rng(60)
    c, _, _ = oracle.lattice(oracle.D3Q19)
    for _ in range(50):
        f, rho, u, _ = random_cell(oracle.D3Q19, rng, False)
        rho = f.sum()
        u = (f @ c) / rho
        got = hc_cell(hc, 19, 1, 11, f, rho, u, np.zeros(3), 0.53, 0)
        ref = oracle.rr_cell(oracle.D3Q19, f, rho, *u, np.zeros(3), 0.53, 0)
        assert np.abs(got - ref).max() < 1e-7


def oracle_bgk_cell(hs, dr, u, F, tau):
    """One cell of the oracle's collision_BGK (BGK/collision.cu:3-33), shifted populations."""
    g = oracle.Fields(oracle.D3Q27_CUDA, 1, 1, 1)
    g.f[:, 0] = hs
    g.rho[:] = dr
    g.ux[:], g.uy[:], g.uz[:] = u
    g.set_force(F)
    g.collision_bgk(tau)
    return g.f[:, 0].copy()


@pytest.mark.parametrize("poly", [0, 1, 2])
def test_bgk_cell_forms_match_oracle(hc, poly):
    """The per-direction BGK update and its moment-space form (FP64, and the
    FP16-storage scaling of the latter) against the oracle's collision_BGK."""
    rng = np.random.default_rng(70 + poly)
    c, w, _ = oracle.lattice(oracle.D3Q27_CUDA)
    for _ in range(50):
        u = 0.1 * rng.uniform(-1, 1, 3)
        dr = 0.02 * rng.standard_normal()
        F = 2e-4 * rng.standard_normal(3)
        hs = w * (dr + 3 * (c @ u) + 4.5 * (c @ u) ** 2 - 1.5 * u @ u) + 3e-3 * w * rng.standard_normal(27)
        tau = rng.uniform(0.51, 1.5)
        out = np.zeros(27)
        hc.hc_bgk_cell(0, poly, p(hs), ctypes.c_double(dr), ctypes.c_double(u[0]), ctypes.c_double(u[1]),
                       ctypes.c_double(u[2]), p(F), ctypes.c_double(tau), p(out))
        assert np.abs(out - oracle_bgk_cell(hs, dr, u, F, tau)).max() < 1e-16


def fp32_mass_residuals(hc, n, seed, fast=1):
    """Post-collision mass minus the cell's density excess, in FP32 arithmetic,
    over n random cells whose velocity is of order 0.1-0.2."""
    rng = np.random.default_rng(seed)
    c, w, _ = oracle.lattice(oracle.D3Q27_CUDA)
    r = np.empty(n)
    for k in range(n):
        u = rng.uniform(0.1, 0.2, 3) * rng.choice([-1.0, 1.0], 3)
        f = oracle.feq(oracle.D3Q27_CUDA, 1 + 0.02 * rng.standard_normal(), *u) \
            + 3e-3 * w * rng.standard_normal(27)
        F = np.array([1e-5, 0.0, 0.0])
        rho = f.sum()
        u = (f @ c + 0.5 * F) / rho
        out = hc_cell(hc, 27, 1, fast, f, rho, u, F, 0.53, 0)
        r[k] = (out - w).sum() - float(np.float32(rho - 1.0))
    return r


def test_fp32_collision_mass_has_no_systematic_residual(hc):
    """The FP32 moment-space collision folds the lattice weights into its
    polynomial's coefficients.  The terms that cancel in the mass sum (for
    example -3/2 rho u^2 of the constant coefficient against the c^2
    coefficients) must carry the same rounding: folded into each constant
    separately they leave a residual of one sign, ~1e-9 per cell at these
    velocities, which a 64-box FP32 ensemble shows as a mass drift.  The mean
    residual over 40 000 cells must be rounding noise (within 5 standard
    errors of zero)."""
    r = fp32_mass_residuals(hc, 40000, 31)
    se = r.std() / np.sqrt(r.size)
    print("FP32 mass residual: mean %.3g, std %.3g, standard error %.3g" % (r.mean(), r.std(), se))
    assert abs(r.mean()) < 5 * se
