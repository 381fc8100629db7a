"""Pins of the oracle (CPU, no GPU).

Each check is fixed by something other than the oracle's own formula: the
reference's committed output (Taylor-Green), the reference's Fortran compiled
from source (D3Q19 Taylor-Green), closed-form lattice moments, conservation
laws, and an independent matrix formulation of WALE in long double.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN

KINDS = [oracle.D3Q27_CUDA, oracle.D3Q27_FORTRAN, oracle.D3Q19]
CS = 1.0 / np.sqrt(3.0)


# ---------------------------------------------------------------- lattice
@pytest.mark.parametrize("kind", KINDS)
def test_lattice_moments(kind):
    c, w, opp = oracle.lattice(kind)
    q = len(w)
    assert q == (19 if kind == oracle.D3Q19 else 27)
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.allclose(w @ c, 0.0, atol=1e-16)
    assert np.allclose(np.einsum("i,ia,ib->ab", w, c, c), np.eye(3) / 3.0, atol=1e-16)
    assert np.allclose(np.einsum("i,ia,ib,ic->abc", w, c, c, c), 0.0, atol=1e-16)
    assert (c[opp] == -c).all()
    assert len({tuple(v) for v in c}) == q
    if q == 27:  # D3Q27 weights factorise over the axes: W(0) = 2/3, W(+-1) = 1/6
        W = np.where(c == 0, 2.0 / 3.0, 1.0 / 6.0).prod(axis=1)
        assert np.allclose(w, W, rtol=1e-15)


def test_channel_ordering_table():
    """The CUDA channel order (initial.cu:67-69): faces 1-6, edges 7-18, corners 19-26,
    with bottom-wall incoming set {2,7,8,12,16,19,20,23,24} (streaming.cu:34-44)."""
    c, _, opp = oracle.lattice(oracle.D3Q27_CUDA)
    assert [int(i) for i in np.where(c[:, 1] > 0)[0]] == [2, 7, 8, 12, 16, 19, 20, 23, 24]
    assert list(opp) == [0, 3, 4, 1, 2, 6, 5, 9, 10, 7, 8, 17, 18, 15, 16, 13, 14, 11, 12, 25, 26, 23, 24, 21,
                         22, 19, 20]


@pytest.mark.parametrize("kind", [oracle.D3Q27_CUDA, oracle.D3Q27_FORTRAN])
def test_hermite_orthogonality_d3q27(kind):
    """sum_i w_i H_a H_b = delta_ab E[H_a^2] of the Gaussian (1, 1, 2, 2, 1) for the
    first/off-diagonal/diagonal second/aab third/xyz orders."""
    _, w, _ = oracle.lattice(kind)
    H = oracle.hermite(kind)
    G = np.einsum("i,ia,ib->ab", w, H, H)
    norms = np.array([1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 1], dtype=float)
    # diagonal second-order polynomials are mutually orthogonal: E[(x^2-1)(y^2-1)] = 0
    assert np.allclose(G, np.diag(norms), atol=1e-14)


# ---------------------------------------------------------------- equilibrium
@pytest.mark.parametrize("kind", KINDS)
def test_feq_moments(kind):
    rng = np.random.default_rng(0)
    c, w, _ = oracle.lattice(kind)
    H = oracle.hermite(kind)
    for _ in range(20):
        rho = 1.0 + 0.05 * rng.standard_normal()
        u = 0.1 * rng.uniform(-1, 1, 3)
        f = oracle.feq(kind, rho, *u)
        assert abs(f.sum() - rho) < 1e-14
        assert np.allclose(f @ c, rho * u, atol=1e-15)
        assert np.allclose(np.einsum("i,ia,ib->ab", f, c, c), rho * (np.eye(3) / 3 + np.outer(u, u)), atol=1e-15)
        ut = u / CS
        a = {"210": ut[0] ** 2 * ut[1], "201": ut[0] ** 2 * ut[2], "021": ut[1] ** 2 * ut[2],
             "120": ut[1] ** 2 * ut[0], "102": ut[2] ** 2 * ut[0], "012": ut[2] ** 2 * ut[1]}
        col = {"210": 9, "201": 10, "021": 11, "120": 12, "102": 13, "012": 14}
        if len(w) == 27:
            # the full third-order Hermite coefficients are reproduced
            for k in a:
                assert abs(f @ H[:, col[k]] - rho * a[k]) < 1e-15
            assert abs(f @ H[:, 15] - rho * ut[0] * ut[1] * ut[2]) < 1e-15
            # hence the raw third moments
            assert abs(f @ (c[:, 0] * c[:, 1] * c[:, 2]) - rho * u.prod()) < 1e-16
            assert abs(f @ (c[:, 0] ** 2 * c[:, 1]) - rho * (u[0] ** 2 * u[1] + u[1] / 3)) < 1e-16
        else:
            # D3Q19: the Malaspinas pairs H+ = H_xxy + H_yzz, H- = H_xxy - H_yzz are reproduced
            for p, q in (("210", "012"), ("102", "120"), ("021", "201")):
                hp, hm = H[:, col[p]] + H[:, col[q]], H[:, col[p]] - H[:, col[q]]
                assert abs(f @ hp - rho * (a[p] + a[q])) < 1e-15
                assert abs(f @ hm - rho * (a[p] - a[q])) < 1e-15


# ---------------------------------------------------------------- collision
def consistent_cell(kind, rng, F):
    """Random populations with (rho, u) consistent with them (macrovar)."""
    c, w, _ = oracle.lattice(kind)
    rho0 = 1 + 0.02 * rng.standard_normal()
    u0 = 0.08 * rng.uniform(-1, 1, 3)
    f = oracle.feq(kind, rho0, *u0) + 2e-3 * w * rng.standard_normal(len(w))
    rho = f.sum()
    u = (f @ c + 0.5 * np.asarray(F)) / rho
    return f, rho, u, c, w


@pytest.mark.parametrize("kind", KINDS)
def test_collision_conservation(kind):
    rng = np.random.default_rng(1)
    for _ in range(20):
        F = 1e-4 * rng.standard_normal(3)
        f, rho, u, c, w = consistent_cell(kind, rng, F)
        tau = rng.uniform(0.51, 2.0)
        out = oracle.rr_cell(kind, f, rho, *u, F, tau, 0)
        assert abs(out.sum() - f.sum()) < 1e-14                           # mass
        # momentum: j* = rho u + rho F/2 = j + F/2 + rho F/2 (He-type term Fbar feq)
        assert np.allclose(out @ c, f @ c + 0.5 * np.asarray(F) * (1 + rho), atol=1e-15)


@pytest.mark.parametrize("kind", KINDS)
def test_collision_special_cases(kind):
    rng = np.random.default_rng(2)
    c, w, _ = oracle.lattice(kind)
    for _ in range(10):
        rho = 1 + 0.02 * rng.standard_normal()
        u = 0.08 * rng.uniform(-1, 1, 3)
        feq = oracle.feq(kind, rho, *u)
        zero = (0.0, 0.0, 0.0)
        # equilibrium in -> equilibrium out (any tau)
        assert np.allclose(oracle.rr_cell(kind, feq, rho, *u, zero, 0.7, 0), feq, atol=1e-15)
        # tau = 1 relaxes any state to feq
        f = feq + 1e-3 * w * rng.standard_normal(len(w))
        assert np.allclose(oracle.rr_cell(kind, f, rho, *u, zero, 1.0, 0), feq, atol=1e-15)
        # a regularised state is a fixed point of the projection (omega' = 1, tau -> inf)
        freg = oracle.rr_cell(kind, f, rho, *u, zero, 1.0, 1)  # RRBC: feq + fneq(A2 of f)
        again = oracle.rr_cell(kind, freg, rho, *u, zero, 1e300, 0)
        assert np.allclose(again, freg, atol=1e-15)
        # RRBC with equilibrium input and a force returns feq - Fbar feq
        F = 1e-4 * rng.standard_normal(3)
        Fbar = 1.5 * (c @ F - u @ F)
        out = oracle.rr_cell(kind, feq, rho, *u, F, 1.0, 1)
        # the projection sees f + Fbar feq, so fneq = w H2:(proj of Fbar feq)/2 + ...: compare moments
        assert abs(out.sum() - rho) < 1e-14
        assert out.shape == feq.shape
        assert np.isfinite(Fbar).all()


# ---------------------------------------------------------------- streaming, BB, macrovar
@pytest.mark.parametrize("kind", [oracle.D3Q27_CUDA, oracle.D3Q19])
def test_streaming_is_a_shift(kind):
    lx, ly, lz = 5, 4, 3
    g = oracle.Fields(kind, lx, ly, lz)
    g.f[:] = np.arange(g.f.size, dtype=float).reshape(g.f.shape)
    g.streaming()
    c, _, _ = oracle.lattice(kind)
    F = g.f.reshape(g.q, lz, ly, lx)
    T = g.ftmp.reshape(g.q, lz, ly, lx)
    for i in range(g.q):
        assert (np.roll(F[i], shift=(c[i, 2], c[i, 1], c[i, 0]), axis=(0, 1, 2)) == T[i]).all()


def test_bounce_back_same_node():
    kind = oracle.D3Q27_CUDA
    g = oracle.Fields(kind, 4, 5, 3)
    g.f[:] = np.random.default_rng(3).standard_normal(g.f.shape)
    g.ftmp[:] = -7.0
    g.bb(walls=(0, 1, 0))
    c, _, opp = oracle.lattice(kind)
    F = g.f.reshape(27, 3, 5, 4)
    T = g.ftmp.reshape(27, 3, 5, 4)
    for i in range(27):
        if c[i, 1] > 0:
            assert (T[i, :, 0, :] == F[opp[i], :, 0, :]).all()
        if c[i, 1] < 0:
            assert (T[i, :, 4, :] == F[opp[i], :, 4, :]).all()
        if c[i, 1] == 0:
            assert (T[i] == -7.0).all()


@pytest.mark.parametrize("kind", [oracle.D3Q27_CUDA, oracle.D3Q19])
def test_closed_box_conserves_mass(kind):
    """Bounce-back on all six faces (no lid, no RR regularisation, no force):
    total mass is conserved step after step."""
    from helpers import random_state
    g = random_state(kind, 6, 5, 4, seed=4, force=(0.0, 0.0, 0.0))
    g.step(1, 0.6, walls=(1, 1, 1), wall_rr=0)   # makes rho, u consistent with f
    m0 = g.f.sum()
    g.step(30, 0.6, walls=(1, 1, 1), wall_rr=0)
    assert abs(g.f.sum() - m0) < 1e-11


def test_macrovar_closed_form():
    from helpers import random_state
    g = random_state(oracle.D3Q27_CUDA, 4, 3, 2, seed=5)
    g.macrovar()
    c, _, _ = oracle.lattice(oracle.D3Q27_CUDA)
    rho = g.f.sum(axis=0)
    j = c.T @ g.f
    assert np.allclose(g.rho, rho, atol=1e-15)
    assert np.allclose(g.ux, (j[0] + 0.5 * g.fx) / rho, atol=1e-16)
    assert np.allclose(g.uz, (j[2] + 0.5 * g.fz) / rho, atol=1e-16)


def test_rrbc_leaves_outgoing_populations():
    from helpers import random_state
    g = random_state(oracle.D3Q27_CUDA, 5, 6, 4, seed=6)
    g.streaming(); g.bb(); g.swap(); g.macrovar()
    before = g.f.copy()
    g.rrbc(walls=(0, 1, 0))
    c, _, _ = oracle.lattice(oracle.D3Q27_CUDA)
    B = before.reshape(27, 4, 6, 5)
    A = g.f.reshape(27, 4, 6, 5)
    assert (A[:, :, 1:5, :] == B[:, :, 1:5, :]).all()          # interior untouched
    for i in range(27):
        if c[i, 1] <= 0:
            assert (A[i, :, 0, :] == B[i, :, 0, :]).all()      # bottom: only c_y > 0 replaced
        if c[i, 1] >= 0:
            assert (A[i, :, 5, :] == B[i, :, 5, :]).all()


# ---------------------------------------------------------------- WALE
def wale_matrix(g, cw=0.5):
    """tau_t from the textbook definition in long double: S = sym(g),
    S^d = sym(g.g) - tr(g.g) I/3, tau_t = 3 Cw^2 (S^d:S^d)^1.5 / ((S:S)^2.5 + (S^d:S^d)^1.25)."""
    g = np.asarray(g, dtype=np.longdouble)
    S = 0.5 * (g + g.T)
    g2 = g @ g
    Sd = 0.5 * (g2 + g2.T) - np.trace(g2) / 3 * np.eye(3, dtype=np.longdouble)
    ss, sd = (S * S).sum(), (Sd * Sd).sum()
    den = ss ** np.longdouble(2.5) + sd ** np.longdouble(1.25)
    return 0.0 if den == 0 else float(3 * np.longdouble(cw) ** 2 * sd ** np.longdouble(1.5) / den)


def test_wale_linear_fields_against_matrix_form():
    rng = np.random.default_rng(7)
    lx, ly, lz = 7, 7, 7
    z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    X = np.stack([x.ravel(), y.ravel(), z.ravel()]).astype(float)
    for _ in range(5):
        A = 1e-3 * rng.standard_normal((3, 3))
        g = oracle.Fields(oracle.D3Q27_CUDA, lx, ly, lz)
        U = A @ X + 1e-2 * rng.standard_normal((3, 1))
        g.ux[:], g.uy[:], g.uz[:] = U
        g.les_wale(cw=0.5, guard=1)
        T = g.taut.reshape(lz, ly, lx)
        expect = wale_matrix(A)
        inner = T[1:-1, 1:-1, 1:-1]
        assert np.allclose(inner, expect, rtol=1e-12, atol=1e-18)


def test_wale_pure_shear_and_rest():
    lx, ly, lz = 6, 6, 6
    g = oracle.Fields(oracle.D3Q27_CUDA, lx, ly, lz)
    y = np.repeat(np.arange(ly), lx)[None, :].repeat(lz, axis=0).ravel().astype(float)
    g.ux[:] = 1e-3 * y
    g.les_wale(guard=1)
    T = g.taut.reshape(lz, ly, lx)
    assert np.abs(T[:, 1:-1, :]).max() < 1e-30      # S^d = 0 for simple shear
    g.ux[:] = 0.0
    g.les_wale(guard=1)
    assert (g.taut == 0.0).all()                     # guarded 0/0
    g.les_wale(guard=0)
    assert np.isnan(g.taut).all()                    # the reference's 0/0 (LES.cu:97)


# ---------------------------------------------------------------- forcing, init
def test_forcing_gp_properties():
    lx, ly, lz = 24, 32, 16
    g = oracle.Fields(oracle.D3Q27_CUDA, lx, ly, lz)
    fin = 1e-6
    g.forcing_gp(fin, 1000, 0.3)        # sin(2 pi 1000/2000) = 0: no perturbation
    assert np.allclose(g.fx, fin, atol=1e-21) and np.abs(g.fy).max() < 1e-20
    g.forcing_gp(fin, 500, 0.3)
    FY = g.fy.reshape(lz, ly, lx)
    assert np.abs(FY).max() > 0
    assert np.allclose(FY.sum(axis=0), 0.0, atol=1e-18)   # cos(2 pi 2 z/LZ) averages out in z
    FX = g.fx.reshape(lz, ly, lx)
    assert np.allclose(FX[:, 0, :], fin) and np.allclose(FX[:, ly // 2, :], fin)  # outside both bands


def test_init_arrays_properties():
    lx, ly, lz = 16, 12, 8
    visc = 0.000622
    u0 = 2.0 * 180.0 * visc / ly
    ystar = visc / u0
    fin = 2.0 * u0 * u0 / ly
    g = oracle.Fields(oracle.D3Q27_CUDA, lx, ly, lz)
    g.init_arrays(u0, ystar, fin)
    c, _, _ = oracle.lattice(oracle.D3Q27_CUDA)
    assert np.allclose(g.f.sum(axis=0), 1.0, atol=1e-15)
    assert np.allclose(c[:, 0] @ g.f, -0.5 * fin, atol=1e-18)       # rest populations shifted by -F/2
    UX = g.ux.reshape(lz, ly, lx)
    iy = np.arange(ly)
    yplus = np.where(iy <= ly // 2, (iy + 0.5) / ystar, (ly - iy - 0.5) / ystar)
    mean = np.where(yplus <= 10.8, u0 * yplus, u0 * (np.log10(yplus) / 0.41 + 5.0))
    # the x-z average removes the sin(x + z) perturbation (initial.cu:300)
    assert np.allclose(UX.mean(axis=(0, 2)), mean, rtol=1e-12)


def _mode_fit(field, basis):
    """Least-squares amplitude of `basis` in `field` over the x-z plane and the
    largest residual |field - a basis|."""
    a = float((field * basis).sum() / (basis * basis).sum())
    return a, float(np.abs(field - a * basis).max())


def test_forcing_gp_bands_are_single_divergence_free_modes():
    """forcing_GP (macrovar.cu:87-112), lower band (h0 < iy < LYq + h0): f_x - F,
    f_y and f_z are each one Fourier mode in x and z -- sin(k_x x) cos(k_z z),
    cos(k_x x) cos(k_z z), cos(k_x x) sin(k_z z) with k_x = 2 pi 3/LX,
    k_z = 2 pi 2/LZ -- whose y-amplitudes are A_x sin(k_y t), A_y (1 - cos(k_y t)),
    A_z sin(k_y t) in t = iy - h0, k_y = 2 pi/LYq; the field is divergence-free,
    A_x k_x + A_y k_y + A_z k_z = 0 (-2 pi + pi + pi), and A(istep) follows
    sin(2 pi istep/2000).  A swapped sin/cos, a wrong wavenumber, a shifted band
    row or a wrong coefficient in any component fails one of these."""
    lx, ly, lz = 48, 32, 16
    h0, lyq = 2, ly // 4
    kx, kz, ky = 2 * np.pi * 3 / lx, 2 * np.pi * 2 / lz, 2 * np.pi / lyq
    X = np.arange(lx)[None, :]
    Z = np.arange(lz)[:, None]
    fin = 1e-5
    g = oracle.Fields(oracle.D3Q27_CUDA, lx, ly, lz)
    amps = {}
    for istep in (250, 500):
        g.forcing_gp(fin, istep, 0.3)
        FX, FY, FZ = (a.reshape(lz, ly, lx) for a in (g.fx, g.fy, g.fz))
        rows = [iy for iy in range(ly) if h0 < iy < lyq + h0]
        assert rows == list(range(h0 + 1, h0 + lyq))
        A = []
        for iy in rows:
            ax, rx = _mode_fit(FX[:, iy, :] - fin, np.sin(kx * X) * np.cos(kz * Z))
            ay, ry = _mode_fit(FY[:, iy, :], np.cos(kx * X) * np.cos(kz * Z))
            az, rz = _mode_fit(FZ[:, iy, :], np.cos(kx * X) * np.sin(kz * Z))
            assert max(rx, ry, rz) < 1e-12 * fin, (iy, rx, ry, rz)
            A.append((ax, ay, az))
        A = np.array(A)
        t = np.array(rows) - h0
        Ax, Ay, Az = (_mode_fit(A[:, k], b)[0] for k, b in
                      ((0, np.sin(ky * t)), (1, 1 - np.cos(ky * t)), (2, np.sin(ky * t))))
        assert np.allclose(A[:, 0], Ax * np.sin(ky * t), rtol=0, atol=1e-12 * fin)
        assert np.allclose(A[:, 1], Ay * (1 - np.cos(ky * t)), rtol=0, atol=1e-12 * fin)
        assert np.allclose(A[:, 2], Az * np.sin(ky * t), rtol=0, atol=1e-12 * fin)
        assert abs(Ax * kx + Ay * ky + Az * kz) < 1e-12 * abs(Ay * ky)      # divergence-free
        # amplitude ratios LX/beta : LYq/2 : LZ/(2 gamma) with beta = 3, gamma = 2 (sign of f_x: -)
        assert np.isclose(-Ax / Ay, (lx / 3.0) / (0.5 * lyq), rtol=1e-13)
        assert np.isclose(Az / Ay, (0.5 * lz / 2.0) / (0.5 * lyq), rtol=1e-13)
        amps[istep] = Ay
        # every other row outside the two bands carries the uniform force
        upper = [iy for iy in range(ly) if 0.75 * ly - h0 < iy < ly - h0]
        assert upper == list(range(ly - h0 - lyq + 1, ly - h0))
        other = [iy for iy in range(ly) if iy not in rows and iy not in upper]
        assert (FX[:, other, :] == fin).all() and (FY[:, other, :] == 0).all() and (FZ[:, other, :] == 0).all()
    # A(istep) = (100 beta / LY) sin(2 pi istep / 2000): A(500) / A(250) = 1 / sin(pi/4)
    assert np.isclose(amps[500] / amps[250], np.sqrt(2.0), rtol=1e-13)
    assert np.isclose(amps[500], fin * (100 * 3.0 / ly) * 0.5 * lyq, rtol=1e-13)


def test_forcing_gp_upper_band_is_x_independent():
    """The upper band (0.75 LY - h0 < iy < LY - h0) takes its x phase as the
    integer ix / LX (always 0) plus the random phase (macrovar.cu:106-112):
    exactly x-independent, the z modes cos(k_z z) / sin(k_z z), the mirrored
    y-profile of the lower band (t = iy - (LY - h0 - LYq)) with the opposite
    signs, and the phase entering as sin / cos(2 pi 3 phase).  Without the x
    term its divergence is -2 pi A sin(k_y t) cos(6 pi phase) cos(k_z z), not 0."""
    lx, ly, lz = 48, 32, 16
    h0, lyq = 2, ly // 4
    kz, ky = 2 * np.pi * 2 / lz, 2 * np.pi / lyq
    Z = np.arange(lz)[:, None] + 0 * np.arange(lx)[None, :]
    fin = 1e-5
    g = oracle.Fields(oracle.D3Q27_CUDA, lx, ly, lz)
    lower = {}
    g.forcing_gp(fin, 500, 0.0)
    FX, FY, FZ = (a.reshape(lz, ly, lx) for a in (g.fx, g.fy, g.fz))
    for iy in range(h0 + 1, h0 + lyq):
        lower[iy - h0] = FY[0, iy, 0]    # x = 0, z = 0: A_y (1 - cos(k_y t))
    for phase in (0.0, 0.13, 0.3):
        g.forcing_gp(fin, 500, phase)
        FX, FY, FZ = (a.reshape(lz, ly, lx) for a in (g.fx, g.fy, g.fz))
        for iy in range(ly - h0 - lyq + 1, ly - h0):
            t = iy - (ly - h0 - lyq)
            for F in (FX, FY, FZ):
                assert (F[:, iy, :] == F[:, iy, :1]).all()                   # x-independent
            ay, ry = _mode_fit(FY[:, iy, :], np.cos(kz * Z))
            az, rz = _mode_fit(FZ[:, iy, :], np.sin(kz * Z))
            ax, rx = _mode_fit(FX[:, iy, :] - fin, np.cos(kz * Z))
            assert max(rx, ry, rz) < 1e-12 * fin
            c6, s6 = np.cos(6 * np.pi * phase), np.sin(6 * np.pi * phase)
            assert np.isclose(ay, -lower[t] * c6, rtol=1e-12, atol=1e-12 * fin)
            # A_z / A_y and A_x / A_y as in the lower band, times the phase factors
            ratio_z = (0.5 * lz / 2.0) / (0.5 * lyq) * np.sin(ky * t) / (1 - np.cos(ky * t))
            ratio_x = (lx / 3.0) / (0.5 * lyq) * np.sin(ky * t) / (1 - np.cos(ky * t))
            assert np.isclose(az, -lower[t] * c6 * ratio_z, rtol=1e-12, atol=1e-12 * fin)
            assert np.isclose(ax, lower[t] * s6 * ratio_x, rtol=1e-12, atol=1e-12 * fin)


def test_init_arrays_perturbation_and_populations():
    """init_arrays (initial.cu:277-324): u_x minus its x-z plane mean is the
    single oblique mode sin(x9 + z9) at the cell centres (x9 = 2 pi (ix + 1/2)/LX,
    no cos part: the half-cell phase), with amplitude ccc1 y+ -- linear in y+
    from the nearer wall, the lower formula up to iy = LY/2 inclusive, with
    ccc1 = -LX A9 u*/(2 pi ystar cc^2) -- and the populations are the rest
    state shifted by the force, w_i (1 - 3 F c_ix / 2), whatever u is.  A y+
    from the wrong wall, a missing half cell, a z-only or x-only mode, or
    populations at the equilibrium of u fail one of these."""
    lx, ly, lz = 16, 12, 8
    visc = 0.000622
    u0 = 2.0 * 180.0 * visc / ly
    ystar = visc / u0
    fin = 2.0 * u0 * u0 / ly
    g = oracle.Fields(oracle.D3Q27_CUDA, lx, ly, lz)
    g.init_arrays(u0, ystar, fin)
    c, w, _ = oracle.lattice(oracle.D3Q27_CUDA)
    expect_f = w[:, None] * (1.0 - 1.5 * fin * c[:, 0:1].astype(float))
    assert np.abs(g.f - expect_f).max() < 1e-17
    UX = g.ux.reshape(lz, ly, lx)
    x9 = 2 * np.pi * (np.arange(lx) + 0.5) / lx
    z9 = 2 * np.pi * (np.arange(lz) + 0.5) / lz
    S = np.sin(x9[None, :] + z9[:, None])
    C = np.cos(x9[None, :] + z9[:, None])
    slope = []
    for iy in range(ly):
        pert = UX[:, iy, :] - UX[:, iy, :].mean()
        a, res = _mode_fit(pert, S)
        b, _ = _mode_fit(pert, C)
        assert res < 1e-15 and abs(b) < 1e-16, (iy, res, b)
        yplus = (iy + 0.5) / ystar if iy <= ly // 2 else (ly - iy - 0.5) / ystar
        slope.append(a / yplus)
    assert np.allclose(slope, slope[0], rtol=1e-12)
    ccc1 = -lx * 0.1 * np.sqrt(2.0) * u0 / (2 * np.pi * ystar * 60.0 ** 2)
    assert np.isclose(slope[0], ccc1, rtol=1e-12)
    assert np.all(g.uy == 0) and np.all(g.uz == 0) and np.all(g.rho == 1.0)


@pytest.mark.parametrize("seed", [1, 2])
def test_diag_and_profiles_match_long_double_sums(seed):
    """orc_diag (diag_flow's sums, output.cu:24-35, with the momentum and kinetic
    energy the library also reports) and orc_profiles (output_profile's plane
    statistics, output.cu:182-215: plane means and the covariances of the
    fluctuations about them, divided by LX*LZ) against numpy long-double sums
    over random fields.  A dropped term, a covariance about zero instead of the
    plane mean, a division by the wrong count or a transposed plane index
    fails at far above the 1e-13 tolerance."""
    lx, ly, lz = 13, 9, 7
    rng = np.random.default_rng(seed)
    g = oracle.Fields(oracle.D3Q27_CUDA, lx, ly, lz)
    n = g.n
    g.rho[:] = 1.0 + 0.01 * rng.standard_normal(n)
    for a in (g.ux, g.uy, g.uz):
        a[:] = 0.3 + 0.05 * rng.standard_normal(n)
    g.taut[:] = rng.uniform(0, 0.1, n)
    L = np.longdouble
    r, u, v, w = (a.astype(L) for a in (g.rho, g.ux, g.uy, g.uz))
    want = [u.sum(), v.sum(), w.sum(), r.sum(), (r - 1).sum(), (r * u).sum(), (r * v).sum(), (r * w).sum(),
            (L(0.5) * r * (u * u + v * v + w * w)).sum()]
    got = g.diag()
    for k in range(9):
        assert abs(got[k] - float(want[k])) <= 1e-13 * max(1.0, abs(float(want[k]))), k
    prof = g.profiles(with_taut=True)
    shape = (lz, ly, lx)
    U, V, W, T = (a.astype(L).reshape(shape) for a in (g.ux, g.uy, g.uz, g.taut))
    for iy in range(ly):
        pu, pv, pw, pt = (A[:, iy, :] for A in (U, V, W, T))
        du, dv, dw = pu - pu.mean(), pv - pv.mean(), pw - pw.mean()
        cnt = lx * lz
        ref = [pu.mean(), pv.mean(), pw.mean(), (du * du).sum() / cnt, (dv * dv).sum() / cnt, (dw * dw).sum() / cnt,
               (du * dv).sum() / cnt, (du * dw).sum() / cnt, (dv * dw).sum() / cnt, pt.mean()]
        for k in range(10):
            assert abs(prof[iy, k] - float(ref[k])) <= 1e-13 * max(1e-3, abs(float(ref[k]))), (iy, k)
    # a plane-constant field has zero covariances: the fluctuations are about the plane mean
    g.ux[:] = np.repeat(np.arange(ly, dtype=float), lx)[None, :].repeat(lz, axis=0).ravel()
    assert np.abs(g.profiles()[:, 3]).max() == 0.0


def test_init_channel_fields_properties():
    """The channel configuration's seeded initial field (the library's
    rrlbm_init_channel; no reference counterpart): rho = 1; without noise, u_x
    is the wall law U = u_tau U+(y+) (y+ in the viscous layer, ln(y+)/0.41 + 5
    above 10.8, the two branches meeting there to 4e-3) plus one streak mode
    0.1 U sin(2 pi k_z (z + 1/2)/LZ) cos(2 pi k_x (x + 1/2)/LX + phi) with
    k_x, k_z in {2, 3, 4}, and u_y = u_z = 0; the noise on each component is
    bounded by 1% of U_b, centred, with the variance of a uniform variable and
    uncorrelated between components."""
    lx, ly, lz = 48, 32, 40
    u_tau, visc = 0.0045, 0.0045 * 16 / 180.0
    ystar = visc / u_tau
    assert abs(np.log(10.8) / 0.41 + 5.0 - 10.8) < 4e-3
    iy = np.arange(ly)
    d = np.where(iy < ly // 2, iy + 0.5, ly - iy - 0.5) / ystar
    U = u_tau * np.where(d <= 10.8, d, np.log(d) / 0.41 + 5.0)
    ub = U.mean()
    X = (np.arange(lx) + 0.5)[None, :]
    Z = (np.arange(lz) + 0.5)[:, None]
    for seed in (3, 8, 21):
        g = oracle.Fields(oracle.D3Q27_CUDA, lx, ly, lz)
        g.init_channel_fields(u_tau, visc, seed, amp_streak=0.1, amp_noise=0.0)
        assert np.all(g.rho == 1.0) and np.all(g.uy == 0) and np.all(g.uz == 0)
        UX = g.ux.reshape(lz, ly, lx)
        assert np.allclose(UX.mean(axis=(0, 2)), U, rtol=1e-12, atol=0)
        found = []
        for kx in (1, 2, 3, 4, 5):
            for kz in (1, 2, 3, 4, 5):
                for bx in (np.cos, np.sin):
                    b = np.sin(2 * np.pi * kz * Z / lz) * bx(2 * np.pi * kx * X / lx)
                    a, _ = _mode_fit(UX[:, 5, :] - U[5], b)
                    if abs(a) > 1e-9 * U[5]:
                        found.append((kx, kz))
        kx, kz = found[0]
        assert set(found) == {(kx, kz)} and kx in (2, 3, 4) and kz in (2, 3, 4), found
        for j in range(ly):
            pert = UX[:, j, :] - U[j]
            ac, _ = _mode_fit(pert, np.sin(2 * np.pi * kz * Z / lz) * np.cos(2 * np.pi * kx * X / lx))
            as_, _ = _mode_fit(pert, np.sin(2 * np.pi * kz * Z / lz) * np.sin(2 * np.pi * kx * X / lx))
            assert np.isclose(np.hypot(ac, as_), 0.1 * U[j], rtol=1e-12)
        g.init_channel_fields(u_tau, visc, seed, amp_streak=0.0, amp_noise=0.01)
        nx_, ny_, nz_ = g.ux - np.repeat(U, lx)[None, :].repeat(lz, axis=0).ravel(), g.uy, g.uz
        for e in (nx_, ny_, nz_):
            assert np.abs(e).max() <= 0.01 * ub
            assert abs(e.mean()) < 4 * 0.01 * ub / np.sqrt(3 * e.size)
            assert abs(e.var() / (0.01 * ub) ** 2 - 1 / 3) < 0.02
        assert abs(np.corrcoef(nx_, ny_)[0, 1]) < 0.02 and abs(np.corrcoef(ny_, nz_)[0, 1]) < 0.02


def splitmix64(x):
    m = (1 << 64) - 1
    x = (x + 0x9E3779B97F4A7C15) & m
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & m
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & m
    return x ^ (x >> 31)


def test_counter_rng():
    """Known-answer: splitmix64(0) = 0xE220A8397B1DCDAF (Vigna's reference generator)."""
    assert splitmix64(0) == 0xE220A8397B1DCDAF
    for seed, stream, ctr in ((1, 10, 5), (2024, 12, 123456)):
        h = splitmix64(seed ^ splitmix64((stream * 0xD1B54A32D192ED03 + ctr) & ((1 << 64) - 1)))
        assert oracle.uniform(seed, stream, ctr) == (h >> 11) / 2.0 ** 53


# ---------------------------------------------------------------- Taylor-Green fixtures
def midline(v, rho, L, LZ):
    V = v.reshape(LZ, L, L)
    R = rho.reshape(LZ, L, L)
    vm = 0.5 * (V[0, L // 2 - 1, :] + V[0, L // 2, :])
    pm = 0.5 * ((R[0, L // 2 - 1, :] - 1) / 3 + (R[0, L // 2, :] - 1) / 3)
    return vm, pm


def test_tg_d3q19_matches_reference_fortran():
    """Oracle with REAL*4 intermediates vs the reference D3Q19RR.f90 built from
    source (L = 50, 10 000 steps); the file prints 10 decimals."""
    ref = np.loadtxt(os.path.join(GOLDEN, "tg_d3q19_L50_it10000_fortran_midline.txt"))
    f, rho, u, v, w, tau = oracle.tg_reference(oracle.D3Q19, 50, 2, 10000, real4=True)
    vm, pm = midline(v, rho, 50, 2)
    assert np.abs(vm - ref[:, 1]).max() < 1.0e-10
    assert np.abs(pm - ref[:, 2]).max() < 1.0e-10


@pytest.mark.skipif(os.environ.get("RRLBM_FAST_TESTS") == "1", reason="long run skipped on request")
def test_tg_d3q27_matches_reference_output():
    """Oracle with REAL*4 intermediates vs the reference's committed
    Taylor_Green/RRBMa0.01L100.txt (100x100x2, 20 000 steps), to its print
    resolution; the pure-FP64 oracle differs from it by the REAL*4 rounding."""
    ref = np.loadtxt(os.path.join(GOLDEN, "tg_d3q27_L100_it20000_fortran_midline.txt"))
    f, rho, u, v, w, tau = oracle.tg_reference(oracle.D3Q27_FORTRAN, 100, 2, 20000, real4=True)
    vm, pm = midline(v, rho, 100, 2)
    assert np.abs(vm - ref[:, 1]).max() < 1.0e-10
    assert np.abs(pm - ref[:, 2]).max() < 1.0e-10
    fp64 = np.loadtxt(os.path.join(GOLDEN, "tg_d3q27_L100_it20000_oracle_fp64_midline.txt"))
    assert np.abs(fp64[:, 1] - ref[:, 1]).max() < 3e-6


def test_laminar_matches_reference_fields():
    """The Laminar forced channel (Laminar/: 60x20x2, walls at both x faces,
    uniform force in y, RR regularisation of the wall-incoming populations)
    restated by the oracle reproduces the reference's committed FP64 fields
    (Laminar/output/uy{5000,10000}.dat, profiles in the golden file) to
    rounding, including the reference's wall-loop Fbar left over from the
    sweep's last node; the same run with each wall node's own Fbar differs by
    1e-11..3e-10 of u_y (the quirk's size, growing with the step count), so the
    fixture tells the two apart."""
    ref = np.loadtxt(os.path.join(GOLDEN, "laminar_60x20x2_uy_profiles.txt"))
    runs = {}
    for stale in (True, False):
        g, prm = oracle.laminar_start()
        assert abs(prm["tau"] - 0.68) < 1e-15
        done, out = 0, []
        for it in (5000, 10000):
            g.step_laminar(it - done, prm["tau"], stale_fbar=stale)
            done = it
            out.append(g.uy.reshape(2, 20, 60).copy())
            assert np.abs(g.uz).max() < 1e-15
        runs[stale] = out
    for k, col in enumerate((1, 2)):
        scale = np.abs(ref[:, col]).max()
        assert np.abs(runs[True][k] - ref[None, None, :, col]).max() < 1e-13 * scale
        diff = np.abs(runs[False][k] - ref[None, None, :, col]).max()
        assert 1e-12 < diff < 1e-8  # measured 9.4e-12 (5000 steps), 3.0e-11 (10000)


def test_cavity_matches_reference_solver():
    """The reference's lid-driven cavity solver (Cavity/*.f90 compiled from
    its sources by oracle/build_oracle.py: 128x128x2, D3Q27 RR, walls on the x
    and y faces, lid x = LX moving in +y at 0.1, Re 1000) after 5000 steps,
    against the oracle's reference-order step with the same walls: equal to
    rounding (measured 4e-15) on the fixture's lines, the corner and edge rows
    included.  This pins the multi-face order of the wall treatment
    (Cavity/streaming.f90:89-292, face by face: bounce-back, then RR
    regularisation with the later faces' populations still periodically
    streamed): doing every face's bounce-back before the first regularisation
    moved the fields by up to 7e-3 (ux, uy) at step 5000, and a lid term on the
    wrong face or a dropped face would too."""
    from helpers import CAVITY_LID, CAVITY_N, CAVITY_STEPS, cavity_fixture_errors, cavity_tau
    n = CAVITY_N
    g = oracle.Fields(oracle.D3Q27_CUDA, n, n, 2)
    g.init_equilibrium()  # rho = 1, u = 0 (Cavity/main.f90:15-21, initial.f90)
    g.step(CAVITY_STEPS, cavity_tau(), walls=(1, 1, 0), wall_rr=1, lid_face=1, lid_u=(0.0, CAVITY_LID, 0.0))
    er, eu, ev = cavity_fixture_errors(g.rho, g.ux, g.uy, GOLDEN)
    assert er < 1e-13 and eu < 1e-13 and ev < 1e-13, (er, eu, ev)
    assert np.abs(g.uz).max() < 1e-15


def _bgk_state(seed, lx=6, ly=5, lz=4, force=(2e-4, -1e-4, 5e-5)):
    rng = np.random.default_rng(seed)
    g = oracle.Fields(oracle.D3Q27_CUDA, lx, ly, lz)
    _, w, _ = oracle.lattice(oracle.D3Q27_CUDA)
    g.f[:] = w[:, None] * 0.02 * rng.standard_normal((27, g.n))   # shifted populations
    g.set_force(force)
    g.macrovar_bgk()
    return g


def test_bgk_collision_conserves_mass_and_adds_the_force():
    """collision_BGK (BGK/collision.cu:3-33) with rho, u from the BGK
    macrovar (u = j + F/2, BGK/macrovar.cu:27): mass exactly conserved, the
    momentum of every cell changes by exactly F (relaxation gives F/(2 tau),
    the forcing term (1 - 1/(2 tau)) F on D3Q27)."""
    c, w, _ = oracle.lattice(oracle.D3Q27_CUDA)
    for tau in (0.52, 0.8, 1.7):
        g = _bgk_state(3)
        f0 = g.f.copy()
        g.collision_bgk(tau)
        assert np.abs(g.f.sum(0) - f0.sum(0)).max() < 1e-15
        dj = c.T.astype(float) @ (g.f - f0)
        F = np.array([g.fx, g.fy, g.fz])
        assert np.abs(dj - F).max() < 1e-16


def test_bgk_fixed_points_and_relaxation_rate():
    """Without force: the rest state and any equilibrium state are fixed
    points; the non-equilibrium second moment relaxes by exactly 1 - 1/tau."""
    c, w, _ = oracle.lattice(oracle.D3Q27_CUDA)
    g = oracle.Fields(oracle.D3Q27_CUDA, 4, 4, 4)
    g.set_force((0, 0, 0))
    g.f[:] = 0.0      # shifted populations of the rest state (BGK/initial.cu:131,188-191)
    g.macrovar_bgk()
    g.collision_bgk(0.7)
    assert np.abs(g.f).max() == 0.0
    # an equilibrium state of the BGK model: feq(rho, u) with u = sum c f
    rng = np.random.default_rng(8)
    rho, u = 0.01 * rng.standard_normal(g.n), 0.05 * rng.standard_normal((3, g.n))
    eu = 3.0 * (c.astype(float) @ u)
    uv = 3.0 * (u * u).sum(0)
    g.f[:] = w[:, None] * (rho[None, :] + eu + 0.5 * (eu * eu - uv[None, :]))
    g.macrovar_bgk()
    f0 = g.f.copy()
    g.collision_bgk(0.7)
    assert np.abs(g.f - f0).max() < 1e-16
    # relaxation of the second moment
    g = _bgk_state(5, force=(0, 0, 0))
    tau = 0.63
    f0 = g.f.copy()
    eu = 3.0 * (c.astype(float) @ np.array([g.ux, g.uy, g.uz]))
    uv = 3.0 * (g.ux ** 2 + g.uy ** 2 + g.uz ** 2)
    feq = w[:, None] * (g.rho[None, :] + eu + 0.5 * (eu * eu - uv[None, :]))
    g.collision_bgk(tau)
    for a in range(3):
        for b in range(3):
            m = (c[:, a] * c[:, b]).astype(float)
            before = m @ (f0 - feq)
            after = m @ (g.f - feq)
            assert np.abs(after - (1 - 1 / tau) * before).max() < 1e-16


def test_bgk_channel_step_conserves_mass():
    """The BGK solver's step (collision, periodic streaming, half-way
    bounce-back at the y walls, macrovar) keeps the total mass of a closed
    channel to rounding."""
    g = _bgk_state(11, lx=8, ly=6, lz=5, force=(1e-4, 0, 0))
    m0 = g.f.sum()
    g.step_bgk(20, 0.56)
    assert abs(g.f.sum() - m0) < 1e-14


# ---------------------------------------------------------------- RR terms of order 4-6
def product_form(c, u):
    """Per-axis tables of the D3Q27 full-order equilibrium at the lattice
    velocities: P = 1 + 3uc + 9/2 u^2 c^2 - 3/2 u^2, T1 = 3uc, T2 = 3/2 u^2 (3c^2 - 1),
    h = 3c^2 - 1 (scaled second Hermite polynomial), D = c + (3c^2 - 1) u
    (derivative of P in the scaled velocity, divided by sqrt(3))."""
    c = c.astype(float)
    P = 1 + 3 * u * c + 4.5 * u * u * c * c - 1.5 * u * u
    T1 = 3 * u * c
    T2 = 1.5 * u * u * (3 * c * c - 1)
    h = 3 * c * c - 1
    D = c + (3 * c * c - 1) * u
    return P, T1, T2, h, D


def sixth_order_cell(f, rho, u, F, tau, mode, c, w):
    """The RR collision (mode 0) / RRBC (mode 1) with all Hermite orders D3Q27
    carries, from the generating function G = prod_a P_a: feq = w rho G, the
    non-equilibrium part w (1/2) sum_ab A_ab d2G/du~_a du~_b, the force term on
    the equilibrium truncated after fourth order."""
    P, T1, T2, h, D = product_form(c, u)
    x, y, z = 0, 1, 2
    feq6 = w * rho * P[:, x] * P[:, y] * P[:, z]
    q4 = (P[:, x] * P[:, y] * P[:, z] - T2[:, x] * T2[:, y] * T1[:, z] - T2[:, x] * T1[:, y] * T2[:, z]
          - T1[:, x] * T2[:, y] * T2[:, z] - T2[:, x] * T2[:, y] * T2[:, z])
    feq4 = w * rho * q4
    Fbar = 1.5 * (c @ F - u @ F)
    g = f + Fbar * feq4
    H2 = 3.0 * np.einsum("ia,ib->iab", c, c) - np.eye(3)[None]
    A = np.einsum("iab,i->ab", H2, g) - 3.0 * rho * np.outer(u, u)
    fneq = w * (0.5 * (A[x, x] * h[:, x] * P[:, y] * P[:, z] + A[y, y] * P[:, x] * h[:, y] * P[:, z]
                       + A[z, z] * P[:, x] * P[:, y] * h[:, z])
                + 3.0 * (A[x, y] * D[:, x] * D[:, y] * P[:, z] + A[x, z] * D[:, x] * P[:, y] * D[:, z]
                         + A[y, z] * P[:, x] * D[:, y] * D[:, z]))
    if mode == 0:
        return feq6 + (1 - 1 / tau) * fneq + Fbar * feq4
    return feq6 + fneq - Fbar * feq4


@pytest.mark.parametrize("kind", [oracle.D3Q27_CUDA, oracle.D3Q27_FORTRAN])
def test_full_hermite_basis_orthogonal(kind):
    """All 27 tensor Hermite polynomials of D3Q27 (orders 0-6): sum_i w H_m H_n =
    delta_mn prod(a!) -- in particular sum_i w H222 = 0, which the reference's
    commented H222 (initial.cu:128-130: + c_y^2 c_z^2 / cs^4) violates."""
    c, w, _ = oracle.lattice(kind)
    H = np.column_stack([np.ones(len(w)), oracle.hermite(kind), oracle.hermite_high(kind)])
    orders = ["000", "100", "010", "001", "110", "101", "011", "200", "020", "002", "210", "201", "021", "120",
              "102", "012", "111", "220", "202", "022", "211", "121", "112", "221", "212", "122", "222"]
    norms = np.array([np.prod([2.0 if d == "2" else 1.0 for d in o]) for o in orders])
    G = np.einsum("i,im,in->mn", w, H, H)
    assert np.allclose(G, np.diag(norms), atol=1e-13)
    # each table is the product of the 1D polynomials 1, c/cs, c^2/cs^2 - 1
    one_d = [lambda v: np.ones_like(v, dtype=float), lambda v: v / CS, lambda v: v * v / CS ** 2 - 1]
    for k, o in enumerate(orders):
        ref = one_d[int(o[0])](c[:, 0]) * one_d[int(o[1])](c[:, 1]) * one_d[int(o[2])](c[:, 2])
        assert np.allclose(H[:, k], ref, atol=1e-13), o


@pytest.mark.parametrize("kind", [oracle.D3Q27_CUDA, oracle.D3Q27_FORTRAN])
def test_sixth_order_equilibrium_is_the_product_form(kind):
    rng = np.random.default_rng(11)
    c, w, _ = oracle.lattice(kind)
    with oracle.rr_order(6):
        for _ in range(20):
            rho = 1 + 0.03 * rng.standard_normal()
            u = 0.15 * rng.uniform(-1, 1, 3)
            P = product_form(c, u)[0]
            assert np.allclose(oracle.feq(kind, rho, *u), w * rho * P.prod(axis=1), rtol=1e-14, atol=1e-17)
    # the third-order equilibrium differs by the fourth-order terms (the test sees them)
    rho, u = 1.0, np.array([0.1, -0.08, 0.06])
    diff = oracle.feq(kind, rho, *u) - w * rho * product_form(c, u)[0].prod(axis=1)
    assert np.abs(diff).max() > 1e-6


@pytest.mark.parametrize("kind", [oracle.D3Q27_CUDA, oracle.D3Q27_FORTRAN])
@pytest.mark.parametrize("mode", [0, 1])
def test_sixth_order_collision_matches_generating_function(kind, mode):
    rng = np.random.default_rng(12 + mode)
    with oracle.rr_order(6):
        for _ in range(20):
            F = 2e-4 * rng.standard_normal(3)
            f, rho, u, c, w = consistent_cell(kind, rng, F)
            f = f + 5e-3 * w * rng.standard_normal(len(w))    # a sizeable non-equilibrium part
            tau = rng.uniform(0.51, 1.5)
            out = oracle.rr_cell(kind, f, rho, *u, F, tau, mode)
            ref = sixth_order_cell(f, rho, u, F, tau, mode, c, w)
            assert np.abs(out - ref).max() < 1e-15
    # at third order the same inputs give a different state (so the test sees the extra terms)
    out3 = oracle.rr_cell(kind, f, rho, *u, F, tau, mode)
    assert np.abs(out3 - ref).max() > 1e-9


@pytest.mark.parametrize("kind", [oracle.D3Q27_CUDA, oracle.D3Q27_FORTRAN])
def test_sixth_order_collision_conservation(kind):
    rng = np.random.default_rng(13)
    with oracle.rr_order(6):
        for _ in range(20):
            F = 1e-4 * rng.standard_normal(3)
            f, rho, u, c, w = consistent_cell(kind, rng, F)
            out = oracle.rr_cell(kind, f, rho, *u, F, rng.uniform(0.51, 2.0), 0)
            assert abs(out.sum() - f.sum()) < 1e-14
            assert np.allclose(out @ c, f @ c + 0.5 * np.asarray(F) * (1 + rho), atol=1e-15)
            # tau = 1: equilibrium plus the force term, whatever f was
            feq = oracle.feq(kind, rho, *u)
            one = oracle.rr_cell(kind, f, rho, *u, (0.0, 0.0, 0.0), 1.0, 0)
            assert np.allclose(one, feq, atol=1e-15)
    assert oracle.get_rr_order() == 3


# ---------------------------------------------------------------- FDBC
def fdbc_field(lx, ly, lz, seed):
    """Oracle Fields whose velocity is, in every (x, z) column, a quadratic in y
    vanishing on both half-way walls (y = -1/2 and LY - 1/2), modulated
    smoothly in x and z; returns the fields and the exact profile functions."""
    rng = np.random.default_rng(seed)
    g = oracle.Fields(oracle.D3Q27_CUDA, lx, ly, lz)
    amp = np.array([0.02, 0.004, -0.01])
    ex, ez = 0.3, 0.2
    X, Y, Z = np.meshgrid(np.arange(lx), np.arange(ly), np.arange(lz), indexing="xy")
    X, Y, Z = (np.transpose(a, (2, 0, 1)).ravel() for a in (X, Y, Z))  # idx = (z*LY + y)*LX + x
    mod = [1 + ex * np.sin(2 * np.pi * (X + k) / lx) + ez * np.cos(2 * np.pi * (Z + 2 * k) / lz) for k in range(3)]
    prof = (Y + 0.5) * (ly - 0.5 - Y)
    dprof = (ly - 1 - 2.0 * Y)
    for k, arr in enumerate((g.ux, g.uy, g.uz)):
        arr[:] = amp[k] * mod[k] * prof
    _, w, _ = oracle.lattice(oracle.D3Q27_CUDA)
    g.f[:] = w[:, None] * (1 + 0.01 * rng.standard_normal((27, g.n)))
    return g, amp, mod, prof, dprof, (X, Y, Z)


def test_fdbc_closed_forms():
    lx, ly, lz, tau = 8, 9, 6, 0.6
    g, amp, mod, prof, dprof, (X, Y, Z) = fdbc_field(lx, ly, lz, 3)
    c, w, _ = oracle.lattice(oracle.D3Q27_CUDA)
    H = oracle.hermite(oracle.D3Q27_CUDA)
    fin = g.f.copy()
    interior = (Y > 0) & (Y < ly - 1)
    u_before = np.stack([g.ux, g.uy, g.uz]).copy()
    g.fdbc(tau)
    u = np.stack([g.ux, g.uy, g.uz])
    exact = np.stack([amp[k] * mod[k] * prof for k in range(3)])
    wall = ~interior
    # quadratic extrapolation is exact for these profiles; the interior is untouched
    assert np.abs(u[:, wall] - exact[:, wall]).max() < 1e-16
    assert np.array_equal(u[:, interior], u_before[:, interior])
    assert np.array_equal(g.f[:, interior], fin[:, interior])
    # strain rate at the wall: exact y derivative (the one-sided difference is
    # exact for quadratics), central differences in x and z of the wall values
    def at(arr, x, y, z):
        return arr[((z % lz) * ly + y) * lx + (x % lx)]
    for idx in np.where(wall)[0]:
        x, y, z = X[idx], Y[idx], Z[idx]
        grad = np.zeros((3, 3))  # grad[a, b] = d u_a / d x_b
        for a in range(3):
            grad[a, 0] = 0.5 * (at(u[a], x + 1, y, z) - at(u[a], x - 1, y, z))
            grad[a, 1] = amp[a] * mod[a][idx] * dprof[idx]
            grad[a, 2] = 0.5 * (at(u[a], x, y, z + 1) - at(u[a], x, y, z - 1))
        S = 0.5 * (grad + grad.T)
        known = c[:, 1] < 0 if y == 0 else c[:, 1] > 0
        fo, fi = g.f[:, idx], fin[:, idx]
        # the density keeps the known populations' sum
        assert abs(fo[known].sum() - fi[known].sum()) < 1e-15
        rho = fo.sum()
        uw = u[:, idx]
        assert np.allclose(fo @ c, rho * uw, atol=1e-16)
        # second moment: rho (I/3 + u u) + the Chapman-Enskog stress -2/3 rho tau S
        P = np.einsum("i,ia,ib->ab", fo, c, c)
        assert np.allclose(P, rho * (np.eye(3) / 3 + np.outer(uw, uw) - 2.0 / 3.0 * tau * S), atol=1e-16)
        # third order: sum H210 f = rho (u~x^2 u~y + 2 A210), A210 = A200 u~y + 2 u~x A110, A = -tau S
        ut = uw / CS
        A200, A110 = -tau * S[0, 0], -tau * S[0, 1]
        assert abs(fo @ H[:, 9] - rho * (ut[0] ** 2 * ut[1] + 2 * (A200 * ut[1] + 2 * ut[0] * A110))) < 1e-15
    # idempotent: a second pass keeps the state
    f1 = g.f.copy()
    g.fdbc(tau)
    assert np.abs(g.f - f1).max() < 1e-15


def test_fdbc_force_term():
    """With a body force the reconstruction carries -Fbar feq: its first moment
    is rho (u_w - F/2)."""
    g, amp, mod, prof, dprof, (X, Y, Z) = fdbc_field(6, 7, 4, 4)
    F = np.array([2e-4, -1e-4, 5e-5])
    g.set_force(F)
    c, _, _ = oracle.lattice(oracle.D3Q27_CUDA)
    g.fdbc(0.7)
    for idx in np.where((Y == 0) | (Y == 6))[0]:
        fo = g.f[:, idx]
        uw = np.array([g.ux[idx], g.uy[idx], g.uz[idx]])
        assert np.allclose(fo @ c, fo.sum() * (uw - F / 2), atol=1e-16)
