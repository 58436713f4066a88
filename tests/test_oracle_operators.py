"""Pins of the oracle's field operators and Poisson solve (rows S1-S11, G8,
G9) against what the mathematics fixes: single-Fourier-mode closed forms on a
non-cubic box, summation by parts, brute-force DFTs and a manufactured
variable-coefficient problem.  Each check is chosen so that a plausible slip
(an F2C/C2F offset in the wrong direction, a dropped factor, a sign) fails
it."""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle, fft3d
from tpx.params import Params

NX, NY, NZ = 16, 8, 32
LX, LY, LZ = 2 * math.pi, 3.0, 5.0


def make(nx=NX, ny=NY, nz=NZ, bs=2, **kw):
    p = Params(Nx_tot=nx, Ny_tot=ny, Nz_tot=nz, Lx=LX, Ly=LY, Lz=LZ, bs=bs, dt=1.0, **kw)
    return Oracle(p.struct_fields(), np_capacity=1), p


def coords(nx=NX, ny=NY, nz=NZ):
    dx, dy, dz = LX / nx, LY / ny, LZ / nz
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return (i + 0.5) * dx, (j + 0.5) * dy, (k + 0.5) * dz, i * dx, j * dy, k * dz, (dx, dy, dz)


def wave(m=(3, 2, 5)):
    return (2 * math.pi * m[0] / LX, 2 * math.pi * m[1] / LY, 2 * math.pi * m[2] / LZ)


def lam(kv, d):
    return sum((2 / d[a] ** 2) * (1 - math.cos(kv[a] * d[a])) for a in range(3))


def test_div_f2c_single_mode():
    o, _ = make()
    xc, yc, zc, xf, yf, zf, d = coords()
    kx, ky, kz = wave()
    # each component lives on its own face: x face at (i dx, yc, zc) ...
    F = [np.sin(kx * xf + ky * yc + kz * zc), np.sin(kx * xc + ky * yf + kz * zc), np.sin(kx * xc + ky * yc + kz * zf)]
    th = kx * xc + ky * yc + kz * zc
    expect = sum((2 / d[a]) * math.sin(wave()[a] * d[a] / 2) for a in range(3)) * np.cos(th)
    got = o.apply("div_f2c", *F)
    assert np.max(np.abs(got - expect)) < 1e-12


def test_grad_c2f_single_mode():
    o, _ = make()
    xc, yc, zc, xf, yf, zf, d = coords()
    kv = wave((1, 3, 2))
    P = np.cos(kv[0] * xc + kv[1] * yc + kv[2] * zc)
    g = o.apply("grad_c2f", P)
    faces = [(xf, yc, zc), (xc, yf, zc), (xc, yc, zf)]
    for a in range(3):
        th = kv[0] * faces[a][0] + kv[1] * faces[a][1] + kv[2] * faces[a][2]
        expect = -(2 / d[a]) * math.sin(kv[a] * d[a] / 2) * np.sin(th)
        assert np.max(np.abs(g[a] - expect)) < 1e-12, a


def test_del2_eigenvalue():
    o, _ = make()
    xc, yc, zc, _, _, _, d = coords()
    kv = wave((4, 1, 7))
    f = np.cos(kv[0] * xc + kv[1] * yc + kv[2] * zc)
    assert np.max(np.abs(o.apply("del2", f) + lam(kv, d) * f)) < 1e-10


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_interpolations(axis):
    o, _ = make()
    xc, yc, zc, xf, yf, zf, d = coords()
    kv = wave((2, 1, 3))
    c = math.cos(kv[axis] * d[axis] / 2)
    cen = [xc, yc, zc]
    fac = [xf, yf, zf]
    # C2F along `axis` of a centred mode lands on that axis' faces
    f_c = np.cos(kv[0] * xc + kv[1] * yc + kv[2] * zc)
    pos = list(cen)
    pos[axis] = fac[axis]
    expect = c * np.cos(kv[0] * pos[0] + kv[1] * pos[1] + kv[2] * pos[2])
    got = o.apply(["ix_c2f", "iy_c2f", "iz_c2f"][axis], f_c)
    assert np.max(np.abs(got - expect)) < 1e-13
    # F2C along `axis` of a face mode lands on the centres
    f_f = np.cos(kv[0] * pos[0] + kv[1] * pos[1] + kv[2] * pos[2])
    got = o.apply(["ix_f2c", "iy_f2c", "iz_f2c"][axis], f_f)
    assert np.max(np.abs(got - c * f_c)) < 1e-13


def test_i_c2f_and_i_f2c_mean():
    o, _ = make()
    rng = np.random.default_rng(1)
    r = rng.standard_normal((NZ, NY, NX))
    v = o.apply("i_c2f", r)
    assert np.allclose(v[0], 0.5 * (r + np.roll(r, 1, axis=2)), atol=0, rtol=0)
    assert np.allclose(v[1], 0.5 * (r + np.roll(r, 1, axis=1)), atol=0, rtol=0)
    assert np.allclose(v[2], 0.5 * (r + np.roll(r, 1, axis=0)), atol=0, rtol=0)
    a, b, c = (np.full((NZ, NY, NX), x) for x in (1.0, 2.0, 6.0))
    assert np.max(np.abs(o.apply("i_f2c", a, b, c) - 3.0)) < 1e-15


def test_summation_by_parts_and_zero_sum():
    o, _ = make()
    rng = np.random.default_rng(2)
    F = [rng.standard_normal((NZ, NY, NX)) for _ in range(3)]
    p = rng.standard_normal((NZ, NY, NX))
    div = o.apply("div_f2c", *F)
    g = o.apply("grad_c2f", p)
    lhs = float(np.sum(div * p))
    rhs = -float(sum(np.sum(F[a] * g[a]) for a in range(3)))
    assert abs(lhs - rhs) < 1e-11 * (abs(lhs) + 1)
    assert abs(float(np.sum(div))) < 1e-11


def test_ghost_fold_equals_modular_deposit():
    o, _ = make(bs=2)
    rng = np.random.default_rng(3)
    bs = 2
    G = rng.standard_normal((NZ + 2 * bs, NY + 2 * bs, NX + 2 * bs))
    got = o.apply("fold", G.ravel())
    expect = np.zeros((NZ, NY, NX))
    for k in range(NZ + 2 * bs):
        for j in range(NY + 2 * bs):
            np.add.at(expect[(k - bs) % NZ, (j - bs) % NY], (np.arange(NX + 2 * bs) - bs) % NX, G[k, j])
    assert np.max(np.abs(got - expect)) < 1e-13


def test_sum_and_max_abs():
    o, _ = make()
    rng = np.random.default_rng(4)
    r = rng.standard_normal((NZ, NY, NX)) * 1e3
    assert abs(o.sum(r) - math.fsum(r.ravel())) <= 1e-12 * np.sum(np.abs(r))
    assert o.max_abs(r) == float(np.max(np.abs(r)))


# ---------------------------------------------------------------------------
# FFT (stand-in for the absent Sandia fft_3d / FFTW 2.1.5)
# ---------------------------------------------------------------------------
def brute_dft(a, sign):
    nz, ny, nx = a.shape
    out = np.zeros_like(a, dtype=complex)
    I = np.arange(nx); J = np.arange(ny); K = np.arange(nz)
    Wx = np.exp(sign * 2j * np.pi * np.outer(I, I) / nx)
    Wy = np.exp(sign * 2j * np.pi * np.outer(J, J) / ny)
    Wz = np.exp(sign * 2j * np.pi * np.outer(K, K) / nz)
    out = np.einsum("zyx,xa->zya", a, Wx)
    out = np.einsum("zya,yb->zba", out, Wy)
    out = np.einsum("zba,zc->cba", out, Wz)
    return out


def test_fft_against_brute_force_dft():
    rng = np.random.default_rng(5)
    a = rng.standard_normal((4, 8, 16)) + 1j * rng.standard_normal((4, 8, 16))
    for sign in (-1, 1):
        assert np.max(np.abs(fft3d(a, sign) - brute_dft(a, sign))) < 1e-12


def test_fft_against_numpy_fixture():
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "fft_8x16x32.npz")
    f = np.load(path)
    a = f["input"]
    assert np.max(np.abs(fft3d(a, -1) - f["forward"])) < 1e-11
    assert np.max(np.abs(fft3d(a, 1) - f["inverse_unnormalised"])) < 1e-11


def test_ccp_fft_single_mode_and_zero_mode():
    o, _ = make()
    xc, yc, zc, _, _, _, d = coords()
    kv = wave((2, 3, 1))
    r = np.cos(kv[0] * xc + kv[1] * yc + kv[2] * zc)
    dp = o.apply("ccp_fft", r)
    assert np.max(np.abs(dp + r / lam(kv, d))) < 1e-13
    # the zero mode passes through undivided: dp = -mean(r)
    dp0 = o.apply("ccp_fft", np.full((NZ, NY, NX), 0.75))
    assert np.max(np.abs(dp0 + 0.75)) < 1e-14


def test_ccp_fft_inverts_the_7_point_laplacian():
    o, _ = make()
    rng = np.random.default_rng(6)
    r = rng.standard_normal((NZ, NY, NX))
    dp = o.apply("ccp_fft", r)
    lap = o.apply("del2", dp)
    assert np.max(np.abs(lap - (r - r.mean()))) < 1e-11
    assert abs(dp.mean() + r.mean()) < 1e-13


def _solve(o, rho_face, rhs, P0=None):
    o.set("Rho_face", rho_face)
    o.set("RHS_Pois", rhs)
    o.set("P", np.zeros(rhs.shape) if P0 is None else P0)
    o.solve_poisson()
    return o.get("P"), o.scalars()["num_iteration"]


def test_poisson_uniform_coefficient_is_exact():
    o, _ = make(Iteration1=50, epsilon=1e-13)
    rng = np.random.default_rng(7)
    rhs = rng.standard_normal((NZ, NY, NX))
    rhs -= rhs.mean()
    rho = 1.7
    P, it = _solve(o, np.full((3, NZ, NY, NX), rho), rhs)
    # C = 1/rho: div(C grad P) = RHS
    g = o.apply("grad_c2f", P)
    res = o.apply("div_f2c", g[0] / rho, g[1] / rho, g[2] / rho) - rhs
    assert np.max(np.abs(res)) < 1e-11 * np.max(np.abs(rhs))
    assert it == 2  # the second iteration computes a round-off correction


def test_poisson_variable_coefficient_residual_decreases():
    """Manufactured problem in the spirit of grid::Test_Poisson
    (grid.cpp:1095-1143): residual of div((1/rho_face) grad P) = RHS falls
    monotonically with the iteration count."""
    n = 16
    res = []
    for iters in (1, 2, 3, 4):
        p = Params(Nx_tot=n, Ny_tot=n, Nz_tot=n, bs=2, dt=1.0, Iteration1=iters, epsilon=0.0)
        o = Oracle(p.struct_fields(), np_capacity=1)
        L = p.Lx
        dx = L / n
        k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        X, Y, Z = (i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx
        rho = 1 + 1 / (10 + np.sin(X) * np.sin(Y) * np.sin(Z))
        rf = o.apply("i_c2f", rho)
        phi = np.sin(3 * X) * np.sin(5 * Y) * np.sin(4 * Z)
        g = o.apply("grad_c2f", phi)
        rhs = o.apply("div_f2c", g[0] / rf[0], g[1] / rf[1], g[2] / rf[2])
        P, it = _solve(o, rf, rhs)
        assert it == iters
        gp = o.apply("grad_c2f", P)
        r = o.apply("div_f2c", gp[0] / rf[0], gp[1] / rf[1], gp[2] / rf[2]) - rhs
        res.append(float(np.max(np.abs(r))))
    assert all(b < a for a, b in zip(res, res[1:])), res
    assert res[-1] < 1e-8 * res[0]
