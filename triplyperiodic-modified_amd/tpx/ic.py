"""Seeded synthetic initial conditions and the reference's restart format.

The reference has no spectral initial-condition generator: it reads
U/V/W.bin or Restart_*.bin (grid.cpp:221-255).  This module builds the
solenoidal random-phase field the benchmark configurations ask for, samples
it on the staggered faces, removes the discrete divergence with the same
modified-wavenumber operator the Poisson solve uses, and writes it in the
reference's restart format so every code path starts from identical bits.

Layout: scalar fields are (nz, ny, nx) arrays, x fastest (the global C order
of communicator.cpp:29-39); vector fields are (3, nz, ny, nx) with the x, y, z
face components.  Face x of cell (i,j,k) sits at (i dx, (j+1/2) dy,
(k+1/2) dz), as in the reference's initial conditions (grid.cpp:139-144).
"""
from __future__ import annotations

import math
import os
from typing import Dict, Optional, Tuple

import numpy as np

from .params import PI, Params


def _wavenumbers(shape: Tuple[int, int, int], L: Tuple[float, float, float]):
    nz, ny, nx = shape
    kx = np.fft.fftfreq(nx, d=1.0 / nx) * (2 * PI / L[0])
    ky = np.fft.fftfreq(ny, d=1.0 / ny) * (2 * PI / L[1])
    kz = np.fft.fftfreq(nz, d=1.0 / nz) * (2 * PI / L[2])
    return np.meshgrid(kz, ky, kx, indexing="ij")  # KZ, KY, KX


def modified_wavenumber(shape, L) -> np.ndarray:
    """lambda(k) = sum_d (2/dd^2)(1 - cos(k_d dd)): minus the eigenvalue of the
    7-point Laplacian, as divided by in poisson.cpp:433-449."""
    nz, ny, nx = shape
    d = (L[0] / nx, L[1] / ny, L[2] / nz)
    KZ, KY, KX = _wavenumbers(shape, L)
    return (2 / d[0] ** 2) * (1 - np.cos(KX * d[0])) + (2 / d[1] ** 2) * (1 - np.cos(KY * d[1])) + \
           (2 / d[2] ** 2) * (1 - np.cos(KZ * d[2]))


def project_discrete(u: np.ndarray, L) -> np.ndarray:
    """Remove the discrete (div_F2C o grad_C2F) divergence of a face field."""
    shape = u.shape[1:]
    nz, ny, nx = shape
    d = (L[0] / nx, L[1] / ny, L[2] / nz)
    KZ, KY, KX = _wavenumbers(shape, L)
    K = (KX, KY, KZ)
    uh = [np.fft.fftn(u[c]) for c in range(3)]
    # div_F2C: (f[i+1]-f[i])/d  ->  (e^{ikd}-1)/d ;  grad_C2F: (1-e^{-ikd})/d
    fwd = [(np.exp(1j * K[c] * d[c]) - 1) / d[c] for c in range(3)]
    bwd = [(1 - np.exp(-1j * K[c] * d[c])) / d[c] for c in range(3)]
    div = fwd[0] * uh[0] + fwd[1] * uh[1] + fwd[2] * uh[2]
    lap = -modified_wavenumber(shape, L)
    lap[0, 0, 0] = 1.0
    phi = div / lap
    phi[0, 0, 0] = 0.0
    return np.stack([np.real(np.fft.ifftn(uh[c] - bwd[c] * phi)) for c in range(3)])


def spectral_velocity(shape: Tuple[int, int, int], L=(2 * PI,) * 3, seed: int = 0,
                      k0: float = 4.0, uprime: float = 1.0) -> np.ndarray:
    """Random-phase solenoidal velocity with E(k) ~ k^4 exp(-2 (k/k0)^2),
    sampled on the staggered faces, discretely divergence free, scaled so that
    sqrt(<u^2+v^2+w^2>/3) = uprime."""
    nz, ny, nx = shape
    rng = np.random.default_rng(seed)
    KZ, KY, KX = _wavenumbers(shape, L)
    K = (KX, KY, KZ)
    kmag = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    # integer-index magnitude for the |k| <= N/2 cut
    nmin = min(nx, ny, nz)
    kk = np.sqrt((KX * L[0] / (2 * PI)) ** 2 + (KY * L[1] / (2 * PI)) ** 2 + (KZ * L[2] / (2 * PI)) ** 2)
    keep = (kk > 0) & (kk <= nmin / 2)
    amp = np.zeros(shape)
    amp[keep] = kmag[keep] * np.exp(-(kmag[keep] / k0) ** 2)  # sqrt(E(k)/(4 pi k^2))
    uh = []
    for _ in range(3):
        phase = rng.uniform(0.0, 2 * PI, size=shape)
        mag = rng.standard_normal(size=shape)
        uh.append(amp * mag * np.exp(1j * phase))
    kdotu = (K[0] * uh[0] + K[1] * uh[1] + K[2] * uh[2])
    k2 = kmag ** 2
    k2[0, 0, 0] = 1.0
    uh = [uh[c] - K[c] * kdotu / k2 for c in range(3)]
    # sample component c on its face: shift by half a cell in the two other directions
    d = (L[0] / nx, L[1] / ny, L[2] / nz)
    u = np.empty((3,) + shape)
    for c in range(3):
        shift = np.ones(shape, dtype=complex)
        for a in range(3):
            if a != c:
                shift = shift * np.exp(1j * K[a] * d[a] / 2)
        u[c] = np.real(np.fft.ifftn(uh[c] * shift))
    u = project_discrete(u, L)
    for c in range(3):
        u[c] -= u[c].mean()
    up = math.sqrt(float(np.mean(u[0] ** 2 + u[1] ** 2 + u[2] ** 2)) / 3.0)
    return u * (uprime / up)


def discrete_divergence(u: np.ndarray, L) -> np.ndarray:
    nz, ny, nx = u.shape[1:]
    d = (L[0] / nx, L[1] / ny, L[2] / nz)
    return ((np.roll(u[0], -1, axis=2) - u[0]) / d[0] + (np.roll(u[1], -1, axis=1) - u[1]) / d[1] +
            (np.roll(u[2], -1, axis=0) - u[2]) / d[2])


def dissipation(u: np.ndarray, nu: float, L) -> float:
    """epsilon = 2 nu <S_ij S_ij> with the staggered differences of the
    reference: normal strains at cell centres (ddx_F2C), shear strains at the
    cell edges (ddy_C2F of u with ddx_C2F of v, ...)."""
    nz, ny, nx = u.shape[1:]
    d = (L[0] / nx, L[1] / ny, L[2] / nz)
    ax = (2, 1, 0)  # array axis of x, y, z
    sq = 0.0
    for c in range(3):
        g = (np.roll(u[c], -1, axis=ax[c]) - u[c]) / d[c]
        sq += float(np.mean(g * g))
    for a in range(3):
        for b in range(a + 1, 3):
            dua_db = (u[a] - np.roll(u[a], 1, axis=ax[b])) / d[b]
            dub_da = (u[b] - np.roll(u[b], 1, axis=ax[a])) / d[a]
            s = 0.5 * (dua_db + dub_da)
            sq += 2.0 * float(np.mean(s * s))
    return 2.0 * nu * sq


def energy_spectrum(u: np.ndarray, L) -> np.ndarray:
    """Shell-binned kinetic energy spectrum E(k) of a velocity field (each
    staggered component transformed on its own grid), shells round(|k|)
    with k in units of 2 pi / L (cubic box).  Normalised so that
    sum_k E(k) = <u.u>/2 (Parseval)."""
    nz, ny, nx = u.shape[1:]
    kx = np.fft.fftfreq(nx, 1.0 / nx)
    ky = np.fft.fftfreq(ny, 1.0 / ny)
    kz = np.fft.fftfreq(nz, 1.0 / nz)
    K = np.sqrt(kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2)
    shell = np.rint(K).astype(int)
    e = np.zeros(shell.max() + 1)
    n = nx * ny * nz
    for c in range(3):
        uh = np.fft.fftn(u[c]) / n
        np.add.at(e, shell.ravel(), 0.5 * (np.abs(uh) ** 2).ravel())
    return e


def particle_rms(parts) -> float:
    """sqrt(<u_p^2 + v_p^2 + w_p^2> / 3) over the particles."""
    return float(np.sqrt(np.mean(parts["u"] ** 2 + parts["v"] ** 2 + parts["w"] ** 2) / 3.0))


def gather_face_velocity(u: np.ndarray, x, y, z, L) -> np.ndarray:
    """P1 interpolation (particle.cpp:340-360): linear along the component's
    own axis from the floor face, piecewise constant across."""
    nz, ny, nx = u.shape[1:]
    idx, idy, idz = nx / L[0], ny / L[1], nz / L[2]
    sx, sy, sz = x * idx, y * idy, z * idz
    ii, jj, kk = sx.astype(np.int64), sy.astype(np.int64), sz.astype(np.int64)
    xl, yl, zl = sx - ii, sy - jj, sz - kk
    i0, j0, k0 = ii % nx, jj % ny, kk % nz
    i1, j1, k1 = (ii + 1) % nx, (jj + 1) % ny, (kk + 1) % nz
    ug = xl * u[0][k0, j0, i1] + (1 - xl) * u[0][k0, j0, i0]
    vg = yl * u[1][k0, j1, i0] + (1 - yl) * u[1][k0, j0, i0]
    wg = zl * u[2][k1, j0, i0] + (1 - zl) * u[2][k0, j0, i0]
    return np.stack([ug, vg, wg])


def uniform_particles(n: int, L, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    pos = np.empty((3, n))
    for c in range(3):
        pos[c] = rng.uniform(0.0, L[c], size=n)
    return pos


def max_cfl_rate(u: np.ndarray, L) -> float:
    """max_c( max(U_c) / d_c ) as in tensor1::max_cfl (tensor1.cpp:365-372)."""
    nz, ny, nx = u.shape[1:]
    d = (L[0] / nx, L[1] / ny, L[2] / nz)
    return max(float(u[c].max()) / d[c] for c in range(3))


# ---------------------------------------------------------------------------
# reference restart format (communicator.cpp:65-106, grid.cpp:221-238,
# particle.cpp:848-989): raw native FP64 in global C order [Nz][Ny][Nx], a
# tensor1 is its x, y, z arrays back to back, particle quantities are flat
# FP64 (flags int32), Restart_numbers.dat holds P0, T_cur, num_timestep.
# ---------------------------------------------------------------------------
PART_Q = ("x", "y", "z", "u", "v", "w", "T")


def write_restart(directory: str, RU: np.ndarray, Rho: np.ndarray, P: np.ndarray,
                  P0: float, T_cur: float, num_timestep: int,
                  particles: Optional[Dict[str, np.ndarray]] = None, prefix: str = "Restart_") -> None:
    os.makedirs(directory, exist_ok=True)
    np.ascontiguousarray(RU, dtype="<f8").tofile(os.path.join(directory, prefix + "RU.bin"))
    np.ascontiguousarray(Rho, dtype="<f8").tofile(os.path.join(directory, prefix + "Rho.bin"))
    np.ascontiguousarray(P, dtype="<f8").tofile(os.path.join(directory, prefix + "P.bin"))
    with open(os.path.join(directory, prefix + "numbers.dat"), "w") as fh:
        fh.write(f"{P0!r}\n{T_cur!r}\n{int(num_timestep)}\n")
    if particles is not None:
        for q in PART_Q:
            np.ascontiguousarray(particles[q], dtype="<f8").tofile(os.path.join(directory, f"{prefix}Part_{q}.bin"))
        np.ascontiguousarray(particles["flag"], dtype="<i4").tofile(os.path.join(directory, f"{prefix}Part_flag.bin"))


def read_restart(directory: str, shape, n_particles: Optional[int] = None, prefix: str = "Restart_"):
    nz, ny, nx = shape
    RU = np.fromfile(os.path.join(directory, prefix + "RU.bin"), dtype="<f8").reshape(3, nz, ny, nx)
    Rho = np.fromfile(os.path.join(directory, prefix + "Rho.bin"), dtype="<f8").reshape(nz, ny, nx)
    P = np.fromfile(os.path.join(directory, prefix + "P.bin"), dtype="<f8").reshape(nz, ny, nx)
    with open(os.path.join(directory, prefix + "numbers.dat")) as fh:
        tok = fh.read().split()
    P0, T_cur, nts = float(tok[0]), float(tok[1]), int(tok[2])
    parts = None
    fpath = os.path.join(directory, f"{prefix}Part_x.bin")
    if os.path.exists(fpath):
        parts = {q: np.fromfile(os.path.join(directory, f"{prefix}Part_{q}.bin"), dtype="<f8") for q in PART_Q}
        parts["flag"] = np.fromfile(os.path.join(directory, f"{prefix}Part_flag.bin"), dtype="<i4")
        if n_particles is not None:
            for k in list(parts):
                parts[k] = parts[k][:n_particles]
    return RU, Rho, P, (P0, T_cur, nts), parts


# ---------------------------------------------------------------------------
# the benchmark configurations
# ---------------------------------------------------------------------------
class Case:
    """A fully specified synthetic run: parameters, initial fields and particles."""

    def __init__(self, params: Params, RU: np.ndarray, Rho: np.ndarray, P: np.ndarray,
                 particles: Dict[str, np.ndarray], tau: Optional[np.ndarray] = None,
                 eps: Optional[np.ndarray] = None, info: Optional[Dict] = None,
                 PS: Optional[np.ndarray] = None):
        self.params, self.RU, self.Rho, self.P = params, RU, Rho, P
        self.particles, self.tau, self.eps = particles, tau, eps
        self.info = info or {}
        self.PS = PS  # the passive scalar's initial field (nz, ny, nx), or None

    @property
    def shape(self):
        return self.Rho.shape


def make_case(n: int, n_particles: int, nu: float, seed: int = 0, St=1.0, phi: float = 0.2,
              heat: bool = False, twc: bool = True, k0: float = 4.0, cfl: float = 0.3,
              eps_range: Optional[Tuple[float, float]] = None, forcing: bool = True,
              L: float = 2 * PI, nsteps: int = 20, drag_limit: bool = False,
              shape: Optional[Tuple[int, int, int]] = None) -> Case:
    """Synthetic case following the benchmark recipe: spectral IC (k0, u'=1),
    CFL-0.3 dt, Lundgren forcing A = eps0/(3 u'^2), particles at seeded
    uniform positions with the gas velocity interpolated to them, Stokes
    number St (scalar, or an array for per-particle tau), mass loading phi.
    The box is n^3 of side L, or shape = (nz, ny, nx) cells of one spacing
    L / nx."""
    if shape is None:
        shape = (n, n, n)
    nz, ny, nx = shape
    Ls = (L, L * ny / nx, L * nz / nx)
    rho0, T0 = 1.0, 1.0
    u = spectral_velocity(shape, Ls, seed=seed, k0=k0, uprime=1.0)
    eps0 = dissipation(u, nu, Ls)
    tau_eta = math.sqrt(nu / eps0)
    dt = cfl / max_cfl_rate(u, Ls)
    p = Params(Nx_tot=nx, Ny_tot=ny, Nz_tot=nz, Lx=Ls[0], Ly=Ls[1], Lz=Ls[2], bs=2)
    p.dt = dt
    p.T_final = dt * nsteps
    p.Mu0 = nu * rho0
    p.Rho0, p.T0 = rho0, T0
    p.Cp, p.Cv = 3.5, 2.5
    p.k = p.Mu0 * p.Cp / 0.7
    p.A1 = p.A2 = (eps0 / 3.0) if forcing else 0.0
    p.TWC = 1 if twc else 0
    p.Iteration1 = p.Iteration2 = 50
    # absolute max|dp| tolerance of poisson.cpp:397.  With |RHS_Pois| ~ 1e2
    # the zero mode of the residual is round-off of order 1e-13 and passes
    # straight into dp (poisson.cpp:442), so a smaller tolerance would never
    # be met and every solve would run all Iteration passes.
    p.epsilon = 1e-10
    vol = Ls[0] * Ls[1] * Ls[2]
    npart = int(n_particles)
    st_arr = np.asarray(St, dtype=float)
    st_ref = float(np.exp(np.mean(np.log(st_arr)))) if st_arr.ndim else float(st_arr)
    tau_ref = st_ref * tau_eta
    if npart > 0:
        mp = phi * rho0 * vol / npart
        Dp = mp / (3 * PI * p.Mu0 * tau_ref)
        p.Dp = Dp
        p.Rhop = 6 * mp / (PI * Dp ** 3)
        p.np0 = (npart + 0.5) / vol
    else:
        p.np0 = 0.0
    p.Cvp = p.Cv
    if drag_limit and npart > 0:
        # Two-way coupled drag relaxes a particle and the gas of its cell
        # together at rate (1 + mp / (rho0 Vcell)) / tau; keep dt times that
        # rate at 1, well inside RK4's real-axis stability limit (2.78) to
        # leave room for cells holding several particles, so boxes whose
        # particles outweigh a cell's gas do not blow up.
        mcell = rho0 * vol / (nx * ny * nz)
        tau_min = float(np.min(st_arr)) * tau_eta if st_arr.ndim else tau_ref
        p.dt = min(p.dt, tau_min / (1.0 + (phi * rho0 * vol / npart) / mcell))
        p.T_final = p.dt * nsteps
    info = dict(eps0=eps0, tau_eta=tau_eta, uprime=1.0, seed=seed, St=st_ref, phi=phi)
    if heat:
        p.Nu = 2.0
        # equilibrium particle excess temperature of 0.5 T0 at the mean
        # absorption efficiency: alpha = 0.5 T0 Nu pi Dp k
        eps_mean = 0.75 if eps_range is None else 0.5 * (eps_range[0] + eps_range[1])
        alpha = 0.5 * T0 * p.Nu * PI * p.Dp * p.k
        p.I01 = p.I02 = 4 * alpha / (eps_mean * PI * p.Dp ** 2)
        info["I0"] = p.I01
    rng = np.random.default_rng(seed + 7919)
    pos = uniform_particles(npart, Ls, seed + 104729)
    vel = gather_face_velocity(u, pos[0], pos[1], pos[2], Ls) if npart else np.zeros((3, 0))
    parts = dict(x=pos[0], y=pos[1], z=pos[2], u=vel[0], v=vel[1], w=vel[2],
                 T=np.full(npart, T0), flag=np.arange(npart, dtype=np.int32))
    tau = None
    if st_arr.ndim and npart:
        tau = st_arr[:npart] * tau_eta
    eps = None
    if eps_range is not None and npart:
        eps = rng.uniform(eps_range[0], eps_range[1], size=npart)
    RU = u * rho0
    Rho = np.full(shape, rho0)
    P = np.zeros(shape)
    return Case(p, RU, Rho, P, parts, tau=tau, eps=eps, info=info)


def loguniform(rng: np.random.Generator, lo: float, hi: float, n: int) -> np.ndarray:
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size=n))


def config_case(name: str, seed: int = 0, scale: int = 1) -> Case:
    """The benchmark configurations by name.  scale > 1 divides the grid size
    and particle count (for tests of the same recipe at small size)."""
    if name == "plain32":        # 32^3, 16,384 particles, nu=0.02, St=1, phi=0.2, heat off
        return make_case(32 // scale, 16384 // scale ** 3, nu=0.02, seed=seed, St=1.0, phi=0.2)
    if name == "plain128":       # 128^3, 1e6 particles, nu=0.005, St ~ logU[0.1,10], phi=0.5
        n, npart = 128 // scale, 1_000_000 // scale ** 3
        st = loguniform(np.random.default_rng(seed + 17), 0.1, 10.0, npart)
        return make_case(n, npart, nu=0.005, seed=seed, St=st, phi=0.5, nsteps=100)
    if name == "heat256":        # 256^3, 1.6e7 particles, St=1, phi=1.0, heat on, eps ~ U[0.5,1]
        n, npart = 256 // scale, 16_000_000 // scale ** 3
        return make_case(n, npart, nu=0.0025 * scale, seed=seed, St=1.0, phi=1.0, heat=True,
                         eps_range=(0.5, 1.0), nsteps=200)
    raise KeyError(name)


def ensemble_cases(n_boxes: int = 64, n: int = 64, seed: int = 0, np_range=(1e4, 2e5),
                   nu_range=(0.01, 0.03), st_range=(0.1, 10.0), phi: float = 0.3,
                   nsteps: int = 50):
    """The ensemble configuration: n_boxes independent n^3 boxes, box b with
    its own seed, viscosity nu ~ U[nu_range], Stokes number St ~
    logU[st_range] and particle count ~ logU[np_range]; mass loading phi,
    heat off.  With few particles each one outweighs the gas of a cell, so
    dt also obeys the coupled-drag limit of make_case(drag_limit=True)."""
    rng = np.random.default_rng(seed + 4242)
    nus = rng.uniform(nu_range[0], nu_range[1], n_boxes)
    sts = loguniform(rng, st_range[0], st_range[1], n_boxes)
    nps = np.round(loguniform(rng, np_range[0], np_range[1], n_boxes)).astype(int)
    return [make_case(n, int(nps[b]), nu=float(nus[b]), seed=seed + 1000 * b, St=float(sts[b]), phi=phi,
                      nsteps=nsteps, drag_limit=True) for b in range(n_boxes)]
