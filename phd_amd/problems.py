"""Synthetic initial data for the benchmark and test workloads.

Each builder returns ``(u, MP, tf, dX, bcs)`` in the reference driver's
convention (code/python/main.py:15-45, problem functions such as
code/python/tests/fluid/newtonian1.py:114-145): ``u`` is C-ordered
``[nx(, ny), V]`` float64, ``MP`` a `Par`, ``bcs`` the boundary codes
(0 transmissive, 1 periodic; code/cpp/src/etc/grid.cpp:5-10) for
(x-left, x-right[, y-left, y-right]).

Workloads (parameters the spec leaves open are fixed here once):
  sod_1d            1D GPR ideal-gas viscous shock tube, 200 cells, V=17
                    (cα=1, κ = μγcv/Pr with Pr = 0.75, objects.py:235-236)
  double_shear_2d   512² periodic double shear layer, V=14, cv=1
                    (cv as in code/python/tests/fluid/newtonian2.py:134)
  shock_tube_batch  256 seeded 1D shock tubes, nx in [100, 4000], V=14, cs=1
  grf_density_2d    256² periodic Gaussian-random-field density, V=14,
                    μ=1e-3, cs=1
  plastic_pulse_2d  1024² elastoplastic velocity pulse, Godunov-Romenski EOS
                    with c0=3.909, α=1, β=3 (code/python/tests/params/solids.py:46-47),
                    transmissive, V=14, p=0
"""

import numpy as np

from .params import Par, material_params, _Γ_MG, internal_energy

TRANSMISSIVE = 0
PERIODIC = 1


def cvec_field(ρ, p, v, MP, A, J=None):
    """Vectorised `Cvec` (structures.py:104-126) over a field of cells.

    ρ, p: [...]; v: [..., 3]; A: [..., 3, 3]; J: [..., 3] or None."""
    ρ = np.asarray(ρ, dtype=np.float64)
    p = np.asarray(p, dtype=np.float64)
    shape = ρ.shape
    V = 14 if J is None else 17
    u = np.zeros(shape + (V,))
    e = np.vectorize(lambda r, pp: internal_energy(r, pp, MP))(ρ, p)
    G = np.einsum("...ki,...kj->...ij", A, A)
    devG = G - (np.trace(G, axis1=-2, axis2=-1) / 3)[..., None, None] * np.eye(3)
    devGsq = np.sum(devG * devG, axis=(-2, -1))
    cs2 = MP.b02 * (ρ / MP.rho0) ** MP.beta
    Γ = np.vectorize(lambda r: _Γ_MG(r, MP))(ρ)
    C0 = (1 - MP.beta / Γ) * cs2
    E = e + C0 / 4 * devGsq + 0.5 * np.sum(v * v, axis=-1)
    if J is not None:
        E = E + MP.calpha2 * np.sum(J * J, axis=-1) / 2
    u[..., 0] = ρ
    u[..., 1] = ρ * E
    u[..., 2:5] = ρ[..., None] * v
    u[..., 5:14] = A.reshape(shape + (9,))
    if J is not None:
        u[..., 14:17] = ρ[..., None] * J
    return u


def sod_1d(nx=200, thermal=True, μ=1e-4):
    """1D Sod shock tube for the viscous, heat-conducting GPR gas."""
    γ, cv, Pr = 1.4, 2.5, 0.75
    if thermal:
        MP = material_params("sg", 1.0, cv=cv, γ=γ, b0=1.0, μ=μ, cα=1.0, Pr=Pr)
    else:
        MP = material_params("sg", 1.0, cv=cv, γ=γ, b0=1.0, μ=μ)
    dx = 1.0 / nx
    x = (np.arange(nx) + 0.5) * dx
    ρ = np.where(x < 0.5, 1.0, 0.125)
    p = np.where(x < 0.5, 1.0, 0.1)
    v = np.zeros((nx, 3))
    A = (ρ / MP.rho0)[:, None, None] ** (1 / 3) * np.eye(3)
    J = np.zeros((nx, 3)) if thermal else None
    u = cvec_field(ρ, p, v, MP, A, J)
    return u, MP, 0.2, [dx], [TRANSMISSIVE, TRANSMISSIVE]


def double_shear_2d(n=512, μ=1e-3, b0=8.0):
    """Periodic double shear layer: v_x = tanh(80(y-1/4)) for y <= 1/2 and
    tanh(80(3/4-y)) above, v_y = 0.05 sin(2πx), ρ=1, p=100/γ, μ=1e-3, cs=8."""
    γ = 1.4
    MP = material_params("sg", 1.0, cv=1.0, γ=γ, b0=b0, μ=μ)
    dx = 1.0 / n
    c = (np.arange(n) + 0.5) * dx
    x, y = np.meshgrid(c, c, indexing="ij")
    vx = np.where(y <= 0.5, np.tanh(80 * (y - 0.25)), np.tanh(80 * (0.75 - y)))
    vy = 0.05 * np.sin(2 * np.pi * x)
    v = np.stack([vx, vy, np.zeros_like(vx)], axis=-1)
    ρ = np.ones((n, n))
    p = np.full((n, n), 100 / γ)
    A = np.broadcast_to(np.eye(3), (n, n, 3, 3)).copy()
    u = cvec_field(ρ, p, v, MP, A)
    return u, MP, None, [dx, dx], [PERIODIC] * 4


def grf_density_2d(n=256, seed=0, ell=0.1, μ=1e-3, p_amp=0.0, thermal=False):
    """ρ = 1 + 0.05·GRF (Gaussian spectrum of length scale ell, unit std),
    v=0, p=1, A = ρ^(1/3) I, periodic.  This start is a steady state (uniform
    pressure, no shear stress); p_amp > 0 sets p = 1 + p_amp·GRF instead, a
    second field drawn after the first, which starts acoustic waves.
    thermal: the heat-conducting gas (V = 17, cα = 1, κ = μγcv/Pr with
    Pr = 0.75), J = 0."""
    γ = 1.4
    if thermal:
        MP = material_params("sg", 1.0, cv=1.0, γ=γ, b0=1.0, μ=μ, cα=1.0, Pr=0.75)
    else:
        MP = material_params("sg", 1.0, cv=1.0, γ=γ, b0=1.0, μ=μ)
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(n, d=1.0 / n)
    kx, ky = np.meshgrid(k, k, indexing="ij")
    spec = np.exp(-0.5 * (2 * np.pi * ell) ** 2 * (kx ** 2 + ky ** 2))

    def field():
        noise = rng.standard_normal((n, n))
        f = np.real(np.fft.ifft2(np.fft.fft2(noise) * np.sqrt(spec)))
        return (f - f.mean()) / f.std()

    ρ = 1 + 0.05 * field()
    p = np.ones((n, n)) if p_amp == 0 else 1 + p_amp * field()
    v = np.zeros((n, n, 3))
    A = (ρ ** (1 / 3))[..., None, None] * np.eye(3)
    u = cvec_field(ρ, p, v, MP, A, np.zeros((n, n, 3)) if thermal else None)
    return u, MP, None, [1.0 / n, 1.0 / n], [PERIODIC] * 4


def plastic_pulse_2d(n=1024):
    """Elastoplastic copper-like solid (Godunov-Romenski EOS, power-law plastic
    relaxation σY=0.4, τ0=1, n=10) with a Gaussian x-velocity pulse."""
    MP = material_params("gr", 8.9, cv=4e-4, γ=2.0, c0=3.909, α=1.0, β=3.0,
                         b0=2.1, σY=0.4, τ0=1.0, n=10)
    dx = 1.0 / n
    c = (np.arange(n) + 0.5) * dx
    x, y = np.meshgrid(c, c, indexing="ij")
    vx = 0.5 * np.exp(-((x - 0.3) ** 2 + (y - 0.5) ** 2) / 0.01)
    v = np.stack([vx, np.zeros_like(vx), np.zeros_like(vx)], axis=-1)
    ρ = np.full((n, n), MP.rho0)
    p = np.zeros((n, n))
    A = np.broadcast_to(np.eye(3), (n, n, 3, 3)).copy()
    u = cvec_field(ρ, p, v, MP, A)
    return u, MP, None, [dx, dx], [TRANSMISSIVE] * 4


def shock_tube_batch(count=256, seed=0, nmin=100, nmax=4000):
    """Independent 1D shock tubes with seeded sizes and parameters.

    Returns a list of (u, MP, tf, dX, bcs)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nx = int(rng.integers(nmin, nmax + 1))
        μ = float(10 ** rng.uniform(-5, -1))
        ρL, ρR = rng.uniform(0.1, 2.0, size=2)
        pL, pR = rng.uniform(0.1, 10.0, size=2)
        γ = float(rng.uniform(1.3, 1.67))
        MP = material_params("sg", 1.0, cv=1.0, γ=γ, b0=1.0, μ=μ)
        dx = 1.0 / nx
        x = (np.arange(nx) + 0.5) * dx
        ρ = np.where(x < 0.5, ρL, ρR)
        p = np.where(x < 0.5, pL, pR)
        v = np.zeros((nx, 3))
        A = (ρ / MP.rho0)[:, None, None] ** (1 / 3) * np.eye(3)
        u = cvec_field(ρ, p, v, MP, A)
        out.append((u, MP, 0.2, [dx], [TRANSMISSIVE, TRANSMISSIVE]))
    return out


def random_states(count, MP, thermal=False, seed=0):
    """Random physically valid states, after the reference's binding-test
    generator (code/cpp/bindings/bindings_tests/test_functions.py:9-17):
    random A with det A > 0, ρ = ρ0 det A, random p, v and J."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        A = np.eye(3) + 0.3 * rng.uniform(-1, 1, size=(3, 3))
        detA = np.linalg.det(A)
        if detA <= 0.1:
            continue
        ρ = MP.rho0 * detA
        p = rng.uniform(0.5, 2.0)
        v = rng.uniform(-1, 1, size=3)
        J = rng.uniform(-1, 1, size=3) if thermal else None
        u = cvec_field(np.array(ρ), np.array(p), v, MP, A, J)
        out.append(u)
    return np.array(out)
