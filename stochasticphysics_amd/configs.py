"""Workload builders for the five benchmark configurations (synthetic, seeded).

Per-particle random inputs come from numpy's PCG64 generator seeded per
configuration, so the same ensemble is rebuilt identically on any machine.
"""
import math

import numpy as np

from .engine import Ensemble

GAMMA = 1.76e11


def unit_sphere(rng, n):
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def to_spherical(m):
    """(theta, phi) of unit vectors (ParticleStateInitializer.h:132-144 convention)."""
    theta = np.arccos(np.clip(m[:, 2], -1.0, 1.0))
    phi = np.arctan2(m[:, 1], m[:, 0])
    return np.stack([theta, phi], axis=1)


def random_axes(rng, n):
    """Body axes = rows of the rotation matrix of a uniform random unit quaternion; [n, 9]."""
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
        2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
        2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], axis=1)
    return R


def easy_axis_axes(z):
    """[n, 9] axes whose z row is the given easy axis (x, y completed orthonormally)."""
    n = len(z)
    a = np.where(np.abs(z[:, :1]) < 0.9, np.array([[1.0, 0.0, 0.0]]), np.array([[0.0, 1.0, 0.0]]))
    x = np.cross(a, z)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    y = np.cross(z, x)
    return np.concatenate([x, y, z], axis=1).reshape(n, 9)


def full(n, v):
    return np.full(n, float(v))


def config1(n=4096, n_steps=200_000, oversampling=100, seed=11, precision="fp64"):
    """Plain Neel relaxation, uniaxial, EulerMaruyama, FP64, zero field, easy axes and m0 along +z."""
    axes = np.tile(np.eye(3).ravel(), (n, 1))
    return Ensemble(problem="Neel", solver="EulerMaruyama", field="none", anisotropy="uniaxial", precision=precision,
                    n_steps=n_steps, oversampling=oversampling, dt=1e-12, seed=seed,
                    radius=full(n, 10e-9), Ms=full(n, 4.8e5), alpha=full(n, 0.1), gamma=full(n, GAMMA),
                    T=full(n, 295.0), K=full(n, 1.0e4), axes=axes, y0=np.tile([0.0, 0.0, 1.0], (n, 1)))


def config2(n=65536, n_steps=200_000, oversampling=50, seed=22, precision="fp64"):
    """Coupled Brown+Neel (full model), Heun_Strong, uniaxial, 12 mT sinusoidal z-field at 100 kHz."""
    rng = np.random.default_rng(seed)
    nn, mm = unit_sphere(rng, n), unit_sphere(rng, n)
    return Ensemble(problem="BrownAndNeel", solver="Heun_Strong", field="sinusoidal", anisotropy="uniaxial",
                    precision=precision, n_steps=n_steps, oversampling=oversampling, dt=1e-11, seed=seed,
                    radius=full(n, 12e-9), Ms=full(n, 4.8e5), alpha=full(n, 0.1), gamma=full(n, GAMMA),
                    T=full(n, 295.0), K=full(n, 5e3), r_h=full(n, 25e-9), viscosity=full(n, 1e-3),
                    y0=np.concatenate([nn, mm], axis=1),
                    field_amplitude=(0.0, 0.0, 0.012), field_freq=(1e5, 0.0, 0.0))


def config3(n=1_048_576, n_steps=500_000, oversampling=100, seed=33, solver="Heun_Strong", precision="fp64"):
    """NeelSpherical, cubic K1 = -1.1e4, log-normal radii, 2D Lissajous field, chart rotation at pi/36."""
    rng = np.random.default_rng(seed)
    radius = 11e-9 * np.exp(0.1 * rng.standard_normal(n))
    axes = random_axes(rng, n)
    y0 = to_spherical(unit_sphere(rng, n))
    return Ensemble(problem="NeelSpherical", solver=solver, field="lissajous", anisotropy="cubic", precision=precision,
                    n_steps=n_steps, oversampling=oversampling, dt=2e-12, seed=seed,
                    radius=radius, Ms=full(n, 4.8e5), alpha=full(n, 0.05), gamma=full(n, GAMMA), T=full(n, 300.0),
                    K=full(n, -1.1e4), axes=axes, y0=y0, coord_transform=True, min_angle=math.pi / 36,
                    field_amplitude=(0.012, 0.012, 0.0), field_freq=(25.25e3, 24.5e3, 0.0))


def config4_sets(n_sets=512, seed=44, precision="mixed", particle_range=(1e3, 1.28e5), step_range=(5e4, 5e5),
                 oversampling=100):
    """Uneven batched ensembles: half Neel (EulerMaruyama / Heun_Strong alternating), half BrownAndNeel."""
    rng = np.random.default_rng(seed)
    sets = []
    for i in range(n_sets):
        n = int(round(math.exp(rng.uniform(math.log(particle_range[0]), math.log(particle_range[1])))))
        steps = int(rng.uniform(*step_range)) // oversampling * oversampling
        r = rng.uniform(8e-9, 16e-9)
        K = rng.uniform(2e3, 2e4)
        alpha = rng.uniform(0.01, 0.5)
        amp = rng.uniform(0.0, 0.025)
        freq = rng.uniform(10e3, 500e3)
        brown = (i % 2) == 1
        sub = np.random.default_rng(seed * 100003 + i)
        common = dict(field="sinusoidal", anisotropy="uniaxial", precision=precision, n_steps=steps,
                      oversampling=oversampling, seed=seed * 1000 + i, radius=full(n, r), Ms=full(n, 4.8e5),
                      alpha=full(n, alpha), gamma=full(n, GAMMA), T=full(n, 295.0), K=full(n, K),
                      field_amplitude=(0.0, 0.0, amp), field_freq=(freq, 0.0, 0.0))
        if brown:
            e = Ensemble(problem="BrownAndNeel", solver="Heun_Strong", dt=1e-11,
                         r_h=full(n, sub.uniform(20e-9, 40e-9)), viscosity=full(n, 1e-3),
                         y0=np.concatenate([unit_sphere(sub, n), unit_sphere(sub, n)], axis=1), **common)
        else:
            solver = "EulerMaruyama" if (i // 2) % 2 == 0 else "Heun_Strong"
            e = Ensemble(problem="Neel", solver=solver, dt=1e-12, axes=easy_axis_axes(unit_sphere(sub, n)),
                         y0=unit_sphere(sub, n), **common)
        sets.append(e)
    return sets


def serving_request(rng, idx, n_range=(256, 4096), step_range=(2_000, 10_000), oversampling=100, precision="mixed"):
    """One request of the serving loop: config-3 physics (NeelSpherical, cubic, Lissajous) with damping
    U[0.01, 0.5]; damping >= 0.3 switches to Implicit_Midpoint (FP64 Newton, tolerance 1e-10)."""
    n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
    steps = int(rng.uniform(*step_range)) // oversampling * oversampling
    alpha = rng.uniform(0.01, 0.5)
    solver = "Implicit_Midpoint" if alpha >= 0.3 else "Heun_Strong"
    radius = 11e-9 * np.exp(0.1 * rng.standard_normal(n))
    return Ensemble(problem="NeelSpherical", solver=solver, field="lissajous", anisotropy="cubic", precision=precision,
                    n_steps=steps, oversampling=oversampling, dt=2e-12, seed=5_000_000 + idx, radius=radius,
                    Ms=full(n, 4.8e5), alpha=full(n, alpha), gamma=full(n, GAMMA), T=full(n, 300.0),
                    K=full(n, -1.1e4), axes=random_axes(rng, n), y0=to_spherical(unit_sphere(rng, n)),
                    coord_transform=True, min_angle=math.pi / 36, newton_tol=1e-10, newton_max_iter=20,
                    field_amplitude=(0.012, 0.012, 0.0), field_freq=(25.25e3, 24.5e3, 0.0))


def uniaxialcubic(n=4096, problem="Neel", solver="Heun_Strong", n_steps=20_000, oversampling=100, seed=66,
                  precision="fp64", Ku=8e3, Kc=-1.1e4, T=300.0):
    """Uniaxialcubic anisotropy (UniaxialCubicAnisotropy.hpp): a uniaxial part rotated against the cube axes by
    per-particle Euler angles (phi U(0, 2 pi), theta U(0, pi), psi U(0, 2 pi), as UseRandomCubicOrientation draws
    them), random cube orientations, a 2D Lissajous field; Neel (Cartesian) or NeelSpherical with chart rotation."""
    rng = np.random.default_rng(seed)
    radius = 11e-9 * np.exp(0.1 * rng.standard_normal(n))
    axes = random_axes(rng, n)
    rot = rng.uniform([0.0, 0.0, 0.0], [2 * math.pi, math.pi, 2 * math.pi], size=(n, 3))
    m0 = unit_sphere(rng, n)
    y0 = to_spherical(m0) if problem == "NeelSpherical" else m0
    return Ensemble(problem=problem, solver=solver, field="lissajous", anisotropy="uniaxialcubic", precision=precision,
                    n_steps=n_steps, oversampling=oversampling, dt=2e-12, seed=seed, radius=radius,
                    Ms=full(n, 4.8e5), alpha=full(n, 0.1), gamma=full(n, GAMMA), T=full(n, T), K=full(n, Ku),
                    K_cubic=full(n, Kc), uniaxial_rotation=rot, axes=axes, y0=y0,
                    coord_transform=problem == "NeelSpherical", min_angle=math.pi / 36,
                    field_amplitude=(0.012, 0.012, 0.0), field_freq=(25.25e3, 24.5e3, 0.0))


def bnes(n=65536, n_steps=200_000, oversampling=50, seed=77, solver="Heun_Strong", precision="fp64", anisotropy="uniaxial"):
    """BrownAndNeelEulerSpherical with the coupled Brown+Neel ensemble's particles (r 12 nm, r_h 25 nm, K 5e3,
    damping 0.1, 295 K, 12 mT 100 kHz sinusoidal z-field, dt 1e-11 s): uniformly random body orientations (Euler
    angles) and moments, both chart rotations on at pi/36."""
    rng = np.random.default_rng(seed)
    eul = np.stack([rng.uniform(0, 2 * math.pi, n), np.arccos(rng.uniform(-1, 1, n)), rng.uniform(0, 2 * math.pi, n)], 1)
    e = Ensemble(problem="BrownAndNeelEulerSpherical", solver=solver, field="sinusoidal", anisotropy=anisotropy,
                 precision=precision, n_steps=n_steps, oversampling=oversampling, dt=1e-11, seed=seed,
                 radius=full(n, 12e-9), Ms=full(n, 4.8e5), alpha=full(n, 0.1), gamma=full(n, GAMMA), T=full(n, 295.0),
                 K=full(n, 5e3 if anisotropy != "cubic" else -1.1e4), r_h=full(n, 25e-9), viscosity=full(n, 1e-3),
                 y0=np.concatenate([eul, to_spherical(unit_sphere(rng, n))], axis=1), coord_transform=True,
                 min_angle=math.pi / 36, euler_transform=True, brown_min_angle=math.pi / 36,
                 field_amplitude=(0.0, 0.0, 0.012), field_freq=(1e5, 0.0, 0.0))
    if anisotropy == "uniaxialcubic":
        e.K_cubic = full(n, -1.1e4)
        e.uniaxial_rotation = rng.uniform([0, 0, 0], [2 * math.pi, math.pi, 2 * math.pi], (n, 3))
    return e
