"""The particle provider on the CPU: the library's host materialisation of a parameter file against the oracle's
restatement of the provider (oracle.provide), and the oracle's draws against the laws they are defined by.

The reference draws its particles from std::random_device with MyCEL distributions, so no reference vector
exists; the provider's definition (Philox stream tag 1, one block per draw) is pinned by (a) the host
materialisation being bit-identical to the oracle's independent restatement, (b) the sample laws: lognormal and
normal means and spreads, discrete type frequencies, exact sequential counts, isotropic random magnetisations and
the ranges of the random angles."""
import math

import numpy as np
import pytest

import oracle
from provider_cases import BASE, CASES, host_particles, host_props, write_run


@pytest.mark.parametrize("name", list(CASES))
def test_host_materialisation_equals_oracle(tmp_path, name):
    """Same type per particle and the same draws: properties and start states bit-identical; body axes within
    2 ulp (the two sides round the products of the Euler-angle sines and cosines in their own order)."""
    problem, aniso, discrete, types = CASES[name]
    n = 700
    path = write_run(tmp_path, name, problem, aniso, types, n, seed=12345, discrete=discrete)
    r, parts = host_particles(path)
    o = oracle.provide(problem, aniso, types, 12345, n, discrete=discrete)
    assert r.n_particles == n and r.anisotropy == aniso
    assert np.array_equal(host_props(parts), o["props"])
    assert np.array_equal(parts["y0"], o["y0"])
    assert np.max(np.abs(parts["axes"] - o["axes"])) <= 2 * 2.0 ** -53 * 2


def test_sequential_selection_counts():
    """Without the discrete draw the types come in list order, each for its count; particles beyond the counts
    take the last type (Provider/ParticleProvider.h:230-295 hands out the list in order)."""
    types = [dict(BASE, count=5), dict(BASE, count=0, K=2e4), dict(BASE, count=7, K=3e4), dict(BASE, count=3, K=4e4)]
    o = oracle.provide(1, 1, types, 9, 20)
    assert o["type"].tolist() == [0] * 5 + [2] * 7 + [3] * 8
    assert np.array_equal(o["props"][:, 5], np.array([1e4] * 5 + [3e4] * 7 + [4e4] * 8))


def test_discrete_selection_frequencies():
    """Discrete selection: type k with probability count_k / total (4 sigma binomial bounds)."""
    types = [dict(BASE, count=1), dict(BASE, count=3), dict(BASE, count=6)]
    n = 40_000
    freq = np.bincount(oracle.provide(1, 1, types, 21, n, discrete=True)["type"], minlength=3) / n
    for k, p in enumerate((0.1, 0.3, 0.6)):
        assert abs(freq[k] - p) < 4 * math.sqrt(p * (1 - p) / n), (k, freq[k], p)


@pytest.mark.parametrize("kind,rel", [("lognormal", True), ("normal", True), ("lognormal", False), ("normal", False)])
def test_distribution_laws(kind, rel):
    """Radius, shell and K distributions: the sample mean and standard deviation are the given mean and width
    (width * mean when relative); a lognormal sample is positive and its log is normal with variance
    ln(1 + sd^2 / mean^2)."""
    n = 40_000
    width_r = 0.15 if rel else 1.5e-9
    width_k = 0.1 if rel else 1e3
    t = dict(BASE, count=n, r_h=3e-8, rad=(kind, width_r, rel), shell=(kind, width_r, rel), K_dist=(kind, width_k, rel))
    o = oracle.provide(3, 1, [t], 33, n)
    radius, rh, K = o["props"][:, 0], o["props"][:, 6], o["props"][:, 5]
    shell = rh - radius
    for x, mean, w in ((radius, BASE["radius"], width_r), (shell, 3e-8 - BASE["radius"], width_r), (K, BASE["K"], width_k)):
        sd = w * mean if rel else w
        assert abs(x.mean() - mean) < 4 * sd / math.sqrt(n), (x.mean(), mean)
        assert abs(x.std() - sd) < 4 * sd / math.sqrt(2 * n) * (3 if kind == "lognormal" else 1.5), (x.std(), sd)
        if kind == "lognormal":
            s2 = math.log1p(sd * sd / (mean * mean))
            assert x.min() > 0 and abs(np.log(x).var() - s2) < 4 * s2 * math.sqrt(2 / n)


def test_random_magnetisation_is_isotropic():
    """Random magnetisations (three normals, normalised; BrownAndNeelEulerSpherical: phi, theta uniform as the
    reference draws them): unit length; the Cartesian draw has mean 0 and second moment I/3."""
    n = 30_000
    o = oracle.provide(1, 1, [dict(BASE, count=n, rand_mag=True)], 5, n)
    m = o["y0"]
    assert np.allclose(np.linalg.norm(m, axis=1), 1.0, atol=1e-15)
    assert np.all(np.abs(m.mean(0)) < 4 * math.sqrt(1 / 3 / n))
    assert np.allclose(m.T @ m / n, np.eye(3) / 3, atol=4 * math.sqrt(4 / 45 / n) + 1e-3)
    e = oracle.provide(4, 1, [dict(BASE, count=n, rand_mag=True, rand_orient=True, r_h=2.5e-8)], 5, n)["y0"]
    assert np.all((e[:, [0, 2]] > 0) & (e[:, [0, 2]] < 2 * math.pi)) and np.all((e[:, 1] > 0) & (e[:, 1] < math.pi))
    assert np.all((e[:, 3] >= 0) & (e[:, 3] <= math.pi)) and np.all(np.abs(e[:, 4]) <= math.pi)
    # theta of the moment U(0, pi) in the reference's parameterisation (not area-uniform): mean pi / 2, var pi^2 / 12
    assert abs(e[:, 3].mean() - math.pi / 2) < 4 * math.pi / math.sqrt(12 * n)


def test_random_orientation_and_rotation_angles():
    """Random orientations: Euler angles U(0, 2 pi) each (NeelRelaxation.h:223-261 builds the axes from them), so
    the axes are orthonormal and the z axis has cos(theta) = z_z with E[z_z] = E[cos U(0, 2 pi)] = 0; random
    uniaxial rotations: U(0, 2 pi), U(0, pi), U(0, 2 pi) (UniaxialCubic_Distribution)."""
    n = 20_000
    o = oracle.provide(1, 4, [dict(BASE, count=n, K_cubic=-1e4, rand_orient=True, rand_urot=True)], 8, n)
    ax = o["axes"].reshape(n, 3, 3)
    assert np.allclose(ax @ ax.transpose(0, 2, 1), np.eye(3), atol=1e-14)
    assert abs(ax[:, 2, 2].mean()) < 4 * math.sqrt(0.5 / n)
    u = o["props"][:, 9:12]
    for k, hi in enumerate((2 * math.pi, math.pi, 2 * math.pi)):
        assert u[:, k].min() > 0 and u[:, k].max() < hi
        assert abs(u[:, k].mean() - hi / 2) < 4 * hi / math.sqrt(12 * n)


def test_fixed_state_from_euler_angles():
    """Fixed orientation and magnetisation: BrownAndNeel's axis n is the first column of the reference's Euler
    matrix (BrownAndNeelRelaxation.inl:297-307), the moment the normalised given direction."""
    a, b, g = 0.7, 0.4, 1.9
    o = oracle.provide(3, 1, [dict(BASE, count=1, euler=(a, b, g), mag=(0.0, 3.0, 4.0))], 1, 1)
    n = [math.cos(a) * math.cos(g) - math.sin(a) * math.cos(b) * math.sin(g),
         -math.cos(a) * math.sin(g) - math.sin(a) * math.cos(b) * math.cos(g), math.sin(a) * math.sin(b)]
    assert np.allclose(o["y0"][0, :3], n, atol=1e-15) and np.allclose(o["y0"][0, 3:], [0.0, 0.6, 0.8], atol=1e-16)
    assert np.allclose(o["axes"][0], oracle.euler_axes(a, b, g).ravel(), atol=0)
