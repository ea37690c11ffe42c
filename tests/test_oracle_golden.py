"""Pin the CPU oracle against the reference's own gtest golden vectors.

Fixture: tests/golden/reference_vectors.json (made by
tests/golden/make_reference_vectors.py from the reference's Tests/ sources);
each case carries its Tests/... file:line citation.  Comparison is Eigen's
isApprox, |a-b| <= rtol*min(|a|,|b|), or the absolute norm bound the reference
test itself uses.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle

DOC = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())
FIX = DOC["fixtures"]
CASES = DOC["cases"]


def is_approx(a, b, rtol):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.linalg.norm(a - b) <= rtol * min(np.linalg.norm(a), np.linalg.norm(b))


def props_of(f):
    return oracle.props_array(f["radius"], f["Ms"], f["alpha"], f["gamma"], f["T"], f["K"],
                              f.get("r_h", 0.0), f.get("viscosity", 0.0))


def ids(kind):
    return [c for c in CASES if c["kind"] == kind]


def cid(c):
    return c["cite"].split("/")[-1]


@pytest.mark.parametrize("c", ids("neel_det"), ids=cid)
def test_neel_det(c):
    f = FIX["neel"]
    ax = oracle.euler_axes(*f["euler"])
    got = oracle.evaluate(1, 1, props_of(f), "det", c["y"], c["h"], axes=ax)
    assert is_approx(got, c["expected"], c["rtol"]), (got, c["expected"])


@pytest.mark.parametrize("c", ids("neel_sto"), ids=cid)
def test_neel_sto(c):
    f = FIX["neel"]
    got = oracle.evaluate(1, 1, props_of(f), "sto", c["y"], axes=oracle.euler_axes(*f["euler"]))
    assert is_approx(got, c["expected"], c["rtol"])


@pytest.mark.parametrize("c", ids("neel_drift"), ids=cid)
def test_neel_drift(c):
    f = FIX["neel"]
    got = oracle.evaluate(1, 1, props_of(f), "drift", c["y"], axes=oracle.euler_axes(*f["euler"]))
    assert is_approx(got, c["expected"], c["rtol"])
    # direction test of the same gtest: drift is anti-parallel to m
    assert is_approx(got / np.linalg.norm(got), -np.asarray(c["y"]), 1e-12)


@pytest.mark.parametrize("c", ids("neel_jacdet_reference") + ids("neel_jacsto_reference"), ids=cid)
def test_neel_reference_jacobians(c):
    f = FIX["neel"]
    ax = oracle.euler_axes(*f["euler"])
    if c["kind"] == "neel_jacdet_reference":
        got = oracle.evaluate(1, 1, props_of(f), "jacdet_ref", c["y"], c["h"], axes=ax, dt=c["dt"])
    else:
        got = oracle.evaluate(1, 1, props_of(f), "jacsto_ref", c["y"], c["dW"], axes=ax)
    assert is_approx(got, c["expected"], c["rtol"])


@pytest.mark.parametrize("c", ids("uniaxial_jacobian"), ids=cid)
def test_uniaxial(c):
    p = c["props"]
    props = oracle.props_array(p["radius"], p["Ms"], 0.1, 1.76e11, 295.0, p["K"])
    axes = np.zeros((3, 3))
    axes[2] = c["axis"]
    field, jac = oracle.aniso_field(1, props, c["m"], axes)
    assert is_approx(jac, c["expected"], c["rtol"])
    # Test_Uniaxial_Anisotropy.cpp:10-31: H = -(2K/Ms)(e.m) e
    exp = -(2 * p["K"] / p["Ms"]) * np.dot(c["axis"], c["m"]) * np.asarray(c["axis"])
    assert is_approx(field, exp, 1e-12)


@pytest.mark.parametrize("c", ids("spherical_det") + ids("spherical_drift"), ids=cid)
def test_spherical_det_drift(c):
    f = FIX["spherical"]
    ax = oracle.euler_axes(*f["euler"])
    what = "det" if c["kind"] == "spherical_det" else "drift"
    kw = dict(axes=ax, rotate=f["rotate"], min_angle=f["min_angle"])
    got = oracle.evaluate(2, 1, props_of(f), what, c["y"], c["h"], **kw)
    if "atol_norm" in c:
        assert np.linalg.norm(got - np.asarray(c["expected"])) < c["atol_norm"]
    else:
        assert is_approx(got, c["expected"], c["rtol"]), (got, c["expected"])
    if c["rotated"] is not None:
        assert bool(oracle.evaluate(2, 1, props_of(f), "rotated", c["y"], **kw)[0]) == c["rotated"]


@pytest.mark.parametrize("c", ids("spherical_sto"), ids=cid)
def test_spherical_sto(c):
    f = FIX["spherical"]
    kw = dict(axes=oracle.euler_axes(*f["euler"]), rotate=f["rotate"], min_angle=f["min_angle"])
    got = oracle.evaluate(2, 1, props_of(f), "sto", c["y"], **kw)
    assert is_approx(got, c["expected"], c["rtol"])
    assert bool(oracle.evaluate(2, 1, props_of(f), "rotated", c["y"], **kw)[0]) == c["rotated"]


def wrap(y):
    """theta in [0, pi], phi in [0, 2 pi) (MyCEL Wrap2DSphericalCoordinates on a point of the sphere)."""
    th, ph = y
    v = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    return np.array([math.acos(max(-1.0, min(1.0, v[2]))), math.atan2(v[1], v[0]) % (2 * math.pi)])


@pytest.mark.parametrize("c", ids("spherical_finish_roundtrip"), ids=cid)
def test_spherical_finish(c):
    f = FIX["spherical"]
    kw = dict(axes=oracle.euler_axes(*f["euler"]), rotate=f["rotate"], min_angle=f["min_angle"])
    got = oracle.evaluate(2, 1, props_of(f), "finish", c["y"], **kw)
    if c["theta_only"]:
        assert wrap(got)[0] == pytest.approx(c["y"][0], abs=1e-15 * 4)
    else:
        assert is_approx(wrap(got), wrap(c["y"]), 1e-12)


def bn_props():
    return props_of(FIX["brown_and_neel"])


@pytest.mark.parametrize("c", ids("bn_det"), ids=cid)
def test_bn_det(c):
    got = oracle.evaluate(3, 1, bn_props(), "det", c["y"], c["h"])
    exp = np.asarray(c["expected"])
    if "atol_norm" in c:
        assert np.linalg.norm(got - exp) < c["atol_norm"]
    else:
        assert is_approx(got, exp, c["rtol"]), (got, exp)


@pytest.mark.parametrize("c", ids("bn_sto") + ids("bn_drift"), ids=cid)
def test_bn_sto_drift(c):
    what = "sto" if c["kind"] == "bn_sto" else "drift"
    got = oracle.evaluate(3, 1, bn_props(), what, c["y"], simple=c["simple"])
    assert is_approx(got, c["expected"], c["rtol"]), (got, c["expected"])


@pytest.mark.parametrize("c", ids("newton"), ids=cid)
def test_newton_roots(c):
    x, it = oracle.newton_testfn(c["fn"], c["x0"], c["tol"], c["max_iter"])
    ok = any(is_approx(x, r, 1e-12) for r in c["roots"])
    if c["known_failing"] and not ok:
        pytest.xfail("excluded from the reference CI (ci.linux.yml:151)")
    assert ok and it > 0, (x, it)


def bnes_kw():
    f = FIX["brown_and_neel_euler_spherical"]
    return dict(rotate=True, min_angle=f["neel_min_angle"], euler_rotate=True, brown_min_angle=f["brown_min_angle"])


@pytest.mark.parametrize("c", ids("bnes_det") + ids("bnes_drift") + ids("bnes_sto"), ids=cid)
def test_bnes_terms(c):
    """BrownAndNeelEulerSpherical det (with and without field), stochastic matrix and drift in all four chart
    combinations (Test_NeelSpherical_BrownEuler_Problem.cpp), with the rotation flags the reference test asserts."""
    props = props_of(FIX["brown_and_neel_euler_spherical"])
    what = {"bnes_det": "det", "bnes_drift": "drift", "bnes_sto": "sto"}[c["kind"]]
    got = oracle.evaluate(4, 1, props, what, c["y"], c.get("h"), **bnes_kw())
    assert is_approx(got, c["expected"], c["rtol"]), (got, c["expected"])
    flags = oracle.evaluate(4, 1, props, "rotated", c["y"], **bnes_kw())
    assert (bool(flags[0]), bool(flags[1])) == (c["brown_rotated"], c["neel_rotated"])


@pytest.mark.parametrize("c", ids("bnes_rotation"), ids=cid)
def test_bnes_rotation_roundtrip(c):
    """prepareCalculations rotates the chart(s) to the Mathematica state; finishCalculations maps it back (the
    reference test adds 2 pi to negative angles of the rotated blocks before comparing)."""
    props = props_of(FIX["brown_and_neel_euler_spherical"])
    out = oracle.evaluate(4, 1, props, "finish", c["y"], **bnes_kw())
    prepared, back = out[:5], out[5:].copy()
    assert is_approx(prepared, c["expected"], c["rtol"]), (prepared, c["expected"])
    blocks = ([0, 1, 2] if c["brown_rotated"] else []) + ([3, 4] if c["neel_rotated"] else [])
    for i in blocks:
        if back[i] < 0:
            back[i] += 2 * math.pi
    assert is_approx(back, c["y"], 1e-12), (back, c["y"])


def test_bnes_output_and_chart_invariance():
    """Output = (body x axis, body y axis, m) (BrownAndNeelRelaxationEulerSpherical.h:674-735): the axes are the
    columns of scipy's extrinsic z-x-z rotation of the Euler angles, m the spherical unit vector.  Closed forms of the
    chart maps: the rotated Euler angles and spherical angles describe the body axes and the moment turned by +90
    degrees about y (Ry R, Ry m), and mapping back returns the same body axes and the same m."""
    from scipy.spatial.transform import Rotation
    props = props_of(FIX["brown_and_neel_euler_spherical"])
    rng = np.random.default_rng(4)
    for _ in range(16):
        y = np.concatenate([rng.uniform(0, 2 * math.pi, 1), rng.uniform(0.05, math.pi - 0.05, 1), rng.uniform(0, 2 * math.pi, 1),
                            rng.uniform(0.05, math.pi - 0.05, 1), rng.uniform(0, 2 * math.pi, 1)])
        o = oracle.evaluate(4, 1, props, "output", y)
        R = Rotation.from_euler("zxz", y[:3]).as_matrix()
        assert np.allclose(o[0:3], R[:, 0], atol=1e-15) and np.allclose(o[3:6], R[:, 1], atol=1e-15)
        th, ph = y[3], y[4]
        assert np.allclose(o[6:9], [math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)], atol=1e-15)
        # force both charts (windows wider than the state): prepared angles describe the same body / moment seen from
        # the frame turned by 90 degrees about y, and finish maps them back
        kw = dict(rotate=True, min_angle=math.pi / 2 + 0.01, euler_rotate=True, brown_min_angle=math.pi / 2 + 0.01)
        out = oracle.evaluate(4, 1, props, "finish", y, **kw)
        Ry = Rotation.from_euler("y", 90, degrees=True).as_matrix()
        Rp = Rotation.from_euler("zxz", out[:3]).as_matrix()
        assert np.allclose(Ry @ R, Rp, atol=1e-12)
        tp, pp = out[3], out[4]
        assert np.allclose(Ry @ o[6:9], [math.sin(tp) * math.cos(pp), math.sin(tp) * math.sin(pp), math.cos(tp)], atol=1e-12)
        ob = oracle.evaluate(4, 1, props, "output", out[5:])
        assert np.allclose(ob, o, atol=1e-12)


def test_neel_param_closed_forms():
    """Tests/Neel_Problem_Test/Test_ParamCalc_Neel.cpp:12-126 (Ms=477464, r=10nm, alpha=0.1, gamma=1.76e11)."""
    Ms, r, al, g, T = 477464.0, 10e-9, 0.1, 1.76e11, 295.0
    kB = 1.3806485279e-23
    p = oracle.neel_params(r, Ms, al, g, T)
    mu = Ms * 4.0 / 3.0 * math.pi * r ** 3
    pre1 = -g / (1 + al ** 2)
    pre2 = abs(g) * al / (1 + al ** 2)
    noise = math.sqrt(2 * kB * T * al / (g * mu))
    assert p["NeelFactor1"] == pytest.approx(pre1, rel=4 * 2.2e-16)
    assert p["NeelFactor1"] < 0.0      # PositivGyromagneticRatioIsElectron
    assert p["NeelFactor2"] == pytest.approx(pre2, rel=4 * 2.2e-16)
    assert p["NoisePrefactor"] == pytest.approx(noise, rel=4 * 2.2e-16)
    assert p["NoisePrefactor"] * p["NeelFactor1"] == pytest.approx(pre1 * noise, rel=4 * 2.2e-16)
    assert p["DriftPrefactor"] == pytest.approx(-(pre1 ** 2 + pre2 ** 2) * noise ** 2, rel=8 * 2.2e-16)
    q = oracle.neel_params(r, Ms, math.inf, g, T)        # InfDampingTest
    assert q["NeelFactor1"] == 0.0 and q["NeelFactor2"] == 0.0
    assert q["NoisePrefactor"] == 0.0 and q["DriftPrefactor"] == 0.0


def test_uniaxial_sign_convention():
    """NegativeKmeansUniaxial (Test_Uniaxial_Anisotropy.cpp:82-99) on its fixture (K = +1e4,
    Test_Uniaxial_Anisotropy.h:41-50); PositiveKmeansUniaxial asserts the opposite ordering and is excluded
    from the reference CI (ci.linux.yml:151)."""
    axes = np.zeros((3, 3))
    axes[2] = [1.0, 0.0, 0.0]
    props = oracle.props_array(10e-9, 477464.0, 0.1, 1.76e11, 295.0, 1e4)
    vals = [-oracle.aniso_field(1, props, e, axes)[0] @ e for e in np.eye(3)]
    assert vals[0] > vals[1] and vals[0] > vals[2] and vals[1] == vals[2]


def test_cubic_field_is_energy_gradient():
    """Cubic anisotropy has no reference test: pin H = K1_MS sum_i c_i (c_j^2 + c_k^2) e_i (CubicAnisotropy.h:109-146,
    K1_MS = -2K/Ms) as -(1/Ms) grad of K sum_{i<j} c_i^2 c_j^2 (central differences) for random cube orientations,
    plus its symmetries: zero along a cube axis, parallel to m along a body diagonal."""
    rng = np.random.default_rng(12)
    Ms, K = 4.8e5, -1.1e4
    props = oracle.props_array(11e-9, Ms, 0.05, 1.76e11, 300.0, K)

    def energy(m, ax):
        c = ax @ m
        return K * (c[0] ** 2 * c[1] ** 2 + c[1] ** 2 * c[2] ** 2 + c[0] ** 2 * c[2] ** 2)

    for _ in range(8):
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        ax = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)],
                       [2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)],
                       [2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)]])
        m = rng.standard_normal(3)
        m /= np.linalg.norm(m)
        H, _ = oracle.aniso_field(2, props, m, ax)
        h = 1e-6
        grad = np.array([(energy(m + h * e, ax) - energy(m - h * e, ax)) / (2 * h) for e in np.eye(3)])
        assert np.allclose(H, -grad / Ms, rtol=1e-7, atol=1e-12 * abs(K) / Ms)
        for i in range(3):                                   # along a cube axis: c = (1, 0, 0) up to rounding -> H = 0
            assert np.allclose(oracle.aniso_field(2, props, ax[i], ax)[0], 0.0, atol=1e-14 * abs(K) / Ms)
        d = ax.sum(axis=0) / math.sqrt(3.0)                  # body diagonal: c_i = 1/sqrt(3) -> H = (-2K/Ms)(2/3) d
        Hd = oracle.aniso_field(2, props, d, ax)[0]
        assert np.allclose(Hd, (-2 * K / Ms) * (2.0 / 3.0) * d, rtol=1e-13, atol=1e-14 * abs(K) / Ms)


def test_fields_closed_forms():
    """The oracle's fields at points where the reference's formulas have closed values (ConstantField.h:43,
    TriangularField.h:51-67, RectangularField.h:26-63, SincField.h:41-45)."""
    A, off, P = np.array([0.0, 0.01, 0.02]), np.array([0.001, 0.0, -0.002]), 1e-6
    h = lambda f, t: oracle.field_at(f, t)
    const = dict(type=2, offset=off)
    assert np.array_equal(h(const, 0.3), off)
    tri = dict(type=5, offset=off, amp=A, period=P)
    for t, s in [(0.0, 3.0), (P / 4, 2.0), (P / 2, 1.0), (3 * P / 4, 2.0), (-P / 4, 2.0), (P + P / 4, 2.0)]:
        assert np.allclose(h(tri, t), s * A + off, rtol=1e-12, atol=1e-18), (t, s)
    tau = P / 10
    rect = dict(type=6, offset=off, amp=A, period=P, tau=tau)
    for pos in (0.1 * P, 0.3 * P, 0.5 * P):
        assert np.allclose(h(rect, pos), A * (1 - math.exp(-pos / tau)) + off, rtol=1e-12, atol=1e-18)
    mx = A * (1 - math.exp(-0.5 * P / tau))
    for pos in (0.6 * P, 0.9 * P):
        assert np.allclose(h(rect, pos), A * math.exp(-(pos - 0.5 * P) / tau) + off - (A - mx), rtol=1e-12, atol=1e-18)
    assert np.allclose(h(rect, 0.5 * P), h(rect, 0.5 * P * (1 + 1e-12)), rtol=1e-9, atol=1e-15)   # continuous
    alt = dict(type=6, offset=off, amp=A, period=P, tau=tau, alternating=True)
    ts = np.linspace(0, P, 20001)
    vals = np.array([h(alt, t) for t in ts])
    # the alternating wave is symmetric about the offset: max - offset = offset - min
    assert np.allclose(vals.max(0) - off, off - vals.min(0), rtol=1e-3, atol=1e-12)
    sinc = dict(type=7, offset=off, amp=A, period=P)
    half = P / 2
    assert np.allclose(h(sinc, half), A + off, rtol=1e-15)
    assert np.allclose(h(sinc, half + half / 11), off, atol=1e-18)            # sinc(pi) = 0
    x = math.pi * 0.5
    assert np.allclose(h(sinc, half + half / 22), A * (math.sin(x) / x) ** 2 + off, rtol=1e-12)


def test_uniaxialcubic_field_is_energy_gradient():
    """Uniaxialcubic anisotropy has no reference test: pin H = (-2K_u/Ms)(m.u)u + cubic field (UniaxialCubicAnisotropy.hpp:114)
    with u = R^T z, R the matrix of calcRotation (:61-90).  R is checked against a library rotation (scipy's
    extrinsic z-x-z Euler rotation), H against -(1/Ms) grad of K_u (m.u)^2 + K_c sum_{i<j} c_i^2 c_j^2 by central
    differences, and the two limits K_c = 0 (the uniaxial field about u) and K_u = 0 (the cubic field)."""
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(31)
    Ms, Ku, Kc = 4.8e5, 7.0e3, -1.1e4
    base = oracle.props_array(11e-9, Ms, 0.05, 1.76e11, 300.0, Ku)
    for _ in range(8):
        ax = Rotation.random(random_state=rng).as_matrix()   # rows: cube axes x, y, z
        e = rng.uniform([0, 0, 0], [2 * math.pi, math.pi, 2 * math.pi])
        R = Rotation.from_euler("zxz", e).as_matrix()
        u = R.T @ ax[2]
        props = np.concatenate([base, [Kc, *e]])
        m = rng.standard_normal(3)
        m /= np.linalg.norm(m)

        def energy(v):
            c = ax @ v
            return Ku * (v @ u) ** 2 + Kc * (c[0] ** 2 * c[1] ** 2 + c[1] ** 2 * c[2] ** 2 + c[0] ** 2 * c[2] ** 2)
        H, _ = oracle.aniso_field(4, props, m, ax)
        h = 1e-6
        grad = np.array([(energy(m + h * d) - energy(m - h * d)) / (2 * h) for d in np.eye(3)])
        scale = (abs(Ku) + abs(Kc)) / Ms
        assert np.allclose(H, -grad / Ms, rtol=1e-7, atol=1e-12 * scale)
        # K_c = 0: the uniaxial field about u; K_u = 0: the cubic field about the particle's axes
        ax_u = ax.copy()
        ax_u[2] = u
        Hu = oracle.aniso_field(4, np.concatenate([base, [0.0, *e]]), m, ax)[0]
        assert np.allclose(Hu, oracle.aniso_field(1, base, m, ax_u)[0], rtol=1e-14, atol=1e-16 * scale)
        Hc = oracle.aniso_field(4, np.concatenate([base[:5], [0.0], base[6:], [Kc, *e]]), m, ax)[0]
        assert np.allclose(Hc, oracle.aniso_field(2, np.concatenate([base[:5], [Kc], base[6:]]), m, ax)[0],
                           rtol=1e-14, atol=1e-16 * scale)
