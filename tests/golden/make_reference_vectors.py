"""Write tests/golden/reference_vectors.json from the reference's own gtest files.

The reference (C++, Eigen/MyCEL/gtest; not buildable here) checks its problem
terms against vectors it states were computed in Mathematica.  This script
reads those test sources as text, extracts the numeric literals of every
``lhs << a, b, ...;`` initialiser, evaluates the small constant expressions
(``pi / 2.5`` ...) and groups them into cases with a file:line citation.  The
JSON it writes is data only (inputs, expected outputs, tolerances).

Run in the development container (the reference is not present on GPU boxes):
    python tests/golden/make_reference_vectors.py
"""
import json
import math
import re
from pathlib import Path

REF = Path("/root/reference/Tests")
OUT = Path(__file__).with_name("reference_vectors.json")


def statements(rel):
    """(first_line, last_line, lhs, values) for each ``lhs << v1, v2, ...;``."""
    lines = (REF / rel).read_text(encoding="latin-1").splitlines()
    out, i = [], 0
    while i < len(lines):
        ln = lines[i].strip()
        if ln.startswith("//") or "<<" not in ln or "cout" in ln:
            i += 1
            continue
        start, buf = i, ln
        while ";" not in buf:
            i += 1
            buf += " " + lines[i].strip()
        lhs, rhs = buf.split("<<", 1)
        rhs = rhs.split(";", 1)[0].replace("math::constants::pi<Precision>", "pi")
        try:
            vals = [eval(e, {"__builtins__": {}}, {"pi": math.pi}) for e in rhs.split(",") if e.strip()]
        except NameError:      # copies of other vectors (e.g. ``Input2 << Input1;``)
            i += 1
            continue
        out.append((start + 1, i + 1, lhs.strip(), [float(v) for v in vals]))
        i += 1
    return out


def test_ranges(rel):
    """name -> (first_line, last_line) of each active TEST_F."""
    lines = (REF / rel).read_text(encoding="latin-1").splitlines()
    heads = [(k + 1, re.search(r"TEST_F\(\s*\w+\s*,\s*(\w+)\)", l).group(1))
             for k, l in enumerate(lines) if l.startswith("TEST_F(")]
    rng = {}
    for j, (ln, name) in enumerate(heads):
        end = heads[j + 1][0] - 1 if j + 1 < len(heads) else len(lines)
        rng[name] = (ln, end)
    return rng


def check_after(rel, line):
    """The gtest assertion that follows an ``Expected <<`` statement: ('norm', bound) for
    ``(Output-Expected).norm() < bound`` or ('approx', prec) for ``isApprox(Expected[, prec])``."""
    lines = (REF / rel).read_text(encoding="latin-1").splitlines()
    for l in lines[line:line + 8]:
        m = re.search(r"\(Output\s*-\s*Expected\)\.norm\(\)\s*<\s*([0-9.Ee+-]+)", l)
        if m:
            return "norm", float(m.group(1))
        m = re.search(r"isApprox\(Expected(?:\s*,\s*([0-9.Ee+-]+))?\)", l)
        if m:
            return "approx", float(m.group(1)) if m.group(1) else EIGEN_PREC
    raise ValueError(f"no check after {rel}:{line}")


def in_test(stmts, rng):
    return [s for s in stmts if rng[0] <= s[0] <= rng[1]]


def unit(v):
    n = math.sqrt(sum(x * x for x in v))
    return [x / n for x in v]


def cite(rel, a, b):
    return f"Tests/{rel}:{a}-{b}" if a != b else f"Tests/{rel}:{a}"


cases = []
EIGEN_PREC = 1e-12   # Eigen isApprox default precision for double


def add(kind, cite_, **kw):
    kw.setdefault("rtol", EIGEN_PREC)
    cases.append(dict(kind=kind, cite=cite_, **kw))


# ---------------------------------------------------------------- Neel
NEEL_FIXTURE = dict(T=295.0, alpha=0.1, gamma=1.76e11, Ms=4.77464e5, radius=10e-9, K=1e4,
                    euler=[0.0, 0.0, 0.0], cite="Tests/Neel_Problem_Test/Test_Problem_Neel.h:58-78")
rel = "Neel_Problem_Test/Test_Problem_Neel.cpp"
st, tr = statements(rel), test_ranges(rel)
for test, field in (("DeterministicVectorTestWithoutField", False), ("DeterministicVectorTestWithField", True)):
    s = in_test(st, tr[test])
    ys = [x for x in s if x[2].startswith("testy")]
    rs = [x for x in s if x[2].startswith("testres")]
    h = unit(ys[0][3]) if field else [0.0, 0.0, 0.0]
    for k, (y, r) in enumerate(zip(ys, rs)):
        add("neel_det", cite(rel, y[0], r[1]), y=unit(y[3]), h=h, expected=r[3],
            rtol=1e-6 if k == 4 else EIGEN_PREC)
s = in_test(st, tr["StochasticMatrixTest"])
add("neel_sto", cite(rel, s[0][0], s[1][1]), y=unit(s[0][3]), expected=s[1][3])
s = in_test(st, tr["DriftTest"])
add("neel_drift", cite(rel, s[0][0], s[1][1]), y=unit(s[0][3]), expected=s[1][3])
s = in_test(st, tr["AllParts"])
d = {x[2]: x for x in s}
y1 = unit(d["Input1"][3])
add("neel_det", cite(rel, d["ExpectedDet"][0], d["ExpectedDet"][1]), y=y1, h=y1, expected=d["ExpectedDet"][3])
add("neel_sto", cite(rel, d["ExpectedSto"][0], d["ExpectedSto"][1]), y=y1, expected=d["ExpectedSto"][3])
add("neel_jacdet_reference", cite(rel, d["ExpectedJacDet"][0], d["ExpectedJacDet"][1]), y=y1, h=y1, dt=1e-6,
    expected=d["ExpectedJacDet"][3])
add("neel_jacsto_reference", cite(rel, d["ExpectedJacSto"][0], d["ExpectedJacSto"][1]), y=y1, dW=y1,
    expected=d["ExpectedJacSto"][3], rtol=2.220446049250313e-16 * 100)

# ---------------------------------------------------------------- uniaxial anisotropy
rel = "Neel_Problem_Test/Test_Uniaxial_Anisotropy.cpp"
st, tr = statements(rel), test_ranges(rel)
s = in_test(st, tr["AnisotropyFieldJacobi"])
d = {x[2]: x for x in s}
add("uniaxial_jacobian", cite(rel, d["TestVector"][0], d["expectedval"][1]),
    props=dict(Ms=477464.0, K=1e4, radius=10e-9, cite="Tests/Neel_Problem_Test/Test_Uniaxial_Anisotropy.h:41-50"),
    m=unit(d["TestVector"][3]), axis=unit(d["EasyAxis"][3]), expected=d["expectedval"][3])

# ---------------------------------------------------------------- NeelSpherical
SPH_FIXTURE = dict(NEEL_FIXTURE, rotate=True, min_angle=math.pi / 6.0,
                   cite="Tests/Neel_Spherical_Test/Test_Neel_Spherical_Problem.h:66-103")
rel = "Neel_Spherical_Test/Test_Neel_Spherical_Problem.cpp"
st, tr = statements(rel), test_ranges(rel)
lines = (REF / rel).read_text(encoding="latin-1").splitlines()


def rotated_flags(rng):
    flags = []
    for l in lines[rng[0] - 1:rng[1]]:
        if "EXPECT_TRUE(isRotated)" in l or "EXPECT_TRUE(!isRotated)" in l or "EXPECT_FALSE(isRotated)" in l:
            flags.append("EXPECT_TRUE(isRotated)" in l)
    return flags


sph_abs_tol = {  # cases the reference checks with an absolute norm instead of isApprox
    ("DeterministicVectorTestWithoutField", 2): 1e-6, ("DeterministicVectorTestWithoutField", 3): 1e-5,
    ("DriftTest", 2): 1e-6, ("DriftTest", 3): 2e-6,
}
for test, field in (("DeterministicVectorTestWithoutField", None), ("DeterministicVectorTestWithField", "Testfield"),
                    ("DriftTest", None)):
    s = in_test(st, tr[test])
    h = [0.0, 0.0, 0.0]
    if field:
        h = [x for x in s if x[2] == field][0][3]
    ys = [x for x in s if x[2].startswith("testy")]
    rs = [x for x in s if x[2].startswith("testres")]
    flags = rotated_flags(tr[test])
    for k, (y, r) in enumerate(zip(ys, rs)):
        c = dict(y=y[3], h=h, expected=r[3])
        if test == "DriftTest":
            c["rotated"] = flags[k - 1] if k > 0 else None   # first drift case has no isRotated check
        else:
            c["rotated"] = flags[k]
        if (test, k) in sph_abs_tol:
            c["atol_norm"] = sph_abs_tol[(test, k)]
        add("spherical_drift" if test == "DriftTest" else "spherical_det", cite(rel, y[0], r[1]), **c)
s = in_test(st, tr["StochasticMatrixTest"])
d = {x[2]: x for x in s}
add("spherical_sto", cite(rel, d["TestInput1"][0], d["Expected1"][1]), y=d["TestInput1"][3],
    expected=d["Expected1"][3], rotated=False)
add("spherical_sto", cite(rel, d["TestInput2"][0], d["Expected2"][1]), y=d["TestInput2"][3],
    expected=d["Expected2"][3], rotated=True)
s = in_test(st, tr["finishCalculations"])
for k, y in enumerate([x for x in s if x[2].startswith("testy")]):
    add("spherical_finish_roundtrip", cite(rel, y[0], y[1]), y=y[3], theta_only=k in (2, 3))
s = in_test(st, tr["CheckRotationFunctions"])
for y in [x for x in s if x[2].startswith("testvec")]:
    add("spherical_rotation_roundtrip", cite(rel, y[0], y[1]), y=y[3])

# ---------------------------------------------------------------- BrownAndNeel
BN_FIXTURE = dict(T=295.0, alpha=0.1, gamma=1.76e11, Ms=4.77464e5, radius=10e-9, K=-1e4, r_h=20e-9, viscosity=1e-3,
                  cite="Tests/BrownAndNeel_Relaxation_Test/Test_BrownAndNeel_Relaxation.h:64-93")
rel = "BrownAndNeel_Relaxation_Test/Test_BrownAndNeel_Relaxation.cpp"
st, tr = statements(rel), test_ranges(rel)
for test, h in (("DeterministicVectorTestWithoutField", [0.0, 0.0, 0.0]),
                ("DeterministicVectorTestWithField", [-0.123, 0.523, -0.872])):
    s = in_test(st, tr[test])
    ins = [x for x in s if x[2] == "Input"]
    exs = [x for x in s if x[2] == "Expected"]
    for y, r in zip(ins, exs):
        c = dict(y=y[3], h=h, expected=r[3])
        kind, v = check_after(rel, r[1])
        if kind == "norm":
            c["atol_norm"] = v
        else:
            c["rtol"] = v
        add("bn_det", cite(rel, y[0], r[1]), **c)
s = in_test(st, tr["StochasticMatrixTest"])
y = [x for x in s if x[2] == "Input"][0]
exs = [x for x in s if x[2] == "Expected"]
add("bn_sto", cite(rel, y[0], exs[0][1]), y=y[3], simple=False, expected=exs[0][3])
add("bn_sto", cite(rel, y[0], exs[1][1]), y=y[3], simple=True, expected=exs[1][3])
s = in_test(st, tr["DriftTest"])
y = [x for x in s if x[2] == "Input"][0]
exs = [x for x in s if x[2] == "Expected"]
add("bn_drift", cite(rel, y[0], exs[0][1]), y=y[3], simple=False, expected=exs[0][3], rtol=1e-6)
add("bn_drift", cite(rel, y[0], exs[1][1]), y=y[3], simple=True, expected=exs[1][3], rtol=1e-6)

# ---------------------------------------------------------------- BrownAndNeelEulerSpherical
BNES_FIXTURE = dict(T=295.0, alpha=0.1, gamma=1.76e11, Ms=4.77464e5, radius=10e-9, K=-1e4, r_h=20e-9, viscosity=1e-3,
                    brown_min_angle=math.pi / 6.0, neel_min_angle=math.pi / 6.0,
                    cite="Tests/NeelSpherical_BrownEuler_Relaxation_Test/Test_NeelSpherical_BrownEuler_Problem.h:68-107")
rel = "NeelSpherical_BrownEuler_Relaxation_Test/Test_NeelSpherical_BrownEuler_Problem.cpp"
st, tr = statements(rel), test_ranges(rel)
lines = (REF / rel).read_text(encoding="latin-1").splitlines()
for test, (lo, hi) in tr.items():
    body = lines[lo - 1:hi]
    s = in_test(st, (lo, hi))
    state = [x for x in s if x[2] == "state"]
    if not state:
        continue
    flags = {}
    for l in body:
        m = re.search(r"EXPECT_(TRUE|FALSE)\((Brown|Neel)Cache\.isRotated\)", l)
        if m:
            flags[m.group(2).lower()] = m.group(1) == "TRUE"
    prec = EIGEN_PREC
    for l in body:
        m = re.search(r"isApprox\(expectedresult(?:\s*,\s*([0-9.Ee+-]+))?\)", l)
        if m and m.group(1):
            prec = float(m.group(1))
    c = dict(y=state[0][3], brown_rotated=flags["brown"], neel_rotated=flags["neel"], rtol=prec)
    exp = [x for x in s if "expectedresult" in x[2]]
    if test.startswith("Deterministic"):
        h = [0.0, 0.0, 0.0]
        if "WithoutField" not in test:
            m = [re.search(r"Testfield\(([^)]*)\)", l) for l in body]
            h = [float(v) for v in [x for x in m if x][0].group(1).split(",")]
        c.update(h=h, expected=exp[0][3])
        kind = "bnes_det"
    elif test.startswith("StochasticMatrix"):
        c.update(expected=exp[0][3])
        kind = "bnes_sto"
    elif test.startswith("Drift"):
        c.update(expected=exp[0][3])
        kind = "bnes_drift"
    else:   # NeelRotation, BrownRot, BothRotated: prepareCalculations, then finishCalculations back
        if exp:
            expected = exp[0][3]
        else:  # ``Vec5D expectedresult(state); expectedresult(i) = v;``
            expected = list(state[0][3])
            for l in body:
                m = re.search(r"expectedresult\((\d)\)\s*=\s*([-0-9.Ee+]+)\s*;", l)
                if m:
                    expected[int(m.group(1))] = float(m.group(2))
        c.update(expected=expected)
        kind = "bnes_rotation"
    add(kind, cite(rel, lo, hi), **c)

# ---------------------------------------------------------------- Newton (Implicit_Solver_Test)
roots = {
    1: ([7.0, 1.0], [17.0 / 6.0, 37.0 / 12.0], "Implicit_Solver_Test/Test_Function_1.h:42-50"),
    2: ([4.171693661908896399197322740218308796100, -3.067844841059378634654464094551742447729],
        [-3.177905973754653339504659865520039419155, 3.104580387803790513959254734612565911544],
        "Implicit_Solver_Test/Test_Function_2.h:43-51"),
    3: ([24.94264094051703985939357751254743582237, 2.748175016536519576254822119664834650189, -14.29609896775523355019361832067594174884],
        [15.18760735359791398897973420200247195760, -1.952310380253357906329267841085196680235, -8.59580163271701364049493070256259835979],
        "Implicit_Solver_Test/Test_Function_3.h:50-57"),
}
for fn in (1, 2, 3):
    rel = f"Implicit_Solver_Test/Test_Function_{fn}.cpp"
    st, tr = statements(rel), test_ranges(rel)
    for test in ("SolverTest1", "SolverTest2", "SolverTest3"):
        if test not in tr:
            continue
        g = [x for x in in_test(st, tr[test]) if x[2] == "InitGuess"][0]
        add("newton", cite(rel, tr[test][0], tr[test][1]), fn=fn, x0=g[3], roots=roots[fn][:2], roots_cite=roots[fn][2],
            tol=2.220446049250313e-16 * 1000, max_iter=1000,
            # TestFunction3.SolverTest1|2 are excluded from the reference CI (ci.linux.yml:151)
            known_failing=(fn == 3 and test in ("SolverTest1", "SolverTest2")))

doc = dict(
    source="Neumann-A/StochasticPhysics Tests/ (gtest, values computed in Mathematica per the test comments)",
    generator="tests/golden/make_reference_vectors.py",
    eigen_isApprox="|a-b| <= rtol * min(|a|, |b|) (Frobenius norms)",
    fixtures=dict(neel=NEEL_FIXTURE, spherical=SPH_FIXTURE, brown_and_neel=BN_FIXTURE,
                  brown_and_neel_euler_spherical=BNES_FIXTURE),
    cases=cases,
)
OUT.write_text(json.dumps(doc, indent=1))
print(f"wrote {len(cases)} cases to {OUT}")
