"""Atom data files (the reference's enrichment XML format) and their radial
splines: the library's host-side reader and spline against the oracle
(xml.etree + scipy's clamped CubicSpline) on the reference's own atom files
(tests/golden/atoms: H.xml and Be1s2s.xml of test/ksdft/src), the reader's
error checks, and pins of the oracle's evaluation.  CPU only."""
import math
import os

import numpy as np
import pytest

import atomdata as A

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "atoms")
FILES = {"H.xml": ["orbital", "vnuclear"],
         "Be1s2s.xml": ["orbital", "vnuclear", "density", "vhartree", "vtotal"]}


def write_atom(path, r, funcs, field="orbital", NR=None, ns=A.NS, extra=""):
    """A synthetic atom file: funcs = [(q, (cutoff, smoothness), values)]."""
    body = "".join(
        f"<{field}><qNumbers>{' '.join(map(str, q))}</qNumbers>"
        f"<values>{' '.join(repr(float(v)) for v in vals)}</values>"
        f"<cutoffInfo>{co[0]!r} {co[1]!r}</cutoffInfo></{field}>\n"
        for q, co, vals in funcs)
    NR = len(r) if NR is None else NR
    path.write_text(f'<?xml version="1.0" ?>\n<atom xmlns="{ns}">\n<symbol>X</symbol><Z>3</Z>'
                    f"<charge>0.0</charge><NR>{NR}</NR>\n"
                    f"<r>{' '.join(repr(float(v)) for v in r)}</r>\n{body}{extra}</atom>\n")
    return str(path)


def piecewise(r_grid, fun, x):
    """The library's spline from its coefficient table (header semantics)."""
    x = np.asarray(x, float)
    i = np.clip(np.searchsorted(r_grid, x, side="right") - 1, 0, len(r_grid) - 1)
    h = x - r_grid[i]
    y, b, c, d = fun.values[i], fun.spline_b[i], fun.spline_c[i], fun.spline_d[i]
    v = ((d * h + c) * h + b) * h + y
    v = np.where(x < r_grid[0], fun.values[0] + fun.spline_b[0] * (x - r_grid[0]), v)
    last = x > r_grid[-1]
    hl = x - r_grid[-1]
    return np.where(last, (fun.spline_c[-1] * hl + fun.spline_b[-1]) * hl + fun.values[-1], v)


@pytest.mark.parametrize("fname", sorted(FILES))
def test_reader_matches_oracle_bitwise(fname):
    import dfx
    path = os.path.join(GOLD, fname)
    fields = FILES[fname]
    a = dfx.AtomSphericalDataEnrichment(path, fields, ["symbol", "Z", "charge", "NR"])
    o = A.read_atom(path, fields)
    for k in ("symbol", "Z", "charge", "NR"):
        assert a.getMetadata(k) == o["meta"][k]
    assert np.array_equal(a.radialGrid(), o["r"])
    for f in fields:
        got = a.getSphericalData(f)
        ref = o["fields"][f]
        assert len(got) == len(ref)
        for g, e in zip(got, ref):
            assert g.qNumbers == e["q"]
            assert (g.cutoff, g.smoothness) == (e["cutoff"], e["smoothness"])
            assert np.array_equal(g.values, e["values"])


def test_reference_files_hold_what_their_tests_use():
    """H: one 1s orbital (cutoff 2, smoothness 0.6) on 1166 points 0.01 apart;
    Be: 1s and 2s orbitals on a 3000-point graded grid."""
    import dfx
    h = dfx.AtomSphericalDataEnrichment(os.path.join(GOLD, "H.xml"), ["orbital"],
                                        ["symbol", "Z"])
    assert (h.getMetadata("symbol"), h.getMetadata("Z"), h.NR) == ("H", "1", 1166)
    orb = h.getSphericalData("orbital")
    assert [f.qNumbers for f in orb] == [[1, 0, 0]]
    assert (orb[0].cutoff, orb[0].smoothness) == (2.0, 0.6)
    be = dfx.AtomSphericalDataEnrichment(os.path.join(GOLD, "Be1s2s.xml"), ["orbital"])
    assert [f.qNumbers for f in be.getSphericalData("orbital")] == [[1, 0, 0], [2, 0, 0]]
    assert be.getSphericalData("orbital", [2, 0, 0]).k == 1


@pytest.mark.parametrize("fname", sorted(FILES))
def test_spline_matches_clamped_cubic_spline(fname):
    """The library's coefficients against scipy's clamped CubicSpline (with the
    reference's extrapolation) at the knots, between them and past both ends."""
    import dfx
    path = os.path.join(GOLD, fname)
    a = dfx.AtomSphericalDataEnrichment(path, ["orbital"])
    r = a.radialGrid()
    rng = np.random.default_rng(7)
    x = np.concatenate([r, 0.5 * (r[1:] + r[:-1]), rng.uniform(r[0], r[-1], 4000),
                        [r[0] * 0.5, 0.0, r[-1] + 0.3, r[-1] + 2.0]])
    for fun in a.getSphericalData("orbital"):
        S = A.Spline(r, fun.values)
        want = S(x)
        got = piecewise(r, fun, x)
        scale = np.abs(fun.values).max()
        assert np.abs(got - want).max() <= 1e-12 * scale


def test_spline_pins(tmp_path):
    """Closed forms the spline must reproduce: a cubic with zero end slopes
    (2t^3 - 3t^2 on [0, 1]) exactly on a graded grid; constants exactly;
    C^1 / C^2 continuity at interior knots; zero end slopes."""
    import dfx
    r = np.sort(np.concatenate([[0.0, 1.0], np.random.default_rng(3).uniform(0, 1, 37)]))
    p = 2 * r ** 3 - 3 * r ** 2
    path = write_atom(tmp_path / "cubic.xml", r, [([1, 0, 0], (0.8, 0.6), p),
                                                 ([2, 0, 0], (0.8, 0.6), np.full(r.size, 1.5))])
    a = dfx.AtomSphericalDataEnrichment(path, ["orbital"])
    f, g = a.getSphericalData("orbital")
    x = np.linspace(0, 1, 1001)
    assert np.abs(piecewise(r, f, x) - (2 * x ** 3 - 3 * x ** 2)).max() < 1e-13
    assert np.abs(A.Spline(r, p)(x) - (2 * x ** 3 - 3 * x ** 2)).max() < 1e-13
    assert np.abs(piecewise(r, g, x) - 1.5).max() == 0.0
    h = np.diff(r)
    # continuity of value, slope and curvature across every interior knot
    v_end = ((f.spline_d[:-1] * h + f.spline_c[:-1]) * h + f.spline_b[:-1]) * h + f.values[:-1]
    s_end = (3 * f.spline_d[:-1] * h + 2 * f.spline_c[:-1]) * h + f.spline_b[:-1]
    c_end = 3 * f.spline_d[:-1] * h + f.spline_c[:-1]
    assert np.abs(v_end - f.values[1:]).max() < 1e-14
    assert np.abs(s_end - f.spline_b[1:]).max() < 1e-12
    assert np.abs(c_end[:-1] - f.spline_c[1:-1]).max() < 1e-10
    assert abs(f.spline_b[0]) < 1e-13 and abs(f.spline_b[-1]) < 1e-12


BAD = {
    "nr_mismatch": (dict(NR=5), "Mismatch in number of radial points"),
    "nr_zero": (dict(NR=0), "Non-positive integer"),
    "m_out_of_range": (dict(q=[2, 1, 2]), "outside of -l and +l"),
    "repeated": (dict(repeat=True), "repeated quantum numbers"),
    "incomplete_shell": (dict(q=[2, 1, 0]), "ascending m"),
    "wrong_namespace": (dict(ns="http://example.org/other"), "/dft-efe:atom not found"),
    "values_count": (dict(short=True), "Mismatch in number of points"),
    "missing_field": (dict(field="density"), "dft-efe:orbital"),
}


@pytest.mark.parametrize("case", sorted(BAD))
def test_reader_errors(tmp_path, case):
    import dfx
    kw, msg = BAD[case]
    r = np.linspace(0, 4, 9)
    vals = np.exp(-r)
    q = kw.get("q", [1, 0, 0])
    funcs = [(q, (2.0, 0.6), vals[:5] if kw.get("short") else vals)]
    if kw.get("repeat"):
        funcs.append((q, (2.0, 0.6), vals))
    path = write_atom(tmp_path / "bad.xml", r, funcs, field=kw.get("field", "orbital"),
                      NR=kw.get("NR"), ns=kw.get("ns", A.NS))
    with pytest.raises(dfx.DfxError) as e:
        dfx.AtomSphericalDataEnrichment(path, ["orbital"])
    assert msg in str(e.value)
    with pytest.raises(A.OracleAtomError):
        A.read_atom(path, ["orbital"])


def test_reader_rejects_malformed_xml(tmp_path):
    import dfx
    p = tmp_path / "broken.xml"
    p.write_text(f'<atom xmlns="{A.NS}"><NR>2</NR><r>0 1</r>')
    with pytest.raises(dfx.DfxError, match="Error parsing the XML file"):
        dfx.AtomSphericalDataEnrichment(str(p), [])


# ------------------------------------------------------------ oracle pins
def _synthetic_p(tmp_path):
    r = np.linspace(0.0, 8.0, 801)
    R = r * np.exp(-1.3 * r)
    funcs = [([1, 0, 0], (2.5, 0.7), np.exp(-r))] + \
        [([2, 1, m], (3.0, 0.6), R) for m in (-1, 0, 1)]
    return write_atom(tmp_path / "p.xml", r, funcs)


def test_oracle_gradient_matches_finite_differences(tmp_path):
    o = A.read_atom(_synthetic_p(tmp_path), ["orbital"])
    rng = np.random.default_rng(11)
    pts = rng.uniform(-3, 3, (200, 3))
    origin = np.array([0.1, -0.2, 0.05])
    eps = 1e-6
    for fun in o["fields"]["orbital"]:
        _, g = A.evaluate(fun, o["r"], origin, pts)
        fd = np.zeros_like(g)
        for k in range(3):
            e = np.zeros(3)
            e[k] = eps
            vp, _ = A.evaluate(fun, o["r"], origin, pts + e)
            vm, _ = A.evaluate(fun, o["r"], origin, pts - e)
            fd[:, k] = (vp - vm) / (2 * eps)
        assert np.abs(g - fd).max() < 1e-6 * max(1.0, np.abs(g).max())


def test_oracle_spherical_harmonics_pins(tmp_path):
    """Y_lm normalized on the sphere; Y_1m proportional to y, z, x for
    m = -1, 0, 1 (real form, no Condon-Shortley phase)."""
    o = A.read_atom(_synthetic_p(tmp_path), ["orbital"])
    xg, wg = np.polynomial.legendre.leggauss(24)
    ph = (np.arange(48) + 0.5) * 2 * math.pi / 48
    ct, P = np.meshgrid(xg, ph, indexing="ij")
    W = np.meshgrid(wg, np.full(48, 2 * math.pi / 48), indexing="ij")
    w = (W[0] * W[1]).ravel()
    st = np.sqrt(1 - ct ** 2)
    dirs = np.stack([st * np.cos(P), st * np.sin(P), ct], axis=-1).reshape(-1, 3)
    rr = 1.0  # inside every cutoff: f = R(1) f_c(1) Y
    for fun, axis in zip(o["fields"]["orbital"], [None, 1, 2, 0]):
        S = A.Spline(o["r"], fun["values"])
        fc, _ = A.cutoff(np.array([rr]), fun["cutoff"], fun["smoothness"])
        v, _ = A.evaluate(fun, o["r"], np.zeros(3), rr * dirs)
        Y = v / (S(np.array([rr]))[0] * fc[0])
        assert abs((w * Y * Y).sum() - 1.0) < 1e-12
        if axis is not None:
            assert np.abs(Y - math.sqrt(3 / (4 * math.pi)) * dirs[:, axis]).max() < 1e-13
        else:
            assert np.abs(Y - 0.5 / math.sqrt(math.pi)).max() < 1e-15


def test_oracle_d_shell_pins(tmp_path):
    """d shell: normalized on the sphere, Y_2m = k (xy, yz, 2z^2 - x^2 - y^2,
    xz, x^2 - y^2) / r^2 for m = -2 .. 2, gradients by finite differences."""
    r = np.linspace(0.0, 8.0, 801)
    path = write_atom(tmp_path / "d.xml", r, [([3, 2, m], (3.0, 0.6), r * r * np.exp(-r))
                                              for m in range(-2, 3)])
    o = A.read_atom(path, ["orbital"])
    xg, wg = np.polynomial.legendre.leggauss(24)
    ph = (np.arange(48) + 0.5) * 2 * math.pi / 48
    ct, P = np.meshgrid(xg, ph, indexing="ij")
    W = np.meshgrid(wg, np.full(48, 2 * math.pi / 48), indexing="ij")
    w = (W[0] * W[1]).ravel()
    st = np.sqrt(1 - ct ** 2)
    d = np.stack([st * np.cos(P), st * np.sin(P), ct], axis=-1).reshape(-1, 3)
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    want = [math.sqrt(15 / (4 * math.pi)) * x * y, math.sqrt(15 / (4 * math.pi)) * y * z,
            math.sqrt(5 / (16 * math.pi)) * (2 * z * z - x * x - y * y),
            math.sqrt(15 / (4 * math.pi)) * x * z, math.sqrt(15 / (16 * math.pi)) * (x * x - y * y)]
    rng = np.random.default_rng(4)
    pts = rng.uniform(-2, 2, (100, 3))
    for fun, Yw in zip(o["fields"]["orbital"], want):
        S = A.Spline(o["r"], fun["values"])
        v, _ = A.evaluate(fun, o["r"], np.zeros(3), d)
        Y = v / (S(np.array([1.0]))[0] * A.cutoff(np.array([1.0]), 3.0, 0.6)[0][0])
        assert abs((w * Y * Y).sum() - 1.0) < 1e-12
        assert np.abs(Y - Yw).max() < 1e-13
        _, g = A.evaluate(fun, o["r"], np.zeros(3), pts)
        fd = np.zeros_like(g)
        for k in range(3):
            e = np.zeros(3)
            e[k] = 1e-6
            fd[:, k] = (A.evaluate(fun, o["r"], np.zeros(3), pts + e)[0] -
                        A.evaluate(fun, o["r"], np.zeros(3), pts - e)[0]) / 2e-6
        assert np.abs(g - fd).max() < 1e-6 * max(1.0, np.abs(g).max())


def test_reader_bounds_nesting(tmp_path):
    import dfx
    p = tmp_path / "deep.xml"
    p.write_text(f'<atom xmlns="{A.NS}">' + "<a>" * 100 + "</a>" * 100 + "</atom>")
    with pytest.raises(dfx.DfxError, match="nested more than 64"):
        dfx.AtomSphericalDataEnrichment(str(p), [])
