"""Pseudopotential files (UPF v2): the library's reader against the oracle
(xml.etree + scipy) on the reference's own Cu.upf (test/atoms/src), what the
physics fixes (the atomic valence density integrates to z_valence, the local
potential approaches -z_valence / r), the spline's end conditions and the
reader's checks.  CPU only."""
import math
import os

import numpy as np
import pytest

import atomdata as A

UPF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "upf", "Cu.upf")
FIELDS = ["vlocal", "beta", "pswfc", "nlcc", "rhoatom"]
META = ["element", "z_valence", "l_max", "number_of_wfc", "number_of_proj", "core_correction",
        "dij"]


@pytest.fixture(scope="module")
def cu():
    import dfx
    return dfx.AtomSphericalDataPSP(UPF, FIELDS, META), A.read_upf(UPF, FIELDS)


def test_reader_matches_oracle(cu):
    a, o = cu
    for k in META[:-1]:
        assert a.getMetadata(k) == o["meta"][k]
    assert np.array_equal(a.radialGrid(), o["r"])
    for f in FIELDS:
        got, ref = a.getSphericalData(f), o["fields"][f]
        assert [g.qNumbers for g in got] == [e["q"] for e in ref]
        for g, e in zip(got, ref):
            assert g.cutoff == e["cutoff"] and g.smoothness == e["smoothness"]
            assert g.mixed == e["mixed"]
            assert np.array_equal(g.values, e["values"])
            if e["mixed"]:
                assert np.array_equal(g.knots, e["knots"])
                assert g.tail == e["tail"]
    dij = np.array([float(t) for t in a.getMetadata("dij").split()])
    import xml.etree.ElementTree as ET
    raw = ET.parse(UPF).getroot().find("PP_NONLOCAL/PP_DIJ").text.split()
    assert dij.size == int(a.getMetadata("number_of_proj")) ** 2
    assert np.abs(dij - 0.5 * np.array([float(t) for t in raw])).max() < 1e-15


def test_reference_structure(cu):
    """Cu: 19 valence electrons; six projectors (two per l = 0, 1, 2) give
    2 (1 + 3 + 5) = 18 beta functions; pseudo-wavefunctions 4S, 4P, 3D ... by
    their labels."""
    a, _ = cu
    assert (a.getMetadata("element"), float(a.getMetadata("z_valence"))) == ("Cu", 19.0)
    beta = a.getSphericalData("beta")
    assert len(beta) == 18
    assert sorted({(q[0], q[1]) for q in (b.qNumbers for b in beta)}) == \
        [(0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (1, 2)]


def test_valence_density_integrates_to_z_valence(cu):
    a, _ = cu
    r = a.radialGrid()
    rho = a.getSphericalData("rhoatom")[0].values * A.Y00
    assert abs(np.trapezoid(4 * math.pi * r * r * rho, r) - 19.0) < 1e-5


def test_local_potential_spline_and_tail(cu):
    a, _ = cu
    v = a.getSphericalData("vlocal")[0]
    x, y = v.knots, v.values
    assert x[-1] <= 10.0001 < a.radialGrid()[x.size]
    # end conditions: f''(x_0) = first divided difference, f'(x_last) = -y / x
    assert abs(2 * v.spline_c[0] - (y[1] - y[0]) / (x[1] - x[0])) < 1e-12 * abs(y).max()
    assert abs(v.spline_b[-1] + y[-1] / x[-1]) < 1e-12 * abs(y).max()
    assert v.cLeft == v.spline_c[0]
    # against scipy's spline with the same end conditions
    S = A.SplineBC(x, y, (y[1] - y[0]) / (x[1] - x[0]), -(y[-1] / x[-1]))
    t = np.concatenate([0.5 * (x[1:] + x[:-1]), x])
    i = np.clip(np.searchsorted(x, t, side="right") - 1, 0, x.size - 1)
    h = t - x[i]
    got = ((v.spline_d[i] * h + v.spline_c[i]) * h + v.spline_b[i]) * h + y[i]
    assert np.abs(got - S(t)).max() < 1e-12 * np.abs(y).max()
    # the potential approaches -z_valence / r: in Ha, times Y_00
    assert abs(y[-1] * A.Y00 * x[-1] + 19.0) < 2e-3
    assert v.tail * A.Y00 == pytest.approx(-19.0, rel=1e-15)


def _edit(tmp_path, old, new):
    s = open(UPF).read()
    assert old in s
    p = tmp_path / "edited.upf"
    p.write_text(s.replace(old, new, 1))
    return str(p)


@pytest.mark.parametrize("case", ["unknown_field", "lmax", "mesh_size", "zval"])
def test_reader_errors(tmp_path, case):
    import dfx
    if case == "unknown_field":
        path, fields, msg = UPF, ["vhartree"], "UPF_PARENT_PATH_LOOKUP"
    elif case == "lmax":
        path, fields, msg = _edit(tmp_path, 'l_max="2"', 'l_max="3"'), ["beta"], "lmax"
    elif case == "mesh_size":
        path, fields, msg = _edit(tmp_path, 'mesh_size="  1492"', 'mesh_size="  1491"'), [], \
            "Mismatch in number of radial points"
    else:
        path, fields, msg = _edit(tmp_path, 'z_valence="   19.00"', 'z_valence="   -1.00"'), [], \
            "z valance"
    with pytest.raises(dfx.DfxError, match=msg):
        dfx.AtomSphericalDataPSP(path, fields)
