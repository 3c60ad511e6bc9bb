"""Numerical radial functions of atom files evaluated on the device against
the oracle (SphericalDataNumerical restated in spherical coordinates with
scipy's clamped spline): values and gradients on the reference's H and Be
files and on a synthetic p-shell file."""
import os

import numpy as np
import pytest

import atomdata as A
from test_atom import GOLD, write_atom

pytestmark = pytest.mark.gpu


def _points(rng, origin, rmax, r0):
    pts = origin + rng.uniform(-rmax, rmax, (3000, 3))
    special = [origin,                                  # the nucleus
               origin + [0, 0, 0.7], origin + [0, 0, -1.1],  # on the polar axis
               origin + [r0 * 0.3, 0, 0],               # below the first grid point
               origin + [rmax * 3, 0, 0]]               # past every support
    return np.vstack([pts, special])


@pytest.mark.parametrize("fname", ["H.xml", "Be1s2s.xml"])
def test_reference_atom_functions_match_oracle(ctx, fname):
    import dfx
    path = os.path.join(GOLD, fname)
    a = dfx.AtomSphericalDataEnrichment(path, ["orbital", "vnuclear"])
    o = A.read_atom(path, ["orbital", "vnuclear"])
    origin = np.array([0.3, -0.4, 1.25])
    rng = np.random.default_rng(5)
    for field in ("orbital", "vnuclear"):
        for fun, ofun in zip(a.getSphericalData(field), o["fields"][field]):
            rs = ofun["cutoff"] * (1 + 1 / ofun["smoothness"])
            pts = _points(rng, origin, rs * 0.8, o["r"][0] if o["r"][0] > 0 else o["r"][1])
            v, g = fun.evaluate(ctx, pts, origin)
            vw, gw = A.evaluate(ofun, o["r"], origin, pts)
            sv = max(np.abs(vw).max(), 1e-300)
            sg = max(np.abs(gw).max(), 1e-300)
            print(f"{fname} {field} {fun.qNumbers}: max |dv| {np.abs(v - vw).max() / sv:.2e} "
                  f"max |dg| {np.abs(g - gw).max() / sg:.2e}", flush=True)
            assert np.abs(v - vw).max() <= 1e-12 * sv
            # gradient: the oracle's spherical form loses a few digits near the
            # nucleus (R / r terms)
            assert np.abs(g - gw).max() <= 1e-10 * sg
            assert v[-1] == 0.0 and not g[-1].any()   # past the support
            assert not g[-5].any()                    # at the nucleus


def test_p_shell_matches_oracle(ctx, tmp_path):
    import dfx
    r = np.linspace(0.0, 8.0, 801)
    R = r * np.exp(-1.3 * r)
    path = write_atom(tmp_path / "p.xml", r, [([2, 1, m], (3.0, 0.6), R) for m in (-1, 0, 1)])
    a = dfx.AtomSphericalDataEnrichment(path, ["orbital"])
    o = A.read_atom(path, ["orbital"])
    origin = np.array([-0.5, 0.25, 0.0])
    pts = _points(np.random.default_rng(9), origin, 4.0, 0.01)
    for fun, ofun in zip(a.getSphericalData("orbital"), o["fields"]["orbital"]):
        v, g = fun.evaluate(ctx, pts, origin)
        vw, gw = A.evaluate(ofun, o["r"], origin, pts)
        assert np.abs(v - vw).max() <= 1e-12 * np.abs(vw).max()
        assert np.abs(g - gw).max() <= 1e-10 * np.abs(gw).max()


def test_d_shell_is_refused(ctx, tmp_path):
    import dfx
    r = np.linspace(0.0, 4.0, 41)
    path = write_atom(tmp_path / "d.xml", r, [([3, 2, m], (2.0, 0.6), np.exp(-r))
                                              for m in range(-2, 3)])
    a = dfx.AtomSphericalDataEnrichment(path, ["orbital"])
    with pytest.raises(dfx.DfxError, match="l > 1"):
        a.getSphericalData("orbital")[0].evaluate(ctx, np.zeros((4, 3)))
