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


def test_d_shell_matches_oracle_and_f_shell_is_refused(ctx, tmp_path):
    import dfx
    r = np.linspace(0.0, 8.0, 801)
    R = r * r * np.exp(-1.1 * r)
    path = write_atom(tmp_path / "d.xml", r, [([3, 2, m], (3.0, 0.6), R) for m in range(-2, 3)])
    a = dfx.AtomSphericalDataEnrichment(path, ["orbital"])
    o = A.read_atom(path, ["orbital"])
    origin = np.array([0.25, 0.0, -0.5])
    pts = _points(np.random.default_rng(21), origin, 4.0, 0.01)
    for fun, ofun in zip(a.getSphericalData("orbital"), o["fields"]["orbital"]):
        v, g = fun.evaluate(ctx, pts, origin)
        vw, gw = A.evaluate(ofun, o["r"], origin, pts)
        assert np.abs(v - vw).max() <= 1e-12 * np.abs(vw).max()
        assert np.abs(g - gw).max() <= 1e-10 * np.abs(gw).max()
    path = write_atom(tmp_path / "f.xml", r, [([4, 3, m], (2.0, 0.6), np.exp(-r))
                                              for m in range(-3, 4)])
    a = dfx.AtomSphericalDataEnrichment(path, ["orbital"])
    with pytest.raises(dfx.DfxError, match="l > 2"):
        a.getSphericalData("orbital")[0].evaluate(ctx, np.zeros((4, 3)))


# ------------------------------------------------ problem builder with atom files
def _slater_file(tmp_path, Z, rcut, smooth):
    """The analytic Slater set (1s: e^-Zr, 2s: r e^-Zr, 2p: r e^-Zr, m = -1, 0, 1)
    written as an atom file on a fine grid."""
    r = np.linspace(0.0, 16.0, 8001)
    e = np.exp(-Z * r)
    funcs = [([1, 0, 0], (rcut, smooth), e), ([2, 0, 0], (rcut, smooth), r * e)] + \
        [([2, 1, m], (rcut, smooth), r * e) for m in (-1, 0, 1)]
    return write_atom(tmp_path / f"slater{Z}.xml", r, funcs)


@pytest.mark.parametrize("orth", [False, True])
def test_atom_file_enrichment_reproduces_the_analytic_set(ctx, tmp_path, orth):
    """An atom file holding the analytic Slater functions gives the problem the
    oracle-pinned analytic builder gives: the same maps bit for bit and the
    same cell matrices up to the spline's interpolation error."""
    import dfx
    atoms = [(-1.5, 0.5, 0.2, 1), (2.0, -1.0, 0.7, 3)]
    kw = dict(p=2, n=(6, 6, 6), lo=(-12.,) * 3, hi=(12.,) * 3, atoms=atoms, enrich=True,
              nq_enr=6, rcut=4.0, smooth=2.0, orthogonalize=orth)
    files = {Z: dfx.AtomSphericalDataEnrichment(_slater_file(tmp_path, Z, 4.0, 2.0),
                                                ["orbital"]) for Z in (1, 3)}
    Pa = dfx.Problem(ctx, **kw)
    Pn = dfx.Problem(ctx, atom_data=[files[a[3]] for a in atoms], **kw)
    assert (Pn.nDofs, Pn.nEnriched) == (Pa.nDofs, Pa.nEnriched)
    assert Pn.nEnriched == 10
    for k in ("cellNdofs", "cellMap", "rows"):
        assert np.array_equal(Pn.get(k), Pa.get(k))
    for k in ("H", "M", "K"):
        a, b = Pa.get(k), Pn.get(k)
        err = np.abs(a - b).max() / np.abs(a).max()
        print(f"orth={orth} {k}: max rel diff {err:.2e}", flush=True)
        assert err < 1e-7
    if orth:
        a, b = Pa.get("alpha"), Pn.get("alpha")
        assert np.abs(a - b).max() < 1e-7 * np.abs(a).max()


def _free_dense(P):
    """Dense H and M over the free dofs (Dirichlet rows removed)."""
    cm, nd = P.get("cellMap"), P.get("cellNdofs")
    H, M = np.zeros((P.nDofs, P.nDofs)), np.zeros((P.nDofs, P.nDofs))
    hs, ms = P.get("H"), P.get("M")
    off = moff = 0
    for n in nd:
        ids = cm[off:off + n]
        H[np.ix_(ids, ids)] += hs[moff:moff + n * n].reshape(n, n, order="F")
        M[np.ix_(ids, ids)] += ms[moff:moff + n * n].reshape(n, n, order="F")
        off += n
        moff += n * n
    free = np.setdiff1d(np.arange(P.nDofs), P.get("rows"))
    return H[np.ix_(free, free)], M[np.ix_(free, free)]


def test_hydrogen_file_enrichment_lowers_the_ground_state(ctx):
    """Rayleigh-Ritz: adding the H.xml 1s orbital (cutoff 2, smoothness 0.6) to
    a coarse classical space can only lower the lowest eigenvalue of the
    smeared-Coulomb Hamiltonian (every cell on the same 8-point rule, so both
    discretizations share the bilinear forms and the spaces are nested)."""
    import dfx
    import scipy.linalg as sl
    Hf = dfx.AtomSphericalDataEnrichment(os.path.join(GOLD, "H.xml"), ["orbital"])
    kw = dict(p=2, n=(5, 5, 5), lo=(-10.,) * 3, hi=(10.,) * 3, atoms=[(0.1, -0.2, 0.3, 1)],
              v_smear=0.2, nq=8, nq_enr=8)
    Pc = dfx.Problem(ctx, **kw)
    Pe = dfx.Problem(ctx, enrich=True, atom_data=[Hf], **kw)
    assert Pe.nEnriched == 1 and Pe.nClassical == Pc.nClassical
    ev = []
    for P in (Pc, Pe):
        H, M = _free_dense(P)
        ev.append(sl.eigh(H, M, eigvals_only=True, subset_by_index=[0, 2]))
    print(f"lowest eigenvalues classical {ev[0]} enriched {ev[1]}", flush=True)
    assert np.all(ev[1] <= ev[0] + 1e-10)
    assert ev[0][0] - ev[1][0] > 0.01          # the orbital carries the cusp region
    assert -0.5 < ev[1][0] < -0.3              # above -1/2 Ha (smeared nucleus)


def test_upf_functions_match_oracle(ctx):
    """Cu.upf on the device: the mixed local potential (spline within 10 bohr,
    -z_valence / r beyond), the valence and core-correction densities and
    the s / p / d projectors and pseudo-wavefunctions against the oracle."""
    import dfx
    path = os.path.join(os.path.dirname(GOLD), "upf", "Cu.upf")
    fields = ["vlocal", "rhoatom", "nlcc", "beta", "pswfc"]
    a = dfx.AtomSphericalDataPSP(path, fields)
    o = A.read_upf(path, fields)
    origin = np.array([0.2, -0.1, 0.4])
    rng = np.random.default_rng(17)
    pts = np.vstack([origin + rng.uniform(-12, 12, (3000, 3)),
                     origin + rng.uniform(-2, 2, (1000, 3)),
                     [origin, origin + [0, 0, 0.005], origin + [10.0, 0, 0],
                      origin + [0, 10.5, 0]]])
    for f in fields:
        for fun, ofun in zip(a.getSphericalData(f), o["fields"][f]):
            v, g = fun.evaluate(ctx, pts, origin)
            vw, gw = A.evaluate_upf(ofun, o["r"], origin, pts)
            sv, sg = np.abs(vw).max(), np.abs(gw).max()
            print(f"{f} {fun.qNumbers}: |dv| {np.abs(v - vw).max() / sv:.2e} "
                  f"|dg| {np.abs(g - gw).max() / max(sg, 1e-300):.2e}", flush=True)
            assert np.abs(v - vw).max() <= 1e-12 * sv
            assert np.abs(g - gw).max() <= 1e-10 * max(sg, 1e-300)
    vl = a.getSphericalData("vlocal")[0]
    far = origin + np.array([[25.0, 0, 0], [0, -40.0, 0]])
    assert np.allclose(vl.evaluate(ctx, far, origin, gradient=False), -19.0 / np.array([25.0, 40.0]),
                       rtol=1e-15)


def test_pseudopotential_local_part_in_the_hamiltonian(ctx):
    """A Cu atom whose local potential comes from Cu.upf: beyond 10 bohr it is
    -z_valence / r, so every cell farther than that from the atom has the H of
    a point charge 19 (smeared Coulomb, erf = 1 there) while M and K do not
    depend on the potential at all; near the atom V_loc is finite and H
    differs."""
    import dfx
    upf = dfx.AtomSphericalDataPSP(os.path.join(os.path.dirname(GOLD), "upf", "Cu.upf"),
                                   ["vlocal"])
    kw = dict(p=2, n=(8, 8, 8), lo=(-16.,) * 3, hi=(16.,) * 3, atoms=[(0.3, -0.2, 0.1, 19)],
              nq=5)
    Pc = dfx.Problem(ctx, **kw)
    Pp = dfx.Problem(ctx, local_potential=[upf], **kw)
    for k in ("cellMap", "cellNdofs"):
        assert np.array_equal(Pc.get(k), Pp.get(k))
    assert np.array_equal(Pc.get("M"), Pp.get("M")) and np.array_equal(Pc.get("K"), Pp.get("K"))
    Hc, Hp = Pc.get("H"), Pp.get("H")
    n = Pc.get("cellNdofs")
    h = 4.0
    far = near = 0
    off = 0
    for c in range(Pc.nCells):
        i, j, k = c % 8, (c // 8) % 8, c // 64
        lo = np.array([-16 + h * i, -16 + h * j, -16 + h * k])
        d = np.maximum(0, np.maximum(lo - [0.3, -0.2, 0.1], [0.3, -0.2, 0.1] - (lo + h)))
        a, b = Hc[off:off + n[c] ** 2], Hp[off:off + n[c] ** 2]
        off += n[c] ** 2
        if np.sqrt((d ** 2).sum()) > 10.0001:
            far += 1
            assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
        elif np.sqrt((d ** 2).sum()) == 0.0:
            near += 1
            assert np.abs(a - b).max() > 1e-3 * np.abs(a).max()
    assert far > 100 and near == 1
