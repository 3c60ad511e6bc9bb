"""CPU restatement of the reference's atom data files and numerical radial
functions (numpy / scipy).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Follows:
  * AtomSphericalDataEnrichment (src/atoms/AtomSphericalDataEnrichment.cpp:
    511-640): XPath reads of /dft-efe:atom/dft-efe:<name> in the namespace
    http://www.dft-efe.com/dft-efe, per field the qNumbers / values /
    cutoffInfo children (:215-286), checks of :347-473;
  * utils::Spline with its defaults (src/utils/Spline.h:87-95: C^2 cubic,
    first-derivative boundary conditions of value 0 at both ends) and its
    evaluation outside the grid (Spline.cpp:490-518: left linear with the
    first slope, right quadratic);
  * SphericalDataNumerical::getValue / getGradientValue (src/atoms/
    SphericalDataNumerical.cpp:40-140): R(r) Y_lm f_c(r) in spherical
    coordinates, Y_lm = B_l|m| D_m P_l^|m|(cos theta) Q_m(phi) without the
    Condon-Shortley phase (SphericalHarmonicFunctions.cpp:15-70, 348-430),
    the gradient from d/dr, d/dtheta and the pole-safe d/dphi / sin(theta) of
    Bosch (2000), zero past cutoff (1 + 1/smoothness);
  * the smooth cutoff of SmoothCutoffFunctions.cpp:25-70.
The spline itself is scipy's clamped CubicSpline (the same unique C^2
interpolant), so the product's own tridiagonal solve is checked against an
independent implementation.
"""
import math
import xml.etree.ElementTree as ET

import numpy as np
from scipy.interpolate import CubicSpline

NS = "http://www.dft-efe.com/dft-efe"


class OracleAtomError(ValueError):
    pass


def read_atom(path, fields):
    """{'meta': {name: text}, 'r': array, 'fields': {name: [fun...]}} with
    fun = {'q': [n, l, m], 'cutoff', 'smoothness', 'values'} in file order."""
    try:
        root = ET.parse(path).getroot()
    except ET.ParseError as e:
        raise OracleAtomError(f"Error parsing the XML file:{path} ({e})")
    if root.tag != f"{{{NS}}}atom":
        raise OracleAtomError("XPath expression /dft-efe:atom not found")
    kids = {}
    for c in root:
        if isinstance(c.tag, str) and c.tag.startswith(f"{{{NS}}}"):
            kids.setdefault(c.tag[len(NS) + 2:], []).append(c)

    def single(name):
        if name not in kids:
            raise OracleAtomError(f"XPath expression /dft-efe:atom/dft-efe:{name} not found.")
        if len(kids[name]) != 1:
            raise OracleAtomError(f"Found more than one /dft-efe:atom/dft-efe:{name} element")
        return kids[name][0].text or ""

    meta = {k: (v[0].text or "").strip() for k, v in kids.items() if len(v) == 1}
    try:
        NR = int(single("NR").strip())
    except ValueError:
        raise OracleAtomError("Error while converting NR to int")
    if NR <= 0:
        raise OracleAtomError("Non-positive integer found for NR")
    r = np.array([float(t) for t in single("r").split()])
    if r.size != NR:
        raise OracleAtomError("Mismatch in number of radial points")
    out = {}
    for f in fields:
        if f not in kids:
            raise OracleAtomError(f"XPath expression /dft-efe:atom/dft-efe:{f}/dft-efe:values not found.")
        funs, qs = [], set()
        for e in kids[f]:
            sub = {}
            for c in e:
                if isinstance(c.tag, str) and c.tag.startswith(f"{{{NS}}}"):
                    sub.setdefault(c.tag[len(NS) + 2:], []).append(c.text or "")
            if any(len(sub.get(t, [])) != 1 for t in ("values", "qNumbers", "cutoffInfo")):
                raise OracleAtomError("Mismatch in number of values / qNumbers / cutoffInfo")
            vals = np.array([float(t) for t in sub["values"][0].split()])
            if vals.size != NR:
                raise OracleAtomError("Mismatch in number of points")
            q = [int(t) for t in sub["qNumbers"][0].split()]
            if len(q) != 3:
                raise OracleAtomError("Expected three quantum number")
            if not -q[1] <= q[2] <= q[1]:
                raise OracleAtomError("Magnetic quantum number (m) found outside of -l and +l")
            co = [float(t) for t in sub["cutoffInfo"][0].split()]
            if len(co) != 2:
                raise OracleAtomError("Expected two values (cutoff and smoothness factor)")
            qs.add(tuple(q))
            funs.append({"q": q, "cutoff": co[0], "smoothness": co[1], "values": vals})
        if len(qs) != len(funs):
            raise OracleAtomError("Found repeated quantum numbers")
        srt = sorted(qs)
        i = 0
        while i < len(srt):
            l = srt[i][1]
            for m in range(-l, l + 1):
                if i >= len(srt) or srt[i][2] != m:
                    raise OracleAtomError("not in correct order (ascending m)")
                i += 1
        out[f] = funs
    return {"meta": meta, "r": r, "fields": out}


class Spline:
    """utils::Spline defaults: clamped (zero end slopes) C^2 cubic, linear to
    the left of the grid, quadratic to the right."""

    def __init__(self, x, y):
        self.x, self.y = np.asarray(x, float), np.asarray(y, float)
        self.cs = CubicSpline(self.x, self.y, bc_type=((1, 0.0), (1, 0.0)))
        self.b0 = float(self.cs(self.x[0], 1))
        self.bn = float(self.cs(self.x[-1], 1))
        self.cn = float(self.cs(self.x[-1], 2)) / 2.0

    def __call__(self, r, nu=0):
        r = np.asarray(r, float)
        v = np.asarray(self.cs(r, nu), float)
        hl, hr = r - self.x[0], r - self.x[-1]
        left, right = r < self.x[0], r > self.x[-1]
        if nu == 0:
            v = np.where(left, self.y[0] + self.b0 * hl, v)
            v = np.where(right, self.y[-1] + (self.bn + self.cn * hr) * hr, v)
        else:
            v = np.where(left, self.b0, v)
            v = np.where(right, self.bn + 2.0 * self.cn * hr, v)
        return v


def cutoff(r, rc, d):
    """SmoothCutoffFunctions.cpp:25-70: value and derivative."""
    def f1(x):
        return np.where(x > 0, np.exp(-1.0 / np.where(x > 0, x, 1.0)), 0.0)

    def f1d(x):
        return np.where(x > 0, f1(x) / np.where(x > 0, x * x, 1.0), 0.0)
    y = 1.0 - d * (r - rc) / rc
    a, b = f1(y), f1(1.0 - y)
    val = a / (a + b)
    der = (f1d(y) * b + a * f1d(1.0 - y)) / (a + b) ** 2 * (-d / rc)
    return val, der


def _blm(l, m):
    v = math.sqrt((2.0 * l + 1) / 2.0)
    for k in range(1, m + 1):
        v /= math.sqrt((l - k + 1.0) * (l + k))
    return v


def _dm(m):
    return 1.0 / math.sqrt(2 * math.pi) if m == 0 else 1.0 / math.sqrt(math.pi)


def _qm(m, phi):
    return np.cos(m * phi) if m > 0 else (np.ones_like(phi) if m == 0 else np.sin(-m * phi))


def _dqm(m, phi):
    return -m * np.sin(m * phi) if m > 0 else (np.zeros_like(phi) if m == 0 else -m * np.cos(-m * phi))


def _plm(l, m, th):
    """P_l^m(cos th) without the Condon-Shortley phase and its first two theta
    derivatives, l <= 2."""
    c, s = np.cos(th), np.sin(th)
    if l == 0:
        return np.ones_like(th), np.zeros_like(th), np.zeros_like(th)
    if l == 1:
        if m == 0:
            return c, -s, -c
        return s, c, -s
    if m == 0:
        return (3 * c * c - 1) / 2, -3 * c * s, -3 * (c * c - s * s)
    if m == 1:
        return 3 * c * s, 3 * (c * c - s * s), -12 * s * c
    return 3 * s * s, 6 * s * c, 6 * (c * c - s * s)


def evaluate(fun, r_grid, origin, points, polar_tol=1e-12):
    """Value and gradient of fun (a read_atom entry) at points (n x 3)."""
    n_, l, m = fun["q"]
    if l > 2:
        raise NotImplementedError("l > 2")
    rc, sm = fun["cutoff"], fun["smoothness"]
    S = Spline(r_grid, fun["values"])
    X = np.asarray(points, float) - np.asarray(origin, float)
    r = np.sqrt((X ** 2).sum(axis=1))
    safe = np.where(r > 0, r, 1.0)
    th = np.where(r > 0, np.arccos(np.clip(X[:, 2] / safe, -1.0, 1.0)), 0.0)
    on_axis = (np.abs(th) < polar_tol) | (np.abs(th - math.pi) < polar_tol)
    ph = np.where(on_axis | (r == 0), 0.0, np.arctan2(X[:, 1], X[:, 0]))
    R, dR = S(r), S(r, 1)
    fc, dfc = cutoff(r, rc, sm)
    am = abs(m)
    P, dP, d2P = _plm(l, am, th)
    const = _blm(l, am) * _dm(m)
    Y = const * P * _qm(m, ph)
    dYdth = const * dP * _qm(m, ph)
    if m != 0:
        dYdph_sin = const * (np.sin(th) * d2P + np.cos(th) * dP + np.sin(th) * l * (l + 1) * P) \
            * (1.0 / (m * m)) * _dqm(m, ph)
    else:
        dYdph_sin = np.zeros_like(r)
    inside = r <= rc + rc / sm
    val = np.where(inside, R * Y * fc, 0.0)
    dVdr = (dR * fc + dfc * R) * Y
    dth = (R / safe) * fc * dYdth
    dph = (R / safe) * fc * dYdph_sin
    st, ct, sp, cp = np.sin(th), np.cos(th), np.sin(ph), np.cos(ph)
    g = np.stack([dVdr * st * cp + dth * ct * cp - sp * dph,
                  dVdr * st * sp + dth * ct * sp + cp * dph,
                  dVdr * ct - dth * st], axis=1)
    g = np.where((inside & (r > 0))[:, None], g, 0.0)
    return val, g


# ------------------------------------------------------------------ UPF
# AtomSphericalDataPSP (src/atoms/AtomSphericalDataPSP.cpp:291-830) restated
# with xml.etree and scipy (test infrastructure only).
Y00 = math.sqrt(1.0 / 2.0) * (1.0 / math.sqrt(2 * math.pi)) * 1.0
UPF_ELEMENT = {"vlocal": "PP_LOCAL", "beta": "PP_NONLOCAL", "pswfc": "PP_PSWFC",
               "nlcc": "PP_NLCC", "rhoatom": "PP_RHOATOM"}


class SplineBC(Spline):
    """utils::Spline with a second-derivative condition at the left end and a
    slope condition at the right (the UPF local potential)."""

    def __init__(self, x, y, left_second, right_slope):
        self.x, self.y = np.asarray(x, float), np.asarray(y, float)
        self.cs = CubicSpline(self.x, self.y, bc_type=((2, left_second), (1, right_slope)))
        self.b0 = float(self.cs(self.x[0], 1))
        self.c0 = float(self.cs(self.x[0], 2)) / 2.0
        self.bn = float(self.cs(self.x[-1], 1))
        self.cn = float(self.cs(self.x[-1], 2)) / 2.0

    def __call__(self, r, nu=0):
        r = np.asarray(r, float)
        v = np.asarray(self.cs(r, nu), float)
        hl, hr = r - self.x[0], r - self.x[-1]
        left, right = r < self.x[0], r > self.x[-1]
        if nu == 0:
            v = np.where(left, self.y[0] + (self.b0 + self.c0 * hl) * hl, v)
            v = np.where(right, self.y[-1] + (self.bn + self.cn * hr) * hr, v)
        else:
            v = np.where(left, self.b0 + 2 * self.c0 * hl, v)
            v = np.where(right, self.bn + 2.0 * self.cn * hr, v)
        return v


def read_upf(path, fields):
    root = ET.parse(path).getroot()
    if root.tag != "UPF":
        raise OracleAtomError("XPath expression /UPF not found")
    hdr = root.find("PP_HEADER").attrib
    meta = {k: v.strip() for k, v in hdr.items()}
    NR = int(float(meta["mesh_size"]))
    zv = float(meta["z_valence"])
    lmax = int(meta.get("l_max", "0"))
    core = meta.get("core_correction", "F").lower() in ("t", "true", ".true.", "1")
    r = np.array([float(t) for t in root.find("PP_MESH/PP_R").text.split()])
    if r.size != NR:
        raise OracleAtomError("Mismatch in number of radial points")
    out = {}
    for f in fields:
        if f not in UPF_ELEMENT:
            raise OracleAtomError(f"No element name '{f}' found in UPF_PARENT_PATH_LOOKUP")
        entries = []
        if f == "beta":
            p = lprev = lmx = 0
            for i, c in enumerate(root.find("PP_NONLOCAL")):
                if "PP_BETA" in c.tag:
                    l = int(c.attrib["angular_momentum"])
                    if i != 0:
                        p = 0 if l != lprev else p + 1
                        lprev = l
                    entries += [(c.text, [p, l, m]) for m in range(-l, l + 1)]
                    lmx = max(lmx, l)
            if lmx != lmax:
                raise OracleAtomError("The lmax from PSP HEADER does not match")
        elif f == "pswfc":
            for c in root.find("PP_PSWFC"):
                if "PP_CHI" in c.tag:
                    l, n = int(c.attrib["l"]), int(c.attrib["label"][0])
                    entries += [(c.text, [n, l, m]) for m in range(-l, l + 1)]
        elif f == "nlcc" and not core:
            entries.append((" ".join(["0"] * NR), [0, 0, 0]))
        else:
            entries.append((root.find(UPF_ELEMENT[f]).text, [0, 0, 0]))
        qs = [tuple(q) for _, q in entries]
        if len(set(qs)) != len(qs):
            raise OracleAtomError("Trying to enter more than one of same quantum number vector")
        funs = []
        for text, q in entries:
            v = np.array([float(t) for t in text.split()])
            if v.size != NR:
                raise OracleAtomError("Mismatch in number of points")
            if f == "vlocal":
                v = v * 0.5
            elif f in ("beta", "pswfc"):
                v[1:] = v[1:] / r[1:]
                v[0] = v[1]
            elif f == "rhoatom":
                v[1:] = v[1:] / (4 * math.pi * r[1:] * r[1:])
                v[0] = v[1]
            cutoff = 1.0e10
            for j in range(NR - 1, 0, -1):
                if abs(v[j]) >= 1e-10:
                    cutoff = r[j] if j != NR - 1 else 1.0e10
                    break
            if f == "rhoatom":
                cutoff = r[-1]
            if f not in ("beta", "pswfc"):
                v = v * (1.0 / Y00)
            fun = {"q": q, "cutoff": cutoff, "smoothness": 1.0e10, "values": v, "mixed": False}
            if f == "vlocal":
                keep = r <= 10.0001
                xs, ys = r[keep], v[keep]
                fun.update(values=ys, knots=xs, mixed=True, tail=-1.0 * (1.0 / Y00) * abs(zv),
                           left=(ys[1] - ys[0]) / (xs[1] - xs[0]), right=-(ys[-1] / xs[-1]))
            funs.append(fun)
        out[f] = funs
    return {"meta": meta, "r": r, "fields": out, "z_valence": zv}


def evaluate_upf(fun, r_grid, origin, points):
    """Value and gradient of a UPF function: the mixed local potential, or a
    numerical function as `evaluate` (l <= 2)."""
    if not fun["mixed"]:
        return evaluate(fun, r_grid, origin, points)
    S = SplineBC(fun["knots"], fun["values"], fun["left"], fun["right"])
    X = np.asarray(points, float) - np.asarray(origin, float)
    r = np.sqrt((X ** 2).sum(axis=1))
    safe = np.where(r > 0, r, 1.0)
    inside = r <= fun["knots"][-1]
    R = np.where(inside, S(r), fun["tail"] / safe)
    dR = np.where(inside, S(r, 1), -fun["tail"] / safe ** 2)
    val = Y00 * R
    g = np.where((r > 0)[:, None], (Y00 * dR / safe)[:, None] * X, 0.0)
    return val, g
