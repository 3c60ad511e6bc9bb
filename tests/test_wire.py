"""The serving wire format: JSON-line requests checked against the schema, and
the footprint of requests that carry atom files (host arithmetic, no GPU)."""
import json
import os

import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "atoms")


def base(**kw):
    r = dict(id="r0", p=3, n=[6, 6, 6], lo=[-8.0] * 3, hi=[8.0] * 3,
             atoms=[[0.0, 0.0, 0.0, 1, "H"], [1.4, 0.0, 0.0, 1, "H"]], states=4, electrons=2.0)
    r.update(kw)
    return r


def test_valid_request_round_trips():
    import dfx
    r = dfx.parse_request(json.dumps(base(tol=1e-7, arrival=0.5, atom_files={"H": "x.xml"})))
    assert r["id"] == "r0" and r["tol"] == 1e-7 and r["atom_files"] == {"H": "x.xml"}
    assert set(dfx.RESPONSE_KEYS) >= {"id", "status", "eigenvalues", "latency_seconds"}


@pytest.mark.parametrize("bad, msg", [
    (dict(states=None), "states"),
    (dict(colour="red"), "unknown request key"),
    (dict(p=3.5), "'p'"),
    (dict(states=True), "'states'"),
    (dict(n=[6, 6]), "'n'"),
    (dict(atoms=[[0.0, 0.0, 1]]), "'atoms'"),
    (dict(atoms=[[0.0, 0.0, 0.0, 1]], atom_files={"H": "x.xml"}), "by symbol"),
])
def test_schema_errors_name_the_key(bad, msg):
    import dfx
    r = base(**bad)
    if bad.get("states", 0) is None:
        del r["states"]
    with pytest.raises(ValueError, match=msg):
        dfx.parse_request(json.dumps(r))


def test_atom_file_requests_count_their_own_functions():
    """One H.xml orbital per atom (support 5.3 bohr) needs less device memory
    than the five analytic functions per atom (support 6 bohr)."""
    import dfx
    r_ana = dfx.parse_request(json.dumps(base()))
    r_num = dfx.parse_request(json.dumps(base(atom_files={"H": os.path.join(GOLD, "H.xml")})))
    r_lst = dfx.parse_request(json.dumps(base(atom_files=[os.path.join(GOLD, "H.xml")] * 2)))
    f_ana, f_num = dfx.request_footprint(r_ana), dfx.request_footprint(r_num)
    assert 0 < f_num < f_ana
    assert dfx.request_footprint(r_lst) == f_num
