"""Serving (continuous batching of independent KS eigenproblems): each result
is the same alone and in the mix, and admission honours the memory budget."""
import numpy as np
import pytest

import dfx
from dfx import configs

pytestmark = pytest.mark.gpu


def small_stream(n=6, seed=11):
    s = configs.Stream(seed)
    out = []
    for i in range(n):
        na = 1 + (i % 2)
        atoms = configs.place_atoms(seed * 100 + i, na, -8.0, 8.0, (1, 6), 2.0, margin=4.0)
        out.append(dict(p=3, n=(4 + i % 3,) * 3, lo=(-8.0,) * 3, hi=(8.0,) * 3, atoms=atoms,
                        enrich=True, states=10, electrons=float(sum(a[3] for a in atoms)),
                        tol=1e-6, max_passes=100, seed=100 + i, arrival=0.02 * i))
    return out


def test_results_are_the_same_alone_and_in_the_mix():
    reqs = small_stream()
    alone_e, alone_r = [], []
    for r in reqs:  # one request at a time, nothing else on the device
        e, res = dfx.serve([dict(r, arrival=0.0)], max_slots=1)
        alone_e.append(e[0]); alone_r.append(res[0])
    mix_e, mix_r = dfx.serve(reqs, max_slots=4)
    for a, ra, m, rm in zip(alone_e, alone_r, mix_e, mix_r):
        assert ra["status"] == 0 and rm["status"] == 0
        assert ra["passes"] == rm["passes"]
        assert np.array_equal(a, m)          # bit for bit
    # the mix really overlapped: some request was admitted before another finished
    spans = sorted((x["admitSeconds"], x["doneSeconds"]) for x in mix_r)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    for r, x in zip(reqs, mix_r):
        assert x["admitSeconds"] >= r["arrival"] - 1e-3


def test_admission_respects_the_memory_budget():
    reqs = [dict(r, n=(5, 5, 5)) for r in small_stream(n=4, seed=12)]
    fps = [dfx.request_footprint(r) for r in reqs]
    budget = max(fps) + min(fps) // 2   # room for one request, never two
    assert budget < 2 * min(fps)
    _, res = dfx.serve([dict(r, arrival=0.0) for r in reqs], max_slots=4,
                       memory_budget=budget)
    spans = sorted((x["admitSeconds"], x["doneSeconds"]) for x in res)
    assert all(x["status"] == 0 for x in res)
    assert all(spans[i + 1][0] >= spans[i][1] - 1e-6 for i in range(len(spans) - 1))


def test_graph_replay_gives_the_eager_results():
    """Requests replay their captured degree loop on every pass (the filter
    bounds change from pass to pass and are read from the device): the
    results are those of eager launches, bit for bit."""
    reqs = [dict(r, arrival=0.0) for r in small_stream(n=3, seed=13)]
    e_graph, r_graph = dfx.serve(reqs, max_slots=2, use_graph=True)
    e_eager, r_eager = dfx.serve(reqs, max_slots=2, use_graph=False)
    for a, ra, b, rb in zip(e_graph, r_graph, e_eager, r_eager):
        assert ra["status"] == 0 and rb["status"] == 0
        assert ra["passes"] == rb["passes"] and ra["passes"] > 2
        assert np.array_equal(a, b)


def test_json_requests_with_atom_files_alone_and_in_the_mix():
    """JSON-line requests, one of them enriched from the reference's H.xml:
    responses come back in input order with the eigenvalues of serve() on the
    same dicts, and the atom-file request gives the same bits alone and in
    the mix."""
    import json
    import os
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "atoms", "H.xml")
    base = dict(p=3, n=[5, 5, 5], lo=[-8.0] * 3, hi=[8.0] * 3, states=6, electrons=2.0,
                tol=1e-6, max_passes=100)
    reqs = [dict(base, id="analytic", atoms=[[-0.7, 0.1, 0.0, 1], [0.7, -0.1, 0.0, 1]],
                 seed=3, arrival=0.0),
            dict(base, id="h-file", atoms=[[-0.7, 0.1, 0.0, 1, "H"], [0.7, -0.1, 0.0, 1, "H"]],
                 atom_files={"H": gold}, seed=4, arrival=0.01),
            dict(base, id="third", n=[4, 4, 4], atoms=[[0.3, 0.2, -0.1, 6]], electrons=6.0,
                 seed=5, arrival=0.02)]
    out = [json.loads(l) for l in dfx.serve_json([json.dumps(r) for r in reqs], max_slots=3)]
    assert [o["id"] for o in out] == ["analytic", "h-file", "third"]
    assert all(o["status"] == 0 and o["latency_seconds"] >= 0 for o in out)
    alone = json.loads(dfx.serve_json([json.dumps(dict(reqs[1], arrival=0.0))], max_slots=1)[0])
    assert alone["eigenvalues"] == out[1]["eigenvalues"]      # bit for bit (JSON floats round-trip)
    assert alone["passes"] == out[1]["passes"]
    # the atom-file request is a different basis from the analytic one
    assert out[1]["eigenvalues"] != out[0]["eigenvalues"]
