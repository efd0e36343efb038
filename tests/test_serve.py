"""The query server (include/gdl_serve.h): every query's result in a mix of
concurrent TC / SG / CSPA queries is the program's result alone -- the same
sorted tuples, per-iteration |delta| and |full|, iteration counts -- and
admission respects the memory budget."""
import numpy as np
import pytest

from conftest import gdl, read_edges


def mix_queries():
    """(qid, kind, a, b) of a small mixed workload."""
    q = []
    tc = [read_edges("hpc_talk.txt"), read_edges("data_7035.txt"), gdl.gen_er(3, 500, 1500),
          gdl.gen_tree(11, 5000, 100), gdl.gen_er(9, 2000, 5000), gdl.pack([1], [2])]
    for e in tc:
        q.append((len(q), gdl.QUERY_TC, e, None))
    for s in range(3):
        q.append((len(q), gdl.QUERY_SG, gdl.gen_rmat(4 + s, 11, 3000), None))
    for s in range(3):
        a, d = gdl.gen_cspa(20 + s, 150, 1.1, 300, 150)
        q.append((len(q), gdl.QUERY_CSPA, a, d))
    return q


def alone(ctx, kind, a, b):
    if kind == gdl.QUERY_TC:
        return ctx.run_tc(a)
    if kind == gdl.QUERY_SG:
        return ctx.run_sg(a)
    return ctx.run_cspa(a, b)


def delta_full(trace):
    return [(t[0], t[1], t[3], t[4]) for t in trace]


def test_query_estimates_grow_with_the_input():
    small = gdl.query_estimate(gdl.QUERY_TC, gdl.gen_er(0, 300, 700))
    big = gdl.query_estimate(gdl.QUERY_TC, gdl.gen_er(0, 2000, 5000))
    assert 0 < small < big
    sg = gdl.query_estimate(gdl.QUERY_SG, gdl.gen_rmat(1, 12, 5000))
    a, d = gdl.gen_cspa(5, 300, 1.1, 500, 250)
    cs = gdl.query_estimate(gdl.QUERY_CSPA, a, d)
    assert sg > 0 and cs > 0


def run_mix(server, queries):
    for qid, kind, a, b in queries:
        server.submit(qid, kind, a, b)
    got = {}
    while len(got) < len(queries):
        r = server.wait(120_000)
        assert r is not None, "server stalled"
        got[r["qid"]] = r
    return got


@pytest.mark.gpu
@pytest.mark.parametrize("batch_tc,take_turns", [(True, False), (False, False), (True, True)])
def test_server_mix_equals_alone(ctx, batch_tc, take_turns):
    queries = mix_queries()
    with gdl.Server(workers=3, batch_tc=batch_tc, keep_keys=True, take_turns=take_turns) as s:
        got = run_mix(s, queries)
        st = s.stats()
    assert st["finished"] == len(queries) and st["reserved"] == 0
    assert (st["batched"] > 0) == batch_tc
    for qid, kind, a, b in queries:
        r, w = got[qid], alone(ctx, kind, a, b)
        assert r["status"] == 0, (qid, r)
        assert r["n_out"] == w["n_out"], qid
        assert np.array_equal(r["keys"], w["out"]), qid
        assert r["iterations"] == w["iterations"], qid
        assert r["derived"] == w["derived"], qid
        assert r["checksum"] == gdl.tuple_checksum(w["out"]), qid
        assert delta_full(r["trace"]) == delta_full(w["trace"]), qid
        if kind != gdl.QUERY_TC:
            assert r["init_iterations"] == w["init_iterations"], qid


@pytest.mark.gpu
def test_server_budget_serialises_admission(ctx):
    queries = mix_queries()[:8]
    with gdl.Server(workers=3, mem_budget=1, keep_keys=False) as s:
        got = run_mix(s, queries)
        st = s.stats()
    # with no room beside a running query, queries run one at a time
    spans = sorted((r["admit_ms"], r["done_ms"]) for r in got.values())
    for (a0, d0), (a1, _) in zip(spans, spans[1:]):
        assert a1 >= d0 - 1e-6
    assert st["peak_reserved"] <= max(r["est_bytes"] for r in got.values())
    for qid, kind, a, b in queries:
        w = alone(ctx, kind, a, b)
        assert got[qid]["n_out"] == w["n_out"] and got[qid]["iterations"] == w["iterations"]
        assert got[qid]["checksum"] == gdl.tuple_checksum(w["out"])
