"""Checkpoint / resume of the programs' fixpoints (gdl_prog_set_checkpoint).

A run stopped right after its snapshot at iteration k (a preempted run) and
then resumed from the snapshot must give the uninterrupted run's sorted
tuples, every trace record and the iteration counts, in both FULL modes, for
the one-LIE TC program (general loop and stream-ordered linear loop) and the
two-LIE SG and CSPA programs, whichever LIE the stop falls in.  A snapshot of
other inputs is refused.
"""
import numpy as np
import pytest

from conftest import gdl

pytestmark = pytest.mark.gpu

MODES = {"auto": gdl.FULL_MODE_AUTO, "merge": gdl.FULL_MODE_MERGE}


@pytest.fixture(autouse=True)
def _checkpoints_off_afterwards():
    # the setting is per thread: a failed test must not leave it on for the next
    yield
    gdl.set_checkpoint(None)


@pytest.fixture(params=sorted(MODES))
def mode(request):
    gdl.set_full_mode(MODES[request.param])
    yield request.param
    gdl.set_full_mode(gdl.FULL_MODE_AUTO)
    gdl.set_checkpoint(None)


def same(got, want):
    assert np.array_equal(got["out"], want["out"])
    for k in ("n_out", "iterations", "init_iterations", "derived"):
        assert got[k] == want[k], k
    assert [tuple(t) for t in got["trace"]] == [tuple(t) for t in want["trace"]]
    assert [tuple(t) for t in got["trace_init"]] == [tuple(t) for t in want["trace_init"]]


def stops(run, *args):
    with pytest.raises(gdl.GdlError) as e:
        run(*args)
    assert e.value.status == gdl.GDL_ERR_BUSY


TC = {
    "tree": lambda: gdl.gen_tree(11, 3000, 60),
    "chain": lambda: gdl.pack(np.arange(39), np.arange(1, 40)),
    "er": lambda: gdl.gen_er(3, 400, 1200),
}


@pytest.mark.parametrize("name,stop,every", [("tree", 7, 0), ("tree", 13, 5), ("chain", 1, 0),
                                             ("chain", 20, 3), ("er", 3, 2)])
def test_tc_resumes_to_the_uninterrupted_result(ctx, mode, tmp_path, name, stop, every):
    e = TC[name]()
    gdl.set_checkpoint(None)
    want = ctx.run_tc(e)
    assert want["iterations"] > stop
    gdl.set_checkpoint(str(tmp_path), every=every, stop_after=stop)
    stops(ctx.run_tc, e)
    gdl.set_checkpoint(str(tmp_path), every=every, resume=True)
    got = ctx.run_tc(e)
    gdl.set_checkpoint(None)
    same(got, want)


def test_tc_max_iteration_snapshot_continues(ctx, mode, tmp_path):
    # max_iteration stops the loop (a normal return with a partial result);
    # its snapshot lets a later run with no bound finish the same fixpoint
    e = TC["tree"]()
    gdl.set_checkpoint(None)
    want = ctx.run_tc(e)
    gdl.set_checkpoint(str(tmp_path))
    part = ctx.run_tc(e, max_iteration=9)
    assert part["iterations"] == 10 and part["n_out"] < want["n_out"]
    gdl.set_checkpoint(str(tmp_path), resume=True)
    got = ctx.run_tc(e)
    gdl.set_checkpoint(None)
    same(got, want)


def test_finished_snapshot_restores_without_a_body(ctx, mode, tmp_path):
    e = TC["chain"]()
    gdl.set_checkpoint(str(tmp_path))
    want = ctx.run_tc(e)
    gdl.set_checkpoint(str(tmp_path), resume=True)
    got = ctx.run_tc(e)
    gdl.set_checkpoint(None)
    same(got, want)


def _two_lie_stops(want):
    # a stop inside the first LIE, and one inside the second (the first then
    # finishes and leaves a finished snapshot)
    out = [1] if want["init_iterations"] > 1 else []
    if want["iterations"] > want["init_iterations"] + 1:
        out.append(want["init_iterations"] + 1)
    return out


def test_sg_resumes_in_either_lie(ctx, mode, tmp_path):
    e = gdl.gen_rmat(1, 11, 4000)
    gdl.set_checkpoint(None)
    want = ctx.run_sg(e)
    plan = _two_lie_stops(want)
    assert plan
    for i, stop in enumerate(plan):
        d = tmp_path / f"s{i}"
        d.mkdir()
        gdl.set_checkpoint(str(d), stop_after=stop)
        stops(ctx.run_sg, e)
        gdl.set_checkpoint(str(d), resume=True)
        got = ctx.run_sg(e)
        gdl.set_checkpoint(None)
        same(got, want)


def test_cspa_resumes_in_either_lie(ctx, mode, tmp_path):
    a, d = gdl.gen_cspa(5, 150, 1.1, 200, 100)
    gdl.set_checkpoint(None)
    want = ctx.run_cspa(a, d)
    plan = _two_lie_stops(want)
    assert plan
    for i, stop in enumerate(plan):
        p = tmp_path / f"c{i}"
        p.mkdir()
        gdl.set_checkpoint(str(p), every=2, stop_after=stop)
        stops(ctx.run_cspa, a, d)
        gdl.set_checkpoint(str(p), every=2, resume=True)
        got = ctx.run_cspa(a, d)
        gdl.set_checkpoint(None)
        same(got, want)


def test_snapshot_of_other_inputs_is_refused(ctx, tmp_path):
    e = TC["tree"]()
    gdl.set_checkpoint(str(tmp_path), stop_after=4)
    stops(ctx.run_tc, e)
    other = e.copy()
    other[0] = gdl.pack([2999], [0])[0]  # one edge changed: another static relation
    gdl.set_checkpoint(str(tmp_path), resume=True)
    with pytest.raises(gdl.GdlError) as err:
        ctx.run_tc(other)
    gdl.set_checkpoint(None)
    assert err.value.status == gdl.GDL_ERR_INVALID_ARG


def same_modulo_newt(got, want):
    # the trace's newt field counts raw rows in append mode and deduplicated
    # rows in merge mode; everything else is mode-independent
    assert np.array_equal(got["out"], want["out"])
    for k in ("n_out", "iterations", "init_iterations", "derived"):
        assert got[k] == want[k], k
    for key in ("trace", "trace_init"):
        assert [(t[0], t[1], t[3], t[4]) for t in got[key]] == \
            [(t[0], t[1], t[3], t[4]) for t in want[key]], key


CROSS = [("merge", "auto"), ("auto", "merge")]


@pytest.mark.parametrize("first,then", CROSS)
@pytest.mark.parametrize("name,stop", [("tree", 7), ("er", 3), ("chain", 20)])
def test_tc_resumes_across_full_modes(ctx, tmp_path, first, then, name, stop):
    # a snapshot written in one FULL mode resumes in the other: a sorted FULL
    # from merge mode becomes an append log whose tail is the delta, which the
    # stream-ordered linear loop reads
    e = TC[name]()
    gdl.set_full_mode(gdl.FULL_MODE_AUTO)
    gdl.set_checkpoint(None)
    want = ctx.run_tc(e)
    try:
        gdl.set_full_mode(MODES[first])
        gdl.set_checkpoint(str(tmp_path), stop_after=stop)
        stops(ctx.run_tc, e)
        gdl.set_full_mode(MODES[then])
        gdl.set_checkpoint(str(tmp_path), resume=True)
        got = ctx.run_tc(e)
    finally:
        gdl.set_checkpoint(None)
        gdl.set_full_mode(gdl.FULL_MODE_AUTO)
    same_modulo_newt(got, want)


@pytest.mark.parametrize("first,then", CROSS)
def test_sg_and_cspa_resume_across_full_modes(ctx, tmp_path, first, then):
    cases = [("sg", ctx.run_sg, (gdl.gen_rmat(1, 11, 4000),)),
             ("cspa", ctx.run_cspa, gdl.gen_cspa(5, 150, 1.1, 200, 100))]
    for label, run, args in cases:
        gdl.set_full_mode(gdl.FULL_MODE_AUTO)
        gdl.set_checkpoint(None)
        want = run(*args)
        for i, stop in enumerate(_two_lie_stops(want)):
            d = tmp_path / f"{label}{i}"
            d.mkdir()
            try:
                gdl.set_full_mode(MODES[first])
                gdl.set_checkpoint(str(d), stop_after=stop)
                stops(run, *args)
                gdl.set_full_mode(MODES[then])
                gdl.set_checkpoint(str(d), resume=True)
                got = run(*args)
            finally:
                gdl.set_checkpoint(None)
                gdl.set_full_mode(gdl.FULL_MODE_AUTO)
            same_modulo_newt(got, want)
