"""Progress reports and abort (src/decompose.ts:62-63, :238-252, :279-382, src/merge.ts:189, :205, :279).

The decomposer is host code shared by the product and the oracle, so these run on the CPU through the
oracle library: the phases arrive in the reference's order with a monotone fraction ending at 1, a set
abort flag stops the call at the next boundary with AbortError, and an abort during the merge stops the
merge rounds."""
import pytest


def _run(cw, oracle, m, **kw):
    return cw.decompose(m, api=oracle, max_iters=4, threshold=0.05, width=12, width2=6, depth=2, **kw)


def test_progress_phases(cw, oracle):
    from fixtures import stairs
    seen = []
    r = _run(cw, oracle, stairs(cw), on_progress=seen.append)
    phases = [p["phase"] for p in seen]
    assert phases[0] == "start" and phases[-1] == "done"
    assert phases[1] == "select"
    # an iteration that cuts reports select, tree, cut, hull, iteration in that order
    first = phases.index("tree")
    assert phases[first - 1: first + 4] == ["select", "tree", "cut", "hull", "iteration"]
    assert "merge" in phases and phases.index("merge") > max(i for i, p in enumerate(phases) if p == "iteration")
    fr = [p["fraction"] for p in seen]
    assert all(b >= a for a, b in zip(fr, fr[1:])) and fr[-1] == 1.0 and 0 < fr[0] <= 0.01 + 1e-12
    assert seen[-1]["n_parts"] == len(r.parts) and seen[-1]["message"] == f"done: {len(r.parts)} parts"
    assert seen[1]["worst_cost"] > 0 and seen[1]["iteration"] == 0 and seen[0]["iteration"] == -1
    sel = [p for p in seen if p["phase"] == "select"]
    assert all(p["message"].startswith(f"iteration {p['iteration']}: ") for p in sel)
    merges = [p for p in seen if p["phase"] == "merge"]
    assert merges[0]["message"].startswith("merging hulls: ") and any("merge round 0" in p["message"] for p in merges)


def test_abort_before_and_during(cw, oracle):
    from fixtures import stairs
    m = stairs(cw)
    flag = cw.AbortFlag()
    flag.set()
    with pytest.raises(cw.AbortError, match="decomposition aborted"):
        _run(cw, oracle, m, abort=flag)
    # set from the progress callback when the first cut phase is announced: the call stops at the next
    # boundary (before the hulls of that iteration)
    flag = cw.AbortFlag()
    seen = []

    def cb(p):
        seen.append(p["phase"])
        if p["phase"] == "cut":
            flag.set()
    with pytest.raises(cw.AbortError):
        _run(cw, oracle, m, abort=flag, on_progress=cb)
    assert seen[-1] == "cut" and "hull" not in seen
    # during the merge: the merge rounds stop (src/merge.ts:189)
    flag = cw.AbortFlag()

    def cb2(p):
        if p["phase"] == "merge" and "round 0" in p["message"]:
            flag.set()
    with pytest.raises(cw.AbortError, match="merge aborted"):
        _run(cw, oracle, m, abort=flag, on_progress=cb2)


def test_progress_batch_lanes(cw, oracle):
    """A batch split over engines reports from each engine (serialized), each ending with done."""
    from fixtures import stairs, square_torus
    seen = []
    cw.decompose_batch([stairs(cw), square_torus(cw), cw.l_shape()], api=oracle, batch_lanes=2, max_iters=3,
                       threshold=0.05, width=12, on_progress=seen.append)
    lanes = {p["lane"] for p in seen}
    assert lanes == {0, 1}
    for ln in lanes:
        mine = [p for p in seen if p["lane"] == ln]
        assert mine[0]["phase"] == "start" and mine[-1]["phase"] == "done"
