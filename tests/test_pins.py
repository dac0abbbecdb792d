"""Pins of the decision arithmetic: the rv cost's cube root against exact decimal roots and
closed-form part volumes, the evaluate rule against a brute force over every leaf path,
the refinement probe rule, and the Hausdorff sample positions against the Wang hash by
hand. Each runs on the oracle (CPU) and again on the HIP path (marked gpu), so the two
are pinned by the same independent checks, not only against each other."""
import math

import numpy as np
import pytest

from fixtures import l_prism, square_torus, stairs
import pins

BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]


@pytest.fixture(params=BACKENDS)
def api(request, cw, oracle):
    return oracle if request.param == "oracle" else cw.load_product()


# ---- rv cost: the library's deterministic cube root ----------------------------------------

def test_rv_cost_against_exact_roots(cw, oracle):
    """cuacd_rv_cost (the cost record evaluates on the device and the oracle on the host) against
    0.3 cbrt(3/(4 pi) gap) in 60-digit decimal arithmetic, over gaps from 1e-300 to 1e3: at most
    2 ulp apart. A wrong constant, a dropped Newton step or a bad exponent split fails this."""
    prod = cw.load_product()
    rng = np.random.default_rng(5)
    gaps = np.concatenate([np.logspace(-300, 3, 1500), 10.0 ** rng.uniform(-300, 3, 1500),
                           rng.uniform(0, 2, 500), [1e-300, 0.5, 0.25, 0.375, 0.125, 1.0, 8.0]])
    worst = 0.0
    for g in gaps:
        v = prod.rv_cost(0.0, float(g))
        assert v == oracle.rv_cost(0.0, float(g))
        assert v == prod.rv_cost(1.0, 1.0 + float(g)) or 1.0 + float(g) - 1.0 != float(g)
        worst = max(worst, pins.ulps(v, pins.exact_rv(float(g))))
    print(f"rv_cost: worst error {worst:.3f} ulp over {len(gaps)} gaps")
    assert worst <= 2.0
    # no gap, negative gap: zero; tiny and subnormal gaps: positive and increasing
    assert prod.rv_cost(1.0, 1.0) == 0.0 and prod.rv_cost(2.0, 1.0) == 0.0
    tiny = [prod.rv_cost(0.0, g) for g in (1e-321, 1e-320, 1e-310, 1e-305, 1e-301)]
    assert all(t > 0 for t in tiny) and tiny == sorted(tiny)
    assert abs(tiny[-1] / float(pins.exact_rv(1e-301)) - 1) < 1e-14


def _l_prism_raw(cw):
    """The L prism [0,2]^2 x [0,1] without the corner [1,2]^2 (tests/fixtures.ts:75-77), unnormalized."""
    return l_prism(cw)


def test_level_costs_closed_form(cw, api):
    """Depth-1 tree of the unnormalized L prism with width 9 and no margin: planes x, y in {0.5, 1, 1.5}
    and z in {0.25, 0.5, 0.75}. Every child is a box or an L prism whose mesh and hull volumes have
    closed forms, and every corner sits on the 2^21 hull grid, so the volumes are exact:
    cut x=0.5: L' of volume 2 and hull 2.5 (gap 0.5); x=1: two boxes (0); x=1.5: L of 2.5 and hull
    2.75 (0.25); z=h: L prisms of volume 3h, hull 3.5h (gap 0.5 h). Each level cost must equal the
    library's rv of that gap bit for bit and the exact 0.3 cbrt(3/(4 pi) gap) within 2 ulp; the best
    cut is x=1 (cost 0, lowest index on the tie with y=1)."""
    t = cw.GpuTree(api=api, width=9, width2=3, depth=1, n_concave_edges=0, min_edge_dist=0.0)
    r = t.run([_l_prism_raw(cw)])
    items = r.levels[1]
    gaps = {0: 0.5, 1: 0.0, 2: 0.25, 3: 0.5, 4: 0.0, 5: 0.25, 6: 0.375, 7: 0.25, 8: 0.375}
    for it in items[:9]:
        assert it["flags"] == 1 and it["n_levels"] == 1
        cut = it["cut_idx"]
        vols = sorted((t.read_part(p, geometry=False)["mesh_vol"], t.read_part(p, geometry=False)["hull_vol"])
                      for p in it["part_ids"][: it["np"]])
        assert abs(sum(v[0] for v in vols) - 3.0) < 1e-12
        g = max(h - m for m, h in vols)
        assert g == gaps[cut], (cut, vols)
        c = it["level_costs"][0]
        assert c == cw.load_product().rv_cost(0.0, gaps[cut])
        assert pins.ulps(c, pins.exact_rv(gaps[cut])) <= 2.0 if c else gaps[cut] == 0.0
    assert r.seeds[0].best_cut == 1 and r.seeds[0].best_cost == 0.0
    t.destroy()


# ---- evaluate: brute force over every leaf path ---------------------------------------------

def _cases(cw):
    norm = lambda m: cw.normalize(m)[0]  # noqa: E731
    return {
        "stairs": [norm(stairs(cw))],
        "l_prism": [norm(l_prism(cw))],
        "torus": [norm(square_torus(cw))],
        "mixed": [norm(stairs(cw)), norm(square_torus(cw)), norm(l_prism(cw))],
    }


def _check_tree(cw, t, seeds, r, depth, width, m, edge_planes=None):
    """best_cut, best_cost and plane of every seed against the brute force over the run's tables."""
    rv = cw.load_product().rv_cost
    cache = {}

    def part_rv(p):
        if p not in cache:
            d = t.read_part(p, geometry=False)
            assert d["hull_vol"] >= 0 and not (d["flags"] & 4)
            cache[p] = rv(d["mesh_vol"], d["hull_vol"])
        return cache[p]
    n = len(seeds)
    found_nonzero = False
    for s in range(n):
        paths = pins.path_costs(r.levels, s, n, depth, part_rv)
        best_c, best_cut = pins.best_of(paths)
        if best_cut == pins.NONE:
            best_c, best_cut = pins.greedy_level0(r.levels, s, n, part_rv)
        got = r.seeds[s]
        assert got.best_cut == (best_cut if best_cut != pins.NONE else -1), (s, got, best_cut)
        if best_cut == pins.NONE:
            continue
        assert got.best_cost == best_c, (s, got.best_cost, best_c)
        found_nonzero |= best_c > 0 and len({round(pins.mean_cost(c), 12) for _, c in paths}) > 2
        if best_cut < width:
            v = seeds[s].verts
            assert got.plane == pins.axis_plane(v.min(0), v.max(0), m, width, best_cut), (got.plane, best_cut)
        else:
            e = np.asarray(edge_planes[s], np.float32).reshape(-1, 4)[best_cut - width]
            assert got.plane == tuple(float(x) for x in e)
    return found_nonzero


@pytest.mark.parametrize("case", ["stairs", "l_prism", "torus", "mixed"])
@pytest.mark.parametrize("depth", [2, 3])
def test_evaluate_brute_force(cw, api, case, depth):
    """The chosen first cut of every seed is the argmin of the mean level cost over every leaf path
    (full-depth items and early leaves padded with zero costs), lowest initial cut on ties, with level
    costs recomputed from the part volumes; the returned plane is that cut's plane. Branch and bound
    does not change the answer: the same run without last-level pruning gives the same cut and cost."""
    seeds = _cases(cw)[case]
    m, width = 0.0125, 6
    planes = [cw.concave_edge_planes(s, n_concave_edges=2, seed=11 + i) for i, s in enumerate(seeds)]
    cfg = dict(width=width, width2=3, depth=depth, n_concave_edges=2, min_edge_dist=m)
    t = cw.GpuTree(api=api, **cfg)
    r = t.run(seeds, planes)
    assert not r.overflow and not r.np_overflow
    nonzero = _check_tree(cw, t, seeds, r, depth, width, m, planes)
    t.destroy()
    u = cw.GpuTree(api=api, prune_last_level=False, **cfg)
    ru = u.run(seeds, planes)
    _check_tree(cw, u, seeds, ru, depth, width, m, planes)
    for a, b in zip(r.seeds, ru.seeds):
        assert (a.best_cut, a.best_cost, a.plane) == (b.best_cut, b.best_cost, b.plane)
    u.destroy()
    if case != "l_prism":  # the L prism has a cut of cost 0 (the junction)
        assert nonzero, "the case must rank several distinct nonzero path costs"


def test_evaluate_early_leaves(cw, api):
    """A margin so large that level-1 parts fall under the leaf rule (every extent <= 2 m + 1e-6):
    those items end early with zero costs for the remaining levels and compete with full-depth paths
    through their mean; the brute force must find the same argmin. The normalized L prism, and the
    same flattened to 0.3 of its height, at margins 0.525-0.6 have both kinds of leaf."""
    base = cw.normalize(l_prism(cw))[0]
    flat = base.verts.copy()
    flat[:, 2] *= 0.3
    found = 0
    for seed in (base, cw.Mesh(flat.astype(np.float32), base.tris)):
        for m in (0.525, 0.55, 0.6):
            t = cw.GpuTree(api=api, width=6, width2=3, depth=3, n_concave_edges=0, min_edge_dist=m)
            r = t.run([seed])
            early = sum(1 for lvl in r.levels[1:-1] for it in lvl if it["flags"] == 2)
            full = sum(1 for it in r.levels[-1] if it["flags"] == 1)
            _check_tree(cw, t, [seed], r, 3, 6, m)
            assert r.seeds[0].best_cost > 0
            t.destroy()
            found += early > 0 and full > 0
    assert found >= 4, "too few cases with both early and full-depth leaves"


@pytest.mark.parametrize("case", ["stairs", "mixed"])
def test_refine_probe_rule(cw, api, case):
    """Refinement rounds (src/gpu/shaders.ts:1489-1514): round r evaluates nProbes planes parallel to
    the current one, offset by half * (2 i / (n - 1) - 1) with half = spacing, then 2 half / (n - 1);
    a probe is taken only when its brute-force path cost is below the current cost - 1e-7. Runs with
    0, 1, 2 and 3 rounds give the state after each round; each round's tables are the probe tree's.
    The cases take probes in several rounds."""
    norm = lambda m: cw.normalize(m)[0]  # noqa: E731
    if case == "stairs":
        seeds, spacing = [norm(stairs(cw))], [0.1]
    else:
        seeds = [norm(square_torus(cw)), norm(l_prism(cw)), norm(cw.generate("staircase", 20260100, [5, 1.0, 0.2, 0.2]))]
        spacing = [0.05, 0.2, 0.1]
    n = len(seeds)
    probes = 8
    cfg = dict(width=6, width2=3, depth=2, n_concave_edges=2, min_edge_dist=0.0125)
    planes = None
    rv = cw.load_product().rv_cost
    prev = None
    half = list(spacing)
    taken = 0
    for rounds in range(4):
        t = cw.GpuTree(api=api, **cfg)
        refine = dict(rounds=rounds, probes=probes, spacing=spacing) if rounds else None
        r = t.run(seeds, planes, refine=refine)
        if rounds == 0:
            _check_tree(cw, t, seeds, r, 2, 6, 0.0125)
        else:
            cache = {}

            def part_rv(p):
                if p not in cache:
                    d = t.read_part(p, geometry=False)
                    cache[p] = rv(d["mesh_vol"], d["hull_vol"])
                return cache[p]
            for s in range(n):
                c, cut = pins.best_of(pins.path_costs(r.levels, s, n, 2, part_rv))
                if cut == pins.NONE:
                    c, cut = pins.greedy_level0(r.levels, s, n, part_rv)
                cur = prev.seeds[s]
                got = r.seeds[s]
                if cut != pins.NONE and cur.best_cut >= 0 and c < cur.best_cost - 1e-7:
                    assert got.best_cut == cut and got.best_cost == c
                    assert got.plane == pins.probe_plane(cur.plane, half[s], probes, cut), (got.plane, cut)
                    taken += 1
                else:
                    assert (got.best_cut, got.best_cost, got.plane) == (cur.best_cut, cur.best_cost, cur.plane)
            half = [pins.next_half(h, probes) for h in half]
        prev = r
        t.destroy()
    print(f"probes taken: {taken}")
    assert taken > 0


# ---- Hausdorff samples -----------------------------------------------------------------------

def test_sample_positions_by_hand(cw, api):
    """One triangle with corners at the origin and the unit x and y axes: sample k of triangle 0 lies at
    (w1, w2, 0), so its coordinates are the barycentric weights from the Wang hash, computed here by
    hand in float32. Also the counts of a many-triangle mesh (area rule, fixed rule, step rule)."""
    tri = cw.Mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.uint32))
    t = cw.GpuTree(api=api, width=3, width2=3, depth=1, n_concave_edges=0, hausdorff_samples=0)
    s = t.sample_mesh(tri)
    assert len(s) == 1000  # area 0.5: max(1000, floor(2000 * 0.5)) samples, all on the one triangle
    for k in (0, 1, 2, 17, 500, 999):
        w0, w1, w2, (ha, hb) = pins.sample_weights(0, k)
        assert tuple(s[k]) == (float(w1), float(w2), 0.0), (k, s[k], w1, w2)
    # known hash values, by hand from src/cpu/hausdorff.ts:10-17: seed 0 -> 61 -> 549 -> 519 ->
    # 519 * 0x27d4eb2d mod 2^32 = 3232286779 -> 0xc0a9496a; seed 1 -> 60 -> 540 -> 573 -> 663905209
    assert pins.wang_hash(0) == 0xC0A9496A and pins.wang_hash(1) == 0x27922C9D
    # counts: a closed mesh of uneven triangles, the default rule and fixed resolutions
    for mesh in (cw.l_shape(), stairs(cw), cw.generate("blobs", 3, [3, 12])):
        for res in (0, 50, 3000):
            u = cw.GpuTree(api=api, width=3, width2=3, depth=1, n_concave_edges=0, hausdorff_samples=res)
            counts = pins.sample_counts(mesh.verts, mesh.tris, res)
            smp = u.sample_mesh(mesh)
            assert len(smp) == sum(counts), (res, len(smp), sum(counts))
            # every sample of triangle tt is w0 a + w1 b + w2 c in float32, in triangle order
            V = mesh.verts.reshape(-1, 3)
            T = mesh.tris.reshape(-1, 3)
            i = 0
            for tt, n in enumerate(counts):
                if n and tt % 7 == 0:
                    for k in range(n):
                        w0, w1, w2, _ = pins.sample_weights(tt, k)
                        a, b, c = V[T[tt, 0]], V[T[tt, 1]], V[T[tt, 2]]
                        want = np.float32(np.float32(w0 * a) + np.float32(w1 * b)) + np.float32(w2 * c)
                        assert np.array_equal(smp[i + k], want), (tt, k)
                i += n
            u.destroy()
    t.destroy()


def test_sample_weights_match_float64_formula():
    """The float32 weights against the reference CPU path's float64 formula (src/cpu/hausdorff.ts:52-57):
    within a few float32 ulps, and the weights sum to 1 within float32 rounding."""
    worst = 0.0
    for t in range(0, 2000, 37):
        for k in range(0, 40, 3):
            w0, w1, w2, (ha, hb) = pins.sample_weights(t, k)
            ra, rb = (ha & 0xFFFFFF) / 0xFFFFFF, (hb & 0xFFFFFF) / 0xFFFFFF
            sq = math.sqrt(ra)
            for got, ref in ((w0, 1 - sq), (w1, sq * (1 - rb)), (w2, sq * rb)):
                worst = max(worst, abs(float(got) - ref))
            assert abs(float(w0) + float(w1) + float(w2) - 1.0) < 4e-7
    assert worst < 3e-7, worst
