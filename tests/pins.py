"""Independent restatements used to pin the oracle's decision arithmetic (and, in the GPU
tests, the HIP path's): exact cube roots in decimal arithmetic, the evaluate rule
recomputed by brute force over every leaf path of a tree run, the refinement probe
rule, and the Hausdorff sample positions from the Wang hash computed by hand.

Nothing here calls the library's own evaluate, record or sampling code: each
function takes only what the mathematics or the reference's published rules fix
(src/gpu/shaders.ts:1404-1516 evaluate, :1519-1544 record, :334-348 leaf rule,
src/part.ts:24-26 rvCost, src/cpu/hausdorff.ts:10-67 wangHash and sampleMesh).
"""
import math
from decimal import Decimal, localcontext

import numpy as np

NONE = 0xFFFFFFFF
PI60 = Decimal("3.141592653589793238462643383279502884197169399375105820974944592")


def exact_rv(gap):
    """0.3 * cbrt(3 / (4 pi) * gap) to 50 significant digits (src/part.ts:24-26), as a Decimal."""
    with localcontext() as ctx:
        ctx.prec = 60
        if not gap > 0:
            return Decimal(0)
        x = Decimal(3) / (4 * PI60) * Decimal(gap)
        y = (x.ln() / 3).exp()
        for _ in range(3):  # Newton polish in 60 digits
            y = y - (y * y * y - x) / (3 * y * y)
        return Decimal("0.3") * y


def ulps(value, exact):
    """|value - exact| in units of the last place of value (a float)."""
    if value == 0.0:
        return 0.0 if exact == 0 else math.inf
    with localcontext() as ctx:
        ctx.prec = 60
        return float(abs(Decimal(value) - exact) / Decimal(math.ulp(value)))


# ---- evaluate by brute force over the leaf paths of one tree run ---------------------------

def path_costs(levels, seed, n_seeds, depth, part_rv):
    """Every leaf path of one seed with its recomputed level costs.

    levels[0] holds the seeds, levels[l + 1] the items of tree level l. A leaf is a full-depth
    item (flags 1 at the last level) or an item that the leaf rule stopped early (flags 2 at an
    earlier level, whose remaining level costs are zero). A level cost is recomputed from the
    part volumes: the largest rv over the item's parts at that level (record: sort ascending,
    take the last). Returns [(initial_cut, [cost per level])]."""
    out = []
    for l in range(depth):
        items = levels[l + 1]
        per = len(items) // n_seeds
        want = 1 if l == depth - 1 else 2
        for i in range(seed * per, (seed + 1) * per):
            it = items[i]
            if it["flags"] != want or it["n_levels"] == 0 or it["initial_cut"] == NONE:
                continue
            costs = []
            lvl, node = l, it
            chain = []
            while lvl >= 0:  # the item and its ancestors, deepest first
                chain.append((lvl, node))
                node = levels[lvl][node["parent_item"]] if lvl > 0 else None
                lvl -= 1
            for lvl, node in reversed(chain):
                costs.append(max(part_rv(p) for p in node["part_ids"][: node["np"]]))
            costs += [0.0] * (min(depth, 4) - len(costs))  # the leaf rule's zero costs
            out.append((it["initial_cut"], costs))
    return out


def mean_cost(costs):
    s = 0.0
    for c in costs:
        s += c
    return s / len(costs)


def best_of(paths):
    """Smallest mean path cost, lowest initial cut on ties; (NONE, 1e30) when there is no leaf."""
    best_c, best_cut = 1e30, NONE
    for cut, costs in paths:
        c = mean_cost(costs)
        if c < best_c or (c == best_c and cut < best_cut):
            best_c, best_cut = c, cut
    return best_c, best_cut


def greedy_level0(levels, seed, n_seeds, part_rv):
    """The fallback when no leaf is valid: the level-0 item with the smallest level-0 cost."""
    items = levels[1]
    per = len(items) // n_seeds
    best_c, best_cut = 1e30, NONE
    for i in range(seed * per, (seed + 1) * per):
        it = items[i]
        if it["flags"] == 1 and it["n_levels"] >= 1 and it["initial_cut"] != NONE:
            c = max(part_rv(p) for p in it["part_ids"][: it["np"]]) if it["n_levels"] == 1 else it["level_costs"][0]
            if c < best_c or (c == best_c and it["initial_cut"] < best_cut):
                best_c, best_cut = c, it["initial_cut"]
    return best_c, best_cut


def axis_plane(lo, hi, m, width, cut):
    """The level-0 axis plane of cut index `cut` (src/gpu/shaders.ts:350-378), in float32 as
    written: x_a = (lo_a + m) + ((s + 1) / (cpa + 1)) * ((hi_a - m) - (lo_a + m))."""
    f = np.float32
    cpa = width // 3
    axis, s = cut // cpa, cut % cpa
    lo_v = f(f(lo[axis]) + f(m))
    hi_v = f(f(hi[axis]) - f(m))
    frac = f(f(f(s) + f(1)) / f(f(cpa) + f(1)))
    pos = f(lo_v + f(frac * f(hi_v - lo_v)))
    pl = [0.0, 0.0, 0.0, float(-pos)]
    pl[axis] = 1.0
    return tuple(float(f(x)) for x in pl)


def probe_plane(plane, half, n_probes, i):
    """Probe i of a refinement round: the plane moved by half * (2 i / (n - 1) - 1) along its normal
    (src/gpu/shaders.ts:1505-1514), float32 as written."""
    f = np.float32
    off = f(f(half) * f(f(f(f(2) * f(i)) / f(n_probes - 1)) - f(1)))
    return (plane[0], plane[1], plane[2], float(f(f(plane[3]) - off)))


def next_half(half, n_probes):
    f = np.float32
    return float(f(f(f(2) * f(half)) / f(n_probes - 1))) if n_probes > 1 else 0.0


# ---- Hausdorff samples from the hash, by hand ----------------------------------------------

def wang_hash(s):
    """src/cpu/hausdorff.ts:10-17, on Python integers reduced mod 2^32."""
    s = ((s ^ 61) ^ (s >> 16)) & 0xFFFFFFFF
    s = (s * 9) & 0xFFFFFFFF
    s = (s ^ (s >> 4)) & 0xFFFFFFFF
    s = (s * 0x27D4EB2D) & 0xFFFFFFFF
    s = (s ^ (s >> 15)) & 0xFFFFFFFF
    return s


def sample_weights(t, k):
    """Barycentric weights (w0, w1, w2) of sample k of triangle t, in float32 as the device computes
    them (the reference's formula, src/cpu/hausdorff.ts:52-57, with float32 operands)."""
    f = np.float32
    ha = wang_hash((t * 65537 + k) & 0xFFFFFFFF)
    hb = wang_hash((ha + 0x9E3779B9) & 0xFFFFFFFF)
    ra = f(f(ha & 0xFFFFFF) / f(16777215.0))
    rb = f(f(hb & 0xFFFFFF) / f(16777215.0))
    sqa = f(np.sqrt(ra))
    return f(f(1) - sqa), f(sqa * f(f(1) - rb)), f(sqa * rb), (ha, hb)


def sample_counts(verts, tris, res_fixed=0):
    """Per-triangle sample counts (src/cpu/hausdorff.ts:27-46): float32 areas from float64 arithmetic,
    float64 total, res = max(1000, floor(2000 area)) or the fixed count, floor(res / area * a_t), else
    one sample on every step-th (or every even) triangle."""
    v = np.asarray(verts, np.float64).reshape(-1, 3)
    t = np.asarray(tris, np.int64).reshape(-1, 3)
    areas = []
    for a, b, c in t:
        ab, ac = v[b] - v[a], v[c] - v[a]
        n = (ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0])
        areas.append(np.float32(0.5 * math.sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2])))
    area = 0.0
    for x in areas:
        area += float(x)
    res = float(res_fixed) if res_fixed else max(float(math.floor(2000 * area)), 1000.0)
    nt = len(t)
    step = math.floor(nt / res) if nt > res else 0
    counts = []
    for i, x in enumerate(areas):
        ac = math.floor(res / area * float(x)) if area > 1e-20 else 0
        if nt > res:
            counts.append(ac if ac > 0 else (1 if step > 0 and i % step == 0 else 0))
        else:
            counts.append(ac if ac > 0 else (1 if i % 2 == 0 else 0))
    return counts
