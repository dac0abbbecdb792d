"""Brute-force restatement of the sampled Hausdorff distance in numpy float32, for checking the
oracle's BVH search: the same sample counts (src/cpu/hausdorff.ts:27-48), Wang-hash barycentrics
(src/gpu/hausdorff_shader.ts:88-99) and point-triangle distance (src/gpu/hausdorff_shader.ts:33-51;
sliver triangles, sin^2 of the angle at a below 1e-2, take src/cpu/hausdorff.ts:79-115 in float64,
where the float32 inside test fails), every float32 operation in the order the kernels use, and the
minimum over ALL triangles (no pruning). numpy rounds each float32 operation on its own (no fused multiply-add), so the result is
bit-comparable. Test infrastructure only."""
import numpy as np

F = np.float32


def wang(s):
    s = s.astype(np.uint32)
    s = (s ^ np.uint32(61)) ^ (s >> np.uint32(16))
    s = s * np.uint32(9)
    s = s ^ (s >> np.uint32(4))
    s = s * np.uint32(0x27D4EB2D)
    s = s ^ (s >> np.uint32(15))
    return s


def sample_offsets(verts, tris, fixed_res):
    V = verts
    nt = len(tris)
    areas = np.zeros(nt, np.float32)
    area = 0.0
    for t in range(nt):
        a, b, c = (V[tris[t, k]] for k in range(3))
        abx, aby, abz = float(b[0]) - float(a[0]), float(b[1]) - float(a[1]), float(b[2]) - float(a[2])
        acx, acy, acz = float(c[0]) - float(a[0]), float(c[1]) - float(a[1]), float(c[2]) - float(a[2])
        nx = aby * acz - abz * acy
        ny = abz * acx - abx * acz
        nz = abx * acy - aby * acx
        areas[t] = np.float32(0.5 * np.sqrt(nx * nx + ny * ny + nz * nz))
        area += float(areas[t])
    if fixed_res > 0:
        res = float(fixed_res)
    else:
        res = max(np.floor(2000.0 * area), 1000.0)
    step = int(np.floor(nt / res)) if nt > res else 0
    off = np.zeros(nt + 1, np.int64)
    tot = 0
    for t in range(nt):
        ac = np.floor(res / area * float(areas[t])) if area > 1e-20 else 0.0
        if nt > res:
            N = int(ac) if ac > 0 else (1 if (step > 0 and t % step == 0) else 0)
        else:
            N = int(ac) if ac > 0 else (1 if t % 2 == 0 else 0)
        off[t] = tot
        tot += N
    off[nt] = tot
    return off


def samples(verts, tris, fixed_res):
    off = sample_offsets(verts, tris, fixed_res)
    cnt = np.diff(off)
    t = np.repeat(np.arange(len(tris), dtype=np.uint32), cnt)
    k = (np.arange(off[-1]) - np.repeat(off[:-1], cnt)).astype(np.uint32)
    ha = wang(t * np.uint32(65537) + k)
    hb = wang(ha + np.uint32(0x9E3779B9))
    ra = (ha & np.uint32(0xFFFFFF)).astype(F) / F(16777215.0)
    rb = (hb & np.uint32(0xFFFFFF)).astype(F) / F(16777215.0)
    sqa = np.sqrt(ra)
    w0 = F(1.0) - sqa
    w1 = sqa * (F(1.0) - rb)
    w2 = sqa * rb
    P = verts[tris[t]]  # (n, 3 corners, 3)
    x = w0[:, None] * P[:, 0]
    x = x + w1[:, None] * P[:, 1]
    x = x + w2[:, None] * P[:, 2]
    return x.astype(F)


def dot3(ax, ay, az, bx, by, bz):
    t = ax * bx
    t = t + ay * by
    return t + az * bz


def seg_sq(px, py, pz, ax, ay, az, bx, by, bz):
    abx, aby, abz = bx - ax, by - ay, bz - az
    apx, apy, apz = px - ax, py - ay, pz - az
    ab2 = dot3(abx, aby, abz, abx, aby, abz)
    with np.errstate(all="ignore"):
        t = np.where(ab2 > F(1e-20), dot3(apx, apy, apz, abx, aby, abz) / ab2, F(0.0))
    t = np.where(t < F(0.0), F(0.0), np.where(t > F(1.0), F(1.0), t)).astype(F)
    dx, dy, dz = apx - t * abx, apy - t * aby, apz - t * abz
    return dot3(dx, dy, dz, dx, dy, dz)


def fminx(a, b):
    return np.where(b < a, b, a)


def tri_sq(px, py, pz, T):
    """px..: (n,1) samples; T: (m, 9) triangles -> (n, m) f32 squared distances."""
    ax, ay, az, bx, by, bz, cx, cy, cz = (T[:, i][None, :] for i in range(9))
    abx, aby, abz = bx - ax, by - ay, bz - az
    acx, acy, acz = cx - ax, cy - ay, cz - az
    apx, apy, apz = px - ax, py - ay, pz - az
    nx = aby * acz - abz * acy
    ny = abz * acx - abx * acz
    nz = abx * acy - aby * acx
    n2 = dot3(nx, ny, nz, nx, ny, nz)
    bpx, bpy, bpz = px - bx, py - by, pz - bz
    cpx, cpy, cpz = px - cx, py - cy, pz - cz
    degen = fminx(dot3(apx, apy, apz, apx, apy, apz), fminx(dot3(bpx, bpy, bpz, bpx, bpy, bpz), dot3(cpx, cpy, cpz, cpx, cpy, cpz)))
    d00 = dot3(abx, aby, abz, abx, aby, abz)
    d01 = dot3(abx, aby, abz, acx, acy, acz)
    d02 = dot3(abx, aby, abz, apx, apy, apz)
    d11 = dot3(acx, acy, acz, acx, acy, acz)
    d12 = dot3(acx, acy, acz, apx, apy, apz)
    with np.errstate(all="ignore"):
        inv = F(1.0) / (d00 * d11 - d01 * d01)
        u = (d11 * d02 - d01 * d12) * inv
        v = (d00 * d12 - d01 * d02) * inv
        dist = dot3(nx, ny, nz, apx, apy, apz)
        inside = (dist * dist) / n2
    edge = fminx(seg_sq(px, py, pz, ax, ay, az, bx, by, bz),
                 fminx(seg_sq(px, py, pz, bx, by, bz, cx, cy, cz), seg_sq(px, py, pz, cx, cy, cz, ax, ay, az)))
    with np.errstate(all="ignore"):
        ins = (u >= F(0.0)) & (v >= F(0.0)) & (u + v <= F(1.0))
    out = np.where(n2 < F(1e-20), degen, np.where(ins, inside, edge)).astype(F)
    sliver = ~(n2 < F(1e-20)) & (n2 < F(1e-2) * (d00 * d11))
    if sliver.any():
        out = np.where(sliver, tri_sq64(px, py, pz, T), out).astype(F)
    return out


def seg_sq64(px, py, pz, ax, ay, az, bx, by, bz):
    """src/cpu/hausdorff.ts:69-77 in float64."""
    abx, aby, abz = bx - ax, by - ay, bz - az
    apx, apy, apz = px - ax, py - ay, pz - az
    ab2 = abx * abx + aby * aby + abz * abz
    with np.errstate(all="ignore"):
        t = np.where(ab2 > 1e-20, (apx * abx + apy * aby + apz * abz) / ab2, 0.0)
    t = np.where(t < 0.0, 0.0, np.where(t > 1.0, 1.0, t))
    dx, dy, dz = apx - t * abx, apy - t * aby, apz - t * abz
    return dx * dx + dy * dy + dz * dz


def tri_sq64(px, py, pz, T):
    """src/cpu/hausdorff.ts:79-115 (past its degenerate test) in float64 from the float32 point and
    corners, rounded to float32: the distance of sliver triangles."""
    D = np.float64
    px, py, pz = px.astype(D), py.astype(D), pz.astype(D)
    ax, ay, az, bx, by, bz, cx, cy, cz = (T[:, i][None, :].astype(D) for i in range(9))
    abx, aby, abz = bx - ax, by - ay, bz - az
    acx, acy, acz = cx - ax, cy - ay, cz - az
    apx, apy, apz = px - ax, py - ay, pz - az
    nx, ny, nz = aby * acz - abz * acy, abz * acx - abx * acz, abx * acy - aby * acx
    n2 = nx * nx + ny * ny + nz * nz
    d00 = abx * abx + aby * aby + abz * abz
    d01 = abx * acx + aby * acy + abz * acz
    d02 = abx * apx + aby * apy + abz * apz
    d11 = acx * acx + acy * acy + acz * acz
    d12 = acx * apx + acy * apy + acz * apz
    with np.errstate(all="ignore"):
        inv = 1.0 / (d00 * d11 - d01 * d01)
        u = (d11 * d02 - d01 * d12) * inv
        v = (d00 * d12 - d01 * d02) * inv
        dist = nx * apx + ny * apy + nz * apz
        inside = (dist * dist) / n2
        ins = (u >= 0.0) & (v >= 0.0) & (u + v <= 1.0)
    e0 = seg_sq64(px, py, pz, ax, ay, az, bx, by, bz)
    e1 = seg_sq64(px, py, pz, bx, by, bz, cx, cy, cz)
    e2 = seg_sq64(px, py, pz, cx, cy, cz, ax, ay, az)
    r = np.where(e1 < e0, e1, e0)
    r = np.where(e2 < r, e2, r)
    return np.where(ins, inside, r).astype(F)


def directed(S, verts, tris):
    T = verts[tris].reshape(-1, 9).astype(F)
    worst = F(0.0)
    for s0 in range(0, len(S), 512):
        P = S[s0:s0 + 512]
        d = tri_sq(P[:, 0:1], P[:, 1:2], P[:, 2:3], T)
        worst = max(worst, F(d.min(axis=1).max()))
    return worst


def hausdorff(a, b, fixed_res=0):
    va, ta = a.verts.astype(F), a.tris.astype(np.int64)
    vb, tb = b.verts.astype(F), b.tris.astype(np.int64)
    if len(ta) == 0 or len(tb) == 0:
        return 0.0
    da = directed(samples(va, ta, fixed_res), vb, tb)
    db = directed(samples(vb, tb, fixed_res), va, ta)
    return float(np.sqrt(np.float64(max(da, db))))
