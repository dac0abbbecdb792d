"""Reference test fixtures restated (tests/fixtures.ts:4-77, src/mesh.ts:19-58)."""
import numpy as np


def square_torus(cw):
    """tests/fixtures.ts:4-20: outer half-width 1, inner 0.5, height 0.5; volume 1.5."""
    R, r, h = 1.0, 0.5, 0.25
    verts = []
    for s, z in [(R, -h), (R, h), (r, -h), (r, h)]:
        verts += [[-s, -s, z], [s, -s, z], [s, s, z], [-s, s, z]]
    tris = []

    def quad(a, b, c, d):
        tris.extend([a, b, c, a, c, d])
    for i in range(4):
        j = (i + 1) % 4
        quad(i, j, 4 + j, 4 + i)
        quad(8 + j, 8 + i, 12 + i, 12 + j)
        quad(4 + i, 4 + j, 12 + j, 12 + i)
        quad(j, i, 8 + i, 8 + j)
    return cw.Mesh(np.array(verts, np.float32), np.array(tris, np.uint32))


def extrude_polygon(cw, poly, height):
    """tests/fixtures.ts:28-62: ear-clipped CCW polygon extruded along +z."""
    n = len(poly)
    verts = [[x, y, z] for z in (0.0, height) for x, y in poly]
    tris = []
    idx = list(range(n))

    def area2(a, b, c):
        return (poly[b][0] - poly[a][0]) * (poly[c][1] - poly[a][1]) - (poly[b][1] - poly[a][1]) * (poly[c][0] - poly[a][0])

    def inside(p, a, b, c):
        return area2(a, b, p) >= 0 and area2(b, c, p) >= 0 and area2(c, a, p) >= 0
    guard = 0
    while len(idx) > 3 and guard < 1000:
        guard += 1
        clipped = False
        for i in range(len(idx)):
            a, b, c = idx[(i - 1) % len(idx)], idx[i], idx[(i + 1) % len(idx)]
            if area2(a, b, c) <= 0:
                continue
            if any(p not in (a, b, c) and inside(p, a, b, c) for p in idx):
                continue
            tris += [a, c, b, n + a, n + b, n + c]
            idx.pop(i)
            clipped = True
            break
        if not clipped:
            break
    tris += [idx[0], idx[2], idx[1], n + idx[0], n + idx[1], n + idx[2]]
    for i in range(n):
        j = (i + 1) % n
        tris += [i, j, n + j, i, n + j, n + i]
    return cw.Mesh(np.array(verts, np.float32), np.array(tris, np.uint32))


def stairs(cw):
    """tests/fixtures.ts:23-25: 3-step welded staircase prism, volume 6."""
    return extrude_polygon(cw, [[0, 0], [3, 0], [3, 1], [2, 1], [2, 2], [1, 2], [1, 3], [0, 3]], 1.0)


def l_prism(cw):
    """tests/fixtures.ts:75-77: welded L prism with one concave vertical edge."""
    return extrude_polygon(cw, [[0, 0], [2, 0], [2, 1], [1, 1], [1, 2], [0, 2]], 1.0)


def park_miller_cloud(n=1000):
    """tests/cpu.test.ts:65-69: 3n floats from the Park-Miller LCG (seed 1, x16807), in [-1, 1)."""
    pts = np.zeros(3 * n, np.float32)
    s = 1
    for i in range(3 * n):
        s = (s * 16807) % 2147483647
        pts[i] = s / 2147483647 * 2 - 1
    return pts.reshape(-1, 3)
