"""Independent statements of the apparent-pair scans, for the tests:

* the scans' predicates written from their definitions (the youngest
  cofacet by (largest edge rank, colex), and, for H2, "the triangle is the
  largest of that tetrahedron's four faces" with faces ordered by (largest
  edge rank, colex3)), with Python sets and dicts;
* the reference's bounded merge loops (kernels/apparent_pair_kernel.py:
  133-156, kernels/apparent_pair_h2_kernel.py:194-253) walked literally,
  one step at a time, for the step-cap cases.
"""
from __future__ import annotations

import numpy as np


def colex3(a, b, c):
    a, b, c = sorted((int(a), int(b), int(c)))
    return c * (c - 1) * (c - 2) // 6 + b * (b - 1) // 2 + a


def colex4(a, b, c, d):
    a, b, c, d = sorted((int(a), int(b), int(c), int(d)))
    return d * (d - 1) * (d - 2) * (d - 3) // 24 + colex3(a, b, c)


def _nbrs(v, ptr, idx, rank):
    s, e = int(ptr[v]), int(ptr[v + 1])
    return dict(zip(idx[s:e].tolist(), rank[s:e].tolist()))


def h1_definition(ei, ej, ranks, ptr, idx, rank):
    """Non-forest edge e = (u, v): a common neighbour w with both edges
    ranked below e; the smallest such w gives the triangle's colex."""
    flag = np.zeros(len(ranks), dtype=np.int32)
    colex = np.zeros(len(ranks), dtype=np.int64)
    for k, re in enumerate(ranks):
        u, v = sorted((int(ei[re]), int(ej[re])))
        nu, nv = _nbrs(u, ptr, idx, rank), _nbrs(v, ptr, idx, rank)
        ws = [w for w in nu if w in nv and nu[w] < re and nv[w] < re]
        if ws:
            flag[k] = 1
            colex[k] = colex3(u, v, min(ws))
    return flag, colex


def h2_definition(tri, ptr, idx, rank):
    """tri: (T, 4) rows (a, b, c, r_t).  The youngest cofacet is the
    tetrahedron t + {w} of least (largest edge rank, colex4) over the common
    neighbours w; t is apparent when it is the largest face of that
    tetrahedron, faces ordered by (largest edge rank, colex3)."""
    T = len(tri)
    flag = np.zeros(T, dtype=np.int32)
    colex = np.zeros(T, dtype=np.int64)
    for k in range(T):
        a, b, c, rt = (int(x) for x in tri[k])
        nb = {v: _nbrs(v, ptr, idx, rank) for v in (a, b, c)}
        common = [w for w in nb[a] if w in nb[b] and w in nb[c]]
        if not common:
            continue
        best = min(common, key=lambda w: (max(rt, nb[a][w], nb[b][w], nb[c][w]),
                                           colex4(a, b, c, w)))
        verts = sorted((a, b, c, best))
        er = {}
        for v in verts:
            nv = nb[v] if v in nb else _nbrs(v, ptr, idx, rank)
            for x in verts:
                if x != v:
                    er[(v, x)] = nv[x]
        faces = [tuple(v for v in verts if v != drop) for drop in verts]
        top = max(faces, key=lambda f: (max(er[(f[0], f[1])], er[(f[0], f[2])],
                                            er[(f[1], f[2])]), colex3(*f)))
        if top == (a, b, c):
            flag[k] = 1
            colex[k] = colex4(*verts)
    return flag, colex


def h1_reference_walk(u, v, re, ptr, idx, rank, md):
    """The reference's per-edge merge loop with its 2 md step bound."""
    iu, iv = int(ptr[u]), int(ptr[v])
    ue, ve = int(ptr[u + 1]), int(ptr[v + 1])
    for _ in range(2 * md):
        if iu >= ue or iv >= ve:
            break
        nu, nv = idx[iu], idx[iv]
        if nu == nv:
            if rank[iu] < re and rank[iv] < re:
                return int(nu)
            iu += 1
            iv += 1
        elif nu < nv:
            iu += 1
        else:
            iv += 1
    return -1


def h2_reference_walk(a, b, c, rt, ptr, idx, rank, md):
    """The reference's per-triangle three-way merge with its 3 md step
    bound: returns the vertex d of the youngest cofacet among the matches
    the loop examines, if that cofacet's largest rank is rt, else -1."""
    p = [int(ptr[a]), int(ptr[b]), int(ptr[c])]
    e = [int(ptr[a + 1]), int(ptr[b + 1]), int(ptr[c + 1])]
    best = None
    for _ in range(3 * md):
        if not all(p[q] < e[q] for q in range(3)):
            break
        n = [int(idx[p[q]]) for q in range(3)]
        m = max(n)
        match = n[0] == n[1] == n[2]
        if match:
            key = (max(rt, *(int(rank[p[q]]) for q in range(3))), colex4(a, b, c, n[0]))
            if best is None or key < best[0]:
                best = (key, n[0])
        for q in range(3):
            if match or n[q] < m:
                p[q] += 1
    if best is None or best[0][0] != rt:
        return -1
    return best[1]


def h2_oldest_face(a, b, c, rt, d, ptr, idx, rank):
    """The reference's d < s test (kernels/apparent_pair_h2_kernel.py:
    255-305) for a youngest cofacet t + {d} of rank rt."""
    if d < a:
        return True
    if d >= c:
        return False
    nb, na = _nbrs(b, ptr, idx, rank), _nbrs(a, ptr, idx, rank)
    if nb[c] == rt:
        return False
    if na[c] == rt:
        return d < b
    return True


def triangles(ei, ej, ptr, idx, rank):
    """All triangles of the flag complex as rows (a, b, c, largest edge
    rank), a < b < c, by brute force over each edge's common neighbours."""
    rows = []
    for e in range(len(ei)):
        u, v = sorted((int(ei[e]), int(ej[e])))
        nu, nv = _nbrs(u, ptr, idx, rank), _nbrs(v, ptr, idx, rank)
        for w in nu:
            if w in nv and nu[w] < e and nv[w] < e:
                rows.append(sorted((u, v, w)) + [e])
    return np.array(rows, dtype=np.int64).reshape(-1, 4)
