"""Graph and matrix builders shared by the tests and the fixture script.
The small graphs are the ones of the reference's colouring tests
(tests/metal/graphColoringTest.cc:222-505) plus a few stencils."""
import numpy as np


def csr_from_edges(n, edges, diag=True):
    rows = [set() for _ in range(n)]
    for a, b in edges:
        rows[a].add(b)
        rows[b].add(a)
    if diag:
        for i in range(n):
            rows[i].add(i)
    ptr = [0]
    col = []
    for r in rows:
        col.extend(sorted(r))
        ptr.append(len(col))
    return np.array(ptr, np.int32), np.array(col, np.int32)


def grid2d(nx, ny):
    e = []
    for y in range(ny):
        for x in range(nx):
            i = y * nx + x
            if x + 1 < nx:
                e.append((i, i + 1))
            if y + 1 < ny:
                e.append((i, i + nx))
    return csr_from_edges(nx * ny, e)


def small_graphs():
    g = {
        "path5": csr_from_edges(5, [(0, 1), (1, 2), (2, 3), (3, 4)]),
        "triangle": csr_from_edges(3, [(0, 1), (1, 2), (0, 2)]),
        "cycle4": csr_from_edges(4, [(0, 1), (1, 2), (2, 3), (3, 0)]),
        "k5": csr_from_edges(5, [(a, b) for a in range(5) for b in range(a + 1, 5)]),
        "grid10x10": grid2d(10, 10),
        "grid8x8": grid2d(8, 8),
        "two_components": csr_from_edges(7, [(0, 1), (1, 2), (3, 4), (4, 5), (5, 6), (6, 3)]),
        "star": csr_from_edges(6, [(0, k) for k in range(1, 6)]),
    }
    # A random sparse symmetric graph (fixed seed).
    rng = np.random.default_rng(1234)
    n = 60
    edges = set()
    while len(edges) < 150:
        a, b = rng.integers(0, n, 2)
        if a != b:
            edges.add((min(a, b), max(a, b)))
    g["random60"] = csr_from_edges(n, sorted(edges))
    return g


def large_patterns():
    """Benchmark-size sparsity patterns, from the product's generators."""
    import gpusolver as gs
    out = {}
    A = gs.poisson2d(256, 256)
    out["poisson2d_256"] = (A.ptr, A.col)
    A = gs.poisson3d(128)
    out["poisson3d_128"] = (A.ptr, A.col)
    E, _, _ = gs.elasticity(24)
    gp, gc = gs.block_graph(E, 3)
    out["elasticity24_blockgraph"] = (gp, gc)
    out["elasticity24_scalar"] = (E.ptr, E.col)
    return out
