"""Seeded input generators shared by the tests and bench.py (the workloads of
the benchmark configurations and smaller parity cases)."""
import numpy as np


def points(rng, n, dist):
    if n == 0:
        return None
    if dist == "cube":
        return rng.random((n, 3))
    if dist == "sphere":
        x = rng.normal(size=(n, 3))
        return x / np.linalg.norm(x, axis=1)[:, None]
    if dist == "shell":
        x = rng.normal(size=(n, 3))
        x /= np.linalg.norm(x, axis=1)[:, None]
        return x * rng.uniform(0.995, 1.005, n)[:, None]
    if dist == "clusters":
        c = rng.random((8, 3)) * 0.8 + 0.1
        return np.clip(c[rng.integers(0, 8, n)] + 0.02 * rng.normal(size=(n, 3)), 0, 1)
    if dist == "plummer":
        out = np.empty((0, 3))
        while out.shape[0] < n:
            m = 2 * (n - out.shape[0]) + 16
            u = rng.uniform(1e-6, 1, m)
            r = 0.1 / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
            d = rng.normal(size=(m, 3))
            d /= np.linalg.norm(d, axis=1)[:, None]
            x = 0.5 + r[:, None] * d
            x = x[np.all((x >= 0) & (x <= 1), axis=1)]
            out = np.concatenate([out, x])
        return out[:n]
    if dist == "plane":
        x = rng.random((n, 3))
        x[:, 2] = 0.25
        return x
    raise ValueError(dist)


def strengths(rng, n, ifcharge=True, ifdipole=False, unit_dipvec=True):
    q = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)) if ifcharge else None
    ds = dv = None
    if ifdipole:
        ds = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
        if unit_dipvec:
            z = rng.uniform(-1, 1, n)
            ph = rng.uniform(0, 2 * np.pi, n)
            s = np.sqrt(1 - z * z)
            dv = np.stack([s * np.cos(ph), s * np.sin(ph), z], 1)
        else:
            dv = rng.normal(size=(n, 3))
    return q, ds, dv


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


# tolerance implied by iprec (src/lfmm3dpart.f:277-285)
IPREC_EPS = {-2: .5e0, -1: .5e-1, 0: .5e-2, 1: .5e-3, 2: .5e-6, 3: .5e-9, 4: .5e-12, 5: .5e-15}


def stream4_requests(seed=4, nreq=512):
    """The Laplace serving stream: lfmm3dpartself requests with N
    log-uniform in [1e4, 2e6], points from the unit cube, a sphere shell
    (r in [0.995, 1.005]) or 8 Gaussian clusters (sigma 0.02) chosen
    uniformly, iprec uniform in {1, 2, 3}, charges with U[-1,1] parts.
    Returns (iprec, points, charges) per request."""
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(nreq):
        n = int(10 ** rng.uniform(4, np.log10(2e6)))
        dist = ["cube", "shell", "clusters"][rng.integers(0, 3)]
        x = points(rng, n, dist)
        q = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
        reqs.append((int(rng.integers(1, 4)), x, q))
    return reqs


def stream5_requests(seed=5, nreq=256):
    """The mixed serving stream: N log-uniform in [5e3, 5e5], Plummer points
    (a = 0.1, centred in the unit cube, points outside rejected), Laplace or
    Helmholtz (zk uniform in [1, 30]) per request, iprec in {2, 3}, charges
    and dipoles (U[-1,1] parts, unit dipole vectors).  Returns (zk or None,
    iprec, points, charge, dipstr, dipvec) per request."""
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(nreq):
        n = int(10 ** rng.uniform(np.log10(5e3), np.log10(5e5)))
        x = points(rng, n, "plummer")
        zk = complex(rng.uniform(1, 30), 0.0) if rng.integers(0, 2) else None
        iprec = int(rng.integers(2, 4))
        q = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
        ds = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
        z = rng.uniform(-1, 1, n)
        ph = rng.uniform(0, 2 * np.pi, n)
        dv = np.stack([np.sqrt(1 - z * z) * np.cos(ph), np.sqrt(1 - z * z) * np.sin(ph), z], axis=1)
        reqs.append((zk, iprec, x, q, ds, dv))
    return reqs


def icosphere(level, radius=1.0):
    """Triangles (n, 3, 3) of the unit icosahedron refined `level` times
    (20 * 4^level triangles) and projected to the sphere, counter-clockwise
    seen from outside (normals (v2-v1)x(v3-v1) point outward)."""
    t = (1.0 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(level):
        cache, nf = {}, []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    V = np.array(verts) * radius
    return V[np.array(f)]

