"""Seeded synthetic oriented point clouds (pure numpy, no GPU).

``fixture_points`` restates the reference's fixture shapes
(tools/make_fixture.py:29-58) so its published golden reports can be
regenerated bit for bit.  The ``workload_*`` generators build the benchmark
workloads: a sphere scan, a non-uniform noisy torus, a six-ellipsoid union
with 10x density variation, and a displaced sphere.
"""
import numpy as np

SEED = 2026


def fixture_points(shape, count, seed=SEED):
    """Shapes of tools/make_fixture.py:29-58 with numpy's default_rng."""
    rng = np.random.default_rng(seed)
    if shape in ("plane", "sheets"):
        p = rng.uniform(-1, 1, (count, 3))
        p[:, 2] = 0
        n = np.zeros_like(p)
        n[:, 2] = 1
        if shape == "sheets":
            sign = np.where(np.arange(count) % 2, 1., -1.)
            p[:, 2] = 0.035 * sign
            n[:, 2] = sign
        return p, n
    if shape in ("sphere", "cluster", "disconnected"):
        n = rng.normal(size=(count, 3))
        if shape == "cluster":
            n[:9 * count // 10] = rng.normal(scale=0.12, size=(9 * count // 10, 3)) + [0, 0, 1]
        n /= np.linalg.norm(n, axis=1)[:, None]
        if shape == "disconnected":
            p = 0.25 * n
            p[:, 0] += np.where(np.arange(count) % 2, 0.65, -0.65)
            return p, n
        return n, n.copy()
    if shape != "torus":
        raise ValueError(f"Unknown fixture shape: {shape}")
    u, v = rng.uniform(0, 2 * np.pi, size=(2, count))
    p = np.array([(0.7 + 0.26 * np.cos(v)) * np.cos(u),
                  (0.7 + 0.26 * np.cos(v)) * np.sin(u), 0.26 * np.sin(v)]).T
    n = np.array([np.cos(v) * np.cos(u), np.cos(v) * np.sin(u), np.sin(v)]).T
    return p, n


def fixture_input(shape, count, seed=SEED, precision="float", input_scale=1.0, translation=(0., 0., 0.)):
    """Points/normals exactly as a reference fixture's input.pts (make_fixture.py:80-88)."""
    p, n = fixture_points(shape, count, seed)
    p = p * input_scale + np.asarray(translation)
    dtype = np.float32 if precision == "float" else np.float64
    return p.astype(dtype), n.astype(dtype)


def ownership_input(dtype):
    """The split-plane ownership probes of the reference's threshold fixtures
    (tools/check_thresholds.py:25-57): two corner points fixing an exact
    scale-2 normalisation, then for every dyadic split of depths 1..4 inside
    (1/4, 3/4), on each axis, a point just below, on and just above the split
    (model = (cube[[2, 0, 1]] - 1/2) * 2), and five points exercising a
    cancelling aggregate, multiplicity and a zero normal.  Returns (points,
    normals) of ``dtype``; the points must be the reference's bit for bit, so
    the construction follows it step for step."""
    positions = [np.full(3, -.5, dtype=dtype), np.full(3, .5, dtype=dtype)]
    normals = [[1., 0., 0.], [0., 1., 0.]]
    for depth in range(1, 5):
        for numerator in range(1, 1 << depth):
            split = dtype(numerator / (1 << depth))
            if not .25 < split < .75:
                continue
            for axis in range(3):
                for coordinate in (np.nextafter(split, dtype(-np.inf)), split, np.nextafter(split, dtype(np.inf))):
                    cube = np.array([.40625, .546875, .59375], dtype=dtype)
                    cube[axis] = coordinate
                    positions.append((cube[[2, 0, 1]] - dtype(.5)) * dtype(2))
                    normals.append([0., 0., 2.])
    for position, normal in [([.21, -.19, .23], [1, 0, 0]), ([.21, -.19, .23], [-3, 0, 0]),
                             ([-.23, .19, -.21], [0, 2, 0]), ([-.23, .19, -.21], [0, 5, 0]),
                             ([.11, .13, -.17], [0, 0, 0])]:
        positions.append(np.array(position, dtype=dtype))
        normals.append(normal)
    data = np.column_stack([positions, normals]).astype(dtype)
    return np.ascontiguousarray(data[:, :3]), np.ascontiguousarray(data[:, 3:])


def fixture_input_text(points, normals):
    """The input.pts text whose SHA-256 the reference reports record."""
    import io
    buf = io.BytesIO()
    np.savetxt(buf, np.column_stack([points, normals]), fmt="%.17g")
    return buf.getvalue()


def cube_queries(points, count, rng, scale=1.1):
    """Uniform queries in the bounding cube (bounding box side max * scale, centred)."""
    lo, hi = points.min(axis=0).astype(np.float64), points.max(axis=0).astype(np.float64)
    width = float(np.max(hi - lo)) * scale
    center = (hi + lo) / 2
    return (rng.uniform(-0.5, 0.5, size=(count, 3)) * width + center).astype(points.dtype)


def workload_sphere(count=20000, queries=5000, seed=SEED, dtype=np.float32):
    """Uniform unit sphere, exact outward normals; queries = inputs + uniform cube points."""
    rng = np.random.default_rng(seed)
    n = rng.normal(size=(count, 3))
    n /= np.linalg.norm(n, axis=1)[:, None]
    p = n.copy()
    q = np.concatenate([p, cube_queries(p, queries, np.random.default_rng(seed + 1))])
    return p.astype(dtype), n.astype(dtype), q.astype(dtype)


def _tangent_tilt(n, max_degrees, rng):
    """Rotate unit normals by a uniform angle <= max_degrees about a random tangent axis."""
    helper = rng.normal(size=n.shape)
    axis = np.cross(n, helper)
    axis /= np.linalg.norm(axis, axis=1)[:, None]
    angle = np.radians(rng.uniform(0, max_degrees, size=len(n)))[:, None]
    tilted = n * np.cos(angle) + np.cross(axis, n) * np.sin(angle)
    return tilted / np.linalg.norm(tilted, axis=1)[:, None]


def workload_torus(count=250000, queries=20000, seed=SEED, dtype=np.float32, R=0.6, r=0.2):
    """Torus, x>0 half sampled 4x denser, position noise 0.002, normals tilted <=5 deg, 1% zeroed."""
    rng = np.random.default_rng(seed)
    pts, nrm = [], []
    have = 0
    while have < count:
        m = 2 * (count - have) + 1024
        u, v = rng.uniform(0, 2 * np.pi, size=(2, m))
        # area element of the torus is (R + r cos v): accept proportionally for uniform area
        area = rng.uniform(0, 1, m) < (R + r * np.cos(v)) / (R + r)
        x = (R + r * np.cos(v)) * np.cos(u)
        keep = area & (rng.uniform(0, 1, m) < np.where(x > 0, 1.0, 0.25))
        u, v = u[keep], v[keep]
        p = np.stack([(R + r * np.cos(v)) * np.cos(u), (R + r * np.cos(v)) * np.sin(u), r * np.sin(v)], axis=1)
        n = np.stack([np.cos(v) * np.cos(u), np.cos(v) * np.sin(u), np.sin(v)], axis=1)
        pts.append(p)
        nrm.append(n)
        have += len(p)
    p = np.concatenate(pts)[:count]
    n = np.concatenate(nrm)[:count]
    p = p + rng.normal(scale=0.002, size=p.shape)
    n = _tangent_tilt(n, 5.0, rng)
    zero = rng.choice(count, size=count // 100, replace=False)
    n[zero] = 0
    q = np.concatenate([p, cube_queries(p, queries, np.random.default_rng(seed + 1))])
    return p.astype(dtype), n.astype(dtype), q.astype(dtype)


class EllipsoidUnion:
    """Smooth union (soft-min, sharpness s) of seeded random ellipsoids."""

    def __init__(self, count=6, seed=SEED, smooth=0.04):
        rng = np.random.default_rng(seed)
        self.centers = rng.uniform(-0.35, 0.35, size=(count, 3))
        self.axes = rng.uniform(0.15, 0.45, size=(count, 3))
        self.rot = np.stack([np.linalg.qr(rng.normal(size=(3, 3)))[0] for _ in range(count)])
        self.smooth = smooth
        # value-noise lattice for the sampling density (Perlin-style smooth weights)
        self.lattice = rng.uniform(0, 1, size=(9, 9, 9))

    def value_and_gradient(self, x):
        """Soft-min (sharpness ``smooth``) of |A^-1 R^T (x-c)| - 1 over the ellipsoids, and its gradient."""
        s = self.smooth
        f = np.empty((len(x), len(self.centers)))
        g = np.empty((len(x), len(self.centers), 3))
        for k in range(len(self.centers)):
            local = (x - self.centers[k]) @ self.rot[k]  # R^T (x - c) as row vectors
            scaled = local / self.axes[k]
            rad = np.sqrt((scaled * scaled).sum(axis=1))
            f[:, k] = rad - 1.0
            # gradient of |A^-1 R^T (x-c)| = R A^-2 R^T (x-c) / rad
            g[:, k] = ((scaled / self.axes[k]) @ self.rot[k].T) / np.maximum(rad, 1e-12)[:, None]
        m = f.min(axis=1, keepdims=True)
        w = np.exp(-(f - m) / s)
        total = w.sum(axis=1)
        value = m[:, 0] - s * np.log(total)
        grad = np.einsum("nk,nka->na", w, g) / total[:, None]
        return value, grad

    def density(self, x):
        """Smooth value noise in [0.1, 1]: 10x variation over the surface."""
        t = np.clip((x + 1.0) * 4.0, 0, 7.999)
        i = np.floor(t).astype(int)
        f = t - i
        f = f * f * (3 - 2 * f)
        out = np.zeros(len(x))
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    w = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * (f[:, 2] if dz else 1 - f[:, 2])
                    out += w * self.lattice[i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz]
        lo, hi = self.lattice.min(), self.lattice.max()
        return 0.1 + 0.9 * (out - lo) / (hi - lo)

    def project(self, x, steps=12, tol=1e-12):
        """Newton projection onto the zero set; iterates only unconverged points."""
        x = x.copy()
        active = np.arange(len(x))
        for _ in range(steps):
            v, g = self.value_and_gradient(x[active])
            done = np.abs(v) < tol
            step = (v / np.maximum((g * g).sum(axis=1), 1e-20))[:, None] * g
            x[active[~done]] -= step[~done]
            active = active[~done]
            if not len(active):
                break
        v, g = self.value_and_gradient(x)
        return x, v, g


def workload_ellipsoids(count=800000, seed=SEED, dtype=np.float32, noise=0.001):
    """Six-ellipsoid smooth union; density varies 10x; exact analytic normals."""
    shape = EllipsoidUnion(seed=seed)
    rng = np.random.default_rng(seed + 7)
    pts, nrm = [], []
    have = 0
    while have < count:
        m = 3 * (count - have) + 4096
        k = rng.integers(0, len(shape.centers), size=m)
        u = rng.normal(size=(m, 3))
        u /= np.linalg.norm(u, axis=1)[:, None]
        x0 = shape.centers[k] + np.einsum("nij,nj->ni", shape.rot[k], shape.axes[k] * u)
        x, v, g = shape.project(x0)
        ok = np.abs(v) < 1e-9
        ok &= rng.uniform(0, 1, m) < shape.density(x)
        x, g = x[ok], g[ok]
        pts.append(x)
        nrm.append(g / np.linalg.norm(g, axis=1)[:, None])
        have += len(x)
    p = np.concatenate(pts)[:count]
    n = np.concatenate(nrm)[:count]
    p = p + rng.normal(scale=noise, size=p.shape)
    return p.astype(dtype), n.astype(dtype)


def workload_batch(clouds=64, seed=SEED, dtype=np.float32, smallest=5000, largest=200000):
    """Many independent clouds: sizes log-uniform in [smallest, largest], shapes
    cycling sphere / torus / six-ellipsoid union (own seeds), scale U[0.2, 2],
    translation U[-5, 5]^3, depth 7 or 8.  Returns [(points, normals, depth)]."""
    rng = np.random.default_rng(seed + 101)
    out = []
    for i in range(clouds):
        count = int(np.exp(rng.uniform(np.log(smallest), np.log(largest))))
        kind = i % 3
        if kind == 0:
            p, n, _ = workload_sphere(count, queries=0, seed=seed + 1000 + i, dtype=np.float64)
        elif kind == 1:
            p, n, _ = workload_torus(count, queries=0, seed=seed + 1000 + i, dtype=np.float64)
        else:
            p, n = workload_ellipsoids(count, seed=seed + 1000 + i, dtype=np.float64)
        p = p * rng.uniform(0.2, 2.0) + rng.uniform(-5.0, 5.0, size=3)
        out.append((p.astype(dtype), n.astype(dtype), int(rng.integers(7, 9))))
    return out


def workload_displaced_sphere(count=100000, seed=SEED, dtype=np.float64, amplitude=0.05):
    """Unit sphere displaced radially by amplitude*sin(6 theta)*cos(4 phi), analytic normals."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(count, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    theta = np.arccos(np.clip(d[:, 2], -1, 1))
    phi = np.arctan2(d[:, 1], d[:, 0])
    r = 1 + amplitude * np.sin(6 * theta) * np.cos(4 * phi)
    dr_dtheta = amplitude * 6 * np.cos(6 * theta) * np.cos(4 * phi)
    dr_dphi = -amplitude * 4 * np.sin(6 * theta) * np.sin(4 * phi)
    e_theta = np.stack([np.cos(theta) * np.cos(phi), np.cos(theta) * np.sin(phi), -np.sin(theta)], axis=1)
    e_phi = np.stack([-np.sin(phi), np.cos(phi), np.zeros_like(phi)], axis=1)
    p = d * r[:, None]
    # surface x = r(theta,phi) d: normal proportional to r d - dr/dtheta e_theta - dr/dphi/sin(theta) e_phi
    sin_t = np.maximum(np.sin(theta), 1e-12)
    n = r[:, None] * d - dr_dtheta[:, None] * e_theta - (dr_dphi / sin_t)[:, None] * e_phi
    n /= np.linalg.norm(n, axis=1)[:, None]
    return p.astype(dtype), n.astype(dtype)


def leaf_depths(keys, cube_points, depth):
    """Deepest level whose cell (ceil(p*2^d)-1, clamped) is an active key (field.py:406-412)."""
    keys = np.sort(np.asarray(keys, dtype=np.int64))
    p = np.asarray(cube_points, dtype=np.float64)
    leaf = np.full(len(p), -1)
    for d in range(depth, -1, -1):
        n = 1 << d
        c = np.clip(np.ceil(p * n).astype(np.int64) - 1, 0, n - 1)
        k = (np.int64(d) << 48) | (c[:, 0] << 32) | (c[:, 1] << 16) | c[:, 2]
        pos = np.minimum(np.searchsorted(keys, k), len(keys) - 1)
        hit = (keys[pos] == k) & (leaf < 0)
        leaf[hit] = d
    leaf[leaf < 0] = 0
    return leaf


def interface_queries(keys, depth, model_to_cube, cube_to_model, points, count=12000, seed=SEED):
    """Pairs of points one finest cell apart whose leaf depths differ (refinement
    interfaces), in model space; ``count`` points (count/2 pairs)."""
    rng = np.random.default_rng(seed)
    m2c = np.asarray(model_to_cube, dtype=np.float64).reshape(4, 4)
    c2m = np.asarray(cube_to_model, dtype=np.float64).reshape(4, 4)
    cube = points.astype(np.float64) @ m2c[:3, :3].T + m2c[:3, 3]
    h = 1.0 / (1 << depth)
    out = []
    have = 0
    for _ in range(64):
        idx = rng.integers(0, len(cube), size=8 * count)
        a = cube[idx] + rng.uniform(-3 * h, 3 * h, size=(len(idx), 3))
        axis = rng.integers(0, 3, size=len(idx))
        b = a.copy()
        b[np.arange(len(idx)), axis] += h
        ok = (a.min(axis=1) > 0) & (a.max(axis=1) < 1) & (b.min(axis=1) > 0) & (b.max(axis=1) < 1)
        a, b = a[ok], b[ok]
        diff = leaf_depths(keys, a, depth) != leaf_depths(keys, b, depth)
        pairs = np.stack([a[diff], b[diff]], axis=1).reshape(-1, 3)
        out.append(pairs)
        have += len(pairs)
        if have >= count:
            break
    cube_q = np.concatenate(out)[:count]
    return (cube_q @ c2m[:3, :3].T + c2m[:3, 3]).astype(points.dtype)
