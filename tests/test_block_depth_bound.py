"""The raster kernel's block-local depth bound (raster_device.h,
block_depth_range), checked by brute force in float32 on the CPU.

The raster kernel computes a triangle's depth at a pixel centre (x, y) as
the reference does (GetFragmentInput, CPVulkan/CommandBuffer.Draw.cpp:874-903):
e_i = fl(fl(fl(x - ax_i) * dy_i) - fl(fl(y - ay_i) * dx_i)), w_i = e_i / area
(correctly rounded), depth = fl(fl(fl(z0 w0) + fl(z1 w1)) + fl(z2 w2)).
Before a wave visits a primitive it drops it when a bound [lo, hi] of that
depth over the pixel rectangle it can reach in the wave's block (its box
clipped to the block) fails the depth test against every stored depth.
Dropping is exact only if every pixel's depth lies in [lo, hi]; this file
checks that for every pixel centre of many rectangles,
with the reciprocal the kernel uses (v_rcp_f32) taken as the correctly
rounded one and one ulp either side.  It also checks that the bound is
tight enough to matter, so a bound made loose (a factor lost, a corner
never evaluated) fails too.  Numpy's float32 operations round each step to
nearest as the kernel does, with no contraction."""
import numpy as np

f32 = np.float32


def centre(p, extent):
    """Pixel centre in NDC as the raster kernel computes it (ProcessTriangles,
    CommandBuffer.Draw.cpp:1571-1580)."""
    half = f32(f32(1.0) / f32(extent)) * f32(0.5)
    return (f32(p) / f32(extent) + half) * f32(2.0) - f32(1.0)


def put_edges(P):
    """Edge records of setup_kernel (bin_kernels.h): counter-clockwise order
    after the swap, area = |EdgeFunction(p0, p1, p2)|."""
    (x0, y0), (x1, y1), (x2, y2) = P
    area = (x2 - x0) * (y1 - y0) - (y2 - y0) * (x1 - x0)
    if area < 0:
        area = -area
        pairs = ((P[2], P[1]), (P[0], P[2]), (P[1], P[0]))
    else:
        pairs = ((P[1], P[2]), (P[2], P[0]), (P[0], P[1]))
    ax = np.array([a[0] for a, b in pairs], f32)
    ay = np.array([a[1] for a, b in pairs], f32)
    dx = np.array([b[0] - a[0] for a, b in pairs], f32)
    dy = np.array([b[1] - a[1] for a, b in pairs], f32)
    return ax, ay, dx, dy, f32(area)


def pixel_depths(ax, ay, dx, dy, area, z, xs, ys):
    """The reference's depth at every centre of the grid xs x ys."""
    X = xs[None, :]
    Y = ys[:, None]
    w = [((X - ax[i]) * dy[i] - (Y - ay[i]) * dx[i]) / area for i in range(3)]
    return (z[0] * w[0] + z[1] * w[1]) + z[2] * w[2]


def block_depth_range(ax, ay, dx, dy, area, z, xlo, xhi, ylo, yhi, r):
    """Restatement of raster_device.h block_depth_range with reciprocal r."""
    if not area >= f32(2.0 ** -40):
        return None
    d = [f32(0.0)] * 4
    K = f32(0.0)
    Z = f32(0.0)
    for i in range(3):
        pl = (xlo - ax[i]) * dy[i]
        ph = (xhi - ax[i]) * dy[i]
        ql = (ylo - ay[i]) * dx[i]
        qh = (yhi - ay[i]) * dx[i]
        e = (pl - ql, ph - ql, pl - qh, ph - qh)
        for c in range(4):
            d[c] = d[c] + z[i] * (e[c] * r)
        M = max(abs(pl), abs(ph)) + max(abs(ql), abs(qh))
        K = K + abs(z[i]) * M
        Z = Z + abs(z[i])
    K = K * r
    if not K < f32(2.0 ** 100):
        return None
    s = (d[0] + d[1]) + (d[2] + d[3])
    if not np.isfinite(s):
        return None
    E = (K * f32(2.0 ** -18) + Z * f32(2.0 ** -100)) + f32(2.0 ** -118)
    lo = min(d) - E
    hi = max(d) + E
    if not (np.isfinite(lo) and np.isfinite(hi)):
        return None
    return lo, hi


def _triangles(rng, n, vp_w, vp_h):
    """Vertices in NDC: soup-sized, large, slivers, near-degenerate, and
    vertices on pixel centres; depth per vertex in [0, 1] or wider."""
    out = []
    for k in range(n):
        kind = k % 5
        c = rng.uniform(-1.0, 1.0, 2)
        if kind == 0:
            P = c + rng.normal(0.0, 8.0, (3, 2)) * (2.0 / np.array([vp_w, vp_h]))
        elif kind == 1:
            P = c + rng.normal(0.0, 200.0, (3, 2)) * (2.0 / np.array([vp_w, vp_h]))
        elif kind == 2:
            a = c
            b = c + rng.uniform(-0.3, 0.3, 2)
            P = np.array([a, b, a + (b - a) * 0.5 + rng.uniform(-1e-4, 1e-4, 2)])
        elif kind == 3:
            P = c + rng.uniform(-1e-3, 1e-3, (3, 2))
        else:
            P = np.array([[centre(rng.integers(0, 64), vp_w), centre(rng.integers(0, 48), vp_h)] for _ in range(3)])
        z = rng.uniform(0.0, 1.0, 3) if k % 7 else rng.uniform(-4.0, 4.0, 3)
        out.append((P.astype(f32), z.astype(f32)))
    return out


def _check(seed, n, vp_w, vp_h):
    rng = np.random.Generator(np.random.PCG64(seed))
    checked = useful = 0
    for P, z in _triangles(rng, n, vp_w, vp_h):
        ax, ay, dx, dy, area = put_edges(P)
        # a rectangle of pixels near the triangle: 8x8 blocks and 16x16 tiles
        sx = ((P[:, 0].astype(np.float64) + 1.0) * 0.5 * vp_w).mean()
        sy = ((P[:, 1].astype(np.float64) + 1.0) * 0.5 * abs(vp_h)).mean()
        size = 8 if rng.random() < 0.7 else 16
        px0 = int(np.clip(sx + rng.integers(-12, 12), 0, 4000))
        py0 = int(np.clip(sy + rng.integers(-12, 12), 0, 4000))
        xs = np.array([centre(p, vp_w) for p in range(px0, px0 + size)], f32)
        ys = np.array([centre(p, vp_h) for p in range(py0, py0 + size)], f32)
        xlo, xhi, ylo, yhi = xs.min(), xs.max(), ys.min(), ys.max()
        with np.errstate(all="ignore"):
            depth = pixel_depths(ax, ay, dx, dy, area, z, xs, ys)
            r0 = f32(1.0) / area
            for r in (r0, np.nextafter(r0, f32(0.0)), np.nextafter(r0, f32(np.inf))):
                b = block_depth_range(ax, ay, dx, dy, area, z, xlo, xhi, ylo, yhi, r)
                if b is None:
                    continue
                lo, hi = b
                fin = np.isfinite(depth)
                assert (depth[fin] >= lo).all() and (depth[fin] <= hi).all(), (P, z, px0, py0, lo, hi, depth.min(), depth.max())
                checked += 1
                if fin.all() and hi - lo <= (depth.max() - depth.min()) + 1e-4 * (1.0 + abs(float(depth.max()))):
                    useful += 1
    return checked, useful


def test_block_depth_bound_holds_at_every_pixel():
    checked, useful = _check(3, 1500, 1920.0, 1080.0)
    assert checked > 3000
    # the bound is a few 1e-6 wider than the range of the computed depths
    assert useful > checked // 2


def test_block_depth_bound_negative_height_and_small_viewport():
    checked, _ = _check(5, 800, 64.0, -48.0)
    assert checked > 1500


def test_block_depth_bound_refuses_degenerate_input():
    P = np.array([[0.0, 0.0], [1e-30, 0.0], [0.0, 1e-30]], f32)
    ax, ay, dx, dy, area = put_edges(P)
    z = np.array([0.1, 0.5, 0.9], f32)
    xs = np.array([centre(p, 64.0) for p in range(8)], f32)
    with np.errstate(all="ignore"):
        assert block_depth_range(ax, ay, dx, dy, area, z, xs.min(), xs.max(), xs.min(), xs.max(), f32(1.0) / area) is None
        nan = np.array([np.nan, 0.0, 0.0], f32)
        ax2 = ax.copy(); ax2[0] = np.nan
        assert block_depth_range(ax2, ay, dx, dy, f32(0.5), z, xs.min(), xs.max(), xs.min(), xs.max(), f32(2.0)) is None
        assert block_depth_range(ax, ay, dx, dy, f32(0.5), nan, xs.min(), xs.max(), xs.min(), xs.max(), f32(2.0)) is None
