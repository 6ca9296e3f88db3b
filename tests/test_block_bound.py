"""The emit kernels' block-coverage bound (bin_kernels.h,
pair_blocks_covered), checked by brute force in float32 on the CPU.

The raster kernel's edge value at a pixel centre is
fl(fl(fl(x - ax) * dy) - fl(fl(y - ay) * dx)), with the centre
fl(fl(fl(p / extent) + half) * 2) - 1 (ProcessTriangles,
CPVulkan/CommandBuffer.Draw.cpp:1571-1580, GetFragmentInput :874-898).  The
emit kernels drop a block when one edge's value, taken at a single corner
chosen by the signs of dy and dx, is negative.  Dropping must be sound:
every pixel of the block then has a negative value (a NaN keeps a pixel).
And the single corner must decide exactly as the maximum over all four
corners does, so that a mistaken corner choice (a flipped sign test, x and
y swapped, a viewport of negative height) fails here.  Numpy's float32
operations round each step to nearest as the kernel does, with no
contraction."""
import numpy as np

f32 = np.float32


def centre(p, extent):
    half = f32(f32(1.0) / f32(extent)) * f32(0.5)
    return (f32(p) / f32(extent) + half) * f32(2.0) - f32(1.0)


def edge(x, y, ax, ay, dx, dy):
    return (x - ax) * dy - (y - ay) * dx


def corner_bound(xs, ys, ax, ay, dx, dy):
    """The kernel's bound: A at x = (dy > 0 ? xhi : xlo) less B at y = (dx < 0 ? yhi : ylo)."""
    xlo, xhi = min(xs), max(xs)
    ylo, yhi = min(ys), max(ys)
    x = xhi if dy > 0 else xlo
    y = yhi if dx < 0 else ylo
    return (x - ax) * dy - (y - ay) * dx


def _edges(rng, n):
    """Random triangle edges in NDC: general, tiny slivers, axis-aligned and
    vertices on pixel centres of a 64x48 viewport."""
    out = []
    for k in range(n):
        kind = k % 4
        if kind == 0:
            a = rng.uniform(-1.2, 1.2, 2)
            b = rng.uniform(-1.2, 1.2, 2)
        elif kind == 1:
            a = rng.uniform(-1, 1, 2)
            b = a + rng.uniform(-1e-3, 1e-3, 2)
        elif kind == 2:
            a = rng.uniform(-1, 1, 2)
            b = a.copy()
            b[rng.integers(0, 2)] += rng.uniform(-1, 1)
        else:
            a = np.array([centre(rng.integers(0, 64), 64.0), centre(rng.integers(0, 48), 48.0)])
            b = np.array([centre(rng.integers(0, 64), 64.0), centre(rng.integers(0, 48), 48.0)])
        a = a.astype(f32)
        b = b.astype(f32)
        out.append((a[0], a[1], b[0] - a[0], b[1] - a[1]))
    return out


def test_single_corner_bound_is_sound_and_exact():
    rng = np.random.Generator(np.random.PCG64(7))
    edges = _edges(rng, 400)
    culled = kept = 0
    for vp_w, vp_h in ((64.0, 48.0), (64.0, -48.0), (37.0, 29.0)):
        for ax, ay, dx, dy in edges:
            for _ in range(6):
                x0 = int(rng.integers(0, 60)); x1 = x0 + int(rng.integers(1, 9))
                y0 = int(rng.integers(0, 44)); y1 = y0 + int(rng.integers(1, 9))
                px = np.arange(x0, x1)
                py = np.arange(y0, y1)
                cx = np.array([centre(p, vp_w) for p in px], f32)
                cy = np.array([centre(p, vp_h) for p in py], f32)
                xs = (cx[0], cx[-1])
                ys = (cy[0], cy[-1])
                with np.errstate(all="ignore"):
                    bound = corner_bound(xs, ys, ax, ay, dx, dy)
                    corners = [edge(x, y, ax, ay, dx, dy) for x in xs for y in ys]
                    every = edge(cx[None, :], cy[:, None], ax, ay, dx, dy)
                drop = bool(bound < 0)
                # the four-corner rule: all corners negative (no NaN)
                assert drop == all(c < 0 for c in corners), (ax, ay, dx, dy, xs, ys)
                if drop:
                    culled += 1
                    assert (every < 0).all(), "a dropped block has a pixel with a non-negative edge value"
                    assert bound >= every.max(), "the bound is below some pixel's value"
                else:
                    kept += 1
    assert culled > 1000 and kept > 1000


def test_non_finite_edges_never_drop():
    """Overflowing products (inf - inf = NaN at the corner) keep the block."""
    big = f32(3e38)
    xs = (f32(-0.5), f32(0.5))
    ys = (f32(-0.5), f32(0.5))
    with np.errstate(all="ignore"):
        b = corner_bound(xs, ys, f32(0.0), f32(0.0), big, big * f32(4))
    assert not (b < 0)


def centre_range(p0, p1, extent):
    """The emit kernels' interval for the centres of pixels p0..p1
    (bin_kernels.h, centre_range): products with fl(1 / extent) and a
    margin of 2^-20 (|c'| + 4), fmaf(|c'|, 2^-20, 2^-18) rounded once."""
    half = f32(f32(1.0) / f32(extent)) * f32(0.5)
    inv = f32(2.0) * half

    def approx(p):
        return (f32(p) * inv + half) * f32(2.0) - f32(1.0)

    def margin(c):
        return f32(np.float64(abs(c)) * 2.0 ** -20 + 2.0 ** -18)

    a, b = approx(p0), approx(p1)
    return min(a - margin(a), b - margin(b)), max(a + margin(a), b + margin(b))


def test_centre_range_holds_every_centre():
    """Every exact centre (division) of the range lies inside the interval
    computed with the reciprocal, for extents down to 0.5 in magnitude,
    negative heights and pixels up to 32767; and the interval is tight
    (a few 1e-6 beyond the exact ends)."""
    rng = np.random.Generator(np.random.PCG64(11))
    checked = 0
    for extent in (0.5, -0.5, 0.75, 1.0, 3.0, 37.0, 64.0, 1080.0, -1080.0, 1920.0, 4096.0, 16384.0, 33.3, -48.0):
        for _ in range(60):
            p0 = int(rng.integers(0, 32760))
            p1 = p0 + int(rng.integers(0, 8))
            lo, hi = centre_range(p0, p1, extent)
            exact = np.array([centre(p, extent) for p in range(p0, p1 + 1)], f32)
            assert (exact >= lo).all() and (exact <= hi).all(), (extent, p0, p1, lo, hi, exact)
            slack = max(float(exact.min() - lo), float(hi - exact.max()))
            assert slack <= 2.0 ** -17 * float(np.abs(exact).max()) + 2.0 ** -15
            checked += 1
    assert checked == 14 * 60
