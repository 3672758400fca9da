"""The camera step's incoherent benchmark rays ("incoherence r1 r2",
setup_ray_incoherent, rt/cuda/kernels.cu:75-156; spheres from
rt/cuda/rni.cpp:36-72), pinned on the CPU.  The reference draws from cuRAND's
MTGP32, which is absent here, so the random numbers are the workload's PCG32
streams and the checks fix the geometry: a recomputation in float64 from an
independent PCG32, the ball radii, the r^3 law of uniform points in a ball,
the zero-radius limits and the sphere placement."""
import numpy as np

import hiprt
import oracle_py as orc
from pcg import Pcg32, stream_id

FLT_MAX = np.float32(3.4028234663852886e38)


def _ball64(g, R):
    xi = [float(g.next_float()) for _ in range(3)]
    phi = xi[0] * 2 * np.pi
    th = np.arccos(xi[1] * 2 - 1)
    r = R * np.cbrt(xi[2])
    return np.array([r * np.sin(th) * np.cos(phi), r * np.sin(th) * np.sin(phi), r * np.cos(th)])


def test_rays_recomputed_in_float64():
    """Ray i from stream (seed, i, 0): three numbers for the origin's ball
    point, then three for the destination's; direction = destination -
    origin, unnormalized; t over [0, FLT_MAX]."""
    s1, s2 = np.array([1, 2, 3], np.float32), np.array([9, 2, 3], np.float32)
    rays = orc.incoherent_rays(s1, s2, 0.5, 0.25, 2.0, 77, 64)
    for i in range(64):
        g = Pcg32(stream_id(77, i, 0))
        o = s1 + _ball64(g, 0.5 * 2.0)
        t = s2 + _ball64(g, 0.25 * 2.0)
        assert np.allclose(rays[0:3, i], o, rtol=0, atol=2e-6)
        assert np.allclose(rays[4:7, i], t - o, rtol=0, atol=4e-6)
    assert (rays[3] == 0).all() and (rays[7] == FLT_MAX).all()


def test_points_fill_their_balls_uniformly():
    s1, s2 = np.array([0, 0, 0], np.float32), np.array([10, 0, 0], np.float32)
    R1, R2 = 0.75 * 4.0, 0.5 * 4.0
    n = 40000
    rays = orc.incoherent_rays(s1, s2, 0.75, 0.5, 4.0, 3, n)
    o = rays[0:3].T.astype(np.float64)
    end = o + rays[4:7].T
    r_o = np.linalg.norm(o - s1, axis=1)
    r_e = np.linalg.norm(end - s2, axis=1)
    assert r_o.max() <= R1 * (1 + 1e-6) and r_e.max() <= R2 * (1 + 1e-5)
    # uniform in volume: P(r < R/2) = 1/8, P(r < R * 0.9) = 0.729
    for r, R in ((r_o, R1), (r_e, R2)):
        assert abs((r < R / 2).mean() - 0.125) < 0.01
        assert abs((r < 0.9 * R).mean() - 0.729) < 0.01
    # no preferred direction: the mean offset is near the centre, each octant ~1/8
    assert np.abs(o.mean(0)).max() < 0.03 * R1
    octant = (o[:, 0] > 0) * 4 + (o[:, 1] > 0) * 2 + (o[:, 2] > 0)
    assert np.abs(np.bincount(octant, minlength=8) / n - 0.125).max() < 0.01
    # the same seed repeats, another seed does not
    assert np.array_equal(rays, orc.incoherent_rays(s1, s2, 0.75, 0.5, 4.0, 3, n))
    assert not np.array_equal(rays[0], orc.incoherent_rays(s1, s2, 0.75, 0.5, 4.0, 4, n)[0])


def test_zero_radii_keep_the_centres():
    s1, s2 = np.array([1, -2, 0.5], np.float32), np.array([-3, 4, 2], np.float32)
    rays = orc.incoherent_rays(s1, s2, 0.0, 0.0, 5.0, 1, 16)
    assert (rays[0:3].T == s1).all()
    assert (rays[4:7].T == (s2 - s1)).all()
    rays = orc.incoherent_rays(s1, s2, 0.0, 0.3, 5.0, 1, 4000)
    assert (rays[0:3].T == s1).all()
    assert np.linalg.norm(rays[4:7].T + s1 - s2, axis=1).max() <= 1.5 * (1 + 1e-5)


def test_spheres_sit_a_quarter_in_along_the_largest_extent():
    """rni.cpp:36-72: the largest extent (x here) carries the spheres a
    quarter in from either end, the others are centred; r_max is a quarter
    of the largest extent.  Ties go to x, then y."""
    pos = np.array([[-2, 0, 1], [6, 3, 2], [0, 1, 5]], np.float32)   # extents 8, 3, 4
    s1, s2, r_max = hiprt.incoherence_spheres(pos)
    assert s1.tolist() == [0, 1.5, 3] and s2.tolist() == [4, 1.5, 3] and r_max == 2
    cube = np.array([[0, 0, 0], [2, 2, 2]], np.float32)
    s1, s2, r_max = hiprt.incoherence_spheres(cube)
    assert s1.tolist() == [0.5, 1, 1] and s2.tolist() == [1.5, 1, 1] and r_max == 0.5
    tall = np.array([[0, 0, 0], [1, 4, 4]], np.float32)   # y and z tie: y
    s1, s2, _ = hiprt.incoherence_spheres(tall)
    assert s1.tolist() == [0.5, 1, 2] and s2.tolist() == [0.5, 3, 2]
