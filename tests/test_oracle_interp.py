"""The oracle's VSQBS and LBB 2-D interpolators and the resize plans that use
them (resize.rs:176-284), pinned by the reference's known answers and by
what the mathematics fixes:

  * VSQBS: a constant neighbourhood is kept, the 0..15 ramp gives 8.25 at
    phases (0.25, 0.75) (kernel.rs:1096-1112), and the 2 x 2 -> 4 x 4 u8
    upscale with copy extend equals the reference's expected bytes
    (affine/tests.rs:698-722);
  * LBB: zero phases return the centre sample exactly (affine/tests.rs:
    360-376), a linear ramp is reproduced (central differences are exact
    and no limiter triggers), every result stays within the stencil's range
    (affine/tests.rs:379-402 and the property at 1113-), and a
    discontinuity is limited where a plain bicubic would overshoot;
  * plans: LBB is one Affine (down- and upscales), VSQBS upscales never
    zoom, VSQBS downscales are refused (reduce_common.rs:107-113).
"""
import numpy as np
import pytest

from oracle import oracle as O


def test_vsqbs_known_answers():
    for xp, yp in ((0.0, 0.0), (0.25, 0.25), (0.25, 0.75), (0.75, 0.75)):
        assert abs(O.interp_2d(O.VSQBS, np.full((4, 4), 7.0), xp, yp) - 7.0) < 1e-12
    assert abs(O.interp_2d(O.VSQBS, np.arange(16.0).reshape(4, 4), 0.25, 0.75) - 8.25) < 1e-12


def test_vsqbs_affine_upscale_golden():
    # affine/tests.rs:698-722: [[0, 64], [128, 255]] with copy extend,
    # matrix 0.5, offsets 0.25 -- exactly resize(2, 2, VSQBS) of that image.
    src = np.array([[0, 64], [128, 255]], np.uint8)[..., None]
    got = O.Pipeline(2, 2, 1).resize(2.0, 2.0, O.VSQBS).run(src).ravel()
    assert list(got) == [57, 96, 115, 118, 124, 169, 198, 201, 159, 214, 245, 249, 164, 219, 251, 255]


def test_vsqbs_is_a_partition_of_unity():
    # The eight weights sum to four at every phase: a constant stays constant
    # and adding c to every sample adds c.
    r = np.random.default_rng(3)
    n = r.normal(size=(4, 4))
    for xp, yp in r.uniform(0, 1, size=(20, 2)):
        a = O.interp_2d(O.VSQBS, n, xp, yp)
        b = O.interp_2d(O.VSQBS, n + 10.0, xp, yp)
        assert abs(b - a - 10.0) < 1e-12


def test_lbb_identity_and_linear():
    r = np.random.default_rng(5)
    n = r.normal(size=(4, 4)) * 100
    assert O.interp_2d(O.LBB, n, 0.0, 0.0) == n[1, 1]
    y, x = np.mgrid[0:4, 0:4].astype(np.float64)
    ramp = 3.0 * x - 2.0 * y + 5.0
    for xp, yp in r.uniform(0, 1, size=(20, 2)):
        exact = 3.0 * (1 + xp) - 2.0 * (1 + yp) + 5.0
        assert abs(O.interp_2d(O.LBB, ramp, xp, yp) - exact) < 1e-12


def test_lbb_locally_bounded():
    # affine/tests.rs:379-402 at (1.35, 1.65), then random stencils.
    data = np.array([[-1.0, 2.0, 9.0, 3.0], [4.0, 6.0, 8.0, 5.0], [7.0, 0.0, 1.0, 11.0], [10.0, 12.0, -3.0, 13.0]])
    v = O.interp_2d(O.LBB, data, 0.35, 0.65)
    assert data.min() - 1e-12 <= v <= data.max() + 1e-12
    r = np.random.default_rng(11)
    for _ in range(300):
        n = r.normal(size=(4, 4)) * r.uniform(0.1, 100)
        xp, yp = r.uniform(0, 1, size=2)
        v = O.interp_2d(O.LBB, n, xp, yp)
        assert n.min() - 1e-9 <= v <= n.max() + 1e-9


def test_lbb_limits_a_step():
    # A step between columns 1 and 2: the bounded interpolant stays inside
    # [0, 1] (a bicubic's slopes would overshoot next to the step).
    n = np.array([[0.0, 0.0, 1.0, 1.0]] * 4)
    for xp in np.linspace(0, 1, 11):
        v = O.interp_2d(O.LBB, n, xp, 0.5)
        assert -1e-12 <= v <= 1 + 1e-12
    # and it is monotone along the step
    vals = [O.interp_2d(O.LBB, n, xp, 0.5) for xp in np.linspace(0, 0.99, 12)]
    assert all(b >= a - 1e-12 for a, b in zip(vals, vals[1:]))


def test_resize_plans():
    # resize.rs:461-516
    p = O.Pipeline(100, 100, 1).resize(0.6, 0.6, O.LBB)
    assert [t for t, _ in p.nodes()] == [15] and p.output()[:2] == (60, 60)  # one Affine
    p = O.Pipeline(25, 25, 1).resize(2.0, 2.0, O.LBB)
    assert [t for t, _ in p.nodes()] == [15] and p.output()[:2] == (50, 50)
    p = O.Pipeline(2, 2, 1).resize(2.0, 2.0, O.VSQBS)
    assert [t for t, _ in p.nodes()] == [15] and p.output()[:2] == (4, 4)  # no Zoom
    with pytest.raises(O.OracleError):
        O.Pipeline(100, 100, 1).resize(0.5, 0.5, O.VSQBS)  # needs a VSQBS reduce
