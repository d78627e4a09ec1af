"""Pins the oracle's solid-obstacle branches against hand-computed cases.

Three branches that only scenes with colliders reach, each checked against numbers worked out by
hand below (not against the oracle's own formula):
  * the solid bisection of trace_rk2 (operators.cu:256-276): 8 halvings toward the start when the
    RK2 end point lies in a solid, end point start + lo (raw - start), derivative lo times the
    domain-clamp mask;
  * collider rasterisation at cell centres (model.cpp:17-26,56-71): ellipsoid and box membership
    with inclusive bounds, later colliders overwriting earlier ones;
  * scalar advection in solid cells writes the collider value and has no adjoint
    (operators.cu:499-507,537-561).

Geometry of the bisection cases (h = 1, dt = 0.5, 16 x 4 x 4 cells, zero-gradient velocity
boundaries): a box collider covers the cells x = 10..15 (every y and z); the uniform velocity
u_x = -6.75 moves every trace end point by dt * 6.75 = 3.375 cells toward +x. A cell centred at
c = x + 0.5 reaches c + 3.375; the bisection fraction t is outside the solid while
c + 3.375 t < 10, i.e. t < (9.5 - x) / 3.375. After 8 halvings lo is the largest k / 256 below
that bound:
  x = 7: t < 0.74074 -> k = 189;  x = 8: t < 0.44444 -> k = 113;  x = 9: t < 0.14815 -> k = 37.
Each k is odd, so 7 halvings (k = 188, 112, 36) or 9 (379/512, 227/512, 75/512) give other
end points; x = 6 ends at 9.875, outside the solid, and is not bisected.
"""
import numpy as np
import pytest

import oracle as O
from ds_smoke import abi, scenes
from ds_smoke.scenes import cell_shape, face_shape
from helpers import collider

H, DT, U = 1.0, 0.5, 6.75
RES = (16, 4, 4)
LO = {7: 189 / 256, 8: 113 / 256, 9: 37 / 256}
COLLIDER_DENSITY = 7.0


def slab_scene(vy=0.0):
    sc = scenes.Scene(resolution=RES, cell_size=H, time_step=DT, name="slab")
    sc.velocity_modes = (abi.VELOCITY_ZERO_GRADIENT,) * 6
    sc.colliders = [collider(abi.SHAPE_BOX, (13.0, 2.0, 2.0), (2.6, 2.1, 2.1), density=COLLIDER_DENSITY, temperature=-3.0)]
    vel = (np.full(face_shape(RES, 0), -U), np.full(face_shape(RES, 1), vy), np.zeros(face_shape(RES, 2)))
    return sc, vel


def linear_field(ax=1.0, ay=0.0):
    """ax * (x + 0.5) + ay * (y + 0.5) at the cell centres: trilinear sampling reproduces it."""
    z, y, x = np.meshgrid(np.arange(RES[2]), np.arange(RES[1]), np.arange(RES[0]), indexing="ij")
    return ax * (x + 0.5) * H + ay * (y + 0.5) * H


def test_slab_is_rasterised_as_written():
    sc, _ = slab_scene()
    m, code, _ = O.Oracle(sc).masks()
    assert m[:, :, 10:].all() and not m[:, :, :10].any()


def test_bisected_end_points():
    sc, vel = slab_scene()
    o = O.Oracle(sc)
    out = o.advect_scalar_forward(0, linear_field(), vel)
    for x, lo in LO.items():
        np.testing.assert_allclose(out[:, :, x], x + 0.5 + lo * 3.375, rtol=0, atol=1e-12, err_msg=f"cell x={x}")
    # not bisected: the full shift
    for x in range(0, 7):
        np.testing.assert_allclose(out[:, :, x], x + 0.5 + 3.375, rtol=0, atol=1e-12)
    # solid cells: the collider's density (operators.cu:505), whatever the source holds
    assert np.all(out[:, :, 10:] == COLLIDER_DENSITY)


def test_other_halving_counts_are_distinguishable():
    """The hand-computed end points separate 8 halvings from 7 and 9 by more than the tolerance."""
    for x, lo in LO.items():
        k = round(lo * 256)
        assert k % 2 == 1
        for other in (np.floor(lo * 128) / 128, np.floor((9.5 - x) / 3.375 * 512 - 1e-12) / 512):
            assert abs(other - lo) * 3.375 > 1e-3


@pytest.mark.parametrize("x", sorted(LO))
def test_bisected_derivative_is_lo(x):
    """Reverse of a bisected trace: the end point moves by lo * d(raw), raw = start - dt v, so one
    unit of output adjoint at the cell puts -lo * dt * (d source / dx) into the x velocity adjoint
    (the trilinear weights of the uniform field sum to 1; its gradient is 0, so the midpoint
    chain adds nothing), and the source adjoint receives the unit spread over the end point's
    trilinear stencil."""
    sc, vel = slab_scene()
    o = O.Oracle(sc)
    oa = np.zeros(cell_shape(RES))
    oa[1, 1, x] = 1.0
    src_adj = o.zeros_cell()
    ua = o.zeros_faces()
    o.advect_scalar_vjp(0, linear_field(), vel, oa, src_adj, ua)
    np.testing.assert_allclose(ua[0].sum(), -LO[x] * DT, rtol=1e-12)
    assert np.abs(ua[1]).max() == 0.0 and np.abs(ua[2]).max() == 0.0
    np.testing.assert_allclose(src_adj.sum(), 1.0, rtol=1e-12)
    # the end point x + 0.5 + 3.375 lo lies between the cells whose centres bracket it
    p = x + 0.5 + 3.375 * LO[x]
    i0 = int(np.floor(p - 0.5))
    np.testing.assert_allclose(src_adj[1, 1, i0 + 1], p - 0.5 - i0, rtol=1e-12)


def test_bisection_with_clamped_axis():
    """u_y = 6 moves the end point of the y = 2 cells to y = 2.5 - 3 = -0.5: clamped to 0 with
    derivative 0 in y (operators.cu:266-268), and the bisection runs on the segment from the start
    to the clamped point, so the end point's y is 2.5 (1 - lo) = 0.654 (above the first cell
    centre, where the field is still linear) and its y derivative lo * 0 = 0."""
    sc, vel = slab_scene(vy=6.0)
    o = O.Oracle(sc)
    src = linear_field(1.0, 1.0)
    out = o.advect_scalar_forward(0, src, vel)
    lo = LO[7]
    np.testing.assert_allclose(out[1, 2, 7], (7.5 + lo * 3.375) + 2.5 * (1.0 - lo), rtol=0, atol=1e-12)
    oa = np.zeros(cell_shape(RES))
    oa[1, 2, 7] = 1.0
    src_adj = o.zeros_cell()
    ua = o.zeros_faces()
    o.advect_scalar_vjp(0, src, vel, oa, src_adj, ua)
    np.testing.assert_allclose(ua[0].sum(), -lo * DT, rtol=1e-12)
    assert np.abs(ua[1]).max() == 0.0


def test_solid_cells_have_no_scalar_adjoint():
    sc, vel = slab_scene()
    o = O.Oracle(sc)
    oa = np.zeros(cell_shape(RES))
    oa[:, :, 10:] = 1.0
    src_adj = o.zeros_cell()
    ua = o.zeros_faces()
    o.advect_scalar_vjp(0, linear_field(), vel, oa, src_adj, ua)
    assert not src_adj.any() and not any(a.any() for a in ua)


# ---- rasterisation (model.cpp:17-26): h = 1/4, so every cell centre and bound is exact in FP32 ---
RAST_RES = (7, 5, 5)
CENTRE = (0.875, 0.625, 0.625)   # the centre of cell (3, 2, 2)


def rast_scene(colliders):
    sc = scenes.Scene(resolution=RAST_RES, cell_size=0.25, time_step=0.125, name="rast")
    sc.colliders = colliders
    return sc


def cells_of(mask):
    z, y, x = np.nonzero(mask)
    return sorted(zip(x.tolist(), y.tolist(), z.tolist()))


def test_ellipsoid_membership_is_inclusive():
    """Radii (2h, h, h) centred on a cell centre: the lattice offsets (i, j, k) with
    (i/2)^2 + j^2 + k^2 <= 1 are the centre, i = +-1, +-2 on x and +-1 on y and z. Five of the
    nine sit exactly on the surface (value 1): a strict '<' would keep 3 cells, swapped radii
    another set."""
    m, _, _ = O.Oracle(rast_scene([collider(abi.SHAPE_ELLIPSOID, CENTRE, (0.5, 0.25, 0.25))])).masks()
    expect = sorted([(1, 2, 2), (2, 2, 2), (3, 2, 2), (4, 2, 2), (5, 2, 2), (3, 1, 2), (3, 3, 2), (3, 2, 1), (3, 2, 3)])
    assert cells_of(m) == expect


def test_box_membership_is_inclusive():
    """Half extents (h, 0, h): |dx| <= h keeps x = 2..4, |dy| <= 0 keeps y = 2 only, |dz| <= h
    keeps z = 1..3."""
    m, _, _ = O.Oracle(rast_scene([collider(abi.SHAPE_BOX, CENTRE, (0.25, 0.0, 0.25))])).masks()
    expect = sorted((x, 2, z) for x in (2, 3, 4) for z in (1, 2, 3))
    assert cells_of(m) == expect


def test_later_collider_overwrites_collider_values():
    """Two overlapping colliders: cells inside both take the second one's density and
    temperature (model.cpp:62-70 assigns in collider order)."""
    first = collider(abi.SHAPE_BOX, CENTRE, (0.25, 0.25, 0.25), density=1.0, temperature=2.0)
    second = collider(abi.SHAPE_ELLIPSOID, CENTRE, (0.5, 0.25, 0.25), density=5.0, temperature=6.0)
    o = O.Oracle(rast_scene([first, second]))
    m, _, _ = o.masks()
    _, d, t = o.collider_fields()
    assert d[2, 2, 3] == 5.0 and t[2, 2, 3] == 6.0           # in both
    assert d[2, 2, 5] == 5.0                                  # second only
    assert d[1, 1, 2] == 1.0 and t[1, 1, 2] == 2.0            # first only (corner of the box)
    assert m.sum() == 27 + 2                                   # the box's 27 cells + (1, 2, 2) and (5, 2, 2)
