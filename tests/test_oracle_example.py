"""The oracle's forward-example functions (write_control_kernel and reduce_metrics_kernel of
examples/smoke/forward/simulation.cu:35-85), pinned by closed forms and an independent
formulation:

* write_control: at step 0 both jet amplitudes are 0.75 (sin 0, sin pi), at a quarter period 1.0
  and 0.5; a jet centred on a cell centre gives rate * amplitude there; the sum of a Gaussian of
  radius r sampled on cell centres far from the walls is (2 pi)^(3/2) (r/h)^3 to rounding (Poisson
  summation: the first correction is exp(-2 pi^2 r^2 / h^2)); the difference of the two halves of
  the domain follows 0.25 (sin phi - sin(phi + pi)); mirrored jets give mirrored fields.
* reduce_metrics: constant fields give count * value and the maxima (solid cells excluded, maxima
  taken with 0); integer-valued linear face fields give a known divergence; a uniform flow gives
  its speed; random fields against numpy's np.diff divergence and face averages.
"""
import numpy as np
import pytest

import oracle as O
from ds_smoke import scenes
from ds_smoke.example import ForwardSimulationOptions
from helpers import scene_solids


def _jet(period, lc, rc, radius, rho_rate, t_rate, la, ra):
    return [period] + list(lc) + list(rc) + [radius, rho_rate, t_rate] + list(la) + list(ra)


N = 64
H = 1.0 / N   # a power of two: cell centres and jet centres are exact in FP32 and FP64
DT = float(np.float32(1.0 / 120.0))   # the time step as the configuration stores it (FP32)


@pytest.fixture(scope="module")
def jet_oracle():
    opts = ForwardSimulationOptions(resolution=(N, N, N), cell_size=H, time_step=DT)
    return O.Oracle(opts.scene(), solver=1, rbgs_iterations=4)


LC = (14.5 * H, 14.5 * H, 16.5 * H)   # centres of cells (14, 14, 16) and (47, 14, 31)
RC = (47.5 * H, 14.5 * H, 31.5 * H)
R = 2.0 * H
LA = (3.5, 5.0, 1.8)
RA = (-3.5, 5.0, -1.8)


def test_write_control_step_zero_and_quarter_period(jet_oracle):
    dt = DT
    period = 4.0 * 7 * dt                      # step 7 is a quarter period
    jet = _jet(period, LC, RC, R, 4.0, 8.0, LA, RA)
    li = 14 + N * (14 + N * 16)
    ri = 47 + N * (14 + N * 31)
    s_rho, s_t, e = jet_oracle.example_write_control(0, jet)
    np.testing.assert_allclose(s_rho.ravel()[li], 4.0 * 0.75, rtol=1e-12)
    np.testing.assert_allclose(s_t.ravel()[li], 8.0 * 0.75, rtol=1e-12)
    np.testing.assert_allclose(s_rho.ravel()[ri], 4.0 * 0.75, rtol=1e-12)
    for a in range(3):
        np.testing.assert_allclose(e[a].ravel()[li], 0.75 * LA[a], rtol=1e-12)
        np.testing.assert_allclose(e[a].ravel()[ri], 0.75 * RA[a], rtol=1e-12)
    s_rho, s_t, e = jet_oracle.example_write_control(7, jet)
    np.testing.assert_allclose(s_rho.ravel()[li], 4.0 * 1.0, rtol=1e-12)
    np.testing.assert_allclose(s_rho.ravel()[ri], 4.0 * 0.5, rtol=1e-12)
    np.testing.assert_allclose(e[1].ravel()[ri], 0.5 * RA[1], rtol=1e-12)


def test_write_control_gaussian_mass_and_pulse(jet_oracle):
    dt = DT
    period = 0.9
    jet = _jet(period, LC, RC, R, 4.0, 8.0, LA, RA)
    g = (2.0 * np.pi) ** 1.5 * (R / H) ** 3
    for step in (0, 5, 31, 100):
        phi = 2.0 * np.pi * step * dt / period
        amp_l, amp_r = 0.75 + 0.25 * np.sin(phi), 0.75 + 0.25 * np.sin(phi + np.pi)
        s_rho, s_t, e = jet_oracle.example_write_control(step, jet)
        np.testing.assert_allclose(s_rho.sum(), 4.0 * (amp_l + amp_r) * g, rtol=1e-11)
        np.testing.assert_allclose(s_t.sum(), 8.0 * (amp_l + amp_r) * g, rtol=1e-11)
        # the left jet lives in x < 32, the right one in x >= 32 (each 7 radii or more from the split)
        left, right = s_rho[:, :, :N // 2].sum(), s_rho[:, :, N // 2:].sum()
        np.testing.assert_allclose(left - right, 4.0 * (amp_l - amp_r) * g, rtol=1e-10, atol=1e-9)
        np.testing.assert_allclose(e[0].sum(), g * (amp_l * LA[0] + amp_r * RA[0]), rtol=1e-10, atol=1e-9)


def test_write_control_mirror_symmetry(jet_oracle):
    # jets mirrored through the box centre in x and z, accelerations mirrored: at step 0 (equal
    # amplitudes) density is symmetric, the x/z accelerations antisymmetric, y symmetric
    lc = (14.5 * H, 20.5 * H, 17.5 * H)
    rc = (1.0 - lc[0], lc[1], 1.0 - lc[2])
    jet = _jet(0.9, lc, rc, 3.0 * H, 4.0, 8.0, LA, (-LA[0], LA[1], -LA[2]))
    s_rho, _, e = jet_oracle.example_write_control(0, jet)
    flip = lambda f: f[::-1, :, ::-1]          # (z, y, x) -> (N-1-z, y, N-1-x)
    tol = lambda f: 1e-13 * np.abs(f).max()   # both jets overlap where the sums cancel
    np.testing.assert_allclose(s_rho, flip(s_rho), rtol=1e-12, atol=tol(s_rho))
    np.testing.assert_allclose(e[0], -flip(e[0]), rtol=1e-12, atol=tol(e[0]))
    np.testing.assert_allclose(e[1], flip(e[1]), rtol=1e-12, atol=tol(e[1]))
    np.testing.assert_allclose(e[2], -flip(e[2]), rtol=1e-12, atol=tol(e[2]))
    assert s_rho.max() > 1.0


@pytest.fixture(scope="module")
def solids_oracle():
    sc = scene_solids()
    orc = O.Oracle(sc)
    mask = orc.masks()[0]
    assert 0 < mask.sum() < mask.size
    return sc, orc, mask


def _face(sc, a, fn):
    shp = scenes.face_shape(sc.resolution, a)
    z, y, x = np.meshgrid(np.arange(shp[0]), np.arange(shp[1]), np.arange(shp[2]), indexing="ij")
    return fn(x, y, z).astype(np.float64)


def test_reduce_metrics_constant_fields_and_exclusion(solids_oracle):
    sc, orc, mask = solids_oracle
    fluid = int((mask == 0).sum())
    rho = np.where(mask == 0, 0.5, 1e6)
    T = np.where(mask == 0, 0.25, 1e6)
    zero = tuple(np.zeros(scenes.face_shape(sc.resolution, a)) for a in range(3))
    out = orc.example_reduce_metrics(rho, T, zero, zero)
    assert out[0] == 0.5 * fluid and out[1] == 0.5 and out[2] == 0.25
    assert out[3] == 0.0 and out[4] == 0.0 and out[5] == 0.0
    # maxima are taken with 0 (the reference zeroes its buffer before the atomic maxima)
    out = orc.example_reduce_metrics(-rho, -T, zero, zero)
    assert out[0] == -0.5 * fluid and out[1] == 0.0 and out[2] == 0.0


def test_reduce_metrics_linear_and_uniform_velocity(solids_oracle):
    sc, orc, mask = solids_oracle
    fluid = int((mask == 0).sum())
    h = sc.cell_size
    # u_x = i, u_y = 2 j, u_z = -k on the faces: every cell has divergence (1 + 2 - 1) / h
    pre = (_face(sc, 0, lambda x, y, z: x), _face(sc, 1, lambda x, y, z: 2 * y), _face(sc, 2, lambda x, y, z: -z))
    # a uniform flow (3, 4, 12): divergence 0, speed 13
    post = tuple(np.full(scenes.face_shape(sc.resolution, a), v) for a, v in enumerate((3.0, 4.0, 12.0)))
    rho = np.zeros(scenes.cell_shape(sc.resolution))
    out = orc.example_reduce_metrics(rho, rho, pre, post)
    assert out[4] == pytest.approx(fluid * (2.0 / h) ** 2, rel=1e-15)
    assert out[5] == 0.0
    assert out[3] == 13.0


def test_reduce_metrics_against_numpy(solids_oracle):
    sc, orc, mask = solids_oracle
    rng = np.random.default_rng(3)
    rho = rng.standard_normal(mask.shape).astype(np.float32).astype(np.float64)
    T = rng.standard_normal(mask.shape).astype(np.float32).astype(np.float64)
    pre = tuple(rng.standard_normal(scenes.face_shape(sc.resolution, a)).astype(np.float32).astype(np.float64) for a in range(3))
    post = tuple(rng.standard_normal(scenes.face_shape(sc.resolution, a)).astype(np.float32).astype(np.float64) for a in range(3))
    out = orc.example_reduce_metrics(rho, T, pre, post)
    fl = mask == 0
    h = np.float64(np.float32(sc.cell_size))

    def div(u):   # (nz, ny, nx) cell divergence from the three face arrays: np.diff along each axis
        return (np.diff(u[0], axis=2) + np.diff(u[1], axis=1) + np.diff(u[2], axis=0)) / h

    def speed(u):
        cx = 0.5 * (u[0][:, :, 1:] + u[0][:, :, :-1])
        cy = 0.5 * (u[1][:, 1:, :] + u[1][:, :-1, :])
        cz = 0.5 * (u[2][1:, :, :] + u[2][:-1, :, :])
        return np.sqrt(cx * cx + cy * cy + cz * cz)

    np.testing.assert_allclose(out[0], rho[fl].sum(), rtol=1e-12)
    assert out[1] == max(0.0, rho[fl].max()) and out[2] == max(0.0, T[fl].max())
    np.testing.assert_allclose(out[3], speed(post)[fl].max(), rtol=1e-6)
    np.testing.assert_allclose(out[4], (div(pre)[fl] ** 2).sum(), rtol=1e-6)
    np.testing.assert_allclose(out[5], (div(post)[fl] ** 2).sum(), rtol=1e-6)
