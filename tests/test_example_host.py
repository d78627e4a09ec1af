"""Host side of the forward example harness (no GPU): the options and their configuration
(simulation.ixx:9-27, simulation.cpp:17-35), the jet control's field order as the C ABI and the
oracle read it, and the derived metrics (simulation.cpp:104-119) from known raw values."""
import ctypes as C
import math

import numpy as np

from ds_smoke import abi
from ds_smoke.example import ForwardSimulationMetrics, ForwardSimulationOptions, finish_metrics, format_metrics


def test_defaults_and_configuration():
    o = ForwardSimulationOptions()
    assert o.resolution == (128, 192, 128) and o.pressure_iterations == 160
    assert np.float32(o.cell_size) == np.float32(1.0 / 128.0) and np.float32(o.time_step) == np.float32(1.0 / 120.0)
    assert (o.density_source_rate, o.temperature_source_rate, o.vorticity_confinement) == (4.0, 8.0, 2.0)
    cfg = o.config()
    # five walls with the normal velocity fixed (mode 2), an open top (zero gradient, mode 1)
    assert list(cfg.velocity_boundary.modes) == [2, 2, 2, 1, 2, 2]
    # pressure, density and temperature fixed (mode 0) on y_max only
    for b in (cfg.pressure_boundary, cfg.density_boundary, cfg.temperature_boundary):
        assert list(b.modes) == [1, 1, 1, 0, 1, 1] and list(b.values) == [0.0] * 6
    assert cfg.vorticity_confinement_enabled == 1 and cfg.projection_solver == abi.PROJECTION_RBGS
    assert cfg.pressure_iterations == 160 and cfg.warm_start == abi.WARM_START_NONE
    pcg = ForwardSimulationOptions(projection_solver=abi.PROJECTION_PCG).config()
    assert pcg.projection_solver == abi.PROJECTION_PCG and pcg.warm_start == abi.WARM_START_PREVIOUS


def test_jet_control_layout_and_values():
    o = ForwardSimulationOptions()
    j = o.jet()
    assert C.sizeof(abi.JetControl) == 16 * 4
    v = j.values()
    assert len(v) == 16
    np.testing.assert_array_equal(np.float32(v[1:4]), np.float32(o.left_source_center))
    np.testing.assert_array_equal(np.float32(v[4:7]), np.float32(o.right_source_center))
    np.testing.assert_array_equal(np.float32(v[7:10]), np.float32([o.source_radius, o.density_source_rate, o.temperature_source_rate]))
    np.testing.assert_array_equal(np.float32(v[10:13]), np.float32(o.left_acceleration))
    np.testing.assert_array_equal(np.float32(v[13:16]), np.float32(o.right_acceleration))
    assert np.float32(v[0]) == np.float32(o.pulse_period)


def test_derived_metrics():
    o = ForwardSimulationOptions(resolution=(4, 6, 4), cell_size=0.25, time_step=0.125)
    m = ForwardSimulationMetrics()
    cells = 4 * 6 * 4
    m.step = 1
    finish_metrics(o, [64.0, 0.5, 0.75, 2.0, 4.0 * cells, 1.0 * cells], 3.0, m)
    assert m.density_mass == 64.0 * 0.25 ** 3
    assert m.cfl == 2.0 * 0.125 / 0.25
    assert m.pre_projection_divergence_rms == 2.0 and m.post_projection_divergence_rms == 1.0
    assert m.divergence_ratio == 0.5
    assert m.step_milliseconds == 3.0 and m.average_step_milliseconds == 3.0
    m.step = 2
    finish_metrics(o, [64.0, 0.5, 0.75, 2.0, 4.0 * cells, 1.0 * cells], 5.0, m)
    assert m.average_step_milliseconds == 4.0   # running mean over the steps
    # zero divergence before the projection: IEEE division, as the reference's C++ does
    finish_metrics(o, [0.0, 0.0, 0.0, 0.0, 0.0, 0.0], 1.0, m)
    assert math.isnan(m.divergence_ratio)
    line = format_metrics(m)
    assert line.startswith("step   2  time") and "divergence" in line and "average" in line
