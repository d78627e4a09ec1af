"""Host-side checks of the warm-start network: the PyTorch definition and the HIP loader agree on
the parameter count and order, parameters round-trip through the flat FP32 form, and the input
normalisation gives unit RMS over the unknowns."""
import ctypes as C
import os

import numpy as np
import pytest


def test_parameter_count_matches_library():
    from ds_smoke import abi, warm_start

    model = warm_start.build_unet()
    flat = warm_start.flatten_parameters(model)
    assert flat.size == 339681
    names = [k.split(".")[0] for k in model.state_dict().keys()]
    assert [n for i, n in enumerate(names) if i % 2 == 0] == list(warm_start.LAYERS)
    path = abi.library_path()
    if not os.path.exists(path):
        pytest.fail(f"{path} is missing: run __graft_entry__.build() first")
    lib = C.CDLL(path)
    lib.ds_warm_start_parameter_count.restype = C.c_uint64
    assert lib.ds_warm_start_parameter_count() == flat.size


def test_flat_round_trip_and_input_scale():
    import torch
    from ds_smoke import warm_start

    torch.manual_seed(0)
    a, b = warm_start.build_unet(), warm_start.build_unet()
    warm_start.load_parameters(b, warm_start.flatten_parameters(a))
    x = torch.randn(1, 2, 8, 8, 8)
    with torch.no_grad():
        assert torch.equal(a(x), b(x))
    rhs = torch.randn(8, 8, 8)
    unknown = torch.rand(8, 8, 8) > 0.3
    rhs = torch.where(unknown, rhs, torch.zeros_like(rhs))
    inp, s = warm_start.network_input(rhs, unknown)
    v = inp[0, 0][unknown]
    assert abs(float((v.double() ** 2).mean()) - 1.0) < 1e-5
    assert torch.equal(inp[0, 1], unknown.float())


def test_training_matrix_and_energy_loss_against_oracle():
    """The training loss's A (torch) equals the oracle's pressure matrix on a closed box; the energy
    loss is 0 for the exact solution, scale-invariant, and in (0, 1] otherwise."""
    import oracle as O
    import torch
    from ds_smoke import scenes, warm_start

    sc = scenes.plume(8, with_target=False)
    orc = O.Oracle(sc, tolerance=1e-13)
    _, code, _ = orc.masks()
    rng = np.random.default_rng(4)
    x = np.where(code > 0, rng.standard_normal(code.shape), 0.0)
    diag = torch.from_numpy(code.astype(np.float64))
    ax = warm_start.apply_matrix(torch.from_numpy(x), diag).numpy()
    np.testing.assert_allclose(ax, orc.apply_matrix(x), rtol=1e-12, atol=1e-12)
    b = np.where(code > 0, rng.standard_normal(code.shape), 0.0)
    p, _ = orc.pcg(b, tol=1e-13)
    bt, pt = torch.from_numpy(b)[None], torch.from_numpy(p)[None]
    assert float(warm_start.energy_loss(pt, bt, pt, diag)) < 1e-9
    assert float(warm_start.energy_loss(-3.0 * pt, bt, pt, diag)) < 1e-9
    r = float(warm_start.energy_loss(torch.from_numpy(x)[None], bt, pt, diag))
    assert 0.0 < r <= 1.0
