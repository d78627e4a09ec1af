"""The C ABI of the HIP library: it loads, exports every function include/ds/ds_smoke.h declares,
and the ctypes mirrors in ds_smoke.abi have the header's layout (checked against gcc's own
sizeof/offsetof of the header). No GPU work is launched."""
import ctypes as C
import os
import re
import subprocess

import pytest

from ds_smoke import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ds", "ds_smoke.h")

MIRRORS = {
    "ds_scalar_boundary": abi.ScalarBoundary, "ds_velocity_boundary": abi.VelocityBoundary,
    "ds_collider": abi.Collider, "ds_config": abi.Config, "ds_state": abi.State,
    "ds_control": abi.Control, "ds_parameters": abi.Parameters, "ds_step_cache": abi.StepCache,
    "ds_state_adjoint": abi.StateAdjoint, "ds_control_adjoint": abi.ControlAdjoint,
    "ds_parameter_adjoint": abi.ParameterAdjoint, "ds_solve_info": abi.SolveInfo,
    "ds_step_info": abi.StepInfo, "ds_trajectory_info": abi.TrajectoryInfo,
    "ds_kernel_timing": abi.KernelTiming, "ds_domain_info": abi.DomainInfo,
    "ds_jet_control": abi.JetControl,
}


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(ds_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_expected_surface():
    names = declared_functions()
    for required in ("ds_context_create", "ds_context_destroy", "ds_forward_step", "ds_vjp_step",
                     "ds_simulate", "ds_loss_gradient", "ds_pcg_solve", "ds_projection_forward",
                     "ds_projection_vjp", "ds_last_error"):
        assert required in names
    assert len(names) >= 40


def test_library_exports_every_declared_function():
    path = abi.library_path()
    if not os.path.exists(path):
        pytest.fail(f"{path} is missing: run __graft_entry__.build() first")
    lib = C.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # host-only entry points work without a device
    lib.ds_config_cell_count.restype = C.c_uint64
    lib.ds_config_face_count.restype = C.c_uint64
    cfg = abi.Config()
    cfg.resolution[:] = [5, 4, 3]
    assert lib.ds_config_cell_count(C.byref(cfg)) == 60
    assert [lib.ds_config_face_count(C.byref(cfg), a) for a in range(3)] == [72, 75, 80]
    lib.ds_last_error.restype = C.c_char_p
    assert lib.ds_last_error() is not None


def test_ctypes_mirrors_match_the_header_layout(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "ds/ds_smoke.h"', "int main(void) {"]
    for cname, cls in MIRRORS.items():
        lines.append(f'printf("{cname} @sizeof %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    c_layout = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, cls in MIRRORS.items():
        assert c_layout[(cname, "@sizeof")] == C.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert c_layout[(cname, field)] == getattr(cls, field).offset, (cname, field)


def test_header_compiles_as_cxx(tmp_path):
    src = tmp_path / "h.cpp"
    src.write_text('#include "ds/ds_smoke.h"\nint main() { ds_config c{}; return (int)c.resolution[0]; }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-o", str(tmp_path / "h")], check=True)
