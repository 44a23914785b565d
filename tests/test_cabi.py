"""The C ABI library loads on a machine without a GPU and exports every
function that include/amsc_mpm.h declares; include/gpu_kernels.h compiles
as C++ against it.  No compute calls (those need the GPU)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "amsc_mpm.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(amsc_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_the_18_reference_entry_points():
    names = declared()
    ref = ["reset_nodal_arrays", "scatter_mass_momentum", "gradZ_to_particles",
           "gravity_force", "friction_particle_force", "scatter_gravity_force",
           "scatter_friction_force", "compute_F_ext", "scatter_internal_forces",
           "nodal_force_and_momentum", "nodal_velocity_and_bc",
           "g2p_velocity_and_advect", "g2pd_gradients_and_height_update",
           "voellmy_friction", "bingham_stress_update", "g2p_height",
           "scatter_plotvars", "normalize_plotvars"]
    for r in ref:
        assert "amsc_" + r in names


def test_library_exports_every_declared_symbol():
    from amsc_amd import native
    lib = native.lib()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # exported with C linkage: nm shows the plain names
    out = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    assert set(declared()) <= exported


def test_error_reporting_without_compute():
    from amsc_amd import native
    lib = native.lib()
    # invalid sizes are rejected before anything touches the device
    rc = lib.amsc_bin_particles(-1, 4, 4, 1.0, 1.0, None, None, None, None,
                                None, None, __import__("numpy").zeros(5, "int32"))
    assert rc == 1
    assert b"bad sizes" in lib.amsc_last_error()


@pytest.mark.parametrize("header", ["amsc_mpm.h", "gpu_kernels.h"])
def test_headers_compile_as_cpp(header, tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(f'#include "{header}"\nint main() {{ return 0; }}\n')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                    "-I", os.path.join(ROOT, "include"), str(src)], check=True)


def test_header_compiles_as_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "amsc_mpm.h"\nint main(void) { return 0; }\n')
    subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-Wall", "-Werror",
                    "-I", os.path.join(ROOT, "include"), str(src)], check=True)
