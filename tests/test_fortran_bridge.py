"""The Fortran side of the drop-in boundary, built on the CPU: the bridge
(mfa_brush_amd/fortran/mfa_hip_bridge.f90) compiled with flang against the
reference's own `commons` module, linked with the reference's md_main and
every reference unit it does not replace, and libmfa_hip.so.  Link only, the
program is not run.  The link succeeding shows that the step routines
md_main.f90:85-198 calls resolve to the bridge; nm shows which object
defines each.  Two builds: the reference's default control_simulation.h
(BIN_TYPE 2, cells rebuilt every step) and the one shipped with
inputs/Filled_Channel (BIN_TYPE 0, Verlet lists), where md_main also calls
check_skin and binning (md_main.f90:92-102).  Needs the reference checkout and flang
(skipped otherwise); everything is built in a temporary directory, the
second build from symbolic links to the reference's sources beside that
header."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

REF = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = "/opt/rocm/llvm/bin/flang"
# the reference's units (Makefile:31-79), modules first
MODULES = ["mfa_common", "functions", "util", "par_ziggurat", "ziggurat"]
UNITS = ["md_main", "init_system", "gen_droplet", "gen_brush", "gen_wall", "make_binning_boxes", "new_dpd_fd",
         "init_force_switch_on", "geom_constraint", "verlet_positions", "verlet_velocities", "init_params",
         "init_config", "conf_default", "init_obser", "fluid_fluid", "wall_wall", "fluid_wall", "intra_molec",
         "intra_wall", "observation", "obser_out", "store_config", "store_obser", "binning", "check_skin", "INR250",
         "SUN", "R250", "constant_force", "messages", "dpd_forces", "ewald_real", "ewald_k", "dipolar_correction",
         "lgv_forces", "my_binning", "viscosity", "bending", "bending_melt", "orientation", "magnetic_force",
         "cell_list"]
# the reference units the bridge takes over (mfa_brush_amd/fortran/mfa_hip_bridge.f90 defines them)
REPLACED = ["verlet_positions", "verlet_velocities", "fluid_fluid", "dpd_forces", "wall_wall", "fluid_wall",
            "intra_molec", "bending", "bending_melt", "constant_force", "new_dpd_fd", "check_skin", "binning",
            "my_binning"]
STEP_SYMBOLS = ["verlet_positions_", "fluid_fluid_", "fluid_wall_", "wall_wall_", "intra_molec_", "bending_",
                "bending_melt_", "constant_force_", "verlet_velocities_", "new_dpd_fd_", "check_skin_", "binning_",
                "my_binning_"]


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, "%s\n%s" % (" ".join(cmd), r.stderr[-2000:])
    return r.stdout


def _sources(d, header):
    """Directory whose .f90 / .c / .h files are the reference's, with the
    control_simulation.h of `header` (None: the reference's own directory)."""
    if header is None:
        return REF
    src = os.path.join(d, "src")
    os.mkdir(src)
    for f in os.listdir(REF):
        if f.endswith((".f90", ".c", ".h")) and f != "control_simulation.h":
            os.symlink(os.path.join(REF, f), os.path.join(src, f))
    os.symlink(os.path.join(REF, header, "control_simulation.h"), os.path.join(src, "control_simulation.h"))
    return src


@pytest.mark.skipif(not (os.path.isdir(REF) and os.path.exists(FLANG) and shutil.which("gcc")),
                    reason="needs the reference checkout and flang")
@pytest.mark.parametrize("header,bin_type", [(None, 2), ("inputs/Filled_Channel", 0)])
def test_bridge_links_into_the_reference_driver(tmp_path, header, bin_type):
    lib = os.path.join(ROOT, "mfa_brush_amd", "libmfa_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libmfa_hip.so not built")
    d = str(tmp_path)
    src = _sources(d, header)
    with open(os.path.join(src, "control_simulation.h")) as f:
        assert any(ln.split()[:3] == ["#define", "BIN_TYPE", str(bin_type)] for ln in f), header
    fl = [FLANG, "-cpp", "-O0", "-fopenmp", "-I" + src, "-c"]
    for m in MODULES:                                      # modules in dependency order
        _run(fl + [os.path.join(src, m + ".f90"), "-o", m + ".o"], d)
    kept = [u for u in UNITS if u not in REPLACED]
    with ThreadPoolExecutor(max_workers=8) as ex:
        list(ex.map(lambda u: _run(fl + [os.path.join(src, u + ".f90"), "-o", u + ".o"], d), kept))
    _run(["gcc", "-c", os.path.join(src, "wall_time.c"), "-o", "wall_time.o"], d)
    _run(fl + ["-I" + d, os.path.join(ROOT, "mfa_brush_amd", "fortran", "mfa_hip_bridge.f90"), "-o", "bridge.o"], d)
    objs = [m + ".o" for m in MODULES] + [u + ".o" for u in kept] + ["wall_time.o", "bridge.o"]
    _run([FLANG, "-fopenmp", "-o", "mfa_prog"] + objs + ["-L" + os.path.dirname(lib), "-lmfa_hip",
                                                         "-Wl,-rpath," + os.path.dirname(lib)], d)
    assert os.path.exists(os.path.join(d, "mfa_prog"))
    defined = {}
    for o in objs:
        for line in _run(["nm", "--defined-only", o], d).splitlines():
            parts = line.split()
            if len(parts) == 3 and parts[1] == "T":
                defined.setdefault(parts[2], []).append(o)
    for sym in STEP_SYMBOLS:
        assert defined.get(sym) == ["bridge.o"], (sym, defined.get(sym))
    # md_main calls them (with Verlet lists also check_skin and binning), and
    # the bridge calls the C ABI
    undef_main = _run(["nm", "--undefined-only", "md_main.o"], d)
    called = ["verlet_positions_", "fluid_fluid_", "verlet_velocities_", "intra_molec_"]
    if bin_type == 0:
        called += ["check_skin_", "binning_"]
    for sym in called:
        assert sym in undef_main, sym
    undef_bridge = _run(["nm", "--undefined-only", "bridge.o"], d)
    for sym in ("mfa_hip_create", "mfa_hip_set_state", "mfa_hip_verlet_positions", "mfa_hip_fluid_fluid",
                "mfa_hip_verlet_velocities", "mfa_hip_new_dpd_fd", "mfa_hip_get_state", "mfa_hip_check_skin"):
        assert sym in undef_bridge, sym
