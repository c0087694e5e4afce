"""Host-side checks that need no GPU: input readers, table derivation,
configuration generators, and that the C-ABI library loads and exports
every symbol include/mfa_hip.h declares (no compute call is made)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from mfa_brush_amd import _lib, configs
from mfa_brush_amd.params import Flags, read_mfa_input, read_system_input, write_conf_new, read_conf_old

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "mfa_hip.h")).read()
    return sorted(set(re.findall(r"\b(mfa_hip_\w+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert sorted(_lib.SIGNATURES) == _declared()


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.mfa_hip_version() == 1


def test_struct_layout_matches_header(tmp_path):
    """Sizes and field offsets of the ctypes structures equal what a C
    compiler gives for include/mfa_hip.h."""
    import shutil
    import subprocess
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    fields = {"mfa_system_params": ("SystemParams", [f for f, _ in _lib.SystemParams._fields_]),
              "mfa_energies": ("Energies", [f for f, _ in _lib.Energies._fields_]),
              "mfa_flags": ("Flags", [f for f, _ in _lib.Flags._fields_])}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mfa_hip.h"', "int main(void) {"]
    for cname, (_, names) in fields.items():
        lines.append('printf("%%s size %%zu\\n", "%s", sizeof(%s));' % (cname, cname))
        for f in names:
            lines.append('printf("%%s %%s %%zu\\n", "%s", "%s", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    seen = 0
    for line in out:
        parts = line.split()
        if len(parts) == 3 and parts[1] == "size":
            assert C.sizeof(getattr(_lib, fields[parts[0]][0])) == int(parts[2]), line
            seen += 1
        elif len(parts) == 3:
            st = getattr(_lib, fields[parts[0]][0])
            assert getattr(st, parts[1]).offset == int(parts[2]), line
            seen += 1
    assert seen == sum(1 + len(n) for _, n in fields.values())


def test_make_binning_host_matches_reference_formula():
    lib = _lib.load()
    b = np.array([30.0, 20.0, 40.0])
    l = np.zeros(3); il = np.zeros(3); nc = np.zeros(3, np.int32); nct = C.c_int32(); nn = C.c_int32()
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    assert lib.mfa_hip_make_binning(1, p(b), 2.2449240684509277, p(l), p(il), p(nc), C.byref(nct), C.byref(nn)) == 0
    assert nc.tolist() == [13, 8, 17] and nct.value == 13 * 8 * 17 and nn.value == 13
    assert lib.mfa_hip_make_binning(2, p(b), 2.2449240684509277, p(l), p(il), p(nc), C.byref(nct), C.byref(nn)) == 0
    assert nc.tolist() == [26, 17, 35] and nn.value == 62
    assert np.allclose(l * nc, b)


def test_readers_on_reference_inputs(try_fixture):
    si, mi = try_fixture["si"], try_fixture["mi"]
    assert (si.n_mon, si.n_chain, si.n_cell_w_x, si.n_cell_w_y) == (10, 300, 30, 20)
    assert (si.n_mon_d, si.n_chain_d, si.a_w, si.sigma_w) == (1, 3000, -3.2, 1.0)
    assert mi.dt == 2e-4 and mi.temp == 0.8 and mi.iseed == 311500357
    assert mi.k_bend == 80.0 and mi.k_chain == 30.0 and mi.r_chain == 1.5 and mi.z_space_wall == 40.0
    assert mi.mass_type == [1.0, 1.0, 1.0, 50.0]
    assert mi.epsil[0] == [0.0, 1.0, 0.333, 1.0]
    p = try_fixture["p"]
    assert p.boundary == (30.0, 20.0, 40.0) and p.n_part == 6000
    assert p.r_2_min_init == 0.0          # s_time=2: continuation, no switch-on ramp


def test_conf_roundtrip(tmp_path, try_fixture):
    p = try_fixture["p"]
    path = tmp_path / "conf_new"
    write_conf_new(str(path), p, try_fixture["a_type"], try_fixture["r0"], try_fixture["v"], try_fixture["a"])
    t, r0, v, a = read_conf_old(path.read_text(), p.n_part)
    assert np.array_equal(t, try_fixture["a_type"])
    assert np.abs(r0 - try_fixture["r0"]).max() < 1e-7 and np.abs(v - try_fixture["v"]).max() < 1e-7


def test_channel_generators():
    p, t, r0, v = configs.channel_c1()
    assert p.n_part == 4096 and (t == 1).sum() == 128 and (t == 3).sum() == 2048
    assert np.all(r0[:, 2] > 1.0) and np.all(r0[:, 2] < 19.0)
    heads = np.arange(128) * 16
    assert np.all(t[heads] == 1) and not v[heads].any()
    d = r0[:, None, :] - r0[None, :, :]
    d[..., :2] -= 16.0 * np.round(d[..., :2] / 16.0)
    dd = np.sqrt((d ** 2).sum(-1)); np.fill_diagonal(dd, 9.0)
    assert dd.min() > 0.9
    ps, ts, rs, vs = configs.channel_scaled(12)
    assert abs(ps.n_part / np.prod(ps.boundary) - 0.8) < 0.01


def test_fp32_divide_and_sqrt_from_fp64_are_correctly_rounded():
    """The device Box-Muller takes its FP32 divide and square root from FP64
    and rounds once (mfa_kernels.h div_rn_f32 / sqrt_rn_f32); the oracle uses
    FP32 IEEE divide and sqrtf.  Both agree when double rounding is
    innocuous, checked here exhaustively over the arguments Box-Muller can
    see: every float m in [1, sqrt 2] for (m-1)/(m+1), and every float in
    [2^-24, 64) for sqrt(-2 ln u)."""
    lo = np.float32(1.0).view(np.uint32)
    hi = np.float32(1.41421356).view(np.uint32)
    m = np.arange(lo, hi + 1, dtype=np.uint32).view(np.float32)
    a, b = m - np.float32(1), m + np.float32(1)
    assert np.array_equal((a.astype(np.float64) / b.astype(np.float64)).astype(np.float32), a / b)
    # half-open [0.5, 1): the (m-1)/(m+1) branch after m*0.5 is in this range too
    m = np.arange(np.float32(0.70710677).view(np.uint32), np.float32(1.0).view(np.uint32),
                  dtype=np.uint32).view(np.float32)
    a, b = m - np.float32(1), m + np.float32(1)
    assert np.array_equal((a.astype(np.float64) / b.astype(np.float64)).astype(np.float32), a / b)
    for e in range(-24, 6):
        x = (np.arange(1 << 23, dtype=np.uint32) | np.uint32((e + 127) << 23)).view(np.float32)
        assert np.array_equal(np.sqrt(x.astype(np.float64)).astype(np.float32), np.sqrt(x)), e


def test_mlp_header_binding_and_exports_agree():
    """include/mfa_mlp.h, the ctypes table of mfa_brush_amd/surrogate.py and
    the library's exports name the same entry points (no compute call)."""
    from mfa_brush_amd import surrogate
    txt = open(os.path.join(ROOT, "include", "mfa_mlp.h")).read()
    declared = sorted(set(re.findall(r"\b(mfa_mlp_\w+)\s*\(", txt)))
    assert sorted(surrogate.SIGNATURES) == declared
    lib = C.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
