"""Host side of the observable files (observation.f90:295-322): running
averages over observation steps, output every fifth step after n_relax, the
reference's line formats, and the Fortran E18.7 field of fort.80."""
from types import SimpleNamespace

import numpy as np
import pytest

from mfa_brush_amd.observe import Observer, fortran_e
from mfa_brush_amd.params import Flags


def test_fortran_e_format():
    assert fortran_e(12.345) == "     0.1234500E+02"
    assert fortran_e(-0.00012345678) == "    -0.1234568E-03"
    assert fortran_e(0.0) == "     0.0000000E+00"
    assert fortran_e(0.99999999) == "     0.1000000E+01"
    assert fortran_e(1.0) == "     0.1000000E+01"
    assert all(len(fortran_e(x)) == 18 for x in (1e-30, -3e12, 7.0))


class _Sim:
    def __init__(self):
        self.k = 0

    def energies(self, isys=0):
        self.k += 1
        return dict(v_fluid_fluid=-10.0 * self.k, v_fluid_wall=1.0, v_intra_molec=2.0, v_bend=0.5,
                    v_bend_melt=0.0, t_fluid=3.0)

    def chain_observables(self, isys=0):
        return dict(rg2_brush=np.array([1.0, 2.0, 3.0]), rg2_melt=np.array([0.5, 0.5, 1.0]),
                    anisotropy=np.array([0.25, 0.25]), momentum=np.array([0.0, 1.0, -1.0]) * self.k)

    def f_on_heads(self, isys=0):
        return np.array([[1.0, 0.0, -2.0], [0.5, 0.0, -1.0], [0.0, 1.0, 2.0], [0.0, 1.0, 4.0]])


def test_observer_averages_and_lines():
    p = SimpleNamespace(flags=Flags(bending=False, bending_melt=False), n_chain=4, dt=0.005,
                        a_w=0.0, a_wall=0.0, sigma_w=1.0, z_space_wall=20.0)
    ob = Observer(_Sim(), p, n_relax=0)
    outs = [ob.observe(i) for i in range(1, 11)]
    assert [bool(o) for o in outs] == [i % 5 == 0 for i in range(1, 11)]
    last = outs[-1]
    # v_total at call k: wall_wall(0) + 1 - 10k + 2; mean over k = 1..10
    f40 = last["fort.40"].split()
    assert int(f40[0]) == 10
    assert float(f40[1]) == pytest.approx(np.mean([3.0 - 10.0 * k for k in range(1, 11)]), abs=1e-5)
    assert float(f40[2]) == pytest.approx(3.0)
    assert float(f40[5]) == pytest.approx(np.mean(range(1, 11)), abs=1e-5)
    assert last["fort.41"] == "%7d%15.5f%15.5f%15.5f" % (10, -100.0, 2.0, 1.0)
    assert [float(x) for x in last["fort.42"].split()[1:]] == [1.0, 2.0, 3.0, 6.0, 0.5, 0.5, 1.0, 2.0]
    assert last["fort.43"] == "%7d%15.7f%15.7f" % (10, 0.25, 0.25)
    f80 = [float(last["fort.80"][i:i + 18]) for i in range(0, 7 * 18, 18)]
    assert f80 == pytest.approx([0.05, -1.5, 0.0, 3.0, 0.0, -2.0, -6.0])
    assert len(last["fort.40"]) == 7 + 6 * 15 and len(last["fort.42"]) == 7 + 8 * 15
