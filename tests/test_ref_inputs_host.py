"""The reference's shipped channel input sets (inputs/Filled_Channel: 10 x 300
brush, 3000 solvent, SOLVENT 3 with good-solvent brushes, bending, DPD cutoff
2.24; inputs/Fully_Flexible_Brush_Channel: 32 x 100 brush, 1086 solvent,
SOLVENT 2 with non-additive sigmas 1.2; inputs/Clog: 108 x 10 brush and a
4823-bead droplet in an 80.5 x 13.4 x 40 channel, SOLVENT 3, bending,
Poiseuille force 0.1 x) read by the host readers, and the oracle's
linked-cell pair energy on them against its brute-force sum
(util.f90:284-330 semantics)."""
import numpy as np
import pytest

from conftest import REF_INPUT_CASES, load_ref_input


# box of each set as its mfa_input reports it ("MD Box dimensions")
BOX = {"Filled_Channel": (30.0, 20.0, 40.0), "Fully_Flexible_Brush_Channel": (30.0, 20.0, 40.0),
       "Clog": (80.49816, 13.41636, 40.0)}


@pytest.mark.parametrize("name", REF_INPUT_CASES)
def test_ref_inputs_read_and_pair_energy(oracle_mod, name):
    p, t, r0, v = load_ref_input(name)
    assert t.size == p.n_part == p.n_mon * p.n_chain + p.n_mon_d * p.n_chain_d
    assert p.boundary == pytest.approx(BOX[name], abs=1e-4)
    assert (t[::p.n_mon][:p.n_chain] == 1).all()           # heads lead their chains
    assert ((r0[:, 2] > 0) & (r0[:, 2] < p.boundary[2])).all()
    o = oracle_mod.Oracle(p)
    _, e_cell = o.fluid_fluid(r0, v, t, step=0, noise=0)
    e_brute = o.check_fluid_fluid(r0, t)
    assert e_cell == pytest.approx(e_brute, rel=1e-12)
