import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def try_fixture():
    """The reference's try/ sample (conf_old + inputs + printed V_ff)."""
    from mfa_brush_amd.params import Flags, derive_tables, read_mfa_input, read_system_input
    z = np.load(os.path.join(ROOT, "tests", "golden", "try_fixture.npz"))
    fl = Flags(solvent=3, brush_in_good_sv=True, dpd_cut_off=2.24, bending=True, poiseuille=True)
    si = read_system_input(str(z["system_input"]), fl)
    mi = read_mfa_input(str(z["mfa_input"]), fl)
    p = derive_tables(fl, si, mi)
    return dict(p=p, si=si, mi=mi, a_type=z["a_type"].astype(np.int32), r0=z["r0"], v=z["v"], a=z["a"],
                v_ff=float(z["v_ff_check"]), fort41=z["fort41"])


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.load()
    return oracle


REF_INPUT_CASES = ("Filled_Channel", "Fully_Flexible_Brush_Channel", "Clog")


def load_ref_input(name):
    """One of the reference's shipped channel input sets (inputs/<name>):
    (params, a_type, r0, v) from tests/golden/ref_inputs.npz."""
    import json
    from mfa_brush_amd.params import Flags, derive_tables, read_mfa_input, read_system_input
    z = np.load(os.path.join(ROOT, "tests", "golden", "ref_inputs.npz"))
    fl = Flags(**json.loads(str(z[name + "/flags"])))
    si = read_system_input(str(z[name + "/system_input"]), fl)
    mi = read_mfa_input(str(z[name + "/mfa_input"]), fl)
    p = derive_tables(fl, si, mi)
    return p, z[name + "/a_type"].astype(np.int32), z[name + "/r0"], z[name + "/v"]
