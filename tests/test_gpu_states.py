"""A context's coefficients through the reference's STATES directory layout
(oa_states_write / oa_states_read): written after a step, read into a fresh
context, identical in the binary encodings; the next step from the read
states equals the next step of the writer."""
import numpy as np
import pytest

import oracle
from openatom_amd import system
from openatom_amd.step import CPContext

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fmt", [1, 3])
def test_states_directory_round_trip(gpu, tmp_path, fmt):
    s = system.build("water8", seed=11, kvec_source=oracle.stored_kvecs)
    d = tmp_path / "STATES"
    (d / "Spin.0_Kpt.0_Bead.0_Temper.0").mkdir(parents=True)
    args = (s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"], s["kz"])
    with CPContext(*args) as a, CPContext(*args) as b:
        a.set_psi(s["psi"])
        a.set_vks(s["vks"])
        a.step()
        a.write_states(d, fmt=fmt, iteration=1)
        assert (d / "Spin.0_Kpt.0_Bead.0_Temper.0" / "TimeStamp").read_text() == \
            "time step = 1\n"
        b.read_states(d, fmt=fmt)
        b.set_vks(s["vks"])
        assert np.array_equal(a.get_psi(), b.get_psi())
        ra, rb = a.step(), b.step()
        assert ra["eelec"] == rb["eelec"]
        assert np.array_equal(a.get_psi(), b.get_psi())
