"""PLUTO .dbl output and restart of a resident domain.

The domain advances the stored initial state with PLUTO's dt control and
writes a .dbl output after every step; the files must equal, byte for byte,
the ones the compiled reference wrote in the same run (tests/golden/dbl_*.npz,
see test_dbl_io.py).  A domain restarted from one of PLUTO's own files must
then continue exactly as PLUTO did.
"""
import struct

import pytest

from test_dbl_io import FIXTURES, _initial, _records, _reference_files
from conftest import golden

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def native():
    from pluto_amd import native as N
    if N.device_count() < 1:
        pytest.fail("no HIP device visible")
    return N


@pytest.mark.parametrize("name", FIXTURES)
def test_domain_writes_plutos_files(native, name, tmp_path):
    _say("case: .dbl output of the domain, %s" % name)
    z = golden(name)
    data, dbl_out, restart = _reference_files(z)
    recs = _records(restart)
    dom = native.Domain.from_state(_initial(z))
    dom.set_dt(float(z["first_dt"]), float(z["cfl"]), float(z["cfl_max_var"]))
    done = 0
    for nfile, (nstep, _, _, _) in enumerate(recs):
        dom.run(nstep - done)
        done = nstep
        dom.write_dbl(str(tmp_path), nfile)
        got = (tmp_path / ("data.%04d.dbl" % nfile)).read_bytes()
        assert got == data[nfile], "data.%04d.dbl differs from PLUTO's" % nfile
    assert (tmp_path / "dbl.out").read_text() == dbl_out
    assert _records((tmp_path / "restart.out").read_bytes()) == recs
    dom.close()


@pytest.mark.parametrize("name", FIXTURES)
def test_restart_from_plutos_files(native, name, tmp_path):
    _say("case: restart of the domain from PLUTO's output #2, %s" % name)
    z = golden(name)
    data, dbl_out, restart = _reference_files(z)
    for nfile, b in enumerate(data):
        (tmp_path / ("data.%04d.dbl" % nfile)).write_bytes(b)
    (tmp_path / "dbl.out").write_text(dbl_out)
    (tmp_path / "restart.out").write_bytes(restart)
    recs = _records(restart)
    init = _initial(z)
    dom = native.Domain.from_state(init)
    dom.set_dt(1.0, float(z["cfl"]), float(z["cfl_max_var"]))   # CFL settings; dt comes from the file
    dom.read_dbl(str(tmp_path), 2)
    sc = dom.scalars()
    assert (sc["time"], sc["dt"], sc["step"]) == (recs[2][2], recs[2][3], recs[2][0])
    dom.run(recs[3][0] - recs[2][0])
    out = tmp_path / "out"
    out.mkdir()
    dom.write_dbl(str(out), 3)
    assert (out / "data.0003.dbl").read_bytes() == data[3]
    nstep, nfile, t, dt = _records((out / "restart.out").read_bytes()[3 * 128:])[0]
    assert (nstep, t, dt) == (recs[3][0], recs[3][2], recs[3][3])
    dom.close()
