"""PLUTO's .dbl checkpoint format: the library's writer and reader against
files PLUTO itself wrote.

tests/golden/dbl_*.npz hold an initial state and the bytes of data.NNNN.dbl,
dbl.out and restart.out that the compiled reference wrote with .dbl output
every step (tests/golden/make_fixtures.py).  The oracle (bit-identical to the
reference) advances the same state; the host writer must then reproduce
PLUTO's data files byte for byte and its dbl.out text exactly, and the
restart.out records field by field (PLUTO leaves the record's padding and
fill bytes unset, so those are not compared).  The reader must recover the
oracle's states and the records' time, dt and step.  Host-only: no device.
"""
import os
import struct

import numpy as np
import pytest

from conftest import assert_states_equal, golden

FIXTURES = ["dbl_3d_8x7x5.npz", "dbl_ot2d_12.npz"]


@pytest.fixture(scope="module")
def native():
    from pluto_amd import native as N
    N.lib()
    return N


def _initial(z):
    from pluto_amd.problems import MHDState
    return MHDState(z["prim"].copy(), z["bx"].copy(), z["by"].copy(), z["bz"].copy(),
                    int(z["dims"]), tuple(float(x) for x in z["length"]))


def _records(raw):
    """(nstep, nfile[16], t, dt) of each 128-byte Restart record."""
    out = []
    for k in range(len(raw) // 128):
        rec = raw[128 * k:128 * (k + 1)]
        nstep = struct.unpack_from("<i", rec, 0)[0]
        nfile = struct.unpack_from("<16i", rec, 4)
        t, dt = struct.unpack_from("<2d", rec, 72)
        out.append((nstep, nfile, t, dt))
    return out


def _reference_files(z):
    files = sorted(k for k in z.keys() if k.startswith("data_") and k.endswith("_dbl"))
    return [bytes(z[k]) for k in files], bytes(z["dbl_out"]).decode(), bytes(z["restart_out"])


def _oracle_states(oracle_lib, z, steps):
    """Oracle states after each step count in ``steps`` and (time, next dt)."""
    st0 = _initial(z)
    nmax = max(steps) + 1
    run = st0.copy()
    dts, _ = oracle_lib.run(run, nmax, float(z["first_dt"]), float(z["cfl"]), float(z["cfl_max_var"]))
    out = {}
    for n in steps:
        st = st0.copy()
        if n:
            oracle_lib.run(st, n, float(z["first_dt"]), float(z["cfl"]), float(z["cfl_max_var"]))
        t = 0.0
        for q in range(n):
            t += dts[q]
        out[n] = (st, t, dts[n])
    return out


@pytest.mark.parametrize("name", FIXTURES)
def test_writer_reproduces_plutos_files(native, oracle_lib, name, tmp_path):
    z = golden(name)
    data, dbl_out, restart = _reference_files(z)
    recs = _records(restart)
    steps = [r[0] for r in recs]
    states = _oracle_states(oracle_lib, z, steps)
    for nfile, (nstep, _, t_ref, dt_ref) in enumerate(recs):
        st, t, dt = states[nstep]
        assert t == t_ref and dt == dt_ref, (nfile, t, t_ref, dt, dt_ref)
        native.write_dbl_host(str(tmp_path), nfile, st, t, dt, nstep)
        got = (tmp_path / ("data.%04d.dbl" % nfile)).read_bytes()
        assert got == data[nfile], "data.%04d.dbl differs from PLUTO's" % nfile
    assert (tmp_path / "dbl.out").read_text() == dbl_out
    mine = _records((tmp_path / "restart.out").read_bytes())
    assert mine == recs


@pytest.mark.parametrize("name", FIXTURES)
def test_reader_reads_plutos_files(native, oracle_lib, name, tmp_path):
    z = golden(name)
    data, dbl_out, restart = _reference_files(z)
    for nfile, b in enumerate(data):
        (tmp_path / ("data.%04d.dbl" % nfile)).write_bytes(b)
    (tmp_path / "dbl.out").write_text(dbl_out)
    (tmp_path / "restart.out").write_bytes(restart)
    recs = _records(restart)
    states = _oracle_states(oracle_lib, z, [r[0] for r in recs])
    init = _initial(z)
    for nfile, (nstep, _, t_ref, dt_ref) in enumerate(recs):
        st, t, dt, ns = native.read_dbl_host(str(tmp_path), nfile, init.shape, init.dims, init.length)
        assert (t, dt, ns) == (t_ref, dt_ref, nstep)
        assert_states_equal(st, states[nstep][0], "%s file %d" % (name, nfile))


def test_round_trip_and_big_endian(native, tmp_path):
    from pluto_amd import problems as P
    st = P.random_mhd_3d(6, seed=9)
    native.write_dbl_host(str(tmp_path), 0, st, 0.25, 1e-3, 17)
    back, t, dt, ns = native.read_dbl_host(str(tmp_path), 0, st.shape, 3)
    assert (t, dt, ns) == (0.25, 1e-3, 17)
    assert_states_equal(back, st, "round trip")
    # the same file written on a big-endian machine
    raw = np.fromfile(tmp_path / "data.0000.dbl", dtype="<f8")
    raw.astype(">f8").tofile(tmp_path / "data.0000.dbl")
    line = (tmp_path / "dbl.out").read_text().replace(" little ", " big ")
    (tmp_path / "dbl.out").write_text(line)
    rec = bytearray((tmp_path / "restart.out").read_bytes())
    for off, fmt in [(0, "i")] + [(4 + 4 * q, "i") for q in range(16)] + [(72, "d"), (80, "d")]:
        v = struct.unpack_from("<" + fmt, rec, off)[0]
        struct.pack_into(">" + fmt, rec, off, v)
    (tmp_path / "restart.out").write_bytes(bytes(rec))
    back, t, dt, ns = native.read_dbl_host(str(tmp_path), 0, st.shape, 3)
    assert (t, dt, ns) == (0.25, 1e-3, 17)
    assert_states_equal(back, st, "big-endian file")


def test_round_trip_2d_and_sequence(native, tmp_path):
    from pluto_amd import problems as P
    st = P.orszag_tang_2d(10)
    for n in range(3):
        native.write_dbl_host(str(tmp_path), n, st, 0.1 * n, 1e-4 * (n + 1), 2 * n)
    lines = (tmp_path / "dbl.out").read_text().splitlines()
    assert [int(x.split()[0]) for x in lines] == [0, 1, 2]
    back, t, dt, ns = native.read_dbl_host(str(tmp_path), 1, st.shape, 2)
    assert (t, dt, ns) == (0.1, 2e-4, 2)
    assert_states_equal(back, st, "2D round trip")
    # rewriting output #1 drops the later lines, as a restarted run would
    native.write_dbl_host(str(tmp_path), 1, st, 0.5, 1e-4, 9)
    assert len((tmp_path / "dbl.out").read_text().splitlines()) == 2


def test_reader_rejects_mismatches(native, tmp_path):
    from pluto_amd import problems as P
    st = P.random_mhd_3d(5, seed=2)
    native.write_dbl_host(str(tmp_path), 0, st, 0.0, 1e-3, 0)
    with pytest.raises(native.PGPUError):
        native.read_dbl_host(str(tmp_path), 0, (5, 5, 6), 3)     # wrong box
    with pytest.raises(native.PGPUError):
        native.read_dbl_host(str(tmp_path), 1, st.shape, 3)      # no output #1
    with pytest.raises(native.PGPUError):
        native.read_dbl_host(str(tmp_path), 0, (1, 5, 5), 2)     # 3D file, 2D box
    with pytest.raises(native.PGPUError):
        native.read_dbl_host(str(tmp_path / "missing"), 0, st.shape, 3)
