"""The reference's ASCII field file (restart / spectral field format) written
by the library (calypso_sph_write_field_file, the writer behind
calypso_mhd_write_restart) against the reference's own writers,
write_step_data (m_time_data_IO.f90:107-119) and write_field_data
(field_data_IO.f90:267-298), compiled from its sources into oracle/_ref:
the two files must be byte for byte the same (host only, no GPU)."""
import ctypes
import os

import numpy as np
import pytest

import oracle
from calypso_amd import _lib

pytestmark = pytest.mark.skipif(not oracle.ref_available(),
                                reason="reference build (oracle/_ref) missing")
KCHARA = 255


def write_both(tmp_path, rank, istep, time, dt, names, ncomp, data):
    nfld = len(names)
    nnod = data.shape[0]
    data_f = np.asfortranarray(data)          # (nnod, ntot), component-major
    mine, theirs = tmp_path / "mine.fst", tmp_path / "ref.fst"
    lib = _lib.load()
    cnames = (ctypes.c_char_p * nfld)(*[n.encode() for n in names])
    nc = np.ascontiguousarray(ncomp, dtype=np.int32)
    _lib.check(lib.calypso_sph_write_field_file(
        str(mine).encode(), rank, istep, time, dt, nnod, nfld, nc.ctypes.data,
        ctypes.cast(cnames, ctypes.c_void_p), data_f.ctypes.data))
    padded = b"".join(n.encode().ljust(KCHARA) for n in names)
    f = oracle.ref().ref_write_field_file
    f.restype = None
    f(ctypes.c_char_p(str(theirs).encode()), ctypes.c_int(rank), ctypes.c_int(istep),
      ctypes.c_double(time), ctypes.c_double(dt), ctypes.c_int(nnod), ctypes.c_int(nfld),
      nc.ctypes.data_as(ctypes.c_void_p), ctypes.c_char_p(padded),
      data_f.ctypes.data_as(ctypes.c_void_p))
    return mine.read_bytes(), theirs.read_bytes()


def edge_values():
    return np.array([0.0, 1.0, -1.0, 0.1, -2.5e-3, 9.999999999999999e99,
                     9.9999999999999995e-1, 1.0e-300, -3.7e-310, 5e-324, 1.7976931348623157e308,
                     123456789.0123456, -0.0, 4.9e-5, 2.0 ** -1022, 1.0 / 3.0])


def test_restart_field_set_bytes(tmp_path):
    """The six restart fields (velocity, temperature, magnetic field, the three
    previous explicit terms) on a random field with edge values mixed in."""
    rng = np.random.default_rng(3)
    names = ["velocity", "temperature", "magnetic_field", "previous_momentum",
             "previous_induction", "previous_heat"]
    ncomp = [3, 1, 3, 3, 3, 1]
    nnod = 40
    data = rng.standard_normal((nnod, sum(ncomp))) * 10.0 ** rng.integers(-40, 40, (nnod, 14))
    e = edge_values()
    data.ravel()[: e.size] = e
    mine, theirs = write_both(tmp_path, 0, 1200, 0.06000000000000001, 5.0e-5, names, ncomp, data)
    assert mine == theirs


def test_many_fields_wrap_the_component_record(tmp_path):
    """More than ten fields: the '(10i5)' record of component counts wraps."""
    rng = np.random.default_rng(4)
    names = [f"field_{i}" for i in range(23)]
    ncomp = [int(x) for x in rng.integers(1, 7, 23)]
    data = rng.standard_normal((7, sum(ncomp)))
    mine, theirs = write_both(tmp_path, 3, 7, 12.5, 1.0e-4, names, ncomp, data)
    assert mine == theirs


def test_rejects_bad_component_counts(tmp_path):
    lib = _lib.load()
    names = (ctypes.c_char_p * 1)(b"x")
    nc = np.array([10], dtype=np.int32)
    data = np.zeros(10)
    rc = lib.calypso_sph_write_field_file(str(tmp_path / "x").encode(), 0, 0, 0.0, 0.0, 1, 1,
                                          nc.ctypes.data, ctypes.cast(names, ctypes.c_void_p),
                                          data.ctypes.data)
    assert rc != 0
