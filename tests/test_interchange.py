"""Interchange files (.pot/.potex, .den, .wvfn): the library's writers and
readers against the oracle's byte layout, a record spelled out byte by byte,
gfortran's subrecord split, and the reference's error behaviour."""
import os
import struct

import numpy as np
import pytest

from derek_amd import lib
from oracle import interchange as oi


def _rng(seed):
    return np.random.default_rng(seed)


def _cplx(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def test_record_spelled_out(tmp_path):
    """pot_write of a 1-point potential: the framing and field order written
    out by hand from the record description (no oracle code involved)."""
    nc = np.zeros((1, 2, 2), dtype=np.complex128)
    nc[0, 0, 0] = 1.5 - 2.0j
    nc[0, 1, 0] = 0.25j
    nc[0, 0, 1] = -3.0
    nc[0, 1, 1] = 1.5 - 2.0j
    p = tmp_path / "t.pot"
    lib.pot_write(str(p), (4, 5, 6), nc)
    payload_len = 12 + 4 * 16 + 4
    expect = (struct.pack(">i", payload_len) + struct.pack(">3i", 4, 5, 6)
              # nc_pot(1,1,1), (1,2,1), (1,1,2), (1,2,2): Fortran order, re then im
              + struct.pack(">8d", 1.5, -2.0, 0.0, 0.25, -3.0, 0.0, 1.5, -2.0)
              + struct.pack(">i", 1) + struct.pack(">i", payload_len))
    assert p.read_bytes() == expect


def test_subrecord_split_spelled_out():
    """10 payload bytes split at 4: markers -4/+4, -4/-4, +2/-2."""
    payload = bytes(range(10))
    expect = (struct.pack(">i", -4) + payload[0:4] + struct.pack(">i", 4)
              + struct.pack(">i", -4) + payload[4:8] + struct.pack(">i", -4)
              + struct.pack(">i", 2) + payload[8:10] + struct.pack(">i", -2))
    assert oi.frame(payload, max_sub=4) == expect
    assert oi.unframe(expect) == payload


@pytest.mark.parametrize("max_sub", [None, 40, 100])
def test_pot_den_bytes_match_oracle(tmp_path, monkeypatch, max_sub):
    if max_sub:
        monkeypatch.setenv("DEREK_IO_SUBRECORD", str(max_sub))
    rng = _rng(1)
    std, nfine = (3, 4, 5), 7
    nc = _cplx(rng, nfine, 2, 2)
    lib.pot_write(str(tmp_path / "a.pot"), std, nc, allocated=True)
    ref = oi.frame(oi.pot_payload(std, nc, True), max_sub or oi.MAX_SUBRECORD)
    assert (tmp_path / "a.pot").read_bytes() == ref
    back, alloc = lib.pot_read(str(tmp_path / "a.pot"), std, nfine)
    assert alloc and np.array_equal(back, nc)

    ch, sp = _cplx(rng, nfine), _cplx(rng, nfine, 3)
    lib.den_write(str(tmp_path / "a.den"), std, ch, sp, allocated=False)
    ref = oi.frame(oi.den_payload(std, ch, sp, False), max_sub or oi.MAX_SUBRECORD)
    assert (tmp_path / "a.den").read_bytes() == ref
    c2, s2, alloc = lib.den_read(str(tmp_path / "a.den"), std, nfine)
    assert not alloc and np.array_equal(c2, ch) and np.array_equal(s2, sp)


def test_std_index_matches_enumeration():
    """Standard-grid positions of a k-shifted sphere against the basis loop."""
    from oracle import basis as ob
    cell = np.diag([7.0, 8.0, 9.0])
    std, fine = ob.grid_dims(cell, 6.0, 2.0)
    inv = lib.inv_cell(cell)
    for kf in ([0.0, 0.0, 0.0], [0.25, -0.25, 0.25]):
        idx, _ = lib.sphere(inv, fine, kf, 6.0)
        got = lib.std_index(fine, std, idx)
        assert np.array_equal(got, oi.std_index(fine, std, idx))
        assert (got >= 0).all()
    # a bin outside the standard grid maps to -1
    far = np.array([(fine[2] // 2)], dtype=np.int32)  # mz = fine_z/2 > std_z range
    assert lib.std_index(fine, std, far)[0] == oi.std_index(fine, std, far)[0]


@pytest.mark.parametrize("max_sub", [None, 1000])
def test_wvfn_bytes_and_round_trip(tmp_path, monkeypatch, max_sub):
    if max_sub:
        monkeypatch.setenv("DEREK_IO_SUBRECORD", str(max_sub))
    rng = _rng(2)
    cell = np.diag([6.0, 6.0, 7.0])
    std, fine = lib.grid_dims(cell, 5.0, 2.0)
    inv = lib.inv_cell(cell)
    kfs = [[0.0, 0.0, 0.0], [0.25, 0.25, -0.25]]
    std_idx, psis = [], []
    nb = 3
    for kf in kfs:
        idx, _ = lib.sphere(inv, fine, kf, 5.0)
        std_idx.append(lib.std_index(fine, std, idx))
        psis.append(_cplx(rng, len(idx), nb))
    p = str(tmp_path / "a.wvfn")
    lib.wvfn_write(p, std, psis, std_idx)
    coeff = oi.sphere_to_coeff(std, psis, std_idx)
    assert open(p, "rb").read() == oi.frame(oi.wvfn_payload(std, coeff), max_sub or oi.MAX_SUBRECORD)
    back, alloc = lib.wvfn_read(p, std, std_idx, nb)
    assert alloc
    for a, b in zip(back, psis):
        assert np.array_equal(a, b)
    with pytest.raises(lib.DerekError) as e:  # wrong band count
        lib.wvfn_read(p, std, std_idx, nb + 1)
    assert e.value.code == lib.DEREK_ERR_ARG


def test_reader_errors(tmp_path):
    rng = _rng(3)
    std, nfine = (2, 2, 2), 5
    p = str(tmp_path / "a.pot")
    lib.pot_write(p, std, _cplx(rng, nfine, 2, 2))
    # pot_read's header check (pot.f90:326-332)
    with pytest.raises(lib.DerekError) as e:
        lib.pot_read(p, (2, 2, 3), nfine)
    assert e.value.code == lib.DEREK_ERR_ARG and "fine grid mismatch" in str(e.value)
    # a record too short for the requested grid
    with pytest.raises(lib.DerekError) as e:
        lib.pot_read(p, std, nfine + 1)
    assert e.value.code == lib.DEREK_ERR_IO
    # truncated file, missing file
    data = open(p, "rb").read()
    open(p, "wb").write(data[:-9])
    with pytest.raises(lib.DerekError) as e:
        lib.pot_read(p, std, nfine)
    assert e.value.code == lib.DEREK_ERR_IO
    with pytest.raises(lib.DerekError) as e:
        lib.pot_read(os.path.join(str(tmp_path), "missing.potex"), std, nfine)
    assert e.value.code == lib.DEREK_ERR_IO
    # a longer record is accepted (the reference reads the leading items)
    nc = _cplx(rng, nfine, 2, 2)
    payload = oi.pot_payload(std, nc) + b"\0" * 24
    open(p, "wb").write(oi.frame(payload))
    back, _ = lib.pot_read(p, std, nfine)
    assert np.array_equal(back, nc)


def test_scalar_potential_helpers(tmp_path):
    from derek_amd import interchange as ix
    v = _rng(4).normal(size=11)
    nc = ix.nc_pot_from_v(v)
    assert np.array_equal(nc[:, 0, 0].real, v) and np.array_equal(nc[:, 1, 1].real, v)
    assert not nc[:, 0, 1].any() and not nc[:, 1, 0].any()
    assert np.array_equal(ix.v_from_nc_pot(nc), v)
