"""The C-ABI library on the host (no GPU calls): it loads, exports every
symbol declared in include/derek_hip.h, and its restatement of the
reference discretisation is bit-identical to the oracle's."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from derek_amd import config
from oracle import basis as ob
from oracle import ks as oks

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "derek_hip.h")


def declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"^int32_t\s+(derek_\w+)\(", txt, flags=re.M)))


def test_exports_every_declared_symbol(lib):
    names = declared()
    assert len(names) >= 30
    L = lib.load()
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(lib.EXPORTS)
    assert L.derek_hip_version() >= 1


@pytest.mark.parametrize("name", ["gamma32", "gamma108", "kmesh72"])
def test_grid_dims_match(lib, name):
    sp = config.workload(name)
    cell = np.diag(sp.lengths)
    assert lib.grid_dims(cell, sp.ecut, sp.g_fine_scale) == tuple(ob.grid_dims(cell, sp.ecut, sp.g_fine_scale)) or \
        list(lib.grid_dims(cell, sp.ecut, sp.g_fine_scale)) == list(ob.grid_dims(cell, sp.ecut, sp.g_fine_scale))


@pytest.mark.parametrize("L,dims", [((10.0, 10.0, 10.0), (32, 32, 32)), ((7.3, 9.1, 11.7), (15, 18, 20))])
def test_maps_bit_exact(lib, L, dims):
    cell = np.diag(L)
    inv = lib.inv_cell(cell)
    assert np.array_equal(inv, ob.inv_cell_orthorhombic(cell))
    g, m = lib.g_map(inv, dims)
    G, M = ob.g_map(ob.inv_cell_orthorhombic(cell), list(dims))
    assert np.array_equal(m, M) and np.array_equal(g, G)
    r, f = lib.real_map(cell, dims)
    R, F = ob.real_map(cell, list(dims))
    assert np.array_equal(r, R) and np.array_equal(f, F)
    assert np.array_equal(lib.gsq(inv, dims), ob.gsq_on_fft_grid(ob.inv_cell_orthorhombic(cell), list(dims)))


@pytest.mark.parametrize("mp", [(1, 1, 1), (2, 2, 2), (4, 4, 4), (3, 1, 2)])
def test_kpoints_bit_exact(lib, mp):
    assert np.array_equal(lib.kpoints(mp), ob.kpoint_grid(mp))


@pytest.mark.parametrize("name", ["gamma32", "kmesh72"])
def test_sphere_bit_exact(lib, name):
    sp = config.workload(name)
    cell = np.diag(sp.lengths)
    inv = ob.inv_cell_orthorhombic(cell)
    _, dims = ob.grid_dims(cell, sp.ecut, sp.g_fine_scale)
    G, M = ob.g_map(inv, dims)
    for kf in ob.kpoint_grid(sp.mp)[:3]:
        idx, kin = lib.sphere(inv, dims, kf, sp.ecut)
        oidx, okin, _ = ob.sphere(G, M, dims, ob.kcart(inv, kf), sp.ecut)
        assert np.array_equal(idx, oidx) and np.array_equal(kin, okin)


def test_gaussian_pot_matches_oracle(lib):
    sp = config.workload("gamma108")
    sp2 = config.small_system(seed=7, cos_amp=0.2)
    cell = np.diag(sp2.lengths)
    _, dims = ob.grid_dims(cell, sp2.ecut, sp2.g_fine_scale)
    v = lib.gaussian_pot(cell, dims, sp2.centres, sp2.heights, sp2.widths, 0.2)
    R, _ = ob.real_map(cell, dims)
    ref = oks.gaussian_wells(R, sp2.lengths, sp2.centres, sp2.heights, sp2.widths, 0.2)
    assert np.max(np.abs(v - ref)) < 1e-13 * np.max(np.abs(ref))
    assert len(sp.centres) == 16


def test_host_argument_errors(lib):
    with pytest.raises(lib.DerekError) as e:
        lib.grid_dims(np.eye(3) * 10, -1.0, 2.0)
    assert e.value.code == lib.DEREK_ERR_ARG and "positive" in str(e.value)
    with pytest.raises(lib.DerekError) as e:
        lib.kpoints((0, 1, 1))
    assert e.value.code == lib.DEREK_ERR_ARG
    cell = np.diag([10.0, 10.0, 10.0])
    cell[0, 1] = 1.0
    with pytest.raises(lib.DerekError) as e:
        lib.inv_cell(cell)
    assert "orthorhombic" in str(e.value)


def test_null_context_refused(lib):
    """Entry points taking a context refuse a null one on the host, before
    any HIP call (here: the launch-shape and GEMM-precision setters)."""
    L = lib.load()
    assert L.derek_hip_set_gemm(None, 2) == lib.DEREK_ERR_ARG
    assert L.derek_hip_set_launch(None, 2, 256, 3) == lib.DEREK_ERR_ARG
    buf = C.create_string_buffer(256)
    L.derek_hip_last_error(buf, C.c_int64(256))
    assert b"null" in buf.value
