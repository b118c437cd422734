"""ctypes binding of ``libderek_hip.so`` (the C-ABI in include/derek_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (make in
derek_amd/csrc).  There is no fallback: if the library is missing every
call raises, naming the build step.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libderek_hip.so")

DEREK_OK, DEREK_ERR_ARG, DEREK_ERR_HIP, DEREK_ERR_NUMERIC, DEREK_ERR_STATE, DEREK_ERR_IO = 0, 1, 2, 3, 4, 5
DEREK_FP64, DEREK_MIXED = 0, 1

# every symbol declared in include/derek_hip.h
EXPORTS = (
    "derek_hip_version", "derek_hip_last_error",
    "derek_basis_grid_dims", "derek_basis_inv_cell", "derek_basis_g_map", "derek_basis_real_map",
    "derek_basis_kpoints", "derek_basis_sphere", "derek_basis_gsq", "derek_basis_std_index", "derek_gaussian_pot",
    "derek_io_pot_write", "derek_io_pot_read", "derek_io_den_write", "derek_io_den_read",
    "derek_io_wvfn_write", "derek_io_wvfn_read",
    "derek_info_read", "derek_info_free", "derek_info_num", "derek_info_str", "derek_info_present",
    "derek_info_cell", "derek_info_wells", "derek_info_log", "derek_info_preset_pot", "derek_info_help",
    "derek_hip_init", "derek_hip_destroy", "derek_hip_set_basis", "derek_hip_set_gsq",
    "derek_hip_set_vloc", "derek_hip_set_vext",
    "derek_hip_fft3d", "derek_hip_fft1d", "derek_hip_wave_axpby", "derek_hip_apply_h", "derek_hip_overlap", "derek_hip_overlap_noalloc",
    "derek_hip_rotate",
    "derek_hip_density", "derek_hip_lda", "derek_hip_hartree", "derek_hip_heev", "derek_hip_hegv",
    "derek_hip_scf_init", "derek_hip_scf_step", "derek_hip_scf_polish", "derek_hip_nscf", "derek_hip_get_eigenvalues",
    "derek_hip_get_residuals", "derek_hip_get_occupations", "derek_hip_get_density", "derek_hip_get_potential", "derek_hip_get_psi", "derek_hip_put_psi",
    "derek_hip_apply_h_resident", "derek_hip_sync", "derek_hip_stream",
    "derek_hip_profile_enable", "derek_hip_profile_reset", "derek_hip_profile_read", "derek_hip_get_tuning",
    "derek_hip_set_launch", "derek_hip_set_gemm", "derek_hip_set_tiles",
)

KERNEL_CLASSES = ("sticks_z_inv", "planes_y_inv", "planes_x_vmul", "planes_y_fwd", "sticks_z_fwd", "planes_x_density",
                  "overlap", "rotate", "fft_z", "fft_y", "fft_x", "dense", "other", "slab_yx_vmul",
                  "slab_yx_density")


class DerekError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Dims(C.Structure):
    _fields_ = [("ngrid", C.c_int32 * 3), ("nk", C.c_int32), ("nbands", C.c_int32),
                ("precision", C.c_int32), ("volume", C.c_double), ("max_npw", C.c_int64)]


class ScfParams(C.Structure):
    _fields_ = [("nelec", C.c_double), ("kT", C.c_double), ("mix_beta", C.c_double),
                ("mix_q0", C.c_double), ("n_david_first", C.c_int32), ("n_david", C.c_int32),
                ("history", C.c_int32), ("reserved", C.c_int32), ("seed", C.c_uint64),
                ("fp32_switch_tol", C.c_double), ("refine_steps", C.c_int32), ("kbatch", C.c_int32),
                ("use_graph", C.c_int32), ("autotune", C.c_int32)]


class ScfReport(C.Structure):
    _fields_ = [("iscf", C.c_int32), ("ndav", C.c_int32), ("etot", C.c_double), ("e_band", C.c_double),
                ("e_hartree", C.c_double), ("e_xc", C.c_double), ("e_dc", C.c_double),
                ("minus_ts", C.c_double), ("fermi", C.c_double), ("drho", C.c_double),
                ("max_res", C.c_double), ("t_dense", C.c_double), ("t_step", C.c_double), ("fp32", C.c_int32),
                ("refine_left", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def load():
    """Load the library once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C derek_amd/csrc)")
        _lib = C.CDLL(LIB_PATH)
        for name in EXPORTS:
            getattr(_lib, name).restype = C.c_int32
    return _lib


def last_error():
    buf = C.create_string_buffer(4096)
    load().derek_hip_last_error(buf, C.c_int64(len(buf)))
    return buf.value.decode(errors="replace")


def check(code):
    if code != DEREK_OK:
        raise DerekError(code, last_error())


def ptr(a, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


def _c128(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------- discretisation (host)
def grid_dims(cell, ecut, g_fine_scale):
    cell9 = np.asfortranarray(np.asarray(cell, dtype=np.float64)).ravel(order="F")
    s = (C.c_int32 * 3)()
    f = (C.c_int32 * 3)()
    check(load().derek_basis_grid_dims(ptr(cell9), C.c_double(ecut), C.c_double(g_fine_scale), s, f))
    return list(s), list(f)


def inv_cell(cell):
    cell9 = np.asarray(cell, dtype=np.float64).ravel(order="F").copy()
    out = np.zeros(9)
    check(load().derek_basis_inv_cell(ptr(cell9), ptr(out)))
    return out.reshape(3, 3, order="F")


def g_map(inv, dims):
    d = (C.c_int32 * 3)(*dims)
    n = int(np.prod(dims))
    g = np.zeros((n, 3), order="F")
    m = np.zeros((n, 3), dtype=np.int32, order="F")
    inv9 = np.asarray(inv, dtype=np.float64).ravel(order="F").copy()
    check(load().derek_basis_g_map(ptr(inv9), d, g.ctypes.data_as(C.POINTER(C.c_double)),
                                   m.ctypes.data_as(C.POINTER(C.c_int32))))
    return g, m


def real_map(cell, dims):
    d = (C.c_int32 * 3)(*dims)
    n = int(np.prod(dims))
    r = np.zeros((n, 3), order="F")
    f = np.zeros((n, 3), order="F")
    cell9 = np.asarray(cell, dtype=np.float64).ravel(order="F").copy()
    check(load().derek_basis_real_map(ptr(cell9), d, r.ctypes.data_as(C.POINTER(C.c_double)),
                                      f.ctypes.data_as(C.POINTER(C.c_double))))
    return r, f


def kpoints(mp):
    m = (C.c_int32 * 3)(*mp)
    nk = C.c_int32()
    check(load().derek_basis_kpoints(m, None, C.byref(nk)))
    out = np.zeros((nk.value, 3), order="F")
    check(load().derek_basis_kpoints(m, out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(nk)))
    return out


def sphere(inv, dims, kfrac, ecut):
    d = (C.c_int32 * 3)(*dims)
    inv9 = np.asarray(inv, dtype=np.float64).ravel(order="F").copy()
    kf = np.ascontiguousarray(kfrac, dtype=np.float64)
    n = C.c_int64()
    check(load().derek_basis_sphere(ptr(inv9), d, ptr(kf), C.c_double(ecut), C.c_int64(0), C.byref(n), None, None))
    idx = np.zeros(n.value, dtype=np.int32)
    kin = np.zeros(n.value)
    check(load().derek_basis_sphere(ptr(inv9), d, ptr(kf), C.c_double(ecut), C.c_int64(n.value), C.byref(n),
                                    idx.ctypes.data_as(C.POINTER(C.c_int32)), ptr(kin)))
    return idx, kin


def std_index(fine_dims, std_dims, fft_index):
    """Standard-grid positions (basis order, -1 outside) of fine-grid FFT bins."""
    idx = np.ascontiguousarray(fft_index, dtype=np.int32)
    out = np.zeros(len(idx), dtype=np.int64)
    check(load().derek_basis_std_index((C.c_int32 * 3)(*fine_dims), (C.c_int32 * 3)(*std_dims),
                                       C.c_int64(len(idx)), idx.ctypes.data_as(C.POINTER(C.c_int32)),
                                       out.ctypes.data_as(C.POINTER(C.c_int64))))
    return out


# ---------------------------------------------------------------- .info input (host)
class Info:
    """A parsed .info file (the reference's input language)."""

    def __init__(self, path, safety="Normal"):
        self.h = C.c_void_p()
        check(load().derek_info_read(_path(path), safety.encode(), C.byref(self.h)))

    def close(self):
        if self.h:
            load().derek_info_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def num(self, key, n=1):
        v = np.zeros(3)
        check(load().derek_info_num(self.h, key.encode(), ptr(v)))
        return v[0] if n == 1 else v[:n].copy()

    def str(self, key):
        buf = C.create_string_buffer(512)
        check(load().derek_info_str(self.h, key.encode(), buf, C.c_int64(len(buf))))
        return buf.value.decode()

    def present(self, key):
        p = C.c_int32()
        check(load().derek_info_present(self.h, key.encode(), C.byref(p)))
        return bool(p.value)

    def cell(self):
        c = np.zeros(9)
        check(load().derek_info_cell(self.h, ptr(c)))
        return c.reshape(3, 3, order="F")

    def wells(self):
        n = C.c_int32(0)
        check(load().derek_info_wells(self.h, C.byref(n), None, None, None))
        c, hh, w = np.zeros(3 * n.value), np.zeros(n.value), np.zeros(n.value)
        if n.value:
            check(load().derek_info_wells(self.h, C.byref(n), ptr(c), ptr(hh), ptr(w)))
        return c.reshape(-1, 3), hh, w

    def log(self):
        buf = C.create_string_buffer(1 << 16)
        check(load().derek_info_log(self.h, buf, C.c_int64(len(buf))))
        return buf.value.decode()

    def preset_potential(self, fine_dims):
        v = np.zeros(int(np.prod(fine_dims)))
        check(load().derek_info_preset_pot(self.h, (C.c_int32 * 3)(*fine_dims), ptr(v)))
        return v


def info_help(mode, term=""):
    buf = C.create_string_buffer(1 << 16)
    check(load().derek_info_help(mode.encode(), term.encode(), buf, C.c_int64(len(buf))))
    return buf.value.decode()


# ---------------------------------------------------------------- interchange files (host)
def _path(path):
    return os.fsencode(path)


def pot_write(path, std_dims, nc_pot, allocated=True):
    """Potential record: nc_pot is complex (nfine, 2, 2)."""
    a = np.asarray(nc_pot, dtype=np.complex128)
    flat = np.ascontiguousarray(a.ravel(order="F"))
    check(load().derek_io_pot_write(_path(path), (C.c_int32 * 3)(*std_dims), C.c_int64(a.shape[0]), ptr(flat),
                                    C.c_int32(1 if allocated else 0)))


def pot_read(path, std_dims, nfine):
    """-> (nc_pot complex (nfine, 2, 2), allocated)."""
    flat = np.zeros(4 * nfine, dtype=np.complex128)
    alloc = C.c_int32()
    check(load().derek_io_pot_read(_path(path), (C.c_int32 * 3)(*std_dims), C.c_int64(nfine), ptr(flat),
                                   C.byref(alloc)))
    return flat.reshape((nfine, 2, 2), order="F"), bool(alloc.value)


def den_write(path, std_dims, charge, spin=None, allocated=True):
    """Density record: charge complex (nfine,), spin complex (nfine, 3) or None (zeros)."""
    ch = np.ascontiguousarray(charge, dtype=np.complex128)
    sp = None if spin is None else np.ascontiguousarray(np.asarray(spin, dtype=np.complex128).ravel(order="F"))
    check(load().derek_io_den_write(_path(path), (C.c_int32 * 3)(*std_dims), C.c_int64(len(ch)), ptr(ch), ptr(sp),
                                    C.c_int32(1 if allocated else 0)))


def den_read(path, std_dims, nfine):
    """-> (charge (nfine,), spin (nfine, 3), allocated)."""
    ch = np.zeros(nfine, dtype=np.complex128)
    sp = np.zeros(3 * nfine, dtype=np.complex128)
    alloc = C.c_int32()
    check(load().derek_io_den_read(_path(path), (C.c_int32 * 3)(*std_dims), C.c_int64(nfine), ptr(ch), ptr(sp),
                                   C.byref(alloc)))
    return ch, sp.reshape((nfine, 3), order="F"), bool(alloc.value)


def _sphere_args(psis, std_idx):
    npw = np.array([len(i) for i in std_idx], dtype=np.int64)
    idx = np.ascontiguousarray(np.concatenate([np.asarray(i, dtype=np.int64) for i in std_idx]))
    return npw, idx


def wvfn_write(path, std_dims, psis, std_idx, allocated=True):
    """Wavefunction record from per-k sphere blocks psis[k] (npw_k x nb)."""
    nb = psis[0].shape[1]
    npw, idx = _sphere_args(psis, std_idx)
    flat = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.complex128).ravel(order="F")
                                                for p in psis]))
    check(load().derek_io_wvfn_write(_path(path), (C.c_int32 * 3)(*std_dims), C.c_int32(len(psis)), C.c_int32(nb),
                                     npw.ctypes.data_as(C.POINTER(C.c_int64)),
                                     idx.ctypes.data_as(C.POINTER(C.c_int64)), ptr(flat),
                                     C.c_int32(1 if allocated else 0)))


def wvfn_read(path, std_dims, std_idx, nb):
    """-> (list of per-k sphere blocks npw_k x nb, allocated)."""
    npw, idx = _sphere_args(None, std_idx)
    flat = np.zeros(int(npw.sum()) * nb, dtype=np.complex128)
    alloc = C.c_int32()
    check(load().derek_io_wvfn_read(_path(path), (C.c_int32 * 3)(*std_dims), C.c_int32(len(std_idx)), C.c_int32(nb),
                                    npw.ctypes.data_as(C.POINTER(C.c_int64)),
                                    idx.ctypes.data_as(C.POINTER(C.c_int64)), ptr(flat), C.byref(alloc)))
    out, o = [], 0
    for n in npw:
        out.append(flat[o:o + n * nb].reshape((n, nb), order="F"))
        o += n * nb
    return out, bool(alloc.value)


def gsq(inv, dims):
    d = (C.c_int32 * 3)(*dims)
    inv9 = np.asarray(inv, dtype=np.float64).ravel(order="F").copy()
    out = np.zeros(int(np.prod(dims)))
    check(load().derek_basis_gsq(ptr(inv9), d, ptr(out)))
    return out


def gaussian_pot(cell, dims, centres, heights, widths, cos_amp=0.0):
    d = (C.c_int32 * 3)(*dims)
    cell9 = np.asarray(cell, dtype=np.float64).ravel(order="F").copy()
    cen = np.ascontiguousarray(np.asarray(centres, dtype=np.float64).reshape(-1, 3))
    h = np.ascontiguousarray(heights, dtype=np.float64)
    w = np.ascontiguousarray(widths, dtype=np.float64)
    v = np.zeros(int(np.prod(dims)))
    check(load().derek_gaussian_pot(ptr(cell9), d, C.c_int32(len(h)), ptr(cen) if len(h) else None,
                                    ptr(h) if len(h) else None, ptr(w) if len(h) else None,
                                    C.c_double(cos_amp), ptr(v)))
    return v


# ---------------------------------------------------------------- context
class Context:
    """Owns one derek_ctx (all device memory of one SCF)."""

    def __init__(self, ngrid, nk, nbands, volume, max_npw, precision=DEREK_FP64):
        self.lib = load()
        d = Dims()
        d.ngrid = (C.c_int32 * 3)(*ngrid)
        d.nk, d.nbands, d.precision = nk, nbands, precision
        d.volume, d.max_npw = volume, max_npw
        self.h = C.c_void_p()
        check(self.lib.derek_hip_init(C.byref(self.h), C.byref(d)))
        self.ngrid, self.N = list(ngrid), int(np.prod(ngrid))
        self.nk, self.nbands, self.precision = nk, nbands, precision
        self.npw = [0] * nk

    def close(self):
        if self.h:
            check(self.lib.derek_hip_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_basis(self, k, fft_index, kin, weight):
        idx = np.ascontiguousarray(fft_index, dtype=np.int32)
        kn = np.ascontiguousarray(kin, dtype=np.float64)
        check(self.lib.derek_hip_set_basis(self.h, C.c_int32(k), C.c_int64(len(idx)),
                                           idx.ctypes.data_as(C.POINTER(C.c_int32)), ptr(kn), C.c_double(weight)))
        self.npw[k] = len(idx)

    def set_gsq(self, g):
        g = np.ascontiguousarray(g, dtype=np.float64)
        check(self.lib.derek_hip_set_gsq(self.h, ptr(g)))

    def set_vloc(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        check(self.lib.derek_hip_set_vloc(self.h, ptr(v)))

    def set_vext(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        check(self.lib.derek_hip_set_vext(self.h, ptr(v)))

    def fft3d(self, grid, direction):
        a = np.array(grid, dtype=np.complex128, copy=True)
        shape = a.shape
        a = np.ascontiguousarray(a.reshape(-1, self.N))
        check(self.lib.derek_hip_fft3d(self.h, a.ctypes.data_as(C.POINTER(C.c_double)),
                                       C.c_int32(a.shape[0]), C.c_int32(direction)))
        return a.reshape(shape)

    def fft1d(self, lines, direction):
        """In-place transforms of each row of ``lines`` (howmany x n complex),
        n one of the grid's lengths; reference sign and normalisation."""
        a = np.array(lines, dtype=np.complex128, copy=True)
        shape = a.shape
        n = shape[-1]
        a = np.ascontiguousarray(a.reshape(-1, n))
        check(self.lib.derek_hip_fft1d(self.h, a.ctypes.data_as(C.POINTER(C.c_double)), C.c_int32(n),
                                       C.c_int64(a.shape[0]), C.c_int32(direction)))
        return a.reshape(shape)

    def wave_axpby(self, a, x, b=0.0, y=None):
        """a x + b y for complex arrays of one shape (b = 0: a x, y unread)."""
        xa = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
        ca = (C.c_double * 2)(complex(a).real, complex(a).imag)
        cb = (C.c_double * 2)(complex(b).real, complex(b).imag)
        yp = None
        if complex(b) != 0:
            ya = np.ascontiguousarray(np.asarray(y, dtype=np.complex128))
            if ya.shape != xa.shape:
                raise ValueError("x and y shapes differ")
            yp = ya.ctypes.data_as(C.POINTER(C.c_double))
        z = np.empty_like(xa)
        check(self.lib.derek_hip_wave_axpby(self.h, C.c_int64(xa.size), ca, xa.ctypes.data_as(C.POINTER(C.c_double)),
                                            cb, yp, z.ctypes.data_as(C.POINTER(C.c_double))))
        return z

    def apply_h(self, k, psi):
        """psi: (nb, npw_k) complex (each band contiguous) -> H psi."""
        p, pp = _c128(np.atleast_2d(psi))
        out = np.zeros_like(p)
        check(self.lib.derek_hip_apply_h(self.h, C.c_int32(k), C.c_int32(p.shape[0]), pp,
                                         out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def overlap(self, X, Y):
        """S[i,j] = <X_i|Y_j> for band-stacked X (m, npw), Y (n, npw)."""
        x, xp = _c128(np.atleast_2d(X))
        if Y is X:
            y, yp = x, xp  # same pointer: Hermitian kernel
        else:
            y, yp = _c128(np.atleast_2d(Y))
        m, n, npw = x.shape[0], y.shape[0], x.shape[1]
        S = np.zeros((n, m), dtype=np.complex128)  # column-major m x n == C-order (n, m)
        check(self.lib.derek_hip_overlap(self.h, C.c_int64(npw), C.c_int32(m), C.c_int32(n), xp, yp,
                                         S.ctypes.data_as(C.POINTER(C.c_double))))
        return S.T.copy()

    def rotate(self, X, U):
        """Z_j = sum_i U[i,j] X_i; X (m, npw), U (m, n) -> (n, npw)."""
        x, xp = _c128(np.atleast_2d(X))
        m, npw = x.shape
        n = U.shape[1]
        u, up = _c128(np.asarray(U).T)  # column-major m x n
        Z = np.zeros((n, npw), dtype=np.complex128)
        check(self.lib.derek_hip_rotate(self.h, C.c_int64(npw), C.c_int32(m), C.c_int32(n), xp, up,
                                        Z.ctypes.data_as(C.POINTER(C.c_double))))
        return Z

    def density(self, k, psi, occ, rho=None):
        p, pp = _c128(np.atleast_2d(psi))
        o = np.ascontiguousarray(occ, dtype=np.float64)
        acc = rho is not None
        r = np.array(rho, dtype=np.float64, copy=True) if acc else np.zeros(self.N)
        check(self.lib.derek_hip_density(self.h, C.c_int32(k), C.c_int32(p.shape[0]), pp, ptr(o), ptr(r),
                                         C.c_int32(1 if acc else 0)))
        return r

    def lda(self, rho):
        r = np.ascontiguousarray(rho, dtype=np.float64)
        e = np.zeros_like(r)
        v = np.zeros_like(r)
        check(self.lib.derek_hip_lda(self.h, C.c_int64(len(r)), ptr(r), ptr(e), ptr(v)))
        return e, v

    def hartree(self, rho):
        r = np.ascontiguousarray(rho, dtype=np.float64)
        v = np.zeros_like(r)
        eh = C.c_double()
        check(self.lib.derek_hip_hartree(self.h, ptr(r), ptr(v), C.byref(eh)))
        return v, eh.value

    def heev(self, A, nev=None):
        """Lowest nev eigenpairs of Hermitian A (n x n) on the GPU."""
        A = np.asarray(A, dtype=np.complex128)
        n = A.shape[0]
        nev = n if nev is None else nev
        a = np.ascontiguousarray(A.T)  # column-major
        w = np.zeros(nev)
        Z = np.zeros((nev, n), dtype=np.complex128)
        check(self.lib.derek_hip_heev(self.h, C.c_int32(n), C.c_int32(nev), a.ctypes.data_as(C.POINTER(C.c_double)),
                                      ptr(w), Z.ctypes.data_as(C.POINTER(C.c_double))))
        return w, Z.T.copy()

    def hegv(self, S, H, nev):
        """Lowest nev solutions of H c = w S c on the GPU; returns (w, C m x nev)."""
        S = np.asarray(S, dtype=np.complex128)
        H = np.asarray(H, dtype=np.complex128)
        m = S.shape[0]
        s_ = np.ascontiguousarray(S.T)
        h_ = np.ascontiguousarray(H.T)
        w = np.zeros(nev)
        Cm = np.zeros((nev, m), dtype=np.complex128)
        check(self.lib.derek_hip_hegv(self.h, C.c_int32(m), C.c_int32(nev), s_.ctypes.data_as(C.POINTER(C.c_double)),
                                      h_.ctypes.data_as(C.POINTER(C.c_double)), ptr(w),
                                      Cm.ctypes.data_as(C.POINTER(C.c_double))))
        return w, Cm.T.copy()

    # ---- resident SCF
    def scf_init(self, nelec, kT=0.0, beta=0.5, q0=1.0, n_david_first=4, n_david=2, history=8, seed=1,
                 switch_tol=0.0, refine_steps=0, kbatch=0, graph=False, autotune=False):
        """kbatch: k-points per batched launch (0 = all); graph: replay each SCF
        step from a hipGraph; autotune: time the launch shapes at init;
        switch_tol / refine_steps: the FP32 -> FP64 schedule of DEREK_MIXED."""
        p = ScfParams(nelec=nelec, kT=kT, mix_beta=beta, mix_q0=q0, n_david_first=n_david_first,
                      n_david=n_david, history=history, reserved=0, seed=seed, fp32_switch_tol=switch_tol,
                      refine_steps=refine_steps, kbatch=kbatch, use_graph=1 if graph else 0,
                      autotune=1 if autotune else 0)
        check(self.lib.derek_hip_scf_init(self.h, C.byref(p)))

    def tuning(self):
        """{bands_per_wg, items_per_launch, kbatch, graph, fused, fused_density, z_sticks, gemm_rounds}"""
        out = (C.c_int32 * 8)()
        check(self.lib.derek_hip_get_tuning(self.h, out))
        return {"bands_per_wg": out[0], "items_per_launch": out[1], "kbatch": out[2], "graph": bool(out[3]),
                "fused": bool(out[4] & 1), "fused_density": bool(out[4] & 2), "z_sticks": out[5],
                "gemm_rounds": out[6], "gemm_tile": out[7]}

    def set_tiles(self, z_sticks=0, gemm_rounds=2, gemm_tile=64):
        """Sticks per z-pass workgroup (16, 32; 0: the precision's default),
        the rounds of resident workgroups the overlaps' row chunks fill (1, 2,
        4) and the C tile edge of the BF16 overlaps (64, 128)."""
        check(self.lib.derek_hip_set_tiles(self.h, int(z_sticks), int(gemm_rounds), int(gemm_tile)))

    def set_launch(self, bands_per_wg=2, items_per_launch=256, fused=True, fused_density=None):
        """Launch shapes; fused / fused_density: the fused y/x passes for H|psi>
        and for the density where the grid allows them (fused_density follows
        fused unless given)."""
        fd = fused if fused_density is None else fused_density
        check(self.lib.derek_hip_set_launch(self.h, bands_per_wg, items_per_launch,
                                            (1 if fused else 0) | (2 if fd else 0)))

    def set_gemm(self, bf16_terms=2):
        """Matrix cores of the complex64 overlaps and rotations: 0 FP32 MFMA;
        1..3 BF16 MFMA with each FP32 value split into that many bf16 terms
        (a DEREK_MIXED context starts at 2)."""
        check(self.lib.derek_hip_set_gemm(self.h, int(bf16_terms)))

    def scf_step(self):
        r = ScfReport()
        check(self.lib.derek_hip_scf_step(self.h, C.byref(r)))
        return r.as_dict()

    def scf_polish(self, max_iter=100, res_tol=1e-9):
        r = C.c_double()
        check(self.lib.derek_hip_scf_polish(self.h, C.c_int32(max_iter), C.c_double(res_tol), C.byref(r)))
        return r.value

    def nscf(self, max_iter=200, res_tol=1e-9):
        """Eigenpairs at the potential given with set_vloc; returns the largest residual."""
        r = C.c_double()
        check(self.lib.derek_hip_nscf(self.h, C.c_int32(max_iter), C.c_double(res_tol), C.byref(r)))
        return r.value

    def eigenvalues(self):
        e = np.zeros((self.nk, self.nbands))
        check(self.lib.derek_hip_get_eigenvalues(self.h, ptr(e)))
        return e

    def residuals(self):
        """Residual norms per k-point and band after the last Davidson step."""
        e = np.zeros((self.nk, self.nbands))
        check(self.lib.derek_hip_get_residuals(self.h, ptr(e)))
        return e

    def occupations(self):
        e = np.zeros((self.nk, self.nbands))
        check(self.lib.derek_hip_get_occupations(self.h, ptr(e)))
        return e

    def get_density(self):
        r = np.zeros(self.N)
        check(self.lib.derek_hip_get_density(self.h, ptr(r)))
        return r

    def get_potential(self):
        out = np.zeros(self.N)
        check(self.lib.derek_hip_get_potential(self.h, ptr(out)))
        return out

    def get_psi(self, k):
        out = np.zeros((self.nbands, self.npw[k]), dtype=np.complex128)
        check(self.lib.derek_hip_get_psi(self.h, C.c_int32(k), out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def put_psi(self, k, psi):
        p, pp = _c128(psi)
        check(self.lib.derek_hip_put_psi(self.h, C.c_int32(k), pp))

    def apply_h_resident(self, k, nb):
        check(self.lib.derek_hip_apply_h_resident(self.h, C.c_int32(k), C.c_int32(nb)))

    def apply_h_batch(self):
        """H on every resident band of every k-point, all k-points in one batch."""
        check(self.lib.derek_hip_apply_h_resident(self.h, C.c_int32(-1), C.c_int32(self.nbands)))

    def profile(self, on=True):
        check(self.lib.derek_hip_profile_enable(self.h, C.c_int32(1 if on else 0)))

    def profile_reset(self):
        check(self.lib.derek_hip_profile_reset(self.h))

    def profile_read(self):
        """{class name: (launches, total_ms, algorithmic bytes, flops)}"""
        out = {}
        for i, name in enumerate(KERNEL_CLASSES):
            n = C.c_int64()
            ms, by, fl = C.c_double(), C.c_double(), C.c_double()
            check(self.lib.derek_hip_profile_read(self.h, C.c_int32(i), C.byref(n), C.byref(ms), C.byref(by),
                                                  C.byref(fl)))
            out[name] = (n.value, ms.value, by.value, fl.value)
        return out

    def sync(self):
        check(self.lib.derek_hip_sync(self.h))

    def stream(self):
        s = C.c_void_p()
        check(self.lib.derek_hip_stream(self.h, C.byref(s)))
        return s.value
