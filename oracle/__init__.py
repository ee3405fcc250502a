"""FP64 oracle of the libconv separable convolutions -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker (or as the CPU baseline that is
timed beside the GPU path).  The product library under boast-lig_amd/ never
loads it.

`liblc_oracle.so` (built by `make -C oracle`) is the C restatement of the
BOAST generator's semantics (BigDFT/GenericConvolution2.rb); the `ref`
submodule wraps the reference's own FP64 Fortran oracles compiled into
oracle/_ref/ by build_ref.py (present only where the reference checkout was
available at build time).

Arrays are numpy float64, Fortran order (dim 0 fastest), flattened the way
the libconv entry points see them.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MD, IMD, D2, D1 = 0, 1, 2, 3
SYM8, SYMICOMB8 = 0, 1
PERIODIC, GROW, SHRINK, FREE = 0, 1, -1, 2

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, 'liblc_oracle.so')
        if not os.path.exists(path):
            raise RuntimeError('oracle not built: run make -C oracle')
        L = ctypes.CDLL(path)
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        i, d = ctypes.c_int, ctypes.c_double
        L.lco_s0s0_1d.argtypes = [i, i, i, ip, i, ip, ip, i, dp, dp, d, d, d, dp]
        L.lco_s0s1_1d.argtypes = [i, i, i, ip, i, ip, ip, i, dp, dp, d, d]
        L.lco_s1s0_1d.argtypes = [i, i, i, ip, i, ip, ip, i, dp, dp, d, d]
        L.lco_s0s0_md.argtypes = [i, i, ip, ip, ip, ip, i, dp, dp, d, d, d]
        L.lco_s0s0_1ds.argtypes = [i, i, ip, ip, ip, ip, i, dp, d, d, dp, dp, dp]
        L.lco_s0s1_md.argtypes = [i, i, ip, ip, ip, ip, i, dp, dp, d, d]
        L.lco_s1s0_md.argtypes = [i, i, ip, ip, ip, ip, i, dp, dp, d, d]
        L.lco_filter_taps.argtypes = [i, ip, dp, i]
        L.lco_poisson_taps.argtypes = [i, dp]
        L.lco_poisson_1d.argtypes = [i, i, ip, i, d, d, dp, dp]
        L.lco_div_u_i.argtypes = [i, ip, dp, dp, i, dp, dp]
        L.lco_nabla_u.argtypes = [i, ip, dp, dp, i, dp]
        L.lco_nabla_u_and_square.argtypes = [i, ip, dp, dp, dp, i, dp]
        L.lco_nabla_u_square.argtypes = [i, ip, dp, dp, i, dp]
        L.lco_nabla_u_epsilon.argtypes = [i, ip, dp, dp, i, dp, dp]
        L.lco_update_rhopol.argtypes = [i, ip, dp, i, dp, d, dp, dp, dp]
        for f in ('lco_s0s0_1d', 'lco_s0s1_1d', 'lco_s1s0_1d', 'lco_s0s0_md', 'lco_s0s0_1ds',
                  'lco_s0s1_md', 'lco_s1s0_md', 'lco_filter_taps', 'lco_num_threads', 'lco_poisson_taps',
                  'lco_poisson_1d', 'lco_div_u_i', 'lco_nabla_u', 'lco_nabla_u_and_square',
                  'lco_nabla_u_square', 'lco_nabla_u_epsilon', 'lco_update_rhopol'):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


def _iarr(v):
    a = np.ascontiguousarray(np.asarray(v, dtype=np.int32))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _darr(a):
    assert a.dtype == np.float64 and (a.flags['C_CONTIGUOUS'] or a.flags['F_CONTIGUOUS'])
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc, what):
    if rc != 0:
        raise ValueError('oracle %s rejected its arguments' % what)


def num_threads():
    return lib().lco_num_threads()


_cpu = None


def cpu():
    """liblc_cpu.so: the CPU baseline (generated sweep structure, C++/OpenMP)."""
    global _cpu
    if _cpu is None:
        path = os.path.join(HERE, 'liblc_cpu.so')
        if not os.path.exists(path):
            raise RuntimeError('CPU baseline not built: run make -C oracle')
        L = ctypes.CDLL(path)
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        L.lcc_magicfilter3d.argtypes = [ip, ip, ip, ip, dp, dp, dp, dp]
        L.lcc_kinetic3d.argtypes = [ip, ip, ip, ip, dp, dp, dp, dp]
        L.lcc_magicfilter3d.restype = L.lcc_kinetic3d.restype = L.lcc_num_threads.restype = ctypes.c_int
        _cpu = L
    return _cpu


def cpu_threads():
    return cpu().lcc_num_threads()


def cpu_magicfilter3d(n, bc, nx, ny, x, y, w1, w2):
    """3D MagicFilter as three sweeps (the CPU baseline); w1 >= ny0 nx1 nx2,
    w2 >= ny0 ny1 nx2 values, owned by the caller."""
    (n_, pn), (b_, pb), (nx_, pnx), (ny_, pny) = _iarr(n), _iarr(bc), _iarr(nx), _iarr(ny)
    _check(cpu().lcc_magicfilter3d(pn, pb, pnx, pny, _darr(x), _darr(y), _darr(w1), _darr(w2)),
           'cpu_magicfilter3d')


def cpu_kinetic3d(n, bc, nx, ny, a, x, y):
    """3D Kinetic sum form with energies (the CPU baseline); returns the dots."""
    (n_, pn), (b_, pb), (nx_, pnx), (ny_, pny) = _iarr(n), _iarr(bc), _iarr(nx), _iarr(ny)
    av = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    dot = np.zeros(3)
    _check(cpu().lcc_kinetic3d(pn, pb, pnx, pny, _darr(av), _darr(x), _darr(y), _darr(dot)),
           'cpu_kinetic3d')
    return dot


def taps(fil):
    out = np.zeros(64)
    lo = ctypes.c_int()
    nt = lib().lco_filter_taps(fil, ctypes.byref(lo), _darr(out), 64)
    return lo.value, out[:nt].copy()


def s0s0_1d(fil, idim, n, bc, nx, ny, narr, x, y, a=1.0, a_x=0.0, a_y=0.0, want_dot=False):
    """Convolution sweep along idim (libconv s0s0_1d); y updated in place."""
    (n_, pn), (nx_, pnx), (ny_, pny) = _iarr(n), _iarr(nx), _iarr(ny)
    dot = ctypes.c_double(0.0)
    rc = lib().lco_s0s0_1d(fil, len(n), idim, pn, bc, pnx, pny, narr, _darr(x), _darr(y),
                           a, a_x, a_y, ctypes.byref(dot) if want_dot else None)
    _check(rc, 's0s0_1d')
    return dot.value if want_dot else None


def s0s1_1d(wav, idim, n, bc, nx, ny, narr, x, y, a=1.0, a_y=0.0):
    (n_, pn), (nx_, pnx), (ny_, pny) = _iarr(n), _iarr(nx), _iarr(ny)
    _check(lib().lco_s0s1_1d(wav, len(n), idim, pn, bc, pnx, pny, narr, _darr(x), _darr(y), a, a_y),
           's0s1_1d')


def s1s0_1d(wav, idim, n, bc, nx, ny, narr, x, y, a=1.0, a_y=0.0):
    (n_, pn), (nx_, pnx), (ny_, pny) = _iarr(n), _iarr(nx), _iarr(ny)
    _check(lib().lco_s1s0_1d(wav, len(n), idim, pn, bc, pnx, pny, narr, _darr(x), _darr(y), a, a_y),
           's1s0_1d')


def s0s0_md(fil, n, bc, nx, ny, narr, x, y, a=1.0, a_x=0.0, a_y=0.0):
    (n_, pn), (b_, pb), (nx_, pnx), (ny_, pny) = _iarr(n), _iarr(bc), _iarr(nx), _iarr(ny)
    _check(lib().lco_s0s0_md(fil, len(n), pn, pb, pnx, pny, narr, _darr(x), _darr(y), a, a_x, a_y),
           's0s0_md')


def s0s0_1ds(fil, n, bc, nx, ny, narr, a, a_x, a_y, x, y):
    """Kinetic-style sum over dims; returns the per-dim dots."""
    (n_, pn), (b_, pb), (nx_, pnx), (ny_, pny) = _iarr(n), _iarr(bc), _iarr(nx), _iarr(ny)
    av = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    dot = np.zeros(len(n))
    _check(lib().lco_s0s0_1ds(fil, len(n), pn, pb, pnx, pny, narr, _darr(av), a_x, a_y,
                              _darr(x), _darr(y), _darr(dot)), 's0s0_1ds')
    return dot


def s0s1_md(wav, n, bc, nx, ny, narr, x, y, a=1.0, a_y=0.0):
    (n_, pn), (b_, pb), (nx_, pnx), (ny_, pny) = _iarr(n), _iarr(bc), _iarr(nx), _iarr(ny)
    _check(lib().lco_s0s1_md(wav, len(n), pn, pb, pnx, pny, narr, _darr(x), _darr(y), a, a_y),
           's0s1_md')


def s1s0_md(wav, n, bc, nx, ny, narr, x, y, a=1.0, a_y=0.0):
    (n_, pn), (b_, pb), (nx_, pnx), (ny_, pny) = _iarr(n), _iarr(bc), _iarr(nx), _iarr(ny)
    _check(lib().lco_s1s0_md(wav, len(n), pn, pb, pnx, pny, narr, _darr(x), _darr(y), a, a_y),
           's1s0_md')


def local_hamiltonian(n, a, psig, pot, rho=None):
    """The local-Hamiltonian chain of BigDFT/libconv.f90:238-295 composed from
    the functions above: IDWT grow, MF grow, rho += psir^2 and psir *= pot,
    IMF shrink, plus the free-BC ('F' nogrow, bc 2) Kinetic sum with a_y = 1
    and energies, DWT shrink.  Returns (hpsig, rho, ekin[3]); rho is updated
    in place when given."""
    n = list(n)
    nf7 = [v + 7 for v in n]
    f = [2 * v for v in nf7]
    r = [v + 15 for v in f]
    fscf = np.zeros(int(np.prod(f)))
    s1s0_md(SYM8, n, [GROW] * 3, n, nf7, 1, psig, fscf)
    psir = np.zeros(int(np.prod(r)))
    s0s0_md(MD, f, [GROW] * 3, f, r, 1, fscf, psir)
    if rho is not None:
        rho += psir * psir
    psir *= pot
    hpsi = np.zeros(fscf.size)
    s0s0_md(IMD, f, [SHRINK] * 3, r, f, 1, psir, hpsi)
    ekin = s0s0_1ds(D2, f, [FREE] * 3, f, f, 1, a, 0.0, 1.0, fscf, hpsi)
    hpsig = np.zeros(int(np.prod([2 * v for v in n])))
    s0s1_md(SYM8, n, [SHRINK] * 3, nf7, n, 1, hpsi, hpsig)
    return hpsig, rho, ekin


# ---------------------------------------------------------------------------
# Poisson nabla operators (BigDFT/Poisson2.rb); bc 0 periodic, -2 non-periodic
NPERIODIC = -2
POISSON_ORDERS = (2, 4, 6, 8, 16)


def poisson_taps(nord):
    """(nord+1, nord+1) table c[k + m, j + m] of row k, tap j."""
    out = np.zeros((nord + 1) ** 2)
    _check(lib().lco_poisson_taps(nord, _darr(out)), 'poisson_taps')
    return out.reshape(nord + 1, nord + 1)


def poisson_1d(nord, idim, n, bc, u, du, a=1.0, a_y=0.0):
    n_, pn = _iarr(n)
    _check(lib().lco_poisson_1d(nord, idim, pn, bc, a, a_y, _darr(u), _darr(du)), 'poisson_1d')


def div_u_i(geocode, n, u, du, nord, h, cc=None):
    n_, pn = _iarr(n)
    hv = np.ascontiguousarray(np.asarray(h, dtype=np.float64))
    _check(lib().lco_div_u_i(geocode, pn, _darr(u), _darr(du), nord, _darr(hv),
                             _darr(cc) if cc is not None else None), 'div_u_i')


def nabla_u(geocode, n, u, du, nord, h):
    n_, pn = _iarr(n)
    hv = np.ascontiguousarray(np.asarray(h, dtype=np.float64))
    _check(lib().lco_nabla_u(geocode, pn, _darr(u), _darr(du), nord, _darr(hv)), 'nabla_u')


def nabla_u_and_square(geocode, n, u, du, du2, nord, h):
    n_, pn = _iarr(n)
    hv = np.ascontiguousarray(np.asarray(h, dtype=np.float64))
    _check(lib().lco_nabla_u_and_square(geocode, pn, _darr(u), _darr(du), _darr(du2), nord, _darr(hv)),
           'nabla_u_and_square')


def nabla_u_square(geocode, n, u, ddu, nord, h):
    n_, pn = _iarr(n)
    hv = np.ascontiguousarray(np.asarray(h, dtype=np.float64))
    _check(lib().lco_nabla_u_square(geocode, pn, _darr(u), _darr(ddu), nord, _darr(hv)), 'nabla_u_square')


def nabla_u_epsilon(geocode, n, u, du, nord, h, eps):
    n_, pn = _iarr(n)
    hv = np.ascontiguousarray(np.asarray(h, dtype=np.float64))
    _check(lib().lco_nabla_u_epsilon(geocode, pn, _darr(u), _darr(du), nord, _darr(hv), _darr(eps)),
           'nabla_u_epsilon')


def update_rhopol(geocode, n, u, nord, h, eta, dlogeps, rhopol):
    """rhopol updated in place; returns rhores2."""
    n_, pn = _iarr(n)
    hv = np.ascontiguousarray(np.asarray(h, dtype=np.float64))
    r = np.zeros(1)
    _check(lib().lco_update_rhopol(geocode, pn, _darr(u), nord, _darr(hv), float(eta), _darr(dlogeps),
                                   _darr(rhopol), _darr(r)), 'update_rhopol')
    return float(r[0])


# ---------------------------------------------------------------------------
# size rules (BigDFT/LibconvBase.rb:98-135 outdim/get_ld; GenericConvolution2
# .rb:1584-1618): extent along a dim of the input / output of an operator.
FILTER_LEN = {MD: 16, IMD: 16, D2: 29, D1: 29}
WAVELET_PAIRS = {SYM8: 8, SYMICOMB8: 16}


def conv_extents(fil, bc, n):
    L = FILTER_LEN[fil]
    if bc == GROW:
        return n, n + L - 1
    if bc == SHRINK:
        return n + L - 1, n
    return n, n


def wavelet_extents(wav, bc, n):
    W = WAVELET_PAIRS[wav]
    if bc == GROW:
        return n, n + W - 1
    if bc == SHRINK:
        return n + W - 1, n
    return n, n
