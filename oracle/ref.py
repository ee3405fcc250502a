"""ctypes wrappers of the reference's own FP64 oracles in oracle/_ref/.

TEST INFRASTRUCTURE ONLY.  The shared objects are compiled from the
read-only reference checkout by oracle/build_ref.py; where they are absent
(no checkout at build time) `available()` is False and callers skip.
Fortran calling convention: every argument by reference.
"""
import ctypes
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), '_ref')
_libs = {}


def _load(name):
    if name not in _libs:
        p = os.path.join(HERE, name)
        _libs[name] = ctypes.CDLL(p) if os.path.exists(p) else None
    return _libs[name]


def available(name='libref_mf.so'):
    """Whether the object was built here.  Checks the file only: deciding a
    skip at collection must not map reference code into the process."""
    return os.path.exists(os.path.join(HERE, name))


def _i(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def _d(v):
    return ctypes.byref(ctypes.c_double(float(v)))


def _p(a):
    assert a.dtype == np.float64 and a.flags['F_CONTIGUOUS']
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def magicfilter(kind, n1, ndat, x):
    """magicfilter_{per,free}[_t]_ref (BigDFT/magicfilter_refs.f90): x is
    (n_in, ndat) Fortran order, returns y (ndat, n_out) -- the transposed
    output of the reference routine."""
    L = _load('libref_mf.so')
    n_out = {'per': n1, 'per_t': n1, 'free': n1 + 15, 'free_t': n1}[kind]
    y = np.zeros((ndat, n_out), order='F')
    getattr(L, 'magicfilter_%s_ref_' % kind)(_i(n1), _i(ndat), _p(x), _p(y))
    return y


def kinetic_per(n, hgrid, x, c):
    """kinetic_per_ref (BigDFT/Kinetic.rb:981-1061): y = c x - 1/2 Lap x."""
    L = _load('libref_kin.so')
    h = np.asfortranarray(np.asarray(hgrid, dtype=np.float64))
    y = np.zeros(n, order='F')
    L.kinetic_per_ref_(_i(n[0]), _i(n[1]), _i(n[2]), _p(h), _p(x), _p(y), _d(c))
    return y


def kinetic_ekin(n, hgrid, x, y):
    """kinetic_per_ref_optim_ekin (BigDFT/Kinetic.rb:21-112): y += -1/2 Lap x
    (periodic) and returns the per-direction <x, -1/2 d2/dxi2 x>."""
    L = _load('libref_kin.so')
    h = np.asfortranarray(np.asarray(hgrid, dtype=np.float64))
    k = np.zeros(6, order='F')
    L.kinetic_per_ref_optim_ekin_(_i(n[0] - 1), _i(n[1] - 1), _i(n[2] - 1), _p(h), _p(x), _p(y), _p(k))
    return k[:3].copy()


def idwt_1d(d, idim, n, bc, nx, ny, narr, x, y, a, a_y):
    """BOAST-generated d_s1s0_1d_sym8 dispatcher (BigDFT/Wavelet1d-nb.ipynb
    cell 4): y := a * IDWT_idim(x) + a_y * y."""
    L = _load('libref_idwt.so')
    na = np.asfortranarray(np.asarray(n, dtype=np.int32))
    nxa = np.asfortranarray(np.asarray(nx, dtype=np.int32))
    nya = np.asfortranarray(np.asarray(ny, dtype=np.int32))
    ip = ctypes.POINTER(ctypes.c_int)
    L.d_s1s0_1d_sym8_(_i(d), _i(idim), na.ctypes.data_as(ip), _i(bc), nxa.ctypes.data_as(ip),
                      nya.ctypes.data_as(ip), _i(narr), _p(x), _p(y), _d(a), _d(a_y))


def bigdft_wavelet(kind, n, ndat, x):
    """BigDFT's own periodic synthesis/analysis (BigDFT/old/Synthesis.rb:100-146,
    BigDFT/old/AnaRotPer-2.rb:86-221) on n pairs: x (2n, ndat) -> y (ndat, 2n).
    analysis_per_ref counts pairs from zero (its argument is n-1)."""
    L = _load('libref_wav.so')
    y = np.zeros((ndat, 2 * n), order='F')
    arg = n if kind == 'synthesis_per' else n - 1
    getattr(L, '%s_ref_' % kind)(_i(arg), _i(ndat), _p(x), _p(y))
    return y


def poisson_filter(nord):
    """generate_filter (BigDFT/Poisson2.rb:20-166): the (nord+1)^2 table
    c1D(-m:m, -m:m), returned as [k + m, j + m] (row k, tap j)."""
    L = _load('libref_poisson.so')
    c = np.zeros((nord + 1, nord + 1), order='F')
    L.generate_filter_(_i(nord), _p(c))
    return np.ascontiguousarray(c.T)  # Fortran (j, k) -> [k][j]


def _c1d(table):
    return np.asfortranarray(np.asarray(table, dtype=np.float64).T)  # [k][j] -> Fortran (j, k)


def div_u_i(geocode, n, u, nord, h, table, with_cc=False):
    """div_u_i_ref / div_u_i_opt (BigDFT/Poisson2.rb:183-335, :901-1247) with
    the given filter table; u is (n0, n1, n2, 3) Fortran order."""
    L = _load('libref_poisson.so')
    hv = np.asfortranarray(np.asarray(h, dtype=np.float64))
    du = np.zeros(tuple(n), order='F')
    if not with_cc:
        L.div_u_i_ref_(_i(geocode), _i(n[0]), _i(n[1]), _i(n[2]), _p(u), _p(du), _i(nord), _p(hv),
                       _p(_c1d(table)))
        return du
    cc = np.zeros(tuple(n), order='F')
    L.div_u_i_opt_(_i(geocode), _i(n[0]), _i(n[1]), _i(n[2]), _p(u), _p(du), _i(nord), _p(hv),
                   _p(_c1d(table)), _p(cc))
    return du, cc


def nabla_u(geocode, n, u, nord, h, table):
    """nabla_u_ref (BigDFT/Poisson2.rb:722-874): du (n0, n1, n2, 3)."""
    L = _load('libref_poisson.so')
    hv = np.asfortranarray(np.asarray(h, dtype=np.float64))
    du = np.zeros(tuple(n) + (3,), order='F')
    L.nabla_u_ref_(_i(geocode), _i(n[0]), _i(n[1]), _i(n[2]), _p(u), _p(du), _i(nord), _p(hv),
                   _p(_c1d(table)))
    return du


def nabla_u_square(geocode, n, u, nord, h, table):
    """nabla_u_square_ref (BigDFT/Poisson2.rb:1690-1884): du2 (n0, n1, n2)."""
    L = _load('libref_poisson.so')
    hv = np.asfortranarray(np.asarray(h, dtype=np.float64))
    du2 = np.zeros(tuple(n), order='F')
    L.nabla_u_square_ref_(_i(geocode), _i(n[0]), _i(n[1]), _i(n[2]), _p(u), _p(du2), _i(nord), _p(hv),
                          _p(_c1d(table)))
    return du2


def nabla_u_epsilon(geocode, n, u, nord, h, eps, table):
    """nabla_u_epsilon_ref (BigDFT/Poisson2.rb:1488-1680): du (n0, n1, n2, 3)."""
    L = _load('libref_poisson.so')
    hv = np.asfortranarray(np.asarray(h, dtype=np.float64))
    du = np.zeros(tuple(n) + (3,), order='F')
    L.nabla_u_epsilon_ref_(_i(geocode), _i(n[0]), _i(n[1]), _i(n[2]), _p(u), _p(du), _i(nord), _p(hv),
                           _p(eps), _p(_c1d(table)))
    return du


def update_rhopol(geocode, n, u, nord, h, eta, dlogeps, rhopol, table):
    """update_rhopol_ref (BigDFT/Poisson2.rb:520-699): the updated rhopol
    (n0, n1, n2) and rhores2; dlogeps is (3, n0, n1, n2) Fortran order."""
    L = _load('libref_poisson.so')
    hv = np.asfortranarray(np.asarray(h, dtype=np.float64))
    rp = np.array(rhopol, dtype=np.float64, order='F', copy=True)
    r2 = np.zeros(1)
    L.update_rhopol_ref_(_i(geocode), _i(n[0]), _i(n[1]), _i(n[2]), _p(u), _i(nord), _p(hv), _d(eta),
                         _p(dlogeps), _p(rp), _p(r2), _p(_c1d(table)))
    return rp, float(r2[0])
