"""Host-side mirror of the reference's kernel interface over libconv_amd.so.

The reference drives each generated libconv procedure through BOAST's
`CKernel#run(*args)`, passing the procedure's arguments in order and reading
back `{:duration, :reference_return => {:dot_in}}` (e.g.
BigDFT/MagicFilter1d.rb:55, BigDFT/Kinetic1d-nb.ipynb cells 15-18).  This
module gives the same shape in Python:

    k = libconv.MagicFilter1d()              # d_s0s0_1d_sym8_md
    stats = k.run(3, 0, n, libconv.BC.PERIODIC, n, n, 1, x, y, 1.0, 0.0, 0.0)
    stats['duration'], stats['reference_return']

Arrays may be numpy arrays (host memory; the library stages them) or torch
tensors on the GPU (device memory; run in place).  Every call goes through
the C ABI of include/libconv.h; there is no other compute path, and a
missing library raises at import.
"""
import ctypes
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(os.path.dirname(HERE), 'libconv_amd.so')


class BC:
    """Boundary conditions (BigDFT/GenericConvolution2.rb:684-740)."""
    PERIODIC = 0
    GROW = 1
    SHRINK = -1


class LibconvError(RuntimeError):
    pass


_lib = None
_I = ctypes.c_int
_IP = ctypes.POINTER(ctypes.c_int)
_VP = ctypes.c_void_p
_D = ctypes.c_double
_F = ctypes.c_float
_DP = ctypes.POINTER(ctypes.c_double)
_FP = ctypes.POINTER(ctypes.c_float)


def lib():
    """The loaded libconv_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise LibconvError('libconv_amd.so not built (run make -C boast-lig_amd)')
        try:  # load torch's HIP runtime first so the process holds a single one
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIBPATH)
        for p, R in (('d', _D), ('s', _F)):
            for f in ('md', 'imd', 'd2', 'd1'):
                fn = getattr(L, '%s_s0s0_1d_sym8_%s' % (p, f))
                fn.argtypes = [_I, _I, _IP, _I, _IP, _IP, _I, _VP, _VP, R, R, R]
                fn.restype = None
            for w in ('s0s1', 's1s0'):
                for fam in ('sym8', 'symicomb8'):
                    fn = getattr(L, '%s_%s_1d_%s' % (p, w, fam))
                    fn.argtypes = [_I, _I, _IP, _I, _IP, _IP, _I, _VP, _VP, R, R]
                    fn.restype = None
            for f in ('md', 'imd'):
                fn = getattr(L, '%s_sym8_%s_ld_narr' % (p, f))
                fn.argtypes = [_I, _IP, _IP, _IP, _IP, _I, _VP, _VP, _VP, _VP]
                fn.restype = None
            fn = getattr(L, '%s_sym8_d2' % p)
            fn.argtypes = [_I, _IP, _IP, _VP, _VP, _VP, _VP, _VP, R, R]
            fn.restype = None
            for w in ('dwt', 'idwt'):
                fn = getattr(L, '%s_%s_sym8' % (p, w))
                fn.argtypes = [_I, _IP, _IP, _VP, _VP, _VP, _VP]
                fn.restype = None
        L.d_s0s0_1d_sym8_d2_dotin.argtypes = [_I, _I, _IP, _I, _IP, _IP, _I, _VP, _VP, _D, _D, _D, _DP]
        L.s_s0s0_1d_sym8_d2_dotin.argtypes = [_I, _I, _IP, _I, _IP, _IP, _I, _VP, _VP, _F, _F, _F, _FP]
        L.s_s0s0_1d_sym8_d2_dotin64.argtypes = [_I, _I, _IP, _I, _IP, _IP, _I, _VP, _VP, _F, _F, _F, _DP]
        L.d_sym8_d2_dotin.argtypes = [_I, _IP, _IP, _VP, _VP, _VP, _VP, _VP, _D, _D, _DP]
        L.s_sym8_d2_dotin64.argtypes = [_I, _IP, _IP, _VP, _VP, _VP, _VP, _VP, _F, _F, _DP]
        for n in ('d_s0s0_1d_sym8_d2_dotin', 's_s0s0_1d_sym8_d2_dotin', 's_s0s0_1d_sym8_d2_dotin64',
                  'd_sym8_d2_dotin', 's_sym8_d2_dotin64'):
            getattr(L, n).restype = None
        L.d_s0s0_1d_sym8_md_cost.argtypes = [_I, _I, _IP, _I, _IP, _IP, _I, _VP, _VP, _D, _D, _D, _IP]
        L.d_s0s1_1d_sym8_cost.argtypes = [_I, _I, _IP, _I, _IP, _IP, _I, _VP, _VP, _D, _D, _IP]
        L.d_s1s0_1d_sym8_cost.argtypes = [_I, _I, _IP, _I, _IP, _IP, _I, _VP, _VP, _D, _D, _IP]
        L.lc_last_error.restype = _I
        L.lc_last_error_message.restype = ctypes.c_char_p
        L.lc_version.restype = ctypes.c_char_p
        L.lc_extent.argtypes = [_I, _I, _I, _I]
        L.lc_extent.restype = _I
        L.lc_conv_1d.argtypes = [_I, _I, _I, _I, _IP, _I, _IP, _IP, _I, _VP, _VP, _D, _D, _D, _VP, _VP]
        L.lc_wavelet_1d.argtypes = [_I, _I, _I, _I, _I, _IP, _I, _IP, _IP, _I, _VP, _VP, _D, _D, _VP]
        L.lc_conv_md.argtypes = [_I, _I, _I, _IP, _IP, _IP, _IP, _I, _VP, _VP, _D, _D, _D, _VP]
        L.lc_conv_1ds.argtypes = [_I, _I, _I, _IP, _IP, _IP, _IP, _I, _DP, _D, _D, _VP, _VP, _VP, _VP]
        L.lc_wavelet_md.argtypes = [_I, _I, _I, _I, _IP, _IP, _IP, _IP, _I, _VP, _VP, _D, _D, _VP]
        L.lc_conv_1ds_noalloc.argtypes = L.lc_conv_1ds.argtypes
        for n in ('lc_conv_1d', 'lc_wavelet_1d', 'lc_conv_md', 'lc_conv_1ds', 'lc_wavelet_md',
                  'lc_conv_1ds_noalloc'):
            getattr(L, n).restype = _I
        _lib = L
    return _lib


def version():
    return lib().lc_version().decode()


def check():
    """Raise LibconvError if the last call of this thread recorded an error."""
    code = lib().lc_last_error()
    if code:
        raise LibconvError('libconv error %d: %s' % (code, lib().lc_last_error_message().decode()))


# ------------------------------------------------------------------ arguments
def _ints(v):
    if isinstance(v, (int, np.integer)):
        v = [int(v)]
    a = np.ascontiguousarray(np.asarray(v, dtype=np.int32))
    return a


def _ptr(a):
    """Address of a numpy array or a torch tensor (device or host)."""
    if a is None:
        return None
    if hasattr(a, 'data_ptr'):
        return ctypes.c_void_p(a.data_ptr())
    if isinstance(a, np.ndarray):
        return ctypes.c_void_p(a.ctypes.data)
    raise TypeError('array argument must be a numpy array or a torch tensor')


def _is_f32(a):
    if hasattr(a, 'dtype') and str(a.dtype) in ('float32', 'torch.float32'):
        return True
    return False


class LibconvKernel:
    """One libconv procedure with the reference's `run` calling shape."""

    def __init__(self, name, fp32=False):
        self.name = ('s_' if fp32 else 'd_') + name
        self.fp32 = fp32
        self.fn = getattr(lib(), self.name)

    def _marshal(self, args):
        keep, out = [], []
        for a in args:
            if isinstance(a, (list, tuple)) and not all(isinstance(v, (int, np.integer)) for v in a):
                arr = np.ascontiguousarray(np.asarray(a, dtype=np.float32 if self.fp32 else np.float64))
                keep.append(arr)
                out.append(ctypes.c_void_p(arr.ctypes.data))
            elif isinstance(a, (list, tuple)) or (isinstance(a, np.ndarray) and a.dtype.kind == 'i'):
                arr = _ints(a)
                keep.append(arr)
                out.append(arr.ctypes.data_as(_IP))
            elif isinstance(a, np.ndarray) or hasattr(a, 'data_ptr'):
                keep.append(a)
                out.append(_ptr(a))
            else:
                out.append(a)
        return keep, out

    def run(self, *args, **kw):
        keep, cargs = self._marshal(args)
        dot = None
        if kw.get('dot'):
            dot = (_F if (self.fp32 and not kw.get('dot64')) else _D)(0.0)
            cargs.append(ctypes.byref(dot))
        sync = kw.get('sync')
        t0 = time.perf_counter()
        self.fn(*cargs)
        if sync:
            sync()
        dt = time.perf_counter() - t0
        check()
        stats = {'duration': dt}
        if dot is not None:
            stats['reference_return'] = {'dot_in': dot.value}
        return stats


# ------------------------------------------------------------------ front-ends
# (BigDFT/LibconvBase.rb:226-262 generate_*, MagicFilter1d.rb, Kinetic1d.rb,
#  Wavelet1d2.rb, MagicFilter.rb, Kinetic.rb, Wavelet2.rb)
class _DotKernel(LibconvKernel):
    def run(self, *args, **kw):
        # the generated procedure takes dot_in as its last argument; the value
        # passed in is ignored and the result comes back in reference_return
        args = args[:-1] if len(args) == self.nargs else args
        kw['dot'] = True
        return LibconvKernel.run(self, *args, **kw)


def MagicFilter1d(fp32=False):
    """d_s0s0_1d_sym8_md: (d, idim, n, bc, nx, ny, narr, x, y, a, a_x, a_y)."""
    return LibconvKernel('s0s0_1d_sym8_md', fp32)


def InverseMagicFilter1d(fp32=False):
    return LibconvKernel('s0s0_1d_sym8_imd', fp32)


def Kinetic1d(fp32=False, dot64=False):
    """d_s0s0_1d_sym8_d2_dotin: (..., a, a_x, a_y, dot_in) -> reference_return."""
    name = 's0s0_1d_sym8_d2_dotin64' if (fp32 and dot64) else 's0s0_1d_sym8_d2_dotin'
    k = _DotKernel(name, fp32)
    if fp32 and dot64:
        k.name = 's_' + name
        k.fn = getattr(lib(), k.name)
        k.fp32 = False  # the energy comes back in FP64
    k.nargs = 13
    return k


def Wavelet1d(direction, fp32=False, combined=False):
    """d_s0s1_1d_sym8 (direction 'decompose') / d_s1s0_1d_sym8 ('recompose');
    combined=True gives the symicomb8 operators rtos1 / s1tor."""
    base = 's0s1' if direction == 'decompose' else 's1s0'
    return LibconvKernel('%s_1d_%s' % (base, 'symicomb8' if combined else 'sym8'), fp32)


def MagicFilter(fp32=False, inverse=False):
    """d_sym8_md_ld_narr: (d, n, bc, nx, ny, narr, x, y, w1, w2)."""
    return LibconvKernel('sym8_%s_ld_narr' % ('imd' if inverse else 'md'), fp32)


def Kinetic(fp32=False, dot=False):
    """d_sym8_d2: (d, n, bc, x, y, w1, w2, a, a_x, a_y); with dot=True the
    per-direction energies come back in reference_return['dot_in']."""
    if not dot:
        return LibconvKernel('sym8_d2', fp32)
    return _KineticDot(fp32)


class _KineticDot(LibconvKernel):
    def __init__(self, fp32):
        self.name = 's_sym8_d2_dotin64' if fp32 else 'd_sym8_d2_dotin'
        self.fp32 = fp32
        self.fn = getattr(lib(), self.name)

    def run(self, d, n, bc, x, y, w1, w2, a, a_x, a_y, **kw):
        dots = np.zeros(int(d))
        av = np.ascontiguousarray(np.asarray(a, dtype=np.float32 if self.fp32 else np.float64))
        keep, cargs = self._marshal([d, n, bc, x, y, w1, w2])
        t0 = time.perf_counter()
        self.fn(*cargs, ctypes.c_void_p(av.ctypes.data), a_x, a_y, dots.ctypes.data_as(_DP))
        dt = time.perf_counter() - t0
        check()
        return {'duration': dt, 'reference_return': {'dot_in': dots}}


def Wavelet(direction, fp32=False):
    """d_dwt_sym8 / d_idwt_sym8: (d, n, bc, x, y, w1, w2)."""
    return LibconvKernel('%s_sym8' % ('dwt' if direction == 'decompose' else 'idwt'), fp32)


# ------------------------------------------------------------------ stream API
def _stream(stream):
    if stream is None:
        return None
    if hasattr(stream, 'cuda_stream'):
        return ctypes.c_void_p(stream.cuda_stream)
    return ctypes.c_void_p(int(stream))


def _prec(t):
    return 4 if _is_f32(t) else 8


def _count(ext, mul=1):
    c = 1
    for v in ext:
        c *= mul * int(v)
    return c


def _dev(t, name, prec, count, optional=False):
    """Checks a device-tensor argument before its address reaches the C ABI:
    a torch tensor of the plan's precision (float64 for dot outputs: prec
    'dot'), contiguous, with at least `count` elements (_resident checks
    that it is on the GPU).  A wrong
    dtype or a short or strided buffer would otherwise be read and written
    with the layout the extents describe."""
    if t is None:
        if optional:
            return
        raise LibconvError('%s: missing' % name)
    if not (hasattr(t, 'data_ptr') and hasattr(t, 'is_cuda') and hasattr(t, 'numel')):
        raise TypeError('%s must be a torch tensor on the GPU' % name)
    want = 'torch.float64' if prec in (8, 'dot') else 'torch.float32'
    if str(t.dtype) != want:
        raise LibconvError('%s: dtype %s, expected %s' % (name, t.dtype, want))
    if not t.is_contiguous():
        raise LibconvError('%s: not contiguous' % name)
    if t.numel() < count:
        raise LibconvError('%s: %d elements, the extents need %d' % (name, t.numel(), count))


def _resident(**ts):
    """Device residence, checked after every tensor's dtype, layout and size."""
    for name, t in ts.items():
        if t is not None and not t.is_cuda:
            raise LibconvError('%s: not in device memory' % name)


def _ext1(n, nx, ny, idim):
    """Input and output extents of a 1D sweep along idim (the other dims
    take nx for both, GenericConvolution2.rb:1701-1766)."""
    out = [int(v) for v in nx]
    out[idim] = int(ny[idim])
    return [int(v) for v in nx], out


def _rc(rc):
    if rc:
        check()
        raise LibconvError('libconv returned %d' % rc)


def conv_1d(fil, idim, n, bc, nx, ny, narr, x, y, a=1.0, a_x=0.0, a_y=0.0, dot=None, stream=None):
    """Device-pointer sweep (lc_conv_1d); dot: a float64 device tensor or None."""
    n_, nx_, ny_ = _ints(n), _ints(nx), _ints(ny)
    pr = _prec(x)
    ei, eo = _ext1(n_, nx_, ny_, idim) if 0 <= idim < len(n_) else (nx_, ny_)
    _dev(x, 'x', pr, _count(ei) * narr)
    _dev(y, 'y', pr, _count(eo) * narr)
    _dev(dot, 'dot', 'dot', 1, optional=True)
    _resident(x=x, y=y, dot=dot)
    _rc(lib().lc_conv_1d(_prec(x), fil, len(n_), idim, n_.ctypes.data_as(_IP), bc,
                         nx_.ctypes.data_as(_IP), ny_.ctypes.data_as(_IP), narr, _ptr(x), _ptr(y),
                         a, a_x, a_y, _ptr(dot), _stream(stream)))


def wavelet_1d(wav, inverse, idim, n, bc, nx, ny, narr, x, y, a=1.0, a_y=0.0, stream=None):
    n_, nx_, ny_ = _ints(n), _ints(nx), _ints(ny)
    pr = _prec(x)
    ei, eo = _ext1(n_, nx_, ny_, idim) if 0 <= idim < len(n_) else (nx_, ny_)
    _dev(x, 'x', pr, _count(ei, 2) * narr)
    _dev(y, 'y', pr, _count(eo, 2) * narr)
    _resident(x=x, y=y)
    _rc(lib().lc_wavelet_1d(_prec(x), wav, int(bool(inverse)), len(n_), idim, n_.ctypes.data_as(_IP),
                            bc, nx_.ctypes.data_as(_IP), ny_.ctypes.data_as(_IP), narr, _ptr(x),
                            _ptr(y), a, a_y, _stream(stream)))


def conv_md(fil, n, bc, nx, ny, narr, x, y, a=1.0, a_x=0.0, a_y=0.0, stream=None):
    n_, b_, nx_, ny_ = _ints(n), _ints(bc), _ints(nx), _ints(ny)
    pr = _prec(x)
    _dev(x, 'x', pr, _count(nx_) * narr)
    _dev(y, 'y', pr, _count(ny_) * narr)
    _resident(x=x, y=y)
    _rc(lib().lc_conv_md(_prec(x), fil, len(n_), n_.ctypes.data_as(_IP), b_.ctypes.data_as(_IP),
                         nx_.ctypes.data_as(_IP), ny_.ctypes.data_as(_IP), narr, _ptr(x), _ptr(y),
                         a, a_x, a_y, _stream(stream)))


def conv_1ds(fil, n, bc, nx, ny, narr, a, a_x, a_y, x, y, dot=None, stream=None):
    n_, b_, nx_, ny_ = _ints(n), _ints(bc), _ints(nx), _ints(ny)
    pr = _prec(x)
    _dev(x, 'x', pr, _count(nx_) * narr)
    _dev(y, 'y', pr, _count(ny_) * narr)
    _dev(dot, 'dot', 'dot', len(n_), optional=True)
    _resident(x=x, y=y, dot=dot)
    av = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    _rc(lib().lc_conv_1ds(_prec(x), fil, len(n_), n_.ctypes.data_as(_IP), b_.ctypes.data_as(_IP),
                          nx_.ctypes.data_as(_IP), ny_.ctypes.data_as(_IP), narr,
                          av.ctypes.data_as(_DP), a_x, a_y, _ptr(x), _ptr(y), _ptr(dot),
                          _stream(stream)))


def conv_1ds_noalloc(fil, n, bc, nx, ny, narr, a, a_x, a_y, x, y, dot=None, stream=None):
    """conv_1ds through lc_conv_1ds_noalloc: the energy partials use the
    library's static device buffer, no allocation (calls on one GPU must not
    overlap)."""
    n_, b_, nx_, ny_ = _ints(n), _ints(bc), _ints(nx), _ints(ny)
    av = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    pr = _prec(x)
    _dev(x, 'x', pr, _count(nx_) * narr)
    _dev(y, 'y', pr, _count(ny_) * narr)
    _dev(dot, 'dot', 'dot', len(n_), optional=True)
    _resident(x=x, y=y, dot=dot)
    _rc(lib().lc_conv_1ds_noalloc(pr, fil, len(n_), n_.ctypes.data_as(_IP), b_.ctypes.data_as(_IP),
                                  nx_.ctypes.data_as(_IP), ny_.ctypes.data_as(_IP), narr,
                                  av.ctypes.data_as(_DP), a_x, a_y, _ptr(x), _ptr(y), _ptr(dot),
                                  _stream(stream)))


def wavelet_md(wav, inverse, n, bc, nx, ny, narr, x, y, a=1.0, a_y=0.0, stream=None):
    n_, b_, nx_, ny_ = _ints(n), _ints(bc), _ints(nx), _ints(ny)
    pr = _prec(x)
    _dev(x, 'x', pr, _count(nx_, 2) * narr)
    _dev(y, 'y', pr, _count(ny_, 2) * narr)
    _resident(x=x, y=y)
    _rc(lib().lc_wavelet_md(_prec(x), wav, int(bool(inverse)), len(n_), n_.ctypes.data_as(_IP),
                            b_.ctypes.data_as(_IP), nx_.ctypes.data_as(_IP), ny_.ctypes.data_as(_IP),
                            narr, _ptr(x), _ptr(y), a, a_y, _stream(stream)))


def local_hamiltonian_work(n):
    """Elements of the work array lc_local_hamiltonian needs for n pairs per dim."""
    L = lib()
    L.lc_local_hamiltonian_work.restype = ctypes.c_long
    w = L.lc_local_hamiltonian_work(_ints(n).ctypes.data_as(_IP))
    check()
    return int(w)


def local_hamiltonian(n, a, psig, hpsig, pot, rho, work, ekin=None, stream=None):
    """lc_local_hamiltonian on device tensors (libconv.f90:238-295): hpsig =
    DWT o (IMF o pot o MF + sum_d a_d D2_d) o IDWT psig, rho += psir^2."""
    n_ = _ints(n)
    av = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    pr = _prec(psig)
    ng = _count([2 * v for v in n_])
    nr = _count([2 * v + 29 for v in n_])
    _dev(psig, 'psig', pr, ng)
    _dev(hpsig, 'hpsig', pr, ng)
    _dev(pot, 'pot', pr, nr)
    _dev(rho, 'rho', pr, nr, optional=True)
    _dev(work, 'work', pr, local_hamiltonian_work(n_))
    _dev(ekin, 'ekin', 'dot', 3, optional=True)
    _resident(psig=psig, hpsig=hpsig, pot=pot, rho=rho, work=work, ekin=ekin)
    _rc(lib().lc_local_hamiltonian(_prec(psig), n_.ctypes.data_as(_IP), av.ctypes.data_as(_DP), _ptr(psig),
                                   _ptr(hpsig), _ptr(pot), _ptr(rho), _ptr(work), _ptr(ekin),
                                   _stream(stream)))


# ------------------------------------------------------------------ Poisson
NPERIODIC = -2


def poisson_1d(nord, idim, n, bc, u, du, a=1.0, a_y=0.0, stream=None):
    """lc_poisson_1d (Poisson_broker) on device tensors: du = a D_idim u (+ a_y du)."""
    pr = _prec(u)
    _dev(u, 'u', pr, _count(n))
    _dev(du, 'du', pr, _count(n))
    _resident(u=u, du=du)
    _rc(lib().lc_poisson_1d(_prec(u), nord, idim, _ints(n).ctypes.data_as(_IP), bc, _ptr(u), _ptr(du),
                            _D(a), _D(a_y), _stream(stream)))


def _h(h):
    return np.ascontiguousarray(np.asarray(h, dtype=np.float64))


def div_u_i(geocode, n, u, du, nord, h, cc=None, stream=None):
    hv = _h(h)
    pr = _prec(u)
    _dev(u, 'u', pr, 3 * _count(n))
    _dev(du, 'du', pr, _count(n))
    _dev(cc, 'cc', pr, _count(n), optional=True)
    _resident(u=u, du=du, cc=cc)
    _rc(lib().lc_div_u_i(_prec(u), geocode, _ints(n).ctypes.data_as(_IP), _ptr(u), _ptr(du), nord,
                         hv.ctypes.data_as(_DP), _ptr(cc), _stream(stream)))


def nabla_u(geocode, n, u, du, nord, h, stream=None):
    hv = _h(h)
    pr = _prec(u)
    _dev(u, 'u', pr, _count(n))
    _dev(du, 'du', pr, 3 * _count(n))
    _resident(u=u, du=du)
    _rc(lib().lc_nabla_u(_prec(u), geocode, _ints(n).ctypes.data_as(_IP), _ptr(u), _ptr(du), nord,
                         hv.ctypes.data_as(_DP), _stream(stream)))


def nabla_u_and_square(geocode, n, u, du, du2, nord, h, stream=None):
    hv = _h(h)
    pr = _prec(u)
    _dev(u, 'u', pr, _count(n))
    _dev(du, 'du', pr, 3 * _count(n))
    _dev(du2, 'du2', pr, _count(n))
    _resident(u=u, du=du, du2=du2)
    _rc(lib().lc_nabla_u_and_square(_prec(u), geocode, _ints(n).ctypes.data_as(_IP), _ptr(u), _ptr(du),
                                    _ptr(du2), nord, hv.ctypes.data_as(_DP), _stream(stream)))


def nabla_u_square(geocode, n, u, ddu, nord, h, stream=None):
    """lc_nabla_u_square (Poisson2.rb:2278-2359): ddu = |grad u|^2."""
    hv = _h(h)
    pr = _prec(u)
    _dev(u, 'u', pr, _count(n))
    _dev(ddu, 'ddu', pr, _count(n))
    _resident(u=u, ddu=ddu)
    _rc(lib().lc_nabla_u_square(_prec(u), geocode, _ints(n).ctypes.data_as(_IP), _ptr(u), _ptr(ddu), nord,
                                hv.ctypes.data_as(_DP), _stream(stream)))


def nabla_u_epsilon(geocode, n, u, du, nord, h, eps, stream=None):
    """lc_nabla_u_epsilon (Poisson2.rb:2362-2443): du_k = eps D_k u / h_k."""
    hv = _h(h)
    pr = _prec(u)
    _dev(u, 'u', pr, _count(n))
    _dev(du, 'du', pr, 3 * _count(n))
    _dev(eps, 'eps', pr, _count(n))
    _resident(u=u, du=du, eps=eps)
    _rc(lib().lc_nabla_u_epsilon(_prec(u), geocode, _ints(n).ctypes.data_as(_IP), _ptr(u), _ptr(du), nord,
                                 hv.ctypes.data_as(_DP), _ptr(eps), _stream(stream)))


def update_rhopol(geocode, n, u, nord, h, eta, dlogeps, rhopol, rhores2, stream=None):
    """lc_update_rhopol (Poisson2.rb:2446-2550): rhopol updated in place;
    rhores2 a one-element FP64 device tensor that receives the sum of res^2."""
    hv = _h(h)
    pr = _prec(u)
    _dev(u, 'u', pr, _count(n))
    _dev(dlogeps, 'dlogeps', pr, 3 * _count(n))
    _dev(rhopol, 'rhopol', pr, _count(n))
    _dev(rhores2, 'rhores2', 'dot', 1)
    _resident(u=u, dlogeps=dlogeps, rhopol=rhopol, rhores2=rhores2)
    _rc(lib().lc_update_rhopol(_prec(u), geocode, _ints(n).ctypes.data_as(_IP), _ptr(u), nord,
                               hv.ctypes.data_as(_DP), _D(eta), _ptr(dlogeps), _ptr(rhopol), _ptr(rhores2),
                               _stream(stream)))


FIL_MD, FIL_IMD, FIL_D2, FIL_D1 = 0, 1, 2, 3
WAV_SYM8, WAV_SYMICOMB8 = 0, 1


def extent(op, bc, n, side):
    """Input (side 0) / output (side 1) extent along a dim (lc_extent)."""
    return lib().lc_extent(op, bc, n, side)


# ------------------------------------------------------------------ batches
BATCH_SEP, BATCH_SUM, BATCH_DWT, BATCH_IDWT = 0, 1, 2, 3


class _BatchItem(ctypes.Structure):
    _fields_ = [('kind', _I), ('op', _I), ('n', _I * 3), ('bc', _I * 3), ('nx', _I * 3),
                ('ny', _I * 3), ('x', _VP), ('y', _VP), ('a', _D * 3), ('a_x', _D), ('a_y', _D),
                ('dot', _VP)]


def _batch_lib():
    L = lib()
    if not getattr(L, '_batch_ready', False):
        L.lc_batch_create.argtypes = [ctypes.POINTER(_VP), _I, _I, ctypes.POINTER(_BatchItem)]
        L.lc_batch_create.restype = _I
        L.lc_batch_run.argtypes = [_VP, _VP]
        L.lc_batch_run.restype = _I
        L.lc_batch_run_graph.argtypes = [_VP, _VP]
        L.lc_batch_run_graph.restype = _I
        L.lc_batch_info.argtypes = [_VP, _IP, ctypes.POINTER(ctypes.c_long)]
        L.lc_batch_info.restype = _I
        L.lc_batch_destroy.argtypes = [_VP]
        L.lc_batch_destroy.restype = None
        L.lc_batch_run_persistent.argtypes = [_VP, _VP]
        L.lc_batch_run_persistent.restype = _I
        L.lc_batch_run_persistent_graph.argtypes = [_VP, _VP]
        L.lc_batch_run_persistent_graph.restype = _I
        L.lc_batch_persistent_info.argtypes = [_VP, _IP, _IP]
        L.lc_batch_persistent_info.restype = _I
        L.lc_batch_tune.argtypes = [_VP, _I, _DP]
        L.lc_batch_tune.restype = _I
        L.lc_batch_stage_variant.argtypes = [_VP, _I]
        L.lc_batch_stage_variant.restype = _I
        L._batch_ready = True
    return L


class Batch:
    """A plan over many independent 3D operator applications (lc_batch_*).

    items: dicts with keys kind (BATCH_*), op (FIL_* / WAV_*), n, bc, nx, ny
    (3 ints each, bc may be one int), x, y (device tensors), and optionally
    a (3 floats or one), a_x, a_y, dot (float64 device tensor of 3).  The
    tensors must stay alive as long as the plan.  The precision follows the
    tensors' dtype (float64: 8, float32: 4); every x and y must share it,
    and `prec`, when given, must agree with it.
    """

    def __init__(self, items, prec=None):
        L = _batch_lib()
        arr = (_BatchItem * max(len(items), 1))()
        self._keep = []
        if items:
            pr = _prec(items[0]['x'])
            if prec is not None and int(prec) != pr:
                raise LibconvError('Batch: prec %s does not match the tensors (%s)'
                                   % (prec, items[0]['x'].dtype))
            prec = pr
        elif prec is None:
            prec = 8
        for k, it in enumerate(items):
            mul = 2 if int(it['kind']) in (BATCH_DWT, BATCH_IDWT) else 1
            _dev(it['x'], 'item %d x' % k, prec, _count(it['nx'], mul))
            _dev(it['y'], 'item %d y' % k, prec, _count(it['ny'], mul))
            _dev(it.get('dot'), 'item %d dot' % k, 'dot', 3, optional=True)
        for k, it in enumerate(items):
            _resident(**{'item %d x' % k: it['x'], 'item %d y' % k: it['y'], 'item %d dot' % k: it.get('dot')})
            b = arr[k]
            b.kind, b.op = int(it['kind']), int(it['op'])
            bc = it['bc']
            bc = [bc] * 3 if isinstance(bc, (int, np.integer)) else list(bc)
            for i in range(3):
                b.n[i], b.bc[i] = int(it['n'][i]), int(bc[i])
                b.nx[i], b.ny[i] = int(it['nx'][i]), int(it['ny'][i])
            a = it.get('a', 1.0)
            a = [a, 0.0, 0.0] if isinstance(a, (int, float)) else list(a)
            for i in range(3):
                b.a[i] = float(a[i])
            b.a_x, b.a_y = float(it.get('a_x', 0.0)), float(it.get('a_y', 0.0))
            b.x, b.y = it['x'].data_ptr(), it['y'].data_ptr()
            dot = it.get('dot')
            b.dot = dot.data_ptr() if dot is not None else None
            self._keep.append(it)
        h = _VP()
        _rc(L.lc_batch_create(ctypes.byref(h), prec, len(items), arr))
        self._h = h

    def run(self, stream=None):
        _rc(_batch_lib().lc_batch_run(self._h, _stream(stream)))

    def run_graph(self, stream=None):
        _rc(_batch_lib().lc_batch_run_graph(self._h, _stream(stream)))

    def info(self):
        s, b = _I(), ctypes.c_long()
        _rc(_batch_lib().lc_batch_info(self._h, ctypes.byref(s), ctypes.byref(b)))
        return {'stages': s.value, 'blocks': b.value}

    def run_persistent(self, stream=None):
        """The whole plan in one persistent launch with device work queues."""
        _rc(_batch_lib().lc_batch_run_persistent(self._h, _stream(stream)))

    def run_persistent_graph(self, stream=None):
        _rc(_batch_lib().lc_batch_run_persistent_graph(self._h, _stream(stream)))

    def persistent_info(self):
        """{'grid': blocks of the persistent launch, 'status': 0 unless a
        dependency wait exceeded its bound (synchronises the device)}."""
        g, st = _I(), _I()
        _rc(_batch_lib().lc_batch_persistent_info(self._h, ctypes.byref(g), ctypes.byref(st)))
        return {'grid': g.value, 'status': st.value}

    def tune(self, repeat=3):
        """Tunes every stage over its kernel variants; returns the best ms per stage."""
        n = self.info()['stages']
        ms = (_D * max(n, 1))()
        _rc(_batch_lib().lc_batch_tune(self._h, repeat, ms))
        return [ms[i] for i in range(n)]

    def stage_variant(self, stage):
        v = _batch_lib().lc_batch_stage_variant(self._h, stage)
        if v < 0:
            check()
        return v

    def close(self):
        if getattr(self, '_h', None):
            _batch_lib().lc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ autotuner
FORM_SEP, FORM_SUM, FORM_DWT, FORM_IDWT = 0, 1, 2, 3


def _tune_lib():
    L = lib()
    if not getattr(L, '_tune_ready', False):
        L.lc_tune_md.argtypes = [_I, _I, _I, _IP, _IP, _IP, _IP, _I, _IP, _DP, _DP, _I]
        L.lc_tune_md.restype = _I
        L.lc_tune_save.argtypes = [ctypes.c_char_p]
        L.lc_tune_save.restype = _I
        L.lc_tune_load.argtypes = [ctypes.c_char_p]
        L.lc_tune_load.restype = _I
        L.lc_tune_clear.argtypes = []
        L.lc_tune_clear.restype = None
        L.lc_variant_count.argtypes = [_I]
        L.lc_variant_count.restype = _I
        L.lc_kernel_name.argtypes = [_I, _I, _I, _IP, _IP, _IP, _IP, ctypes.c_char_p, _I]
        L.lc_kernel_name.restype = _I
        L._tune_ready = True
    return L


def _dims3(n, bc, nx, ny):
    bc = [bc] * 3 if isinstance(bc, (int, np.integer)) else bc
    return [_ints(v) for v in (n, bc, nx, ny)]


def tune_md(form, op, n, bc, nx, ny, repeat=3, prec=8):
    """Times every kernel variant of a 3D form on random data of this shape,
    records the fastest for the shape; returns {'best', 'best_ms', 'times'}."""
    L = _tune_lib()
    n_, b_, nx_, ny_ = _dims3(n, bc, nx, ny)
    nv = L.lc_variant_count(form)
    best, best_ms = _I(), _D()
    times = (_D * nv)()
    _rc(L.lc_tune_md(prec, form, op, n_.ctypes.data_as(_IP), b_.ctypes.data_as(_IP),
                     nx_.ctypes.data_as(_IP), ny_.ctypes.data_as(_IP), repeat, ctypes.byref(best),
                     ctypes.byref(best_ms), times, nv))
    return {'best': best.value, 'best_ms': best_ms.value, 'times': [times[i] for i in range(nv)]}


def tune_save(path):
    _rc(_tune_lib().lc_tune_save(os.fsencode(path)))


def tune_load(path):
    _rc(_tune_lib().lc_tune_load(os.fsencode(path)))


def tune_clear():
    _tune_lib().lc_tune_clear()


def variant_count(form):
    return _tune_lib().lc_variant_count(form)


def kernel_name(prec, form, op, n, bc, nx, ny):
    """Name of the kernel a call of this shape launches (as profilers show it)."""
    L = _tune_lib()
    n_, b_, nx_, ny_ = _dims3(n, bc, nx, ny)
    buf = ctypes.create_string_buffer(256)
    _rc(L.lc_kernel_name(prec, form, op, n_.ctypes.data_as(_IP), b_.ctypes.data_as(_IP),
                         nx_.ctypes.data_as(_IP), ny_.ctypes.data_as(_IP), buf, 256))
    return buf.value.decode()
