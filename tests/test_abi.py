"""The C ABI library: builds, loads, exports every symbol include/libconv.h
declares, and rejects invalid arguments before touching a device (CPU only:
no compute call is made here)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'libconv.h')
LIB = os.path.join(ROOT, 'boast-lig_amd', 'libconv_amd.so')


def declared():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    names = re.findall(r'^\s*(?:const\s+)?(?:void|int|char\s*\*|const char\s*\*)\s*\**\s*(\w+)\s*\(', text, re.M)
    return sorted(set(names))


@pytest.fixture(scope='module')
def lib():
    if not os.path.exists(LIB):
        pytest.fail('libconv_amd.so missing: run __graft_entry__.build()')
    import libconv
    return libconv.lib()


def test_header_declares_the_drop_in_names():
    names = declared()
    for must in ('d_s0s0_1d_sym8_md', 'd_s0s0_1d_sym8_d2_dotin', 'd_s0s1_1d_sym8', 'd_s1s0_1d_sym8',
                 'd_sym8_md_ld_narr', 'd_sym8_d2', 'd_dwt_sym8', 'd_idwt_sym8', 's0s0_1d_sym8_md_d_',
                 's0s0_1ds_dot_sym8_d2_d_', 'lc_conv_1d', 'lc_last_error', 'nabla_u', 'div_u_i',
                 'nabla_u_and_square', 'nabla_u_square', 'nabla_u_epsilon', 'update_rhopol',
                 'lc_nabla_u_square', 'lc_nabla_u_epsilon', 'lc_update_rhopol'):
        assert must in names, must


def test_library_exports_every_declared_symbol(lib):
    out = subprocess.run(['nm', '-D', '--defined-only', LIB], capture_output=True, text=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if ' T ' in l)
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_version_and_extents(lib):
    import libconv
    assert 'gfx950' in libconv.version()
    # MF grow +15, kinetic shrink +28, sym8 grow +7 pairs, symicomb8 +15
    assert libconv.extent(0, 1, 100, 1) == 115 and libconv.extent(0, 1, 100, 0) == 100
    assert libconv.extent(2, -1, 100, 0) == 128 and libconv.extent(2, -1, 100, 1) == 100
    assert libconv.extent(10, 1, 50, 1) == 57 and libconv.extent(21, -1, 50, 0) == 65
    assert libconv.extent(0, 5, 10, 0) == -1


@pytest.mark.parametrize('case', ['d0', 'idim', 'bc', 'nx_small', 'ny_grow', 'narr'])
def test_invalid_arguments_are_reported(lib, case):
    import libconv
    n = np.array([8, 8, 8], np.int32)
    nx, ny = n.copy(), n.copy()
    d, idim, bc, narr = 3, 0, 0, 1
    if case == 'd0':
        d = 0
    elif case == 'idim':
        idim = 3
    elif case == 'bc':
        bc = 7
    elif case == 'nx_small':
        bc = -1  # shrink needs nx >= n + 15
    elif case == 'ny_grow':
        bc = 1   # grow needs ny >= n + 15
    elif case == 'narr':
        narr = -1
    x = np.zeros(512)
    y = np.full(512, 3.0)
    k = libconv.MagicFilter1d()
    with pytest.raises(libconv.LibconvError):
        k.run(d, idim, n, bc, nx, ny, narr, x, y, 1.0, 0.0, 0.0)
    assert (y == 3.0).all()  # untouched on error


def test_stream_api_rejects_bad_precision(lib):
    import libconv
    with pytest.raises(libconv.LibconvError):
        libconv._rc(libconv.lib().lc_conv_1d(3, 0, 1, 0, (ctypes.c_int * 1)(8), 0, (ctypes.c_int * 1)(8),
                                             (ctypes.c_int * 1)(8), 1, None, None, 1.0, 0.0, 0.0, None, None))


@pytest.mark.parametrize('case', ['dtype', 'strided', 'short', 'host', 'dot_dtype', 'dot_short'])
def test_stream_api_checks_tensors_before_the_call(lib, case):
    """The Python stream API checks dtype, contiguity, size against the
    extents, device residence and the FP64 energy buffer before any address
    reaches the C ABI (CPU tensors here: every case must be refused before
    the library is asked to compute)."""
    torch = pytest.importorskip('torch')
    import libconv
    n = [8, 9, 10]
    x = torch.zeros(720, dtype=torch.float64)
    y = torch.zeros(720, dtype=torch.float64)
    dot = torch.zeros(3, dtype=torch.float64)
    if case == 'dtype':
        y = y.float()
    elif case == 'strided':
        y = torch.zeros(1440, dtype=torch.float64)[::2]
    elif case == 'short':
        y = y[:719]
    elif case == 'dot_dtype':
        dot = dot.float()
    elif case == 'dot_short':
        dot = dot[:2]
    msg = {'dtype': 'dtype', 'strided': 'contiguous', 'short': 'elements', 'host': 'device',
           'dot_dtype': 'dot: dtype', 'dot_short': 'dot: 2 elements'}[case]
    with pytest.raises(libconv.LibconvError, match=msg):
        libconv.conv_1ds(2, n, [0] * 3, n, n, 1, [-1.0] * 3, 0.0, 0.0, x, y, dot)
    if case in ('dtype', 'strided', 'short', 'host'):
        with pytest.raises(libconv.LibconvError):
            libconv.conv_md(0, n, [0] * 3, n, n, 1, x, y)
        with pytest.raises(libconv.LibconvError):
            libconv.conv_1d(0, 0, n, 0, n, n, 1, x, y)


def test_batch_precision_follows_the_tensors(lib):
    """Batch infers FP64 / FP32 from x and refuses a prec that disagrees or
    mixed dtypes, before creating a plan (CPU tensors: refused either way)."""
    torch = pytest.importorskip('torch')
    import libconv
    n = [10, 10, 10]
    xf = torch.zeros(1000, dtype=torch.float32)
    item = dict(kind=0, op=0, n=n, bc=0, nx=n, ny=n, x=xf, y=torch.zeros(1000, dtype=torch.float32))
    with pytest.raises(libconv.LibconvError, match='prec'):
        libconv.Batch([item], prec=8)
    with pytest.raises(libconv.LibconvError, match='dtype'):
        libconv.Batch([dict(item, y=torch.zeros(1000, dtype=torch.float64))])
    with pytest.raises(libconv.LibconvError, match='elements'):
        libconv.Batch([dict(item, y=torch.zeros(999, dtype=torch.float32))])
    with pytest.raises(libconv.LibconvError, match='device'):
        libconv.Batch([item])


def test_noalloc_entry_point_refuses_oversized_partials_before_the_device(lib):
    """lc_conv_1ds_noalloc keeps its energy partials in a static 16 MiB device
    array: a 480^3 grid with energies needs more and is refused with
    LC_ERR_UNSUPPORTED before any device call (so this runs without a GPU;
    the host buffers passed are never touched)."""
    import libconv
    L = libconv.lib()
    n = (ctypes.c_int * 3)(480, 480, 480)
    bc = (ctypes.c_int * 3)(0, 0, 0)
    a = (ctypes.c_double * 3)(-1.0, -1.0, -1.0)
    buf = np.zeros(4)
    p = ctypes.c_void_p(buf.ctypes.data)
    rc = L.lc_conv_1ds_noalloc(8, 2, 3, n, bc, n, n, 1, a, 0.0, 0.0, p, p, buf.ctypes.data_as(
        ctypes.POINTER(ctypes.c_double)), None)
    assert rc == 2  # LC_ERR_UNSUPPORTED
    assert L.lc_last_error() == 2
    assert (buf == 0).all()


@pytest.mark.parametrize('case', ['nord', 'small_n', 'null_h', 'null_rhores2', 'host_rhores2', 'prec'])
def test_poisson_dielectric_entry_points_validate_before_the_device(lib, case):
    """lc_nabla_u_square, lc_nabla_u_epsilon and lc_update_rhopol refuse a
    bad order, a grid no larger than the stencil, a missing hgrids or output
    and a bad precision before any device call (no GPU here)."""
    import libconv
    L = libconv.lib()
    n = (ctypes.c_int * 3)(9, 9, 9)
    h = (ctypes.c_double * 3)(0.1, 0.1, 0.1)
    nord, prec, hp = 4, 8, h
    dummy = ctypes.c_void_p(16)  # never dereferenced: validation fails first
    r2 = ctypes.cast(dummy, ctypes.POINTER(ctypes.c_double))
    if case == 'nord':
        nord = 5
    elif case == 'small_n':
        nord = 8  # n = 9 is not > 9
        n = (ctypes.c_int * 3)(9, 8, 9)
    elif case == 'null_h':
        hp = None
    elif case == 'null_rhores2':
        r2 = None
    elif case == 'host_rhores2':  # every other argument valid: refused for the host address alone
        host = (ctypes.c_double * 1)()
        r2 = ctypes.cast(host, ctypes.POINTER(ctypes.c_double))
    elif case == 'prec':
        prec = 2
    calls = [lambda: L.lc_update_rhopol(prec, 1, n, dummy, nord, hp, ctypes.c_double(0.5), dummy, dummy, r2, None)]
    if case not in ('null_rhores2', 'host_rhores2'):
        calls += [lambda: L.lc_nabla_u_square(prec, 1, n, dummy, dummy, nord, hp, None),
                  lambda: L.lc_nabla_u_epsilon(prec, 1, n, dummy, dummy, nord, hp, dummy, None)]
    for call in calls:
        with pytest.raises(libconv.LibconvError):
            libconv._rc(call())


@pytest.mark.parametrize('case', ['rhores2_dtype', 'dlogeps_short', 'eps_dtype', 'host'])
def test_poisson_dielectric_wrappers_check_tensors(lib, case):
    torch = pytest.importorskip('torch')
    import libconv
    n = [9, 9, 9]
    np_ = 729
    u = torch.zeros(np_, dtype=torch.float64)
    with pytest.raises(libconv.LibconvError):
        if case == 'rhores2_dtype':
            libconv.update_rhopol(1, n, u, 4, [0.1] * 3, 0.5, torch.zeros(3 * np_, dtype=torch.float64),
                                  torch.zeros(np_, dtype=torch.float64), torch.zeros(1, dtype=torch.float32))
        elif case == 'dlogeps_short':
            libconv.update_rhopol(1, n, u, 4, [0.1] * 3, 0.5, torch.zeros(3 * np_ - 1, dtype=torch.float64),
                                  torch.zeros(np_, dtype=torch.float64), torch.zeros(1, dtype=torch.float64))
        elif case == 'eps_dtype':
            libconv.nabla_u_epsilon(1, n, u, torch.zeros(3 * np_, dtype=torch.float64), 4, [0.1] * 3,
                                    torch.zeros(np_, dtype=torch.float32))
        else:  # CPU tensors of the right shape: refused as not on the device
            libconv.nabla_u_square(1, n, u, torch.zeros(np_, dtype=torch.float64), 4, [0.1] * 3)


def test_static_archive_links_and_runs(tmp_path):
    """libconv_amd.a (make static) links into a C program with the HIP
    runtime and serves the calls that need no device."""
    pkg = os.path.join(ROOT, 'boast-lig_amd')
    subprocess.run(['make', '-s', '-C', pkg, 'static'], check=True, capture_output=True)
    exe = str(tmp_path / 'lc_static_check')
    subprocess.run(['gcc', '-O1', '-I', os.path.join(ROOT, 'include'), os.path.join(ROOT, 'tests', 'c',
                    'lc_static_check.c'), os.path.join(pkg, 'libconv_amd.a'), '-L/opt/rocm/lib', '-lamdhip64',
                    '-lstdc++', '-lm', '-Wl,-rpath,/opt/rocm/lib', '-o', exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip().endswith('OK'), out.stdout + out.stderr
