"""GPU parity: the HIP library through its C ABI against the FP64 oracle.

Bar: FP64 within 1e-13 relative max-norm of the oracle (the path's stated
tolerance; the reference's own tests use 1e-14 for MagicFilter, 1e-12 for
Kinetic, 1e-11 for wavelets); FP32 storage and arithmetic within 1e-6 of the
FP64 oracle on the unrounded FP64 input (so input quantisation included),
relative to the oracle's max-norm (measured on MI355X: <= 3.7e-7 MagicFilter,
1.5e-7 Kinetic, 2.7e-7 local Hamiltonian); FP64 energies within 1e-12
relative, energies over FP32 data within 1e-6.  Inputs are seeded U[-1, 1].
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-13
TOL32 = 1e-6


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope='module')
def lc(gpu):
    import libconv
    libconv.lib()
    return libconv


def conv_sizes(fil, bc, n, idim):
    nin, nout = O.conv_extents(fil, bc, n[idim])
    nx, ny = list(n), list(n)
    nx[idim], ny[idim] = nin, nout
    return nx, ny


FILS = {'md': O.MD, 'imd': O.IMD, 'd2': O.D2, 'd1': O.D1}


# ------------------------------------------------------------------ config 1
def test_config1_magicfilter_dim0_32cube(lc):
    """MagicFilter sweep along dim 0 of a 32^3 FP64 grid, periodic BC."""
    rng = np.random.default_rng(1)
    n = [32, 32, 32]
    x = rng.uniform(-1, 1, 32 ** 3)
    y = np.zeros_like(x)
    lc.MagicFilter1d().run(3, 0, n, lc.BC.PERIODIC, n, n, 1, x, y, 1.0, 0.0, 0.0)
    ref = np.zeros_like(x)
    O.s0s0_1d(O.MD, 0, n, 0, n, n, 1, x, ref)
    assert rel(y, ref) < TOL64


# ------------------------------------------------------------------ 1D sweeps
@pytest.mark.parametrize('fil', ['md', 'imd', 'd2', 'd1'])
@pytest.mark.parametrize('bc', [0, 1, -1])
@pytest.mark.parametrize('idim', [0, 1, 2])
def test_conv_1d_sweeps(lc, fil, bc, idim):
    rng = np.random.default_rng(100 + 7 * idim + bc)
    n = [19, 23, 17]
    nx, ny = conv_sizes(FILS[fil], bc, n, idim)
    nx[idim] += 1  # padded leading dims
    ny[idim] += 2
    narr = 2
    nin = int(np.prod(nx)) * narr
    nout = int(np.prod(ny)) * narr
    x = rng.uniform(-1, 1, nin)
    y0 = rng.uniform(-1, 1, nout)
    for a, a_x, a_y in ((1.0, 0.0, 0.0), (-2.5, 0.75, 0.5)):
        y = y0.copy()
        ref = y0.copy()
        if fil == 'd2':
            st = lc.Kinetic1d().run(3, idim, n, bc, nx, ny, narr, x, y, a, a_x, a_y, 0.0)
            dot = st['reference_return']['dot_in']
            rdot = O.s0s0_1d(O.D2, idim, n, bc, nx, ny, narr, x, ref, a, a_x, a_y, want_dot=True)
            assert abs(dot - rdot) <= 1e-12 * abs(rdot)
        else:
            k = {'md': lc.MagicFilter1d, 'imd': lc.InverseMagicFilter1d}.get(fil)
            if k is None:
                k = lambda: lc.LibconvKernel('s0s0_1d_sym8_d1')
            k().run(3, idim, n, bc, nx, ny, narr, x, y, a, a_x, a_y)
            O.s0s0_1d(FILS[fil], idim, n, bc, nx, ny, narr, x, ref, a, a_x, a_y)
        assert rel(y, ref) < TOL64, (fil, bc, idim, a)


def test_a_y_zero_does_not_read_y(lc):
    n = [16, 9, 5]
    x = np.random.default_rng(2).uniform(-1, 1, int(np.prod(n)))
    y = np.full_like(x, np.nan)
    lc.MagicFilter1d().run(3, 1, n, 0, n, n, 1, x, y, 1.0, 0.0, 0.0)
    assert np.isfinite(y).all()


def test_small_n_below_filter_width(lc):
    """n < filter width: true periodic wrap and both-sided grow clamping."""
    rng = np.random.default_rng(3)
    for nn in (1, 3, 9, 14):
        for bc in (0, 1):
            n = [nn, 6]
            nx, ny = conv_sizes(O.D2, bc, n, 0)
            x = rng.uniform(-1, 1, int(np.prod(nx)))
            y = np.zeros(int(np.prod(ny)))
            ref = y.copy()
            lc.LibconvKernel('s0s0_1d_sym8_d2').run(2, 0, n, bc, nx, ny, 1, x, y, 1.0, 0.0, 0.0)
            O.s0s0_1d(O.D2, 0, n, bc, nx, ny, 1, x, ref)
            # n = 1 periodic maps x to x * sum(D2) ~ 1e-16: compare on the input scale
            scale = max(np.abs(ref).max(), np.abs(x).max())
            assert np.abs(y - ref).max() < TOL64 * scale, (nn, bc)


@pytest.mark.parametrize('comb', [False, True])
@pytest.mark.parametrize('direction', ['decompose', 'recompose'])
@pytest.mark.parametrize('bc', [0, 1, -1])
@pytest.mark.parametrize('idim', [0, 1, 2])
def test_wavelet_1d(lc, comb, direction, bc, idim):
    rng = np.random.default_rng(200 + idim + 5 * bc)
    wav = O.SYMICOMB8 if comb else O.SYM8
    n = [10, 13, 9]
    nin, nout = O.wavelet_extents(wav, bc, n[idim])
    nx, ny = list(n), list(n)
    nx[idim], ny[idim] = nin + 1, nout
    narr = 2
    x = rng.uniform(-1, 1, int(np.prod([2 * v for v in nx])) * narr)
    y0 = rng.uniform(-1, 1, int(np.prod([2 * v for v in ny])) * narr)
    for a, a_y in ((1.0, 0.0), (1.3, -0.4)):
        y, ref = y0.copy(), y0.copy()
        lc.Wavelet1d(direction, combined=comb).run(3, idim, n, bc, nx, ny, narr, x, y, a, a_y)
        f = O.s0s1_1d if direction == 'decompose' else O.s1s0_1d
        f(wav, idim, n, bc, nx, ny, narr, x, ref, a, a_y)
        assert rel(y, ref) < TOL64


# ------------------------------------------------------------------ 3D operators
@pytest.mark.parametrize('bcs', [(0, 0, 0), (1, 1, 1), (-1, -1, -1), (0, 1, -1)])
@pytest.mark.parametrize('inverse', [False, True])
def test_magicfilter_3d(lc, bcs, inverse):
    rng = np.random.default_rng(300)
    fil = O.IMD if inverse else O.MD
    n = [21, 18, 23]
    nx, ny = [], []
    for i in range(3):
        a, b = O.conv_extents(fil, bcs[i], n[i])
        nx.append(a)
        ny.append(b + (2 if bcs[i] == 0 else 0))  # padded output as MagicFilter_test2.rb
    narr = 2
    x = rng.uniform(-1, 1, int(np.prod(nx)) * narr)
    y = np.zeros(int(np.prod(ny)) * narr)
    ref = y.copy()
    lc.MagicFilter(inverse=inverse).run(3, n, list(bcs), nx, ny, narr, x, y, None, None)
    O.s0s0_md(fil, n, list(bcs), nx, ny, narr, x, ref)
    # storage beyond the logical output box (padding) is unspecified
    nout = [O.conv_extents(fil, b, v)[1] for b, v in zip(bcs, n)]
    shape = (narr, ny[2], ny[1], ny[0])
    box = (slice(None), slice(0, nout[2]), slice(0, nout[1]), slice(0, nout[0]))
    assert rel(y.reshape(shape)[box], ref.reshape(shape)[box]) < TOL64


@pytest.mark.parametrize('bcs', [(0, 0, 0), (1, 1, 1), (-1, -1, -1), (1, 0, -1)])
def test_kinetic_3d_with_energies(lc, bcs):
    rng = np.random.default_rng(400)
    n = [17, 20, 15]
    nx = [v + 28 if b == -1 else v for v, b in zip(n, bcs)]
    ny = [v + 28 if b == 1 else v for v, b in zip(n, bcs)]
    x = rng.uniform(-1, 1, int(np.prod(nx)))
    y0 = rng.uniform(-1, 1, int(np.prod(ny)))
    h = (0.3, 0.35, 0.4)
    a = [-0.5 / v ** 2 for v in h]
    y, ref = y0.copy(), y0.copy()
    st = lc.Kinetic(dot=True).run(3, n, list(bcs), x, y, None, None, a, 0.5, 0.25)
    rdot = O.s0s0_1ds(O.D2, n, list(bcs), nx, ny, 1, a, 0.5, 0.25, x, ref)
    assert rel(y, ref) < TOL64
    assert np.abs(st['reference_return']['dot_in'] - rdot).max() <= 1e-12 * np.abs(rdot).max()


@pytest.mark.parametrize('bcs', [(2, 2, 2), (2, 0, -1), (1, 2, 0)])
def test_free_bc_3d_operators(lc, gpu, bcs):
    """bc 2 (free, output on the input box; libconv.f90:245 'F' nogrow) in the
    single-pass 3D kernels: MagicFilter product and Kinetic sum with
    energies, a_x and a_y, through the drop-in names and the stream API."""
    torch = gpu
    rng = np.random.default_rng(700)
    n = [19, 22, 17]
    nx, ny = [], []
    for i in range(3):
        a, b = O.conv_extents(O.MD, bcs[i], n[i])
        nx.append(a)
        ny.append(b)
    x = rng.uniform(-1, 1, int(np.prod(nx)))
    y, ref = np.zeros(int(np.prod(ny))), np.zeros(int(np.prod(ny)))
    lc.MagicFilter().run(3, n, list(bcs), nx, ny, 1, x, y, None, None)
    O.s0s0_md(O.MD, n, list(bcs), nx, ny, 1, x, ref)
    assert rel(y, ref) < TOL64
    xt, yt = torch.from_numpy(x).cuda(), torch.zeros(ref.size, dtype=torch.float64, device='cuda')
    lc.conv_md(lc.FIL_IMD, n, list(bcs), nx, ny, 1, xt, yt)
    O.s0s0_md(O.IMD, n, list(bcs), nx, ny, 1, x, ref)
    torch.cuda.synchronize()
    assert rel(yt.cpu().numpy(), ref) < TOL64
    kx = [v + 28 if b == -1 else v for v, b in zip(n, bcs)]
    ky = [v + 28 if b == 1 else v for v, b in zip(n, bcs)]
    x = rng.uniform(-1, 1, int(np.prod(kx)))
    y0 = rng.uniform(-1, 1, int(np.prod(ky)))
    a = [-0.5 / v ** 2 for v in (0.3, 0.35, 0.4)]
    ax = 0.5 if kx == ky else 0.0
    y, ref = y0.copy(), y0.copy()
    st = lc.Kinetic(dot=True).run(3, n, list(bcs), x, y, None, None, a, ax, 0.25)
    rdot = O.s0s0_1ds(O.D2, n, list(bcs), kx, ky, 1, a, ax, 0.25, x, ref)
    assert rel(y, ref) < TOL64
    assert np.abs(st['reference_return']['dot_in'] - rdot).max() <= 1e-12 * np.abs(rdot).max()


def test_free_bc_rejected_where_no_kernel_takes_it(lc):
    """1D sweeps and wavelets have no free form: a clear error, no launch."""
    n = [12, 12, 12]
    x, y = np.zeros(12 ** 3), np.zeros(12 ** 3)
    with pytest.raises(lc.LibconvError):
        lc.MagicFilter1d().run(3, 0, n, 2, n, n, 1, x, y, 1.0, 0.0, 0.0)
        lc.check()
    with pytest.raises(lc.LibconvError):
        lc.Wavelet('decompose').run(3, [6, 6, 6], [2, 2, 2], np.zeros(12 ** 3), y, None, None)
        lc.check()


@pytest.mark.parametrize('bc', [0, 1, -1])
def test_wavelet_3d_and_round_trip(lc, bc):
    rng = np.random.default_rng(500)
    n = [9, 11, 8]
    nin = [v + 7 if bc == -1 else v for v in n]
    nout = [v + 7 if bc == 1 else v for v in n]
    s = rng.uniform(-1, 1, int(np.prod([2 * v for v in nin])))
    c = np.zeros(int(np.prod([2 * v for v in nout])))
    ref = c.copy()
    lc.Wavelet('decompose').run(3, n, [bc] * 3, s, c, None, None)
    O.s0s1_md(O.SYM8, n, [bc] * 3, nin, nout, 1, s, ref)
    assert rel(c, ref) < TOL64
    if bc == 0:
        back = np.zeros_like(s)
        lc.Wavelet('recompose').run(3, n, [0, 0, 0], c, back, None, None)
        assert rel(back, s) < TOL64


def test_combined_operators_match_compositions(lc, gpu):
    """symicomb8 through the library against the library's own sym8 and
    magic-filter operators (BigDFT/TestOperations.ipynb cells 8-19):
    s1tor = MF o IDWT and rtos1 = DWT o IMF along one dim of a 3D array
    (periodic), the grow / shrink pair along dim 0, and the 3D periodic
    s1tor through the device API."""
    torch = gpu
    rng = np.random.default_rng(610)
    n = [7, 10, 6]
    n2 = [2 * v for v in n]
    x = rng.uniform(-1, 1, int(np.prod(n2)))
    # periodic, idim 1
    w, comp, comb = np.zeros(x.size), np.zeros(x.size), np.zeros(x.size)
    lc.Wavelet1d('recompose').run(3, 1, n, 0, n, n, 1, x, w, 1.0, 0.0)
    lc.MagicFilter1d().run(3, 1, n2, 0, n2, n2, 1, w, comp, 1.0, 0.0, 0.0)
    lc.Wavelet1d('recompose', combined=True).run(3, 1, n, 0, n, n, 1, x, comb, 1.0, 0.0)
    assert rel(comb, comp) < TOL64
    w, comp, comb = np.zeros(x.size), np.zeros(x.size), np.zeros(x.size)
    lc.InverseMagicFilter1d().run(3, 1, n2, 0, n2, n2, 1, x, w, 1.0, 0.0, 0.0)
    lc.Wavelet1d('decompose').run(3, 1, n, 0, n, n, 1, w, comp, 1.0, 0.0)
    lc.Wavelet1d('decompose', combined=True).run(3, 1, n, 0, n, n, 1, x, comb, 1.0, 0.0)
    assert rel(comb, comp) < TOL64
    # grow then shrink along dim 0 (1D arrays)
    m = 13
    v = rng.uniform(-1, 1, 2 * m)
    s0 = np.zeros(2 * (m + 7))
    lc.Wavelet1d('recompose').run(1, 0, [m], 1, [m], [m + 7], 1, v, s0, 1.0, 0.0)
    sr = np.zeros(2 * m + 29)
    lc.MagicFilter1d().run(1, 0, [2 * (m + 7)], 1, [2 * (m + 7)], [2 * m + 29], 1, s0, sr, 1.0, 0.0, 0.0)
    sr1 = np.zeros(2 * (m + 15))
    lc.Wavelet1d('recompose', combined=True).run(1, 0, [m], 1, [m], [m + 15], 1, v, sr1, 1.0, 0.0)
    assert rel(sr1[:2 * m + 29], sr) < TOL64 and sr1[-1] == 0.0
    s0b = np.zeros(2 * (m + 7))
    lc.InverseMagicFilter1d().run(1, 0, [2 * (m + 7)], -1, [2 * m + 29], [2 * (m + 7)], 1, sr, s0b,
                                  1.0, 0.0, 0.0)
    p = np.zeros(2 * m)
    lc.Wavelet1d('decompose').run(1, 0, [m], -1, [m + 7], [m], 1, s0b, p, 1.0, 0.0)
    p2 = np.zeros(2 * m)
    lc.Wavelet1d('decompose', combined=True).run(1, 0, [m], -1, [m + 15], [m], 1, sr1, p2, 1.0, 0.0)
    assert rel(p2, p) < TOL64
    # 3D periodic s1tor on device tensors (three sweeps) vs MF3D o IDWT3D
    xt = torch.from_numpy(x).cuda()
    w, comp, comb = torch.empty_like(xt), torch.empty_like(xt), torch.empty_like(xt)
    lc.wavelet_md(lc.WAV_SYM8, True, n, [0] * 3, n, n, 1, xt, w)
    lc.conv_md(lc.FIL_MD, n2, [0] * 3, n2, n2, 1, w, comp)
    lc.wavelet_md(lc.WAV_SYMICOMB8, True, n, [0] * 3, n, n, 1, xt, comb)
    torch.cuda.synchronize()
    assert rel(comb.cpu().numpy(), comp.cpu().numpy()) < TOL64


def test_wavelet_round_trip_config3_shape(lc, gpu):
    """sym8 DWT then IDWT on n = (96, 96, 96) pairs, periodic: identity."""
    torch = gpu
    n = [96, 96, 96]
    g = torch.Generator(device='cuda').manual_seed(3)
    s = torch.rand(192 ** 3, dtype=torch.float64, device='cuda', generator=g) * 2 - 1
    c = torch.empty_like(s)
    back = torch.empty_like(s)
    lc.wavelet_md(0, False, n, [0, 0, 0], n, n, 1, s, c)
    lc.wavelet_md(0, True, n, [0, 0, 0], n, n, 1, c, back)
    torch.cuda.synchronize()
    err = ((back - s).abs().max() / s.abs().max()).item()
    assert err < TOL64
    # orthonormal: the analysis keeps the 2-norm
    assert abs(c.norm().item() / s.norm().item() - 1) < 1e-13


# ------------------------------------------------------------------ FP32 form
def test_fp32_forms_against_fp64_oracle(lc):
    rng = np.random.default_rng(600)
    n = [24, 20, 22]
    x = rng.uniform(-1, 1, int(np.prod(n)))
    y = np.zeros(int(np.prod(n)), np.float32)
    ref = np.zeros(int(np.prod(n)))
    lc.MagicFilter(fp32=True).run(3, n, [0, 0, 0], n, n, 1, x.astype(np.float32), y, None, None)
    O.s0s0_md(O.MD, n, [0, 0, 0], n, n, 1, x, ref)
    assert rel(y, ref) < TOL32
    a = [-5.5, -4.0, -3.0]
    yk = np.zeros(int(np.prod(n)), np.float32)
    st = lc.Kinetic(fp32=True, dot=True).run(3, n, [0, 0, 0], x.astype(np.float32), yk, None, None,
                                            a, 0.0, 0.0)
    refk = np.zeros(int(np.prod(n)))
    rdot = O.s0s0_1ds(O.D2, n, [0, 0, 0], n, n, 1, a, 0.0, 0.0, x, refk)
    assert rel(yk, refk) < TOL32
    # FP64 energies over FP32 data: the error is the data's rounding only
    assert np.abs(st['reference_return']['dot_in'] - rdot).max() <= TOL32 * np.abs(rdot).max()


# ------------------------------------------------------------------ stream API
def test_stream_api_device_tensors(lc, gpu):
    torch = gpu
    rng = np.random.default_rng(700)
    n = [40, 36, 33]
    bc = [1, 1, 1]
    ny1 = [v + 15 for v in n]
    ny2 = [v + 28 for v in ny1]
    x = rng.uniform(-1, 1, int(np.prod(n)))
    xd = torch.from_numpy(x).cuda()
    y1 = torch.zeros(int(np.prod(ny1)), dtype=torch.float64, device='cuda')
    y2 = torch.zeros(int(np.prod(ny2)), dtype=torch.float64, device='cuda')
    dot = torch.zeros(3, dtype=torch.float64, device='cuda')
    a = [-0.5 / 0.3 ** 2] * 3
    s = torch.cuda.current_stream()
    lc.conv_md(0, n, bc, n, ny1, 1, xd, y1, stream=s)
    lc.conv_1ds(2, ny1, bc, ny1, ny2, 1, a, 0.0, 0.0, y1, y2, dot=dot, stream=s)
    torch.cuda.synchronize()
    r1 = np.zeros(int(np.prod(ny1)))
    O.s0s0_md(O.MD, n, bc, n, ny1, 1, x, r1)
    r2 = np.zeros(int(np.prod(ny2)))
    rdot = O.s0s0_1ds(O.D2, ny1, bc, ny1, ny2, 1, a, 0.0, 0.0, r1, r2)
    assert rel(y1.cpu().numpy(), r1) < TOL64
    assert rel(y2.cpu().numpy(), r2) < TOL64
    assert np.abs(dot.cpu().numpy() - rdot).max() <= 1e-12 * np.abs(rdot).max()


def test_stream_api_two_streams_keep_their_scratch(lc, gpu):
    """Energy partials and sweep intermediates live in per-(device, stream)
    scratch: calls of one host thread on two streams, queued without a
    synchronisation between them, give the bits of the same calls run one
    after the other on one stream."""
    torch = gpu
    rng = np.random.default_rng(710)
    n = [96, 80, 88]
    a = [-0.5 / 0.3 ** 2, -0.5 / 0.25 ** 2, -0.5 / 0.2 ** 2]
    xs = [torch.from_numpy(rng.uniform(-1, 1, int(np.prod(n)))).cuda() for _ in range(2)]
    nsw = [n[0] + 28, n[1], n[2]]

    def work(x, s):
        y = torch.empty_like(x)
        dot = torch.zeros(3, dtype=torch.float64, device='cuda')
        w = torch.empty(int(np.prod(nsw)), dtype=torch.float64, device='cuda')
        d1 = torch.zeros(1, dtype=torch.float64, device='cuda')
        c = torch.empty_like(x)
        pairs = [v // 2 for v in n]
        lc.conv_1ds(2, n, [0] * 3, n, n, 1, a, 0.0, 0.0, x, y, dot=dot, stream=s)   # energy partials
        lc.conv_1d(2, 0, n, 1, n, nsw, 1, x, w, a=-2.0, dot=d1, stream=s)           # 1D partials
        lc.wavelet_md(1, False, pairs, [0] * 3, pairs, pairs, 1, x, c, stream=s)    # intermediates
        return y, dot, w, d1, c

    s0 = torch.cuda.current_stream()
    ref = [work(x, s0) for x in xs]
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    got = [work(xs[0], s1), work(xs[1], s2)]
    torch.cuda.synchronize()
    for r, g in zip(ref, got):
        for u, v in zip(r, g):
            assert torch.equal(u, v)


# ------------------------------------------------------------ local Hamiltonian
def _ham_inputs(torch, rng, n, dtype):
    ng = int(np.prod([2 * v for v in n]))
    nr = int(np.prod([2 * (v + 7) + 15 for v in n]))
    psig = rng.uniform(-1, 1, ng)
    pot = rng.uniform(-1, 1, nr)
    rho = rng.uniform(0, 1, nr)
    return psig, pot, rho


@pytest.mark.parametrize('dtype_name', ['float64', 'float32'])
def test_local_hamiltonian_matches_oracle(lc, gpu, dtype_name):
    """lc_local_hamiltonian (libconv.f90:238-295) against the oracle's
    composition: hpsig, the updated density and the three energies."""
    torch = gpu
    dtype = getattr(torch, dtype_name)
    rng = np.random.default_rng(900)
    n = [7, 9, 6]
    a = [-0.5 / 0.2 ** 2, -0.5 / 0.25 ** 2, -0.5 / 0.3 ** 2]
    psig, pot, rho0 = _ham_inputs(torch, rng, n, dtype)
    cast = (lambda v: v.astype(np.float32).astype(np.float64)) if dtype_name == 'float32' else (lambda v: v)
    ref_h, ref_rho, ref_e = O.local_hamiltonian(n, a, cast(psig), cast(pot), cast(rho0).copy())
    dev = lambda v: torch.from_numpy(v).to('cuda', dtype)
    tp, tpot, trho = dev(psig), dev(pot), dev(rho0)
    th = torch.full_like(tp, 9.0)
    work = torch.empty(lc.local_hamiltonian_work(n), dtype=dtype, device='cuda')
    ekin = torch.zeros(3, dtype=torch.float64, device='cuda')
    lc.local_hamiltonian(n, a, tp, th, tpot, trho, work, ekin)
    torch.cuda.synchronize()
    tol = TOL64 if dtype_name == 'float64' else TOL32
    assert rel(th.cpu().numpy(), ref_h) < tol
    assert rel(trho.cpu().numpy(), ref_rho) < tol
    assert np.abs(ekin.cpu().numpy() - ref_e).max() <= (1e-12 if dtype_name == 'float64' else TOL32) * np.abs(ref_e).max()


def test_local_hamiltonian_is_symmetric(lc, gpu):
    """At a realistic orbital size (48 x 40 x 44 pairs): <phi, H psi> =
    <H phi, psi> for a real potential, and with pot = 0, <psi, H psi> equals
    the sum of the returned energies (the Kinetic term in the wavelet basis)."""
    torch = gpu
    rng = np.random.default_rng(901)
    n = [48, 40, 44]
    a = [-0.5 / 0.2 ** 2] * 3
    psig, pot, _ = _ham_inputs(torch, rng, n, torch.float64)
    phig = rng.uniform(-1, 1, psig.size)
    dev = lambda v: torch.from_numpy(v).cuda()
    work = torch.empty(lc.local_hamiltonian_work(n), dtype=torch.float64, device='cuda')
    tpot = dev(pot)
    hp, hf = torch.empty(psig.size, dtype=torch.float64, device='cuda'), torch.empty(psig.size, dtype=torch.float64, device='cuda')
    tpsi, tphi = dev(psig), dev(phig)
    lc.local_hamiltonian(n, a, tpsi, hp, tpot, None, work)
    lc.local_hamiltonian(n, a, tphi, hf, tpot, None, work)
    lhs, rhs = float(torch.dot(tphi, hp)), float(torch.dot(hf, tpsi))
    assert abs(lhs - rhs) < 1e-11 * max(abs(lhs), 1.0)
    ekin = torch.zeros(3, dtype=torch.float64, device='cuda')
    lc.local_hamiltonian(n, a, tpsi, hp, torch.zeros_like(tpot), None, work, ekin)
    e = float(ekin.sum())
    assert abs(float(torch.dot(tpsi, hp)) - e) < 1e-11 * abs(e)


@pytest.mark.parametrize('v', [0, 6, 12, 17, 22, 25, 28, 33, 37, 40])
def test_local_hamiltonian_potential_step_in_every_form(lc, gpu, monkeypatch, v):
    """The potential / density step rides in the magic filter's store: the
    single-plane (0, 12), multi-plane (17, 22, 25, 28, 33) and pipelined (37, 40)
    separable kernels take it; a wave-specialised choice (6) falls back to the
    default."""
    torch = gpu
    monkeypatch.setenv('LIBCONV_SEP3D_VARIANT', str(v))
    rng = np.random.default_rng(950 + v)
    n = [6, 8, 7]
    a = [-0.5 / 0.2 ** 2, -0.5 / 0.25 ** 2, -0.5 / 0.3 ** 2]
    psig, pot, rho0 = _ham_inputs(torch, rng, n, torch.float64)
    ref_h, ref_rho, _ = O.local_hamiltonian(n, a, psig, pot, rho0.copy())
    tp, tpot, trho = (torch.from_numpy(u).cuda() for u in (psig, pot, rho0))
    th = torch.empty_like(tp)
    work = torch.empty(lc.local_hamiltonian_work(n), dtype=torch.float64, device='cuda')
    lc.local_hamiltonian(n, a, tp, th, tpot, trho, work)
    torch.cuda.synchronize()
    assert rel(th.cpu().numpy(), ref_h) < TOL64
    assert rel(trho.cpu().numpy(), ref_rho) < TOL64
