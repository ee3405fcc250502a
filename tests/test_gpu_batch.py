"""GPU parity of the single-pass 3D wavelet transform and of batch plans.

The single-pass 3D DWT/IDWT is checked against the FP64 oracle's three-sweep
composition (1e-13 relative max-norm; FP32 within 1e-6 on the unrounded
input; measured 4.1e-7).  A batch plan runs
the same kernels over many items from a device argument table: its outputs
must equal, bit for bit, the outputs of one call per item (the z chunking of
a launch does not change any output's arithmetic); its FP64 energies sum the
same per-block partials in another grouping, so they agree to 1e-13.  Graph
replays are checked after eager work between them, each against an eager
run on the same inputs.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-13


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope='module')
def lc(gpu):
    import libconv
    libconv.lib()
    return libconv


def wav_sizes(wav, bcs, n):
    nx, ny = [], []
    for b, v in zip(bcs, n):
        i, o = O.wavelet_extents(wav, b, v)
        nx.append(i)
        ny.append(o)
    return nx, ny


# ------------------------------------------------------- single-pass wavelets
# wav 0 (sym8) runs the single-pass kernel; wav 1 (symicomb8) runs as three
# 1D sweeps, because its single-pass form is switched off in the library
@pytest.mark.parametrize('wav', [0, 1])
@pytest.mark.parametrize('bcs', [(0, 0, 0), (1, 1, 1), (-1, -1, -1), (0, 1, -1), (1, -1, 1)])
@pytest.mark.parametrize('inverse', [False, True])
def test_wavelet_3d_single_pass(lc, gpu, wav, bcs, inverse):
    torch = gpu
    rng = np.random.default_rng(800 + 10 * wav + (3 if inverse else 0) + sum(bcs))
    n = [19, 13, 11]
    nx, ny = wav_sizes(wav, bcs, n)
    bc_call = list(bcs)
    narr = 2
    tin = int(np.prod([2 * v for v in nx]))
    tout = int(np.prod([2 * v for v in ny]))
    x = rng.uniform(-1, 1, tin * narr)
    y0 = rng.uniform(-1, 1, tout * narr)
    a, a_y = 0.75, -0.5
    ref = y0.copy()
    (O.s1s0_md if inverse else O.s0s1_md)(wav, n, bc_call, nx, ny, narr, x, ref, a, a_y)
    xd = torch.from_numpy(x).cuda()
    yd = torch.from_numpy(y0).cuda()
    lc.wavelet_md(wav, inverse, n, bc_call, nx, ny, narr, xd, yd, a=a, a_y=a_y)
    torch.cuda.synchronize()
    assert rel(yd.cpu().numpy(), ref) < TOL64, bc_call
    # FP32 storage and arithmetic
    y32 = torch.from_numpy(y0.astype(np.float32)).cuda()
    lc.wavelet_md(wav, inverse, n, bc_call, nx, ny, narr, xd.float(), y32, a=a, a_y=a_y)
    torch.cuda.synchronize()
    assert rel(y32.cpu().numpy(), ref) < 1e-6


@pytest.mark.skip(reason='the single-pass symicomb8 kernel gave wrong high bands and then faulted '
                         'on the GPU; it is not built and symicomb8 3D runs as three sweeps')
def test_symicomb8_3d_single_pass():
    pass


def test_wavelet_3d_single_pass_equals_sweeps(lc, gpu, monkeypatch):
    """The single-pass kernel and the three 1D sweeps agree to rounding."""
    torch = gpu
    n = [40, 24, 33]
    rng = np.random.default_rng(900)
    x = torch.from_numpy(rng.uniform(-1, 1, 8 * int(np.prod(n)))).cuda()
    y1 = torch.zeros_like(x)
    y2 = torch.zeros_like(x)
    lc.wavelet_md(0, False, n, [0, 0, 0], n, n, 1, x, y1)
    monkeypatch.setenv('LIBCONV_DIRECT', '1')
    lc.wavelet_md(0, False, n, [0, 0, 0], n, n, 1, x, y2)
    torch.cuda.synchronize()
    assert rel(y1.cpu().numpy(), y2.cpu().numpy()) < 1e-14


# ------------------------------------------------------------------ batches
def _orbitals(rng, count, smin, smax):
    orbs = []
    for _ in range(count):
        s = [int(v) for v in rng.integers(smin, smax + 1, 3)]
        bc = int(rng.choice([0, 1, -1]))
        orbs.append((s, bc))
    return orbs


def _chain_sizes(s, bc):
    """x -> MF -> y1 -> Kinetic -> y2 extents for a nominal grid s."""
    if bc == 0:
        return s, s, s, s, s
    if bc == 1:
        y1 = [v + 15 for v in s]
        return s, s, y1, y1, [v + 28 for v in y1]   # n_mf, x, y1, n_k, y2
    y1 = [v + 28 for v in s]
    return y1, [v + 15 for v in y1], y1, s, s


def _build(torch, orbs, rng, dtype):
    """Device buffers and batch items for MF, Kinetic (energies), DWT per orbital."""
    a_k = [-0.5 / 0.3 ** 2, -0.5 / 0.35 ** 2, -0.5 / 0.25 ** 2]
    mf, kin, dwt, bufs = [], [], [], []
    for s, bc in orbs:
        n_mf, xs, y1s, n_k, y2s = _chain_sizes(s, bc)
        x = torch.from_numpy(rng.uniform(-1, 1, int(np.prod(xs)))).to('cuda', dtype)
        y1 = torch.zeros(int(np.prod(y1s)), dtype=dtype, device='cuda')
        y2 = torch.zeros(int(np.prod(y2s)), dtype=dtype, device='cuda')
        dot = torch.zeros(3, dtype=torch.float64, device='cuda')
        wnx, wny = wav_sizes(0, [bc] * 3, s)
        wx = torch.from_numpy(rng.uniform(-1, 1, 8 * int(np.prod(wnx)))).to('cuda', dtype)
        wy = torch.zeros(8 * int(np.prod(wny)), dtype=dtype, device='cuda')
        mf.append(dict(kind=0, op=0, n=n_mf, bc=bc, nx=xs, ny=y1s, x=x, y=y1))
        kin.append(dict(kind=1, op=2, n=n_k, bc=bc, nx=y1s, ny=y2s, x=y1, y=y2, a=a_k, dot=dot))
        dwt.append(dict(kind=2, op=0, n=s, bc=bc, nx=wnx, ny=wny, x=wx, y=wy))
        bufs.append((x, y1, y2, dot, wx, wy))
    return mf + kin + dwt, bufs


def _single_calls(lc, items):
    """The same work, one stream-API call per item."""
    for it in items:
        bc = [it['bc']] * 3
        if it['kind'] == 0:
            lc.conv_md(it['op'], it['n'], bc, it['nx'], it['ny'], 1, it['x'], it['y'])
        elif it['kind'] == 1:
            lc.conv_1ds(it['op'], it['n'], bc, it['nx'], it['ny'], 1, it['a'], 0.0, 0.0, it['x'],
                        it['y'], dot=it['dot'])
        else:
            lc.wavelet_md(it['op'], it['kind'] == 3, it['n'], bc, it['nx'], it['ny'], 1, it['x'],
                          it['y'])


@pytest.mark.parametrize('dtype_name', ['float64', 'float32'])
def test_batch_equals_single_calls_and_oracle(lc, gpu, dtype_name):
    torch = gpu
    dtype = getattr(torch, dtype_name)
    rng = np.random.default_rng(1000)
    orbs = _orbitals(rng, 12, 5, 40)
    items, bufs = _build(torch, orbs, rng, dtype)
    plan = lc.Batch(items, prec=8 if dtype_name == 'float64' else 4)
    assert plan.info()['stages'] == 3
    plan.run()
    torch.cuda.synchronize()
    got = [[b.clone() for b in (y1, y2, dot, wy)] for (_, y1, y2, dot, _, wy) in bufs]
    for (_, y1, y2, dot, _, wy) in bufs:
        for b in (y1, y2, dot, wy):
            b.fill_(7.0)
    _single_calls(lc, items)
    torch.cuda.synchronize()
    for k, (_, y1, y2, dot, _, wy) in enumerate(bufs):
        assert torch.equal(got[k][0], y1), k
        assert torch.equal(got[k][1], y2), k
        assert torch.equal(got[k][3], wy), k
        d_b, d_s = got[k][2].cpu().numpy(), dot.cpu().numpy()
        assert np.abs(d_b - d_s).max() <= 1e-13 * np.abs(d_s).max(), k
    if dtype_name != 'float64':
        return
    # three orbitals against the oracle (one per boundary condition)
    seen = set()
    for k, (s, bc) in enumerate(orbs):
        if bc in seen:
            continue
        seen.add(bc)
        mf, kin, dw = items[k], items[len(orbs) + k], items[2 * len(orbs) + k]
        x = mf['x'].cpu().numpy()
        r1 = np.zeros(mf['y'].numel())
        O.s0s0_md(O.MD, mf['n'], [bc] * 3, mf['nx'], mf['ny'], 1, x, r1)
        r2 = np.zeros(kin['y'].numel())
        rdot = O.s0s0_1ds(O.D2, kin['n'], [bc] * 3, kin['nx'], kin['ny'], 1, kin['a'], 0.0, 0.0, r1, r2)
        rw = np.zeros(dw['y'].numel())
        O.s0s1_md(O.SYM8, dw['n'], [bc] * 3, dw['nx'], dw['ny'], 1, dw['x'].cpu().numpy(), rw)
        assert rel(got[k][0].cpu().numpy(), r1) < TOL64
        assert rel(got[k][1].cpu().numpy(), r2) < TOL64
        assert np.abs(got[k][2].cpu().numpy() - rdot).max() <= 1e-12 * np.abs(rdot).max()
        assert rel(got[k][3].cpu().numpy(), rw) < TOL64


def test_batch_free_bc_items(lc, gpu):
    """bc 2 (free, output on the input box) in SEP and SUM items: the stage
    launches and the persistent launch equal one call per item bit for bit,
    and the oracle to 1e-13."""
    torch = gpu
    rng = np.random.default_rng(1050)
    a_k = [-0.5 / 0.3 ** 2, -0.5 / 0.35 ** 2, -0.5 / 0.25 ** 2]
    mf, kin, bufs = [], [], []
    for _ in range(6):
        s = [int(v) for v in rng.integers(8, 36, 3)]
        bcs = [int(b) for b in rng.choice([2, 0], 3)]
        x = torch.from_numpy(rng.uniform(-1, 1, int(np.prod(s)))).cuda()
        y1 = torch.zeros(x.numel(), dtype=torch.float64, device='cuda')
        y2 = torch.zeros(x.numel(), dtype=torch.float64, device='cuda')
        dot = torch.zeros(3, dtype=torch.float64, device='cuda')
        mf.append(dict(kind=0, op=0, n=s, bc=bcs, nx=s, ny=s, x=x, y=y1))
        kin.append(dict(kind=1, op=2, n=s, bc=bcs, nx=s, ny=s, x=y1, y=y2, a=a_k, dot=dot))
        bufs.append((x, y1, y2, dot))
    items = mf + kin
    plan = lc.Batch(items)
    plan.run()
    torch.cuda.synchronize()
    got = [[b.clone() for b in (y1, y2, dot)] for (_, y1, y2, dot) in bufs]
    for (_, y1, y2, dot) in bufs:
        for b in (y1, y2, dot):
            b.fill_(3.0)
    for it in items:
        if it['kind'] == 0:
            lc.conv_md(it['op'], it['n'], it['bc'], it['nx'], it['ny'], 1, it['x'], it['y'])
        else:
            lc.conv_1ds(it['op'], it['n'], it['bc'], it['nx'], it['ny'], 1, it['a'], 0.0, 0.0,
                        it['x'], it['y'], dot=it['dot'])
    torch.cuda.synchronize()
    for k, (x, y1, y2, dot) in enumerate(bufs):
        assert torch.equal(got[k][0], y1) and torch.equal(got[k][1], y2), k
        r1 = np.zeros(y1.numel())
        O.s0s0_md(O.MD, mf[k]['n'], mf[k]['bc'], mf[k]['n'], mf[k]['n'], 1, x.cpu().numpy(), r1)
        r2 = np.zeros(y2.numel())
        rdot = O.s0s0_1ds(O.D2, kin[k]['n'], kin[k]['bc'], kin[k]['n'], kin[k]['n'], 1, a_k, 0.0, 0.0,
                          r1, r2)
        assert rel(got[k][0].cpu().numpy(), r1) < TOL64
        assert rel(got[k][1].cpu().numpy(), r2) < TOL64
        assert np.abs(got[k][2].cpu().numpy() - rdot).max() <= 1e-12 * np.abs(rdot).max()
    plan.run_persistent()
    assert plan.persistent_info()['status'] == 0
    for k, (_, y1, y2, dot) in enumerate(bufs):
        assert torch.equal(got[k][0], y1) and torch.equal(got[k][1], y2), k
    plan.close()


def test_batch_graph_replays(lc, gpu):
    """Each replay of the captured plan matches an eager run on the same inputs."""
    torch = gpu
    rng = np.random.default_rng(1100)
    orbs = _orbitals(rng, 9, 6, 30)
    items, bufs = _build(torch, orbs, rng, torch.float64)
    plan = lc.Batch(items)
    for rep in range(3):
        # eager work between replays: new inputs in place
        for (x, _, _, _, wx, _) in bufs:
            x.uniform_(-1, 1)
            wx.uniform_(-1, 1)
        plan.run_graph()
        torch.cuda.synchronize()
        got = [[b.clone() for b in (y1, y2, dot, wy)] for (_, y1, y2, dot, _, wy) in bufs]
        plan.run()
        torch.cuda.synchronize()
        for k, (_, y1, y2, dot, _, wy) in enumerate(bufs):
            assert torch.equal(got[k][0], y1), (rep, k)
            assert torch.equal(got[k][1], y2), (rep, k)
            assert torch.equal(got[k][2], dot), (rep, k)
            assert torch.equal(got[k][3], wy), (rep, k)


def test_batch_rejects_bad_items(lc, gpu):
    torch = gpu
    x = torch.zeros(1000, dtype=torch.float64, device='cuda')
    y = torch.zeros(1000, dtype=torch.float64, device='cuda')
    good = dict(kind=0, op=0, n=[10, 10, 10], bc=0, nx=[10, 10, 10], ny=[10, 10, 10], x=x, y=y)
    lc.Batch([good]).close()
    with pytest.raises(lc.LibconvError):
        lc.Batch([dict(good, nx=[9, 10, 10], bc=-1)])          # shrink input too short
    with pytest.raises(lc.LibconvError):
        lc.Batch([dict(good, a_x=0.5)])                        # a_x on a SEP item
    with pytest.raises(lc.LibconvError):
        lc.Batch([dict(good, kind=7)])
    w = dict(good, kind=2, op=1, n=[5, 5, 5], nx=[5, 5, 5], ny=[5, 5, 5])
    with pytest.raises(lc.LibconvError):
        lc.Batch([w])                                          # symicomb8 in a batch


# ------------------------------------------------------------------ persistent
def test_persistent_equals_stage_launches(lc, gpu):
    """One persistent launch (device work queues, dependency counters)
    gives the stage launches' outputs bit for bit; energies to 1e-13."""
    torch = gpu
    rng = np.random.default_rng(1500)
    orbs = _orbitals(rng, 24, 5, 44)
    items, bufs = _build(torch, orbs, rng, torch.float64)
    plan = lc.Batch(items)
    plan.run()
    torch.cuda.synchronize()
    ref = [[b.clone() for b in (y1, y2, dot, wy)] for (_, y1, y2, dot, _, wy) in bufs]
    for (_, y1, y2, dot, _, wy) in bufs:
        for b in (y1, y2, dot, wy):
            b.fill_(5.0)
    plan.run_persistent()
    info = plan.persistent_info()
    assert info['status'] == 0 and info['grid'] >= 8
    for k, (_, y1, y2, dot, _, wy) in enumerate(bufs):
        assert torch.equal(ref[k][0], y1), k
        assert torch.equal(ref[k][1], y2), k
        assert torch.equal(ref[k][3], wy), k
        d_p, d_s = dot.cpu().numpy(), ref[k][2].cpu().numpy()
        assert np.abs(d_p - d_s).max() <= 1e-13 * np.abs(d_s).max(), k


@pytest.mark.parametrize('dtype_name', ['float64', 'float32'])
def test_persistent_unbalanced_queues(lc, gpu, dtype_name):
    """One large orbital among tiny ones: most queues run dry early and their
    blocks take the large items' units from the other queues, Kinetic units
    there waiting for magic-filter units run on other XCDs.  Outputs equal
    the stage launches bit for bit, over replays."""
    torch = gpu
    dtype = getattr(torch, dtype_name)
    rng = np.random.default_rng(1550)
    orbs = [([88, 80, 72], -1)] + [([6, 7, 6], int(rng.choice([0, 1, -1]))) for _ in range(15)]
    items, bufs = _build(torch, orbs, rng, dtype)
    plan = lc.Batch(items, prec=8 if dtype_name == 'float64' else 4)
    plan.run()
    torch.cuda.synchronize()
    ref = [[b.clone() for b in (y1, y2, dot, wy)] for (_, y1, y2, dot, _, wy) in bufs]
    for rep in range(3):
        for (_, y1, y2, dot, _, wy) in bufs:
            for b in (y1, y2, dot, wy):
                b.fill_(3.0 + rep)
        plan.run_persistent()
        torch.cuda.synchronize()
        assert plan.persistent_info()['status'] == 0
        for k, (_, y1, y2, dot, _, wy) in enumerate(bufs):
            assert torch.equal(ref[k][0], y1), (rep, k)
            assert torch.equal(ref[k][1], y2), (rep, k)
            assert torch.equal(ref[k][3], wy), (rep, k)
            d_p, d_s = dot.cpu().numpy(), ref[k][2].cpu().numpy()
            assert np.abs(d_p - d_s).max() <= 1e-13 * np.abs(d_s).max(), (rep, k)


@pytest.mark.parametrize('dtype_name', ['float64', 'float32'])
def test_persistent_graph_replays(lc, gpu, dtype_name):
    torch = gpu
    dtype = getattr(torch, dtype_name)
    rng = np.random.default_rng(1600)
    orbs = _orbitals(rng, 16, 6, 36)
    items, bufs = _build(torch, orbs, rng, dtype)
    plan = lc.Batch(items, prec=8 if dtype_name == 'float64' else 4)
    for rep in range(3):
        for (x, _, _, _, wx, _) in bufs:
            x.uniform_(-1, 1)
            wx.uniform_(-1, 1)
        plan.run_persistent_graph()
        torch.cuda.synchronize()
        got = [[b.clone() for b in (y1, y2, dot, wy)] for (_, y1, y2, dot, _, wy) in bufs]
        plan.run()
        torch.cuda.synchronize()
        for k, (_, y1, y2, dot, _, wy) in enumerate(bufs):
            assert torch.equal(got[k][0], y1), (rep, k)
            assert torch.equal(got[k][1], y2), (rep, k)
            assert torch.equal(got[k][3], wy), (rep, k)
            assert np.abs(got[k][2].cpu().numpy() - dot.cpu().numpy()).max() <= \
                1e-13 * np.abs(dot.cpu().numpy()).max(), (rep, k)
    assert plan.persistent_info()['status'] == 0


def test_persistent_bounded_wait_ends(lc, gpu, monkeypatch):
    """With a wait bound of one sleep, dependent units give up at once: the
    launch still ends, and the status word tells whether any wait did."""
    torch = gpu
    rng = np.random.default_rng(1700)
    items, bufs = _build(torch, _orbitals(rng, 6, 8, 24), rng, torch.float64)
    monkeypatch.setenv('LIBCONV_PERSIST_SPIN_LIMIT', '1')
    plan = lc.Batch(items)
    plan.run_persistent()
    assert plan.persistent_info()['status'] in (0, 1)


def test_persistent_rejects_items_without_a_persistent_form(lc, gpu):
    torch = gpu
    x = torch.zeros(1000, dtype=torch.float64, device='cuda')
    y = torch.zeros(1000, dtype=torch.float64, device='cuda')
    it = dict(kind=1, op=0, n=[10, 10, 10], bc=0, nx=[10, 10, 10], ny=[10, 10, 10], x=x, y=y)  # SUM md
    plan = lc.Batch([it])
    with pytest.raises(lc.LibconvError):
        plan.run_persistent()
