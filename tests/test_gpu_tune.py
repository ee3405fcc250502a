"""GPU: every compiled kernel variant against the FP64 oracle, the tuner's
on-device search, and batch-plan tuning (outputs stay bit-identical to one
call per item, since the batched variants differ only in tiling)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-13


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture()
def lc(gpu):
    import libconv
    libconv.lib()
    libconv.tune_clear()
    yield libconv
    libconv.tune_clear()


@pytest.mark.parametrize('v', range(73))
@pytest.mark.parametrize('bc', [(1, -1, 0), (2, 0, 2)])
def test_every_fused_variant_matches_oracle(lc, gpu, monkeypatch, v, bc):
    torch = gpu
    assert max(lc.variant_count(lc.FORM_SEP), lc.variant_count(lc.FORM_SUM)) == 73
    rng = np.random.default_rng(1200 + v)
    n = [37, 21, 45]
    bc = list(bc)
    # MagicFilter, separable form
    nx = [O.conv_extents(O.MD, b, k)[0] for b, k in zip(bc, n)]
    ny = [O.conv_extents(O.MD, b, k)[1] for b, k in zip(bc, n)]
    x = rng.uniform(-1, 1, int(np.prod(nx)))
    ref = np.zeros(int(np.prod(ny)))
    O.s0s0_md(O.MD, n, bc, nx, ny, 1, x, ref)
    monkeypatch.setenv('LIBCONV_SEP3D_VARIANT', str(v))
    monkeypatch.setenv('LIBCONV_SUM3D_VARIANT', str(v))
    if v < lc.variant_count(lc.FORM_SEP):
        y = torch.zeros(ref.size, dtype=torch.float64, device='cuda')
        lc.conv_md(lc.FIL_MD, n, bc, nx, ny, 1, torch.from_numpy(x).cuda(), y)
        torch.cuda.synchronize()
        assert rel(y.cpu().numpy(), ref) < TOL64
    if v >= lc.variant_count(lc.FORM_SUM):
        return
    # Kinetic, sum form with energies, a_x and a_y
    nx = [O.conv_extents(O.D2, b, k)[0] for b, k in zip(bc, n)]
    ny = [O.conv_extents(O.D2, b, k)[1] for b, k in zip(bc, n)]
    x = rng.uniform(-1, 1, int(np.prod(nx)))
    y0 = rng.uniform(-1, 1, int(np.prod(ny)))
    a = [-5.5, -4.0, -3.0]
    ref = y0.copy()
    rdot = O.s0s0_1ds(O.D2, n, bc, nx, ny, 1, a, 0.0, 0.25, x, ref)
    y = torch.from_numpy(y0).cuda()
    dot = torch.zeros(3, dtype=torch.float64, device='cuda')
    lc.conv_1ds(lc.FIL_D2, n, bc, nx, ny, 1, a, 0.0, 0.25, torch.from_numpy(x).cuda(), y, dot=dot)
    torch.cuda.synchronize()
    assert rel(y.cpu().numpy(), ref) < TOL64
    assert np.abs(dot.cpu().numpy() - rdot).max() <= 1e-12 * np.abs(rdot).max()


@pytest.mark.parametrize('v', range(18))
@pytest.mark.parametrize('inverse', [False, True])
def test_every_wavelet_variant_matches_oracle(lc, gpu, monkeypatch, v, inverse):
    torch = gpu
    rng = np.random.default_rng(1300 + v)
    n = [13, 17, 11]
    bc = [0, 1, -1]
    nx = [n[0], n[1], n[2] + 7]
    ny = [n[0], n[1] + 7, n[2]]
    x = rng.uniform(-1, 1, 8 * int(np.prod(nx)))
    ref = np.zeros(8 * int(np.prod(ny)))
    (O.s1s0_md if inverse else O.s0s1_md)(O.SYM8, n, bc, nx, ny, 1, x, ref)
    monkeypatch.setenv('LIBCONV_WAV3D_VARIANT', str(v))
    y = torch.zeros(ref.size, dtype=torch.float64, device='cuda')
    lc.wavelet_md(lc.WAV_SYM8, inverse, n, bc, nx, ny, 1, torch.from_numpy(x).cuda(), y)
    torch.cuda.synchronize()
    assert rel(y.cpu().numpy(), ref) < TOL64


def test_all_separable_variants_give_identical_bits(lc, gpu, monkeypatch):
    torch = gpu
    rng = np.random.default_rng(1260)
    n = [41, 30, 35]
    bc = [1, 0, -1]
    nx = [n[0], n[1], n[2] + 15]
    ny = [n[0] + 15, n[1], n[2]]
    for dt in (torch.float64, torch.float32):
        x = torch.from_numpy(rng.uniform(-1, 1, int(np.prod(nx)))).to('cuda', dt)
        outs = []
        prec = 8 if dt == torch.float64 else 4
        for v in range(lc.variant_count(lc.FORM_SEP)):
            monkeypatch.setenv('LIBCONV_SEP3D_VARIANT', str(v))
            y = torch.zeros(int(np.prod(ny)), dtype=dt, device='cuda')
            try:
                lc.conv_md(lc.FIL_MD, n, bc, nx, ny, 1, x, y, a=0.5)
            except lc.LibconvError:
                # the 128-bit-LDS kernels take whole 16-byte pieces per task:
                # some of their shapes exist in one precision only
                name = lc.kernel_name(prec, lc.FORM_SEP, lc.FIL_MD, n, bc, nx, ny)
                assert name.startswith('sepq_fused'), (v, name)
                continue
            outs.append(y)
        torch.cuda.synchronize()
        for v, y in enumerate(outs):
            assert torch.equal(y, outs[0]), (str(dt), v)


@pytest.mark.parametrize('v', range(18))
def test_wavelet_variants_on_small_and_multi_array_shapes(lc, gpu, monkeypatch, v):
    """Every wavelet variant on grids smaller than a tile and than the
    filter (periodic dims of 2 and 3 steps wrap several times; a free dim of
    2 steps), with two arrays per call, against the FP64 oracle."""
    torch = gpu
    rng = np.random.default_rng(1350 + v)
    monkeypatch.setenv('LIBCONV_WAV3D_VARIANT', str(v))
    for n, bc in (([3, 5, 2], [0, 0, 0]), ([2, 9, 3], [1, -1, 0]), ([11, 3, 2], [-1, 0, 1])):
        nx = [k + 7 if b == -1 else k for k, b in zip(n, bc)]
        ny = [k + 7 if b == 1 else k for k, b in zip(n, bc)]
        narr = 2
        for inverse in (False, True):
            x = rng.uniform(-1, 1, 8 * int(np.prod(nx)) * narr)
            ref = np.zeros(8 * int(np.prod(ny)) * narr)
            (O.s1s0_md if inverse else O.s0s1_md)(O.SYM8, n, bc, nx, ny, narr, x, ref)
            y = torch.zeros(ref.size, dtype=torch.float64, device='cuda')
            lc.wavelet_md(lc.WAV_SYM8, inverse, n, bc, nx, ny, narr, torch.from_numpy(x).cuda(), y)
            torch.cuda.synchronize()
            assert rel(y.cpu().numpy(), ref) < TOL64, (n, bc, inverse)


@pytest.mark.parametrize('inverse', [False, True])
def test_all_wavelet_variants_give_identical_bits(lc, gpu, monkeypatch, inverse):
    """The step-at-a-time and plane-at-a-time wavelet kernels sum in the same
    order: every variant gives the same bits, in both precisions, with and
    without a and a_y (a shape with partial tiles, several z chunks and every
    boundary condition)."""
    torch = gpu
    rng = np.random.default_rng(1270 + inverse)
    n = [21, 19, 30]
    bc = [1, 0, -1]
    nx = [n[0], n[1], n[2] + 7]
    ny = [n[0] + 7, n[1], n[2]]
    if inverse:
        nx, ny = ny, nx
        bc = [-1, 0, 1]
    for dt in (torch.float64, torch.float32):
        x = torch.from_numpy(rng.uniform(-1, 1, 8 * int(np.prod(nx)))).to('cuda', dt)
        y0 = torch.from_numpy(rng.uniform(-1, 1, 8 * int(np.prod(ny)))).to('cuda', dt)
        for a, a_y in ((1.0, 0.0), (0.75, -0.5)):
            outs = []
            prec = 8 if dt == torch.float64 else 4
            for v in range(lc.variant_count(lc.FORM_DWT)):
                monkeypatch.setenv('LIBCONV_WAV3D_VARIANT', str(v))
                y = y0.clone()
                try:
                    lc.wavelet_md(lc.WAV_SYM8, inverse, n, bc, nx, ny, 1, x, y, a=a, a_y=a_y)
                except lc.LibconvError:
                    # plane-at-a-time shapes with x-tasks of part of a 16-byte
                    # piece exist in FP64 only
                    name = lc.kernel_name(prec, lc.FORM_IDWT if inverse else lc.FORM_DWT, lc.WAV_SYM8, n, bc, nx, ny)
                    assert name.startswith('wavq_fused') and prec == 4, (v, name)
                    continue
                outs.append(y)
            torch.cuda.synchronize()
            for v, y in enumerate(outs):
                assert torch.equal(y, outs[0]), (str(dt), a, a_y, v)


def test_all_variants_give_identical_bits(lc, gpu, monkeypatch):
    """Every variant sums in the same order, so tuning never changes results."""
    torch = gpu
    rng = np.random.default_rng(1250)
    n = [40, 33, 37]
    bc = [0, -1, 1]
    nx = [n[0], n[1] + 28, n[2]]
    ny = [n[0], n[1], n[2] + 28]
    for dt in (torch.float64, torch.float32):
        x = torch.from_numpy(rng.uniform(-1, 1, int(np.prod(nx)))).to('cuda', dt)
        outs = []
        for v in range(lc.variant_count(lc.FORM_SUM)):
            monkeypatch.setenv('LIBCONV_SUM3D_VARIANT', str(v))
            y = torch.zeros(int(np.prod(ny)), dtype=dt, device='cuda')
            lc.conv_1ds(lc.FIL_D2, n, bc, nx, ny, 1, [-5.5, -4.0, -3.0], 0.25, 0.0, x, y)
            outs.append(y)
        torch.cuda.synchronize()
        for v, y in enumerate(outs):
            assert torch.equal(y, outs[0]), (str(dt), v)


def test_tune_md_records_fastest(lc, gpu):
    shape = dict(n=[64, 48, 40], bc=0, nx=[64, 48, 40], ny=[64, 48, 40])
    r = lc.tune_md(lc.FORM_SUM, lc.FIL_D2, repeat=2, **shape)
    assert 0 <= r['best'] < lc.variant_count(lc.FORM_SUM) and np.isfinite(r['times'][:16]).all()
    assert r['best_ms'] == min(r['times'])
    assert lc.kernel_name(8, lc.FORM_SUM, lc.FIL_D2, **shape) != ''
    w = lc.tune_md(lc.FORM_IDWT, lc.WAV_SYM8, repeat=2, prec=4, **shape)
    # FP32: every variant but the FP64-only plane-at-a-time shapes (16, 17)
    assert 0 <= w['best'] < lc.variant_count(lc.FORM_IDWT) and np.isfinite(w['times'][:16]).all()
    assert not np.isfinite(w['times'][16:]).any()


def test_batch_tune_keeps_outputs(lc, gpu):
    torch = gpu
    rng = np.random.default_rng(1400)
    items, ys = [], []
    for k in range(6):
        s = [int(v) for v in rng.integers(8, 40, 3)]
        bc = int(rng.choice([0, 1, -1]))
        nx = [v + 15 if bc == -1 else v for v in s]
        ny = [v + 15 if bc == 1 else v for v in s]
        x = torch.from_numpy(rng.uniform(-1, 1, int(np.prod(nx)))).cuda()
        y = torch.zeros(int(np.prod(ny)), dtype=torch.float64, device='cuda')
        items.append(dict(kind=lc.BATCH_SEP, op=lc.FIL_MD, n=s, bc=bc, nx=nx, ny=ny, x=x, y=y))
        wnx = [v + 7 if bc == -1 else v for v in s]
        wny = [v + 7 if bc == 1 else v for v in s]
        wx = torch.from_numpy(rng.uniform(-1, 1, 8 * int(np.prod(wnx)))).cuda()
        wy = torch.zeros(8 * int(np.prod(wny)), dtype=torch.float64, device='cuda')
        items.append(dict(kind=lc.BATCH_DWT, op=lc.WAV_SYM8, n=s, bc=bc, nx=wnx, ny=wny, x=wx, y=wy))
    items = items[0::2] + items[1::2]
    plan = lc.Batch(items)
    plan.run()
    torch.cuda.synchronize()
    before = [it['y'].clone() for it in items]
    ms = plan.tune(repeat=2)
    assert len(ms) == 2 and all(np.isfinite(ms))
    assert plan.stage_variant(0) not in (5, 6, 7) and 0 <= plan.stage_variant(1) < lc.variant_count(lc.FORM_DWT)
    for it in items:
        it['y'].fill_(3.0)
    plan.run_graph()
    torch.cuda.synchronize()
    for b, it in zip(before, items):
        assert torch.equal(b, it['y'])
