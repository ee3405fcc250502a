"""GPU parity at the working sizes of the benchmark configurations.

The reference's own tests run at the working size (BigDFT/Kinetic_test.rb:
55-104, BigDFT/MagicFilter_test.rb:60-87); these run the HIP library through
its stream API on the configurations it is measured on and compare with the
FP64 oracle on the same seeded inputs:

  C2  MF3D -> K3D on 128^3, periodic and free grow (y1 143^3, y2 171^3),
      U[-1, 1] plus a Gaussian bump (sigma 10 points), h = 0.3, FP64 energies;
  C4  MF3D -> K3D with a 256^3 output, free shrink (x 299^3, y1 284^3), FP64
      and FP32 storage;
  C5  512 orbitals, sizes U{24..96} per dim, BC U{periodic, grow, shrink}
      (seed 5): MF3D -> K3D with energies and a sym8 DWT of (2 s)^3 values
      each, through the graph replay and the persistent launch of one batch
      plan; one orbital per BC and the largest against the oracle, every
      output bit-equal to one stream-API call per item.

Bars: FP64 1e-13 relative max-norm on every grid, energies 1e-12; FP32 1e-6
relative max-norm on y1, y2 and the energies, both against the oracle run on
the FP32-rounded input (the arithmetic error) and on the FP64 input (adding
the input quantisation); RMS errors are reported in the failure messages.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-13
TOLE64 = 1e-12
TOL32 = 1e-6
H = 0.3


def errs(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    d = a - b
    return (float(np.abs(d).max() / max(np.abs(b).max(), 1e-300)),
            float(np.sqrt((d * d).sum() / max((b * b).sum(), 1e-300))))


@pytest.fixture(scope='module')
def lc(gpu):
    import libconv
    libconv.lib()
    return libconv


def bump_input(n, seed):
    """U[-1, 1] (seeded) plus exp(-r^2 / (2 sigma^2)), sigma = 10 points,
    centred; Fortran order flattened (dim 0 fastest; the cube is symmetric)."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, n ** 3)
    ax = np.arange(n, dtype=np.float64) - (n - 1) / 2
    r2 = ax[:, None, None] ** 2 + ax[None, :, None] ** 2 + ax[None, None, :] ** 2
    x += np.exp(-r2 / (2 * 10.0 ** 2)).reshape(-1)
    return x


def chain_shapes(nout, bc):
    """(MF n, x, y1, K n, y2) extents per dim of the chain K3D o MF3D."""
    if bc == 0:
        return nout, nout, nout, nout, nout
    if bc == 1:  # nout is the input grid
        return nout, nout, nout + 15, nout + 15, nout + 43
    return nout + 28, nout + 43, nout + 28, nout, nout  # shrink: nout is the output grid


def oracle_chain(x, nout, bc, a):
    n_mf, nx, ny1, n_k, ny2 = chain_shapes(nout, bc)
    y1 = np.zeros(ny1 ** 3)
    O.s0s0_md(O.MD, [n_mf] * 3, [bc] * 3, [nx] * 3, [ny1] * 3, 1, x, y1)
    y2 = np.zeros(ny2 ** 3)
    dot = O.s0s0_1ds(O.D2, [n_k] * 3, [bc] * 3, [ny1] * 3, [ny2] * 3, 1, a, 0.0, 0.0, y1, y2)
    return y1, y2, dot


def gpu_chain(torch, lc, x, nout, bc, a, dtype):
    n_mf, nx, ny1, n_k, ny2 = chain_shapes(nout, bc)
    xd = torch.from_numpy(x).to('cuda', dtype)
    y1 = torch.full((ny1 ** 3,), float('nan'), dtype=dtype, device='cuda')
    y2 = torch.full((ny2 ** 3,), float('nan'), dtype=dtype, device='cuda')
    dot = torch.zeros(3, dtype=torch.float64, device='cuda')
    lc.conv_md(lc.FIL_MD, [n_mf] * 3, [bc] * 3, [nx] * 3, [ny1] * 3, 1, xd, y1)
    lc.conv_1ds(lc.FIL_D2, [n_k] * 3, [bc] * 3, [ny1] * 3, [ny2] * 3, 1, a, 0.0, 0.0, y1, y2, dot=dot)
    torch.cuda.synchronize()
    return y1.double().cpu().numpy(), y2.double().cpu().numpy(), dot.cpu().numpy()


# ------------------------------------------------------------------ C2
@pytest.mark.parametrize('bc', [0, 1])
def test_config2_chain_128(lc, gpu, bc):
    torch = gpu
    a = [-0.5 / H ** 2] * 3
    x = bump_input(128, 2002)
    y1, y2, dot = gpu_chain(torch, lc, x, 128, bc, a, torch.float64)
    r1, r2, rd = oracle_chain(x, 128, bc, a)
    assert errs(y1, r1)[0] < TOL64, errs(y1, r1)
    assert errs(y2, r2)[0] < TOL64, errs(y2, r2)
    assert np.abs(dot - rd).max() <= TOLE64 * np.abs(rd).max(), (dot, rd)


# ------------------------------------------------------------------ C4
@pytest.fixture(scope='module')
def c4():
    """Input and FP64 oracle of the 256^3 free-shrink chain (shared by the
    FP64 and FP32 tests), plus the oracle on the FP32-rounded input."""
    a = [-0.5 / H ** 2] * 3
    x = bump_input(299, 2004)
    ref = oracle_chain(x, 256, -1, a)
    x32 = x.astype(np.float32).astype(np.float64)
    ref32 = oracle_chain(x32, 256, -1, a)
    return a, x, ref, ref32


def test_config4_chain_256_shrink_fp64(lc, gpu, c4):
    torch = gpu
    a, x, (r1, r2, rd), _ = c4
    y1, y2, dot = gpu_chain(torch, lc, x, 256, -1, a, torch.float64)
    assert errs(y1, r1)[0] < TOL64, errs(y1, r1)
    assert errs(y2, r2)[0] < TOL64, errs(y2, r2)
    assert np.abs(dot - rd).max() <= TOLE64 * np.abs(rd).max(), (dot, rd)
    e = abs(dot.sum() - rd.sum()) / abs(rd.sum())
    assert e < TOLE64, e


def test_config4_chain_256_shrink_fp32(lc, gpu, c4):
    torch = gpu
    a, x, ref, ref32 = c4
    y1, y2, dot = gpu_chain(torch, lc, x, 256, -1, a, torch.float32)
    for which, (r1, r2, rd) in (('rounded input', ref32), ('FP64 input', ref)):
        e1, e2 = errs(y1, r1), errs(y2, r2)
        ed = np.abs(dot - rd).max() / np.abs(rd).max()
        assert e1[0] < TOL32, (which, 'y1 max/rms', e1)
        assert e2[0] < TOL32, (which, 'y2 max/rms', e2)
        assert ed < TOL32, (which, 'ekin', dot, rd)


# ------------------------------------------------------------------ C5
def c5_orbitals(count=512, seed=5):
    rng = np.random.default_rng(seed)
    return [([int(v) for v in rng.integers(24, 97, 3)], int(rng.choice([0, 1, -1])))
            for _ in range(count)]


def c5_shapes(s, bc):
    """(MF n, x, y1, K n, y2) per dim of an orbital's chain, s its nominal grid."""
    if bc == 0:
        return s, s, s, s, s
    if bc == 1:
        y1 = [v + 15 for v in s]
        return s, s, y1, y1, [v + 28 for v in y1]
    y1 = [v + 28 for v in s]
    return y1, [v + 15 for v in y1], y1, s, s


def c5_items(torch, lc, orbs, dtype):
    a = [-0.5 / H ** 2] * 3
    g = torch.Generator(device='cuda').manual_seed(55)
    dots = torch.zeros((len(orbs), 3), dtype=torch.float64, device='cuda')
    pr = lambda v: v[0] * v[1] * v[2]
    mf, kin, dwt = [], [], []
    for k, (s, bc) in enumerate(orbs):
        n_mf, xs, y1s, n_k, y2s = c5_shapes(s, bc)
        x = torch.rand(pr(xs), dtype=torch.float64, device='cuda', generator=g).mul_(2).sub_(1).to(dtype)
        y1 = torch.empty(pr(y1s), dtype=dtype, device='cuda')
        y2 = torch.empty(pr(y2s), dtype=dtype, device='cuda')
        wi = [v + 7 if bc == -1 else v for v in s]
        wo = [v + 7 if bc == 1 else v for v in s]
        wx = torch.rand(8 * pr(wi), dtype=torch.float64, device='cuda', generator=g).mul_(2).sub_(1).to(dtype)
        wy = torch.empty(8 * pr(wo), dtype=dtype, device='cuda')
        mf.append(dict(kind=lc.BATCH_SEP, op=0, n=n_mf, bc=bc, nx=xs, ny=y1s, x=x, y=y1))
        kin.append(dict(kind=lc.BATCH_SUM, op=2, n=n_k, bc=bc, nx=y1s, ny=y2s, x=y1, y=y2, a=a,
                        dot=dots[k]))
        dwt.append(dict(kind=lc.BATCH_DWT, op=0, n=s, bc=bc, nx=wi, ny=wo, x=wx, y=wy))
    return mf, kin, dwt, dots


def c5_singles(lc, items):
    for it in items:
        bc = [it['bc']] * 3
        if it['kind'] == lc.BATCH_SEP:
            lc.conv_md(it['op'], it['n'], bc, it['nx'], it['ny'], 1, it['x'], it['y'])
        elif it['kind'] == lc.BATCH_SUM:
            lc.conv_1ds(it['op'], it['n'], bc, it['nx'], it['ny'], 1, it['a'], 0.0, 0.0, it['x'],
                        it['y'], dot=it['dot'])
        else:
            lc.wavelet_md(it['op'], False, it['n'], bc, it['nx'], it['ny'], 1, it['x'], it['y'])


@pytest.mark.parametrize('dtype_name', ['float64', 'float32'])
def test_config5_batch_512(lc, gpu, dtype_name):
    torch = gpu
    dtype = getattr(torch, dtype_name)
    orbs = c5_orbitals()
    mf, kin, dwt, dots = c5_items(torch, lc, orbs, dtype)
    items = mf + kin + dwt
    outs = [it['y'] for it in items]
    # reference: one stream-API call per item
    c5_singles(lc, items)
    torch.cuda.synchronize()
    single = [y.clone() for y in outs]
    single_dots = dots.clone()
    plan = lc.Batch(items)
    try:
        for run in ('run_graph', 'run_persistent_graph'):
            for y in outs:
                y.fill_(float('nan'))
            dots.zero_()
            getattr(plan, run)()
            torch.cuda.synchronize()
            if run == 'run_persistent_graph':
                assert plan.persistent_info()['status'] == 0
            bad = [k for k, (y, r) in enumerate(zip(outs, single)) if not torch.equal(y, r)]
            assert not bad, (run, bad[:10])
            de = ((dots - single_dots).abs().max() / single_dots.abs().max()).item()
            assert de < TOL64, (run, de)
    finally:
        plan.close()
    # one orbital per BC and the largest against the oracle
    a = kin[0]['a']
    pick = {}
    for k, (s, bc) in enumerate(orbs):
        pick.setdefault(bc, k)
    pick['largest'] = max(range(len(orbs)), key=lambda k: np.prod(orbs[k][0]))
    tol = TOL64 if dtype_name == 'float64' else TOL32
    tole = TOLE64 if dtype_name == 'float64' else TOL32
    for tag, k in pick.items():
        s, bc = orbs[k]
        n_mf, xs, y1s, n_k, y2s = c5_shapes(s, bc)
        x = mf[k]['x'].double().cpu().numpy()
        r1 = np.zeros(int(np.prod(y1s)))
        O.s0s0_md(O.MD, n_mf, [bc] * 3, xs, y1s, 1, x, r1)
        r2 = np.zeros(int(np.prod(y2s)))
        rd = O.s0s0_1ds(O.D2, n_k, [bc] * 3, y1s, y2s, 1, a, 0.0, 0.0, r1, r2)
        e1 = errs(mf[k]['y'].double().cpu().numpy(), r1)
        # the Kinetic step of the GPU chain reads the GPU's y1: compare it with
        # the oracle on that same input, so that each step is checked alone
        y1g = kin[k]['x'].double().cpu().numpy()
        r2g = np.zeros(r2.size)
        rdg = O.s0s0_1ds(O.D2, n_k, [bc] * 3, y1s, y2s, 1, a, 0.0, 0.0, y1g, r2g)
        e2 = errs(kin[k]['y'].double().cpu().numpy(), r2g)
        dg = dots[k].cpu().numpy()
        assert e1[0] < tol, (tag, s, bc, 'y1', e1)
        assert e2[0] < tol, (tag, s, bc, 'y2', e2)
        assert np.abs(dg - rdg).max() <= tole * np.abs(rdg).max(), (tag, dg, rdg)
        # the chain end to end against the oracle on the FP64 data
        assert errs(kin[k]['y'].double().cpu().numpy(), r2)[0] < tol, (tag, 'chain y2')
        assert np.abs(dg - rd).max() <= tole * np.abs(rd).max(), (tag, 'chain ekin')
        wi, wo = dwt[k]['nx'], dwt[k]['ny']
        rw = np.zeros(8 * int(np.prod(wo)))
        O.s0s1_md(O.SYM8, s, [bc] * 3, wi, wo, 1, dwt[k]['x'].double().cpu().numpy(), rw)
        ew = errs(dwt[k]['y'].double().cpu().numpy(), rw)
        assert ew[0] < tol, (tag, s, bc, 'dwt', ew)
