"""GPU parity of the Poisson nabla operators (BigDFT/Poisson2.rb) against the
FP64 oracle and the reference's own outputs (tests/golden/poisson.npz).

Bar: FP64 within 1e-13 relative max-norm of the oracle; FP32 within 1e-6 of
the FP64 oracle on the FP32-rounded input (measured on MI355X: <= 3.3e-7)."""
import ctypes
import os

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


@pytest.mark.parametrize('nord', [2, 4, 6, 8, 16])
@pytest.mark.parametrize('geocode', [1, 2, 3])
@pytest.mark.parametrize('dtype_name', ['float64', 'float32'])
def test_div_and_nabla_match_oracle(lc, gpu, nord, geocode, dtype_name):
    torch = gpu
    dtype = getattr(torch, dtype_name)
    rng = np.random.default_rng(1000 * nord + 10 * geocode)
    n = [nord + 5, nord + 3, nord + 8]
    np_ = int(np.prod(n))
    h = [0.15, 0.2, 0.25]
    u = rng.uniform(-1, 1, 3 * np_)
    if dtype_name == 'float32':
        u = u.astype(np.float32).astype(np.float64)
    tol = TOL64 if dtype_name == 'float64' else 1e-6
    du, cc = np.zeros(np_), np.zeros(np_)
    O.div_u_i(geocode, n, u, du, nord, h, cc)
    tu = torch.from_numpy(u).to('cuda', dtype)
    tdu, tcc = torch.empty(np_, dtype=dtype, device='cuda'), torch.empty(np_, dtype=dtype, device='cuda')
    lc.div_u_i(geocode, n, tu, tdu, nord, h, tcc)
    lc.div_u_i(geocode, n, tu, tdu, nord, h)              # without cc
    gr, g2 = np.zeros(3 * np_), np.zeros(np_)
    O.nabla_u_and_square(geocode, n, u[:np_].copy(), gr, g2, nord, h)
    tgr, tg2 = torch.empty(3 * np_, dtype=dtype, device='cuda'), torch.empty(np_, dtype=dtype, device='cuda')
    lc.nabla_u_and_square(geocode, n, tu[:np_], tgr, tg2, nord, h)
    tgr1 = torch.empty(3 * np_, dtype=dtype, device='cuda')
    lc.nabla_u(geocode, n, tu[:np_], tgr1, nord, h)
    torch.cuda.synchronize()
    assert rel(tdu.cpu().numpy(), du) < tol
    assert rel(tcc.cpu().numpy(), cc) < tol
    assert rel(tgr.cpu().numpy(), gr) < tol
    assert rel(tg2.cpu().numpy(), g2) < tol
    assert torch.equal(tgr, tgr1)


@pytest.mark.parametrize('nord', [2, 8, 16])
@pytest.mark.parametrize('idim', [0, 1, 2])
@pytest.mark.parametrize('bc', [0, -2])
def test_poisson_broker_drop_in(lc, nord, idim, bc):
    """Poisson_broker(nord, idim, nn, bc, u, du, a) on host arrays, and the
    stream form with a_y accumulation, against the oracle."""
    import torch
    rng = np.random.default_rng(2000 + nord + 7 * idim + bc)
    n = [nord + 6, nord + 4, nord + 3]
    u = rng.uniform(-1, 1, int(np.prod(n)))
    du = np.zeros(u.size)
    lc.lib().Poisson_broker(nord, idim, np.asarray(n, dtype=np.int32).ctypes.data_as(lc._IP), bc,
                            u.ctypes.data_as(lc._DP), du.ctypes.data_as(lc._DP), lc._D(2.5))
    lc.check()
    ref = np.zeros(u.size)
    O.poisson_1d(nord, idim, n, bc, u, ref, a=2.5)
    assert rel(du, ref) < TOL64
    y0 = rng.uniform(-1, 1, u.size)
    ty = torch.from_numpy(y0).cuda()
    lc.poisson_1d(nord, idim, n, bc, torch.from_numpy(u).cuda(), ty, a=-1.5, a_y=0.5)
    ref = y0.copy()
    O.poisson_1d(nord, idim, n, bc, u, ref, a=-1.5, a_y=0.5)
    torch.cuda.synchronize()
    assert rel(ty.cpu().numpy(), ref) < TOL64


def test_against_reference_outputs(lc, gpu):
    """The library's div_u_i / nabla_u on the fields of tests/golden/poisson.npz
    against the reference's outputs for the same fields (div_u_i_ref,
    div_u_i_opt, nabla_u_ref); nord 8 with the exact table."""
    torch = gpu
    g = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'poisson.npz'))
    h = [float(v) for v in g['h']]
    for case in g['cases']:
        nord, geo, tab = str(case).split('_')
        nord, geo = int(nord), int(geo)
        if nord == 8 and tab == 'ref':
            continue  # the reference's nord-8 table carries the typo
        key = 'big' if nord == 16 else 'small'
        n = [int(v) for v in g['n_' + key]]
        np_ = int(np.prod(n))
        u = g['u_' + key].reshape(-1, order='F').copy()
        tu = torch.from_numpy(u).cuda()
        tdu, tcc = torch.empty(np_, dtype=torch.float64, device='cuda'), torch.empty(np_, dtype=torch.float64,
                                                                                    device='cuda')
        lc.div_u_i(geo, n, tu, tdu, nord, h, tcc)
        tgr = torch.empty(3 * np_, dtype=torch.float64, device='cuda')
        lc.nabla_u(geo, n, tu[:np_], tgr, nord, h)
        torch.cuda.synchronize()
        assert rel(tdu.cpu().numpy(), g['div_' + str(case)].reshape(-1, order='F')) < TOL64, case
        assert rel(tcc.cpu().numpy(), g['cc_' + str(case)].reshape(-1, order='F')) < TOL64, case
        assert rel(tgr.cpu().numpy(), g['nabla_' + str(case)].reshape(-1, order='F')) < TOL64, case


def test_rejects_bad_orders_and_small_grids(lc, gpu):
    torch = gpu
    u = torch.zeros(3 * 10 ** 3, dtype=torch.float64, device='cuda')
    du = torch.zeros(10 ** 3, dtype=torch.float64, device='cuda')
    with pytest.raises(lc.LibconvError):
        lc.div_u_i(1, [10, 10, 10], u, du, 3, [1.0] * 3)       # order 3
    with pytest.raises(lc.LibconvError):
        lc.div_u_i(1, [10, 10, 10], u, du, 16, [1.0] * 3)      # grid 10 < 17
    with pytest.raises(lc.LibconvError):
        lc.poisson_1d(4, 0, [10, 10, 10], 1, u, du)            # bc 1


@pytest.mark.parametrize('nord', [4, 16])
@pytest.mark.parametrize('geocode', [1, 2, 3])
@pytest.mark.parametrize('dtype_name', ['float64', 'float32'])
def test_z_march_equals_point_kernels(lc, gpu, monkeypatch, nord, geocode, dtype_name):
    """The z-marching kernels (tiles over several blocks, partial edge tiles,
    several z chunks, border lanes and planes) give the bits of the one-point
    kernels (LIBCONV_POISSON_POINT) for div_u_i with and without cc and for
    nabla_u with du2."""
    torch = gpu
    dtype = getattr(torch, dtype_name)
    rng = np.random.default_rng(3000 + nord + geocode)
    n = [70, 37, 2 * nord + 45]
    np_ = int(np.prod(n))
    h = [0.15, 0.2, 0.25]
    tu = torch.from_numpy(rng.uniform(-1, 1, 3 * np_)).to('cuda', dtype)

    def run():
        out = [torch.empty(k * np_, dtype=dtype, device='cuda') for k in (1, 1, 1, 3, 1)]
        lc.div_u_i(geocode, n, tu, out[0], nord, h, out[1])
        lc.div_u_i(geocode, n, tu, out[2], nord, h)
        lc.nabla_u_and_square(geocode, n, tu[:np_], out[3], out[4], nord, h)
        torch.cuda.synchronize()
        return out

    zm = run()
    monkeypatch.setenv('LIBCONV_POISSON_POINT', '1')
    pt = run()
    for a, b in zip(zm, pt):
        assert torch.equal(a, b)


def _dielectric_inputs(rng, n, f32):
    np_ = int(np.prod(n))
    u = rng.uniform(-1, 1, np_)
    eps = rng.uniform(0.5, 2.0, np_)
    dl = rng.uniform(-1, 1, 3 * np_)
    rp = rng.uniform(-1, 1, np_)
    if f32:
        u, eps, dl, rp = (a.astype(np.float32).astype(np.float64) for a in (u, eps, dl, rp))
    return u, eps, dl, rp


def _run_dielectric(lc, torch, geocode, n, nord, h, eta, u, eps, dl, rp, dtype):
    """The three operators on the device; returns ddu, du_eps, rhopol (numpy
    FP64) and rhores2."""
    np_ = int(np.prod(n))
    tu = torch.from_numpy(u).to('cuda', dtype)
    tsq = torch.empty(np_, dtype=dtype, device='cuda')
    lc.nabla_u_square(geocode, n, tu, tsq, nord, h)
    tde = torch.empty(3 * np_, dtype=dtype, device='cuda')
    lc.nabla_u_epsilon(geocode, n, tu, tde, nord, h, torch.from_numpy(eps).to('cuda', dtype))
    trp = torch.from_numpy(rp).to('cuda', dtype)
    r2 = torch.full((1,), -1.0, dtype=torch.float64, device='cuda')
    lc.update_rhopol(geocode, n, tu, nord, h, eta, torch.from_numpy(dl).to('cuda', dtype), trp, r2)
    torch.cuda.synchronize()
    return (tsq.cpu().double().numpy(), tde.cpu().double().numpy(), trp.cpu().double().numpy(),
            float(r2.cpu()[0]))


@pytest.mark.parametrize('nord', [2, 4, 6, 8, 16])
@pytest.mark.parametrize('geocode', [1, 2, 3])
@pytest.mark.parametrize('dtype_name', ['float64', 'float32'])
def test_dielectric_operators_match_oracle(lc, gpu, nord, geocode, dtype_name):
    """nabla_u_square, nabla_u_epsilon and update_rhopol (rhopol in place and
    rhores2) against the oracle.  rhores2 is summed in another order than the
    oracle's sequential loop: FP64 within 1e-12 relative, FP32 (squares of
    FP32 residues, summed in FP64) within 1e-5."""
    torch = gpu
    f32 = dtype_name == 'float32'
    dtype = getattr(torch, dtype_name)
    rng = np.random.default_rng(4000 + 10 * nord + geocode)
    n = [nord + 7, nord + 4, nord + 9]
    np_ = int(np.prod(n))
    h = [0.15, 0.2, 0.25]
    eta = 0.7
    u, eps, dl, rp = _dielectric_inputs(rng, n, f32)
    sq, de, rpo = np.zeros(np_), np.zeros(3 * np_), rp.copy()
    O.nabla_u_square(geocode, n, u, sq, nord, h)
    O.nabla_u_epsilon(geocode, n, u, de, nord, h, eps)
    r2 = O.update_rhopol(geocode, n, u, nord, h, eta, dl, rpo)
    gsq, gde, grp, gr2 = _run_dielectric(lc, torch, geocode, n, nord, h, eta, u, eps, dl, rp, dtype)
    tol = 1e-6 if f32 else TOL64
    assert rel(gsq, sq) < tol
    assert rel(gde, de) < tol
    assert rel(grp, rpo) < tol
    assert abs(gr2 - r2) <= (1e-5 if f32 else 1e-12) * r2


def test_dielectric_against_reference_outputs(lc, gpu):
    """The library's three operators on the fields of
    tests/golden/poisson_dielectric.npz against nabla_u_square_ref,
    nabla_u_epsilon_ref and update_rhopol_ref's outputs for the same fields."""
    torch = gpu
    g = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'poisson_dielectric.npz'))
    h = [float(v) for v in g['h']]
    eta = float(g['eta'])
    f = lambda a: np.ascontiguousarray(a.reshape(-1, order='F'))  # noqa: E731
    for case in g['cases']:
        c = str(case)
        nord, geo = int(c.split('_')[0]), int(c.split('_')[1])
        key = 'big' if nord == 16 else 'small'
        n = [int(v) for v in g['n_' + key]]
        gsq, gde, grp, gr2 = _run_dielectric(lc, torch, geo, n, nord, h, eta, f(g['u_' + key]),
                                             f(g['eps_' + key]), f(g['dlogeps_' + key]), f(g['rhopol_' + key]),
                                             torch.float64)
        assert rel(gsq, f(g['square_' + c])) < TOL64, c
        assert rel(gde, f(g['epsilon_' + c])) < TOL64, c
        assert rel(grp, f(g['rhopol_' + c])) < TOL64, c
        assert abs(gr2 - float(g['rhores2_' + c])) <= 1e-12 * abs(gr2), c


@pytest.mark.parametrize('nord', [4, 16])
def test_dielectric_drop_in_names(lc, gpu, nord):
    """nabla_u_square, nabla_u_epsilon and update_rhopol under the generated
    procedure names, on host arrays (staged, synchronous; rhores2 on the
    host) and with device arrays, against the oracle."""
    torch = gpu
    rng = np.random.default_rng(4100 + nord)
    n = [nord + 6, nord + 5, nord + 3]
    np_ = int(np.prod(n))
    h = np.array([0.3, 0.2, 0.1])
    u, eps, dl, rp = _dielectric_inputs(rng, n, False)
    L, P = lc.lib(), lc._DP
    sq, de, rpo = np.zeros(np_), np.zeros(3 * np_), rp.copy()
    O.nabla_u_square(3, n, u, sq, nord, h)
    O.nabla_u_epsilon(3, n, u, de, nord, h, eps)
    r2 = O.update_rhopol(3, n, u, nord, h, 0.5, dl, rpo)
    hsq, hde, hrp, hr2 = np.zeros(np_), np.zeros(3 * np_), rp.copy(), np.zeros(1)
    L.nabla_u_square(3, n[0], n[1], n[2], u.ctypes.data_as(P), hsq.ctypes.data_as(P), nord, h.ctypes.data_as(P))
    lc.check()
    L.nabla_u_epsilon(3, n[0], n[1], n[2], u.ctypes.data_as(P), hde.ctypes.data_as(P), nord, h.ctypes.data_as(P),
                      eps.ctypes.data_as(P))
    lc.check()
    L.update_rhopol(3, n[0], n[1], n[2], u.ctypes.data_as(P), nord, h.ctypes.data_as(P), lc._D(0.5),
                    dl.ctypes.data_as(P), hrp.ctypes.data_as(P), hr2.ctypes.data_as(P))
    lc.check()
    assert rel(hsq, sq) < TOL64 and rel(hde, de) < TOL64 and rel(hrp, rpo) < TOL64
    assert abs(hr2[0] - r2) <= 1e-12 * r2
    # device arrays through the same names (rhores2 on the device)
    tu, tdl = torch.from_numpy(u).cuda(), torch.from_numpy(dl).cuda()
    trp, tr2 = torch.from_numpy(rp).cuda(), torch.zeros(1, dtype=torch.float64, device='cuda')
    L.update_rhopol(3, n[0], n[1], n[2], lc._ptr(tu), nord, h.ctypes.data_as(P), lc._D(0.5), lc._ptr(tdl),
                    lc._ptr(trp), ctypes.cast(lc._ptr(tr2), P))
    lc.check()
    torch.cuda.synchronize()
    assert np.array_equal(trp.cpu().numpy(), hrp) and float(tr2.cpu()[0]) == hr2[0]


@pytest.mark.parametrize('nord', [4, 16])
@pytest.mark.parametrize('geocode', [1, 2])
@pytest.mark.parametrize('dtype_name', ['float64', 'float32'])
def test_dielectric_z_march_equals_point_kernels(lc, gpu, monkeypatch, nord, geocode, dtype_name):
    """The z-marching and one-point forms of the three operators give the same
    bits (rhores2: the same value within rounding, the partial sums being
    laid out per block), and rhores2 repeats bit for bit from call to call."""
    torch = gpu
    dtype = getattr(torch, dtype_name)
    rng = np.random.default_rng(4200 + nord + geocode)
    n = [70, 37, 2 * nord + 45]
    h = [0.15, 0.2, 0.25]
    u, eps, dl, rp = _dielectric_inputs(rng, n, dtype_name == 'float32')
    zm = _run_dielectric(lc, torch, geocode, n, nord, h, 0.3, u, eps, dl, rp, dtype)
    again = _run_dielectric(lc, torch, geocode, n, nord, h, 0.3, u, eps, dl, rp, dtype)
    assert again[3] == zm[3]
    monkeypatch.setenv('LIBCONV_POISSON_POINT', '1')
    pt = _run_dielectric(lc, torch, geocode, n, nord, h, 0.3, u, eps, dl, rp, dtype)
    for a, b in zip(zm[:3], pt[:3]):
        assert np.array_equal(a, b)
    assert abs(zm[3] - pt[3]) <= 1e-12 * zm[3]


@pytest.mark.parametrize('nord', [2, 8, 16])
@pytest.mark.parametrize('geocode', [1, 2, 3])
def test_dielectric_smallest_grids(lc, gpu, nord, geocode):
    """The three operators on grids of nord + 1 points along every dim (the
    smallest div_u_i_ref accepts: on a free dim every point is a border
    point, and the z march covers a single plane) and on a mixed one."""
    torch = gpu
    for n in ([nord + 1] * 3, [nord + 1, nord + 2, nord + 3]):
        rng = np.random.default_rng(4400 + 10 * nord + geocode + n[2])
        h = [0.2, 0.3, 0.4]
        u, eps, dl, rp = _dielectric_inputs(rng, n, False)
        np_ = int(np.prod(n))
        sq, de, rpo = np.zeros(np_), np.zeros(3 * np_), rp.copy()
        O.nabla_u_square(geocode, n, u, sq, nord, h)
        O.nabla_u_epsilon(geocode, n, u, de, nord, h, eps)
        r2 = O.update_rhopol(geocode, n, u, nord, h, 0.9, dl, rpo)
        gsq, gde, grp, gr2 = _run_dielectric(lc, torch, geocode, n, nord, h, 0.9, u, eps, dl, rp, torch.float64)
        assert rel(gsq, sq) < TOL64 and rel(gde, de) < TOL64 and rel(grp, rpo) < TOL64, n
        assert abs(gr2 - r2) <= 1e-12 * r2, n


def test_dielectric_large_grid(lc, gpu):
    """A 200 x 180 x 160 grid (5.8 M points, many tiles and z chunks) in FP64
    against the oracle, nord 16, x and z periodic, y free."""
    torch = gpu
    rng = np.random.default_rng(4300)
    n = [200, 180, 160]
    h = [0.1, 0.12, 0.14]
    u, eps, dl, rp = _dielectric_inputs(rng, n, False)
    np_ = int(np.prod(n))
    sq, de, rpo = np.zeros(np_), np.zeros(3 * np_), rp.copy()
    O.nabla_u_square(3, n, u, sq, 16, h)
    O.nabla_u_epsilon(3, n, u, de, 16, h, eps)
    r2 = O.update_rhopol(3, n, u, 16, h, 0.5, dl, rpo)
    gsq, gde, grp, gr2 = _run_dielectric(lc, torch, 3, n, 16, h, 0.5, u, eps, dl, rp, torch.float64)
    assert rel(gsq, sq) < TOL64 and rel(gde, de) < TOL64 and rel(grp, rpo) < TOL64
    assert abs(gr2 - r2) <= 1e-12 * r2
