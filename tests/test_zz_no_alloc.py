"""The allocation-free Kinetic entry point (lc_conv_1ds_noalloc) against the
pooled one (lc_conv_1ds), through the C ABI.

Both launch the same kernels and sum the energy partials in the same fixed
order, so outputs and energies must be bit-identical.  Two calls in a row,
the second with a smaller input, check that the static scratch carries
nothing from one call into the next; FP64 and FP32 storage; and a shape whose
partials do not fit is refused with y untouched.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def lc(gpu):
    import libconv
    libconv.lib()
    return libconv


def _case(torch, rng, n, bc, dtype):
    nx = [v + 28 if b == -1 else v for v, b in zip(n, bc)]
    ny = [v + 28 if b == 1 else v for v, b in zip(n, bc)]
    x = torch.from_numpy(rng.uniform(-1, 1, int(np.prod(nx)))).to('cuda', dtype)
    y0 = torch.from_numpy(rng.uniform(-1, 1, int(np.prod(ny)))).to('cuda', dtype)
    return nx, ny, x, y0


@pytest.mark.parametrize('dtype_name', ['float64', 'float32'])
def test_noalloc_equals_pooled_entry_point(lc, gpu, dtype_name):
    torch = gpu
    dtype = getattr(torch, dtype_name)
    rng = np.random.default_rng(4242)
    a = [-0.5 / 0.3 ** 2, -0.5 / 0.25 ** 2, -0.5 / 0.35 ** 2]
    # second call smaller than the first
    for n, bc, a_x, a_y in (([96, 88, 104], [-1, -1, -1], 0.0, 0.0), ([40, 33, 29], [1, 0, -1], 0.5, 0.25)):
        nx, ny, x, y0 = _case(torch, rng, n, bc, dtype)
        ya, yb = y0.clone(), y0.clone()
        da = torch.full((3,), float('nan'), dtype=torch.float64, device='cuda')
        db = torch.full((3,), float('nan'), dtype=torch.float64, device='cuda')
        lc.conv_1ds(2, n, bc, nx, ny, 1, a, a_x, a_y, x, ya, dot=da)
        lc.conv_1ds_noalloc(2, n, bc, nx, ny, 1, a, a_x, a_y, x, yb, dot=db)
        torch.cuda.synchronize()
        assert torch.equal(ya, yb), (n, bc)
        assert torch.equal(da, db), (n, bc, da, db)
        # without energies (the static scratch is not needed)
        yc = y0.clone()
        lc.conv_1ds_noalloc(2, n, bc, nx, ny, 1, a, a_x, a_y, x, yc)
        torch.cuda.synchronize()
        assert torch.equal(ya, yc), (n, bc)


def test_noalloc_refuses_shapes_beyond_its_scratch(lc, gpu):
    torch = gpu
    n = [480, 480, 480]
    x = torch.zeros(480 ** 3, dtype=torch.float32, device='cuda')
    y = torch.full((480 ** 3,), 7.0, dtype=torch.float32, device='cuda')
    dot = torch.zeros(3, dtype=torch.float64, device='cuda')
    with pytest.raises(lc.LibconvError):
        lc.conv_1ds_noalloc(2, n, [0] * 3, n, n, 1, [-1.0] * 3, 0.0, 0.0, x, y, dot=dot)
    torch.cuda.synchronize()
    assert bool((y == 7.0).all())
