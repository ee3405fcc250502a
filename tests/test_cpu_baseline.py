"""The CPU baseline (oracle/liblc_cpu.so, the generated sweep structure in
C++/OpenMP) computes what the oracle computes: the 3D MagicFilter and the 3D
Kinetic sum form with its energies, under periodic, grow, shrink and mixed
BCs, n below the filter width included, with padded (ld) outputs for the
MagicFilter.  It is timed beside the GPU, so it must be right first."""
import numpy as np
import pytest

import oracle as O


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


@pytest.mark.parametrize('bc', [(0, 0, 0), (1, 1, 1), (-1, -1, -1), (0, 1, -1), (-1, 0, 1)])
def test_cpu_magicfilter3d_matches_oracle(bc):
    rng = np.random.default_rng(80 + sum(bc))
    for n in ([21, 18, 23], [40, 9, 33]):
        nx = [v + 15 if b == -1 else v for v, b in zip(n, bc)]
        ny = [(v + 15 if b == 1 else v) + (1 if i == 0 else 0) for i, (v, b) in enumerate(zip(n, bc))]
        x = rng.uniform(-1, 1, int(np.prod(nx)))
        ref = np.zeros(int(np.prod(ny)))
        O.s0s0_md(O.MD, n, list(bc), nx, ny, 1, x, ref)
        y = np.zeros_like(ref)
        w1 = np.zeros(ny[0] * nx[1] * nx[2])
        w2 = np.zeros(ny[0] * ny[1] * nx[2])
        O.cpu_magicfilter3d(n, list(bc), nx, ny, x, y, w1, w2)
        nout = [v + 15 if b == 1 else v for v, b in zip(n, bc)]
        box = (slice(0, nout[2]), slice(0, nout[1]), slice(0, nout[0]))
        shape = (ny[2], ny[1], ny[0])
        assert rel(y.reshape(shape)[box], ref.reshape(shape)[box]) < 1e-13, (n, bc)


@pytest.mark.parametrize('bc', [(0, 0, 0), (1, 1, 1), (-1, -1, -1), (0, 1, -1), (-1, 0, 1)])
def test_cpu_kinetic3d_matches_oracle(bc):
    rng = np.random.default_rng(90 + sum(bc))
    a = [-2.3, -0.9, -1.7]
    for n in ([17, 20, 15], [40, 9, 33]):
        nx = [v + 28 if b == -1 else v for v, b in zip(n, bc)]
        ny = [v + 28 if b == 1 else v for v, b in zip(n, bc)]
        x = rng.uniform(-1, 1, int(np.prod(nx)))
        ref = np.zeros(int(np.prod(ny)))
        rdot = O.s0s0_1ds(O.D2, n, list(bc), nx, ny, 1, a, 0.0, 0.0, x, ref)
        y = np.full_like(ref, np.nan)
        dot = O.cpu_kinetic3d(n, list(bc), nx, ny, a, x, y)
        assert rel(y, ref) < 1e-13, (n, bc)
        assert np.abs(dot - rdot).max() <= 1e-12 * np.abs(rdot).max(), (n, bc, dot, rdot)
