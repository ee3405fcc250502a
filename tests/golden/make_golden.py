#!/usr/bin/env python3
"""Write the golden vectors of tests/golden/ from the reference's own oracles.

Runs only where oracle/_ref/ was built from the reference checkout
(oracle/build_ref.py).  Every fixture holds seeded inputs and the outputs
that the reference's FP64 code produced for them:

  mf.npz      magicfilter_{per,per_t,free,free_t}_ref   BigDFT/magicfilter_refs.f90
  kinetic.npz kinetic_per_ref (c = 0.5, h = (0.5, 0.6, 0.7)) and the
              per-direction energies of kinetic_per_ref_optim_ekin
                                                        BigDFT/Kinetic.rb:21-112,981-1061
  idwt.npz    the BOAST-generated d_s1s0_1d_sym8 (every idim x BC x a/a_y,
              padded leading dims, narr = 2)            BigDFT/Wavelet1d-nb.ipynb cell 4
  bigdft_wavelet.npz  BigDFT's own synthesis_per_ref / analysis_per_ref
                                                        BigDFT/old/Synthesis.rb,
                                                        BigDFT/old/AnaRotPer-2.rb
  poisson.npz generate_filter tables (nord 2..16), div_u_i_ref, div_u_i_opt
              (du, cc), nabla_u_ref                     BigDFT/Poisson2.rb
  poisson_dielectric.npz  nabla_u_square_ref, nabla_u_epsilon_ref,
              update_rhopol_ref (rhopol and rhores2)    BigDFT/Poisson2.rb

Usage: python3 tests/golden/make_golden.py  (from the repository root).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import ref as R  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
SYM8_PAIRS = 8


def mf():
    rng = np.random.default_rng(101)
    n1, ndat = 37, 11
    x = np.asfortranarray(rng.uniform(-1, 1, (n1, ndat)))
    xs = np.asfortranarray(rng.uniform(-1, 1, (n1 + 15, ndat)))
    np.savez_compressed(os.path.join(OUT, 'mf.npz'), n1=n1, ndat=ndat, x=x, x_shrink=xs,
                        per=R.magicfilter('per', n1, ndat, x),
                        per_t=R.magicfilter('per_t', n1, ndat, x),
                        free=R.magicfilter('free', n1, ndat, x),
                        free_t=R.magicfilter('free_t', n1, ndat, xs))


def kinetic():
    rng = np.random.default_rng(202)
    n = (12, 14, 13)
    h = (0.5, 0.6, 0.7)
    x = np.asfortranarray(rng.uniform(-1, 1, n))
    y = R.kinetic_per(n, h, x, 0.5)
    yk = np.zeros(n, order='F')
    kstr = R.kinetic_ekin(n, h, x, yk)
    np.savez_compressed(os.path.join(OUT, 'kinetic.npz'), n=np.array(n), h=np.array(h), c=0.5,
                        x=x, y=y, y_ekin=yk, kstrten=kstr)


def idwt():
    # the processed dim holds 9 pairs: the generated free-BC sweeps read out
    # of bounds below the filter width (8 pairs), so the fixtures stay above it
    rng = np.random.default_rng(303)
    cases = {}
    k = 0
    for idim in range(3):
        n = [2, 3, 2]
        n[idim] = 9
        for bc in (0, 1, -1):
            for a, a_y in ((1.0, 0.0), (1.3, -0.4)):
                nx, ny = list(n), list(n)
                w = SYM8_PAIRS - 1
                ein = n[idim] + (w if bc == -1 else 0)
                eout = n[idim] + (w if bc == 1 else 0)
                nx[idim], ny[idim] = ein + 1, eout + 1   # padded leading dims
                narr = 2
                nin = int(np.prod([2 * v for v in nx]))
                nout = int(np.prod([2 * v for v in ny]))
                x = np.asfortranarray(rng.uniform(-1, 1, nin * narr))
                y0 = np.asfortranarray(rng.uniform(-1, 1, nout * narr))
                y = y0.copy(order='F')
                R.idwt_1d(3, idim, n, bc, nx, ny, narr, x, y, a, a_y)
                cases['c%d' % k] = np.array([idim, bc, narr] + n + nx + ny, dtype=np.int32)
                cases['s%d' % k] = np.array([a, a_y])
                cases['x%d' % k] = x
                cases['y0_%d' % k] = y0
                cases['y%d' % k] = y
                k += 1
    np.savez_compressed(os.path.join(OUT, 'idwt.npz'), ncases=k, **cases)


def bigdft_wavelet():
    rng = np.random.default_rng(404)
    n, ndat = 10, 7
    x = np.asfortranarray(rng.uniform(-1, 1, (2 * n, ndat)))
    np.savez_compressed(os.path.join(OUT, 'bigdft_wavelet.npz'), n=n, ndat=ndat, x=x,
                        synthesis_per=R.bigdft_wavelet('synthesis_per', n, ndat, x),
                        analysis_per=R.bigdft_wavelet('analysis_per', n, ndat, x))


def poisson():
    """generate_filter's tables for every order, and div_u_i_ref, div_u_i_opt
    (du and cc) and nabla_u_ref on seeded vector fields, each with the
    reference's own table; for nord 8 also with the exact Lagrange table
    (the reference's c1D(0, -4) of nord 8 carries a sign typo)."""
    from oracle import poisson_taps
    rng = np.random.default_rng(505)
    small, big = (11, 12, 13), (17, 17, 17)
    us = np.asfortranarray(rng.uniform(-1, 1, small + (3,)))
    ub = np.asfortranarray(rng.uniform(-1, 1, big + (3,)))
    h = np.array([0.15, 0.2, 0.25])
    out = dict(h=h, u_small=us, u_big=ub, n_small=np.array(small), n_big=np.array(big))
    for nord in (2, 4, 6, 8, 16):
        out['table_%d' % nord] = R.poisson_filter(nord)
    cases = [(2, 1, 'ref'), (4, 3, 'ref'), (6, 2, 'ref'), (8, 1, 'ref'), (8, 1, 'lagrange'), (16, 3, 'ref')]
    for nord, geo, tab in cases:
        n = big if nord == 16 else small
        u = ub if nord == 16 else us
        table = R.poisson_filter(nord) if tab == 'ref' else poisson_taps(nord)
        key = '%d_%d_%s' % (nord, geo, tab)
        out['div_' + key] = R.div_u_i(geo, n, u, nord, h, table)
        du, cc = R.div_u_i(geo, n, u, nord, h, table, with_cc=True)
        out['divopt_' + key], out['cc_' + key] = du, cc
        out['nabla_' + key] = R.nabla_u(geo, n, np.asfortranarray(u[..., 0]), nord, h, table)
    out['cases'] = np.array(['%d_%d_%s' % c for c in cases])
    np.savez_compressed(os.path.join(OUT, 'poisson.npz'), **out)


def poisson_dielectric():
    """nabla_u_square_ref, nabla_u_epsilon_ref and update_rhopol_ref (eta
    0.5, as Poisson_test.rb runs it; rhopol starts from a seeded field, not
    zeros, so that its update is tested too) on seeded fields, each with the
    reference's own table, and for nord 8 with the exact Lagrange table."""
    from oracle import poisson_taps
    rng = np.random.default_rng(606)
    h = np.array([0.15, 0.2, 0.25])
    out = dict(h=h, eta=0.5)
    for key, n in (('small', (11, 12, 13)), ('big', (17, 17, 17))):
        out['n_' + key] = np.array(n)
        out['u_' + key] = np.asfortranarray(rng.uniform(-1, 1, n))
        out['eps_' + key] = np.asfortranarray(rng.uniform(0.5, 2.0, n))
        out['dlogeps_' + key] = np.asfortranarray(rng.uniform(-1, 1, (3,) + n))
        out['rhopol_' + key] = np.asfortranarray(rng.uniform(-1, 1, n))
    cases = [(2, 1, 'ref'), (4, 3, 'ref'), (6, 2, 'ref'), (8, 1, 'lagrange'), (16, 3, 'ref')]
    for nord, geo, tab in cases:
        key = 'big' if nord == 16 else 'small'
        n = tuple(int(v) for v in out['n_' + key])
        u = out['u_' + key]
        table = R.poisson_filter(nord) if tab == 'ref' else poisson_taps(nord)
        c = '%d_%d_%s' % (nord, geo, tab)
        out['square_' + c] = R.nabla_u_square(geo, n, u, nord, h, table)
        out['epsilon_' + c] = R.nabla_u_epsilon(geo, n, u, nord, h, out['eps_' + key], table)
        rp, r2 = R.update_rhopol(geo, n, u, nord, h, out['eta'], out['dlogeps_' + key],
                                 out['rhopol_' + key], table)
        out['rhopol_' + c], out['rhores2_' + c] = rp, r2
    out['cases'] = np.array(['%d_%d_%s' % c for c in cases])
    np.savez_compressed(os.path.join(OUT, 'poisson_dielectric.npz'), **out)


def main():
    for lib in ('libref_mf.so', 'libref_kin.so', 'libref_idwt.so', 'libref_wav.so', 'libref_poisson.so'):
        if not R.available(lib):
            print('make_golden: oracle/_ref/%s missing (run make -C oracle ref)' % lib)
            return 1
    mf()
    kinetic()
    idwt()
    bigdft_wavelet()
    poisson()
    poisson_dielectric()
    print('make_golden: wrote', sorted(f for f in os.listdir(OUT) if f.endswith('.npz')))
    return 0


if __name__ == '__main__':
    sys.exit(main())
