"""Fortran binding: the interface module include/libconv_amd.f90 compiles
with amdflang and its declarations link against libconv_amd.so (CPU); the
check program tests/fortran/lc_fortran_check.f90, built by the library's
Makefile, calls the generated C name, the libconv.f90 spelling and the 3D
Kinetic composer from Fortran on the GPU, and its output matches the
oracle (GPU)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, 'boast-lig_amd')
FC = shutil.which('amdflang') or ('/opt/rocm/bin/amdflang' if os.path.exists('/opt/rocm/bin/amdflang') else None)
PROG = os.path.join(PKG, 'lc_fortran_check')


@pytest.mark.skipif(FC is None, reason='amdflang not available')
def test_interface_module_compiles_and_links(tmp_path):
    subprocess.run([FC, '-O0', '-J', str(tmp_path), '-c', os.path.join(ROOT, 'include', 'libconv_amd.f90'),
                    '-o', str(tmp_path / 'm.o')], check=True)
    subprocess.run([FC, '-O0', '-J', str(tmp_path), os.path.join(ROOT, 'tests', 'fortran', 'lc_fortran_check.f90'),
                    str(tmp_path / 'm.o'), '-L', PKG, '-lconv_amd', '-o', str(tmp_path / 'prog')], check=True)
    assert (tmp_path / 'prog').exists()


def _sections(text):
    out, key = {}, None
    for line in text.splitlines():
        parts = line.split()
        if not parts:
            continue
        if parts[0] in ('same_bits', 'bad_bc_error'):
            out[parts[0]] = int(parts[1])
        elif parts[0] in ('mf', 'kin', 'dots'):
            key = parts[0]
            out[key] = []
        elif key:
            out[key].extend(float(v) for v in parts)
    return out


@pytest.mark.gpu
def test_fortran_program_matches_oracle(gpu):
    if not os.path.exists(PROG):
        pytest.skip('lc_fortran_check not built (no amdflang at build time)')
    r = subprocess.run([PROG], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-500:] + r.stderr[-500:]
    s = _sections(r.stdout)
    n = [12, 10, 8]
    i = np.arange(1, int(np.prod(n)) + 1, dtype=np.float64)
    x = np.sin(0.37 * i) + 0.25 * np.cos(1.3 * i)
    ref = np.zeros(x.size)
    O.s0s0_1d(O.MD, 1, n, 0, n, n, 1, x, ref)
    mf = np.asarray(s['mf'])
    assert s['same_bits'] == 1
    assert np.abs(mf - ref).max() <= 1e-13 * np.abs(ref).max()
    ny = [v + 28 for v in n]
    a = [-0.5 / 0.3 ** 2, -0.5 / 0.35 ** 2, -0.5 / 0.4 ** 2]
    rk = np.zeros(int(np.prod(ny)))
    rdot = O.s0s0_1ds(O.D2, n, [1, 1, 1], n, ny, 1, a, 0.0, 0.0, x, rk)
    kin = np.asarray(s['kin'])
    assert np.abs(kin - rk).max() <= 1e-13 * np.abs(rk).max()
    assert np.abs(np.asarray(s['dots']) - rdot).max() <= 1e-12 * np.abs(rdot).max()
    assert s['bad_bc_error'] != 0
