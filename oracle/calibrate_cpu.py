#!/usr/bin/env python3
"""Calibrate the CPU baseline against the reference's own CPU code.

Runs only where the reference checkout is (this build container): compiles
the reference's FP64 Fortran with amdflang -O3 -fopenmp from the sources where
they lie (the same heredoc extraction as build_ref.py, into a temporary
directory; the object goes to oracle/_ref/libref_cal.so, git-ignored) and
times, at the same sizes and thread counts, min of 5 after one warm-up
(BigDFT/MagicFilter_test.rb:69-79):

  * 3D Kinetic, periodic, 256^3: kinetic_per_ref_optim (BigDFT/Kinetic.rb
    :395-958, the reference's optimised OpenMP routine) against
    oracle/liblc_cpu.so lcc_kinetic3d (energies off), all host threads;
  * 3D MagicFilter, periodic, 256^3: three transposing magicfilter_per_ref
    calls (BigDFT/magicfilter_refs.f90:92-133, serial) against
    lcc_magicfilter3d on one thread.

Each side runs in a process of its own: the reference's objects bring
LLVM's OpenMP runtime, whose spinning workers would otherwise compete with
the port's GNU OpenMP threads.  Prints one JSON line per case with both
times and their ratio; the CPU baseline is not weakened when the ratio
(port / reference) is <= 1.
"""
import ctypes
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)
import build_ref  # noqa: E402

REF = os.environ.get('LC_REF', '/root/reference')
OUT = os.path.join(HERE, '_ref', 'libref_cal.so')


def build():
    bd = os.path.join(REF, 'BigDFT')
    if not os.path.isdir(bd):
        return False
    tmp = tempfile.mkdtemp(prefix='lc_cal_')
    kin = build_ref.heredoc_after(os.path.join(bd, 'Kinetic.rb'), 'function_name = "kinetic_per_ref_optim"')
    src = os.path.join(tmp, 'kin_optim.f90')
    with open(src, 'w') as f:
        f.write(kin)
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    subprocess.run([build_ref.FC, '-O3', '-fopenmp', '-fPIC', '-shared', '-module-dir', tmp, '-o', OUT,
                    src, os.path.join(bd, 'magicfilter_refs.f90')], check=True, cwd=tmp)
    return True


CHILD = r'''
import ctypes, json, os, sys, time
import numpy as np
sys.path.insert(0, %(root)r)
case, n, side = sys.argv[1], int(sys.argv[2]), sys.argv[3]
ip = lambda v: ctypes.byref(ctypes.c_int(v))
dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
rng = np.random.default_rng(5)
x = np.asfortranarray(rng.uniform(-1, 1, n ** 3))
def best(f):
    f(); t = 1e9
    for _ in range(5):
        t0 = time.perf_counter(); f(); t = min(t, time.perf_counter() - t0)
    return t
y = np.zeros(n ** 3)
if side == 'ref':
    L = ctypes.CDLL(%(lib)r)
    if case == 'kinetic':
        h = np.array([0.3, 0.3, 0.3]); c = ctypes.c_double(0.0)
        t = best(lambda: L.kinetic_per_ref_optim_(ip(n - 1), ip(n - 1), ip(n - 1), dp(h), dp(x), dp(y), ctypes.byref(c)))
    else:
        w = np.zeros(n ** 3)
        def ref3():
            L.magicfilter_per_ref_(ip(n), ip(n * n), dp(x), dp(y))
            L.magicfilter_per_ref_(ip(n), ip(n * n), dp(y), dp(w))
            L.magicfilter_per_ref_(ip(n), ip(n * n), dp(w), dp(y))
        t = best(ref3)
else:
    import oracle as O
    if case == 'kinetic':
        nn = (ctypes.c_int * 3)(n, n, n); bc = (ctypes.c_int * 3)(0, 0, 0)
        av = (ctypes.c_double * 3)(*([-0.5 / 0.09] * 3))
        C = O.cpu()
        t = best(lambda: C.lcc_kinetic3d(nn, bc, nn, nn, av, dp(x), dp(y), None))
    else:
        w1 = np.zeros(n ** 3); w2 = np.zeros(n ** 3)
        t = best(lambda: O.cpu_magicfilter3d([n] * 3, [0] * 3, [n] * 3, [n] * 3, x, y, w1, w2))
np.save(sys.argv[4], y)
print(json.dumps({'t': t}))
'''


def run_side(code, case, n, side, env, tmp):
    out = os.path.join(tmp, '%s_%s.npy' % (case, side))
    r = subprocess.run([sys.executable, '-c', code, case, str(n), side, out], check=True, env=env,
                       capture_output=True, text=True)
    import numpy as np
    return json.loads(r.stdout.strip().splitlines()[-1])['t'], np.load(out)


def main():
    if not build():
        print('calibrate_cpu: no reference checkout at %s' % REF)
        return 0
    import numpy as np
    code = CHILD % {'root': os.path.dirname(HERE), 'lib': OUT}
    nthreads = str(len(os.sched_getaffinity(0)))
    tmp = tempfile.mkdtemp(prefix='lc_cal_run_')
    for case, n, threads in (('kinetic', 256, nthreads), ('magicfilter', 256, '1')):
        env = dict(os.environ, OMP_NUM_THREADS=threads, OMP_PROC_BIND='close', OMP_PLACES='cores')
        tr, yr = run_side(code, case, n, 'ref', env, tmp)
        tp, yp = run_side(code, case, n, 'port', env, tmp)
        # the reference's MF leaves its output transposed per sweep: after
        # three sweeps of a cube the layout is back in place
        err = float(np.abs(yr - yp).max() / np.abs(yr).max())
        print(json.dumps({'case': case, 'n': n, 'threads': int(threads), 'reference_s': round(tr, 4),
                          'port_s': round(tp, 4), 'port_over_reference': round(tp / tr, 3),
                          'max_rel_diff': err}))
    return 0


if __name__ == '__main__':
    sys.exit(main())
