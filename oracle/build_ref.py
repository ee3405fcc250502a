#!/usr/bin/env python3
"""Compile the reference's own FP64 oracles into oracle/_ref/ (test use only).

The reference checkout is read-only and never copied into this repository:
this recipe reads the sources where they lie, writes any text it has to
extract (Fortran held in Ruby heredocs or in notebook outputs) to a
temporary directory outside the repository, and leaves only the compiled
shared objects in oracle/_ref/ (git-ignored).  On a machine without the
reference checkout it does nothing and the tests that need oracle/_ref skip.

Built objects (Fortran symbols, gfortran/flang convention: lower case with
a trailing underscore, every argument by reference):
  libref_mf.so    magicfilter_{per,free}[_t]_ref      BigDFT/magicfilter_refs.f90
  libref_kin.so   kinetic_per_ref                     BigDFT/Kinetic.rb:981-1061
                  kinetic_per_ref_optim_ekin (+helpers) BigDFT/Kinetic.rb:21-372
  libref_idwt.so  d_s1s0_1d_sym8 and its 72 sweeps    BigDFT/Wavelet1d-nb.ipynb,
                  (the BOAST-generated FP64 IDWT family, cell 4 output)
  libref_wav.so   synthesis_{free,per}_ref,           BigDFT/old/Synthesis.rb:6-147
                  analysis_{free,per}_ref             BigDFT/old/AnaRotPer-2.rb:6-251
  libref_poisson.so generate_filter, div_u_i_ref,     BigDFT/Poisson2.rb (heredocs)
                  div_u_i_opt, nabla_u_ref, nabla_u_square_ref,
                  nabla_u_epsilon_ref, update_rhopol_ref
                  (nabla_u_and_square_ref needs BigDFT's f_err_throw:
                  not built)
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile

FC = os.environ.get('LC_REF_FC', 'amdflang')


def heredoc_after(path, marker):
    """Text of the first <<EOF heredoc following the line containing marker."""
    with open(path) as f:
        lines = f.read().split('\n')
    start = None
    for i, line in enumerate(lines):
        if marker in line:
            start = i
            break
    if start is None:
        raise RuntimeError('%s: marker %r not found' % (path, marker))
    for i in range(start, len(lines)):
        if lines[i].rstrip().endswith('<<EOF'):
            body = []
            for j in range(i + 1, len(lines)):
                if lines[j].strip() == 'EOF':
                    return '\n'.join(body) + '\n'
                body.append(lines[j])
    raise RuntimeError('%s: no heredoc after %r' % (path, marker))


def notebook_result(path, cell):
    nb = json.load(open(path))
    c = nb['cells'][cell]
    for o in c.get('outputs', []):
        if o.get('output_type') == 'execute_result':
            return ''.join(o['data']['text/plain'])
    raise RuntimeError('%s: cell %d has no execute_result' % (path, cell))


FORCE = False  # --force: rebuild libraries that already exist


def compile_so(sources, out, tmp):
    if not FORCE and os.path.exists(out) and os.path.getmtime(out) >= os.path.getmtime(__file__):
        return  # built from the same recipe: the reference checkout is read-only
    paths = []
    for name, text in sources:
        if text is None:
            paths.append(name)
        else:
            p = os.path.join(tmp, name)
            with open(p, 'w') as f:
                f.write(text)
            paths.append(p)
    cmd = [FC, '-O2', '-fPIC', '-shared', '-module-dir', tmp, '-o', out] + paths
    subprocess.run(cmd, check=True, cwd=tmp)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--ref', default='/root/reference')
    ap.add_argument('--out', default=os.path.join(os.path.dirname(os.path.abspath(__file__)), '_ref'))
    ap.add_argument('--force', action='store_true', help='rebuild every library')
    args = ap.parse_args()
    global FORCE
    FORCE = args.force
    bd = os.path.join(args.ref, 'BigDFT')
    if not os.path.isdir(bd):
        print('build_ref: no reference checkout at %s; skipping' % args.ref)
        return 0
    if shutil.which(FC) is None:
        print('build_ref: %s not found; skipping' % FC)
        return 0
    os.makedirs(args.out, exist_ok=True)
    tmp = tempfile.mkdtemp(prefix='lc_ref_')
    try:
        compile_so([(os.path.join(bd, 'magicfilter_refs.f90'), None)],
                   os.path.join(args.out, 'libref_mf.so'), tmp)
        kin = os.path.join(bd, 'Kinetic.rb')
        compile_so([('kin_ref.f90', heredoc_after(kin, 'function_name = "kinetic_per_ref"')),
                    ('kin_ekin.f90', heredoc_after(kin, 'function_name = "kinetic_per_ref_optim_ekin"'))],
                   os.path.join(args.out, 'libref_kin.so'), tmp)
        compile_so([('idwt_gen.f90', notebook_result(os.path.join(bd, 'Wavelet1d-nb.ipynb'), 4))],
                   os.path.join(args.out, 'libref_idwt.so'), tmp)
        syn = os.path.join(bd, 'old', 'Synthesis.rb')
        ana = os.path.join(bd, 'old', 'AnaRotPer-2.rb')
        compile_so([('synth_f.f90', heredoc_after(syn, 'function_name = "synthesis_free_ref"')),
                    ('synth_p.f90', heredoc_after(syn, 'function_name = "synthesis_per_ref"')),
                    ('anal_f.f90', heredoc_after(ana, 'function_name = "analysis_free_ref"')),
                    ('anal_p.f90', heredoc_after(ana, 'function_name = "analysis_per_ref"'))],
                   os.path.join(args.out, 'libref_wav.so'), tmp)
        poi = os.path.join(bd, 'Poisson2.rb')
        # nabla_u_and_square_ref calls BigDFT's f_err_throw, which the
        # reference does not carry: that routine is not built here
        compile_so([('poi_%s.f90' % f, heredoc_after(poi, 'function_name = "%s"' % f))
                    for f in ('generate_filter', 'div_u_i_ref', 'div_u_i_opt', 'nabla_u_ref',
                              'nabla_u_square_ref', 'nabla_u_epsilon_ref', 'update_rhopol_ref')],
                   os.path.join(args.out, 'libref_poisson.so'), tmp)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    print('build_ref: built', sorted(os.listdir(args.out)))
    return 0


if __name__ == '__main__':
    sys.exit(main())
