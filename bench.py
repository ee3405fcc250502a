#!/usr/bin/env python3
"""Benchmark of the libconv hot path on one MI355X.

Step = the 3D MagicFilter followed by the 3D Kinetic operator with its FP64
kinetic energy (y1 = MF3D(x), y2 = K3D(y1), ekin_d = <y1, a_d D2_d y1>),
a_d = -0.5/h^2 with h = 0.3, through the library's stream API on device-
resident data.  Default workload: 256^3 output grid, free BC in shrink mode
(x 299^3 -> y1 284^3 -> y2 256^3), FP64 and FP32 storage.  Synthetic input:
U[-1, 1] (seeded) plus a Gaussian bump of amplitude 1 and sigma = 10 points.

value = nominal output points (256^3) per second of the FP64 chain, in Gpts/s.
Algorithmic bytes per step = s (|x| + 2 |y1| + |y2|) (each operator reads its
input once and writes its output once).  Several buffer sets are rotated so
that the working set of consecutive steps exceeds the 256 MB Infinity Cache.
"""
import argparse
import contextlib
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'boast-lig_amd'))

HBM_PEAK_GBS = 8000.0
# HBM traffic per launch measured by rocprofv3 (FETCH_SIZE, WRITE_SIZE, scaled by
# a plain copy's readings) for the kernel each chain operator launches
TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'chain_256_shrink_hbm_traffic.txt')
MALL_BYTES = 256 << 20
SETTLE_MS = 60.0  # untimed GPU work before each timed region (settle_gpu)
H = 0.3


def workload(name):
    """(mf n, bc, x extent, y1 extent, kinetic n, y2 extent, nominal points)."""
    if name == 'c4':      # 256^3 output, free shrink
        n2 = 256
        return dict(bc=-1, nx=n2 + 15 + 28, n1=n2 + 28, ny1=n2 + 28, n2=n2, ny2=n2, pts=n2 ** 3,
                    desc='MF3D->K3D 256^3 output, free BC shrink (x 299^3, y1 284^3, y2 256^3)')
    if name == 'c2p':     # 128^3 periodic
        return dict(bc=0, nx=128, n1=128, ny1=128, n2=128, ny2=128, pts=128 ** 3,
                    desc='MF3D->K3D 128^3, periodic BC')
    if name == 'c2g':     # 128^3 free grow
        return dict(bc=1, nx=128, n1=128, ny1=143, n2=143, ny2=171, pts=128 ** 3,
                    desc='MF3D->K3D 128^3 input, free BC grow (y1 143^3, y2 171^3)')
    raise SystemExit('unknown workload ' + name)


def make_input(torch, w, dtype, seed=20250):
    n = w['nx']
    g = torch.Generator(device='cuda').manual_seed(seed)
    x = torch.rand((n, n, n), dtype=torch.float64, device='cuda', generator=g) * 2 - 1
    ax = torch.arange(n, dtype=torch.float64, device='cuda') - (n - 1) / 2
    r2 = ax.view(-1, 1, 1) ** 2 + ax.view(1, -1, 1) ** 2 + ax.view(1, 1, -1) ** 2
    x += torch.exp(-r2 / (2 * 10.0 ** 2))
    # storage: Fortran order (dim 0 fastest) == the flat buffer of this tensor
    return x.reshape(-1).to(dtype).contiguous()


def chain_args(w):
    bc = [w['bc']] * 3
    mf = dict(n=[w['n1'] if w['bc'] != 1 else w['nx']] * 3, bc=bc, nx=[w['nx']] * 3, ny=[w['ny1']] * 3)
    if w['bc'] == -1:
        mf['n'] = [w['ny1']] * 3
    kin = dict(n=[w['n2'] if w['bc'] != 1 else w['ny1']] * 3, bc=bc, nx=[w['ny1']] * 3, ny=[w['ny2']] * 3)
    if w['bc'] == -1:
        kin['n'] = [w['ny2']] * 3
    return mf, kin


@contextlib.contextmanager
def quiet_host():
    """Timed regions run without the cyclic garbage collector: a full
    collection of the interpreter's heap (tens of ms) inside a region stalls
    the launches and leaves the GPU idle."""
    gc.collect()
    gc.disable()
    try:
        yield
    finally:
        gc.enable()


def settle_gpu(torch, step, st, target_ms=SETTLE_MS, chunk=8, cap=2000):
    """Untimed work before the warm-up steps: `step` (called with a running
    index) until `target_ms` of GPU time has passed.  After an idle period the
    part first boosts, then throttles for ~10-30 ms (chain steps of 0.26 ms
    slow to 0.34 ms from about the 6th and recover after ~80), so a short
    timed region that follows idle host work measures the throttle, not the
    operator.  Returns the steps run and their GPU time."""
    done, ms = 0, 0.0
    while ms < target_ms and done < cap:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for i in range(chunk):
            step(done + i)
        e1.record(st)
        e1.synchronize()
        ms += e0.elapsed_time(e1)
        done += chunk
    return {'steps': done, 'ms': round(ms, 2)}


def run_gpu(torch, lc, w, dtype, steps, warmup):
    mf, kin = chain_args(w)
    a = [-0.5 / H ** 2] * 3
    s = 4 if dtype == torch.float32 else 8
    set_bytes = s * (w['nx'] ** 3 + w['ny1'] ** 3 + w['ny2'] ** 3)
    nsets = max(1, -(-3 * MALL_BYTES // set_bytes))
    x0 = make_input(torch, w, dtype)
    sets = []
    for k in range(nsets):
        x = x0 if k == 0 else x0.clone()
        y1 = torch.empty(w['ny1'] ** 3, dtype=dtype, device='cuda')
        y2 = torch.empty(w['ny2'] ** 3, dtype=dtype, device='cuda')
        dot = torch.zeros(3, dtype=torch.float64, device='cuda')
        sets.append((x, y1, y2, dot))
    st = torch.cuda.current_stream()

    def step(i, ev=None):
        x, y1, y2, dot = sets[i % nsets]
        if ev:
            ev[0].record(st)
        lc.conv_md(lc.FIL_MD, mf['n'], mf['bc'], mf['nx'], mf['ny'], 1, x, y1, stream=st)
        if ev:
            ev[1].record(st)
        lc.conv_1ds(lc.FIL_D2, kin['n'], kin['bc'], kin['nx'], kin['ny'], 1, a, 0.0, 0.0, y1, y2,
                    dot=dot, stream=st)
        if ev:
            ev[2].record(st)

    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    with quiet_host():
        # the first `steps` steps after start-up, before the clocks settle
        torch.cuda.synchronize()
        e0.record(st)
        for i in range(steps):
            step(i)
        e1.record(st)
        torch.cuda.synchronize()
        ms_cold = e0.elapsed_time(e1) / steps
        settle = settle_gpu(torch, step, st)
        for i in range(warmup):
            step(i)
        torch.cuda.synchronize()
        # whole-step timing (no per-op events inside)
        e0.record(st)
        for i in range(steps):
            step(i)
        e1.record(st)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        # per-operator timing (events between the operators, same stream)
        for i in range(steps):
            step(i, evs[i])
        torch.cuda.synchronize()
    t_mf = sum(e[0].elapsed_time(e[1]) for e in evs) / steps
    t_k = sum(e[1].elapsed_time(e[2]) for e in evs) / steps
    x, y1, y2, dot = sets[0]
    step(0)
    torch.cuda.synchronize()
    return dict(ms=ms, t_mf=t_mf, t_k=t_k, y1=y1.double().cpu().numpy(), y2=y2.double().cpu().numpy(),
                ekin=dot.cpu().numpy(), x=x.double().cpu().numpy(), nsets=nsets, ms_cold=ms_cold,
                settle=settle)


def run_dwt(torch, lc, dtype, steps, warmup, n=96):
    """3D sym8 DWT then IDWT on n^3 pairs ((2n)^3 values), periodic: the
    round trip of the wavelet configuration.  Returns per-op times and the
    round-trip error against the input."""
    g = torch.Generator(device='cuda').manual_seed(3)
    tot = (2 * n) ** 3
    s = torch.rand(tot, dtype=torch.float64, device='cuda', generator=g).mul_(2).sub_(1).to(dtype)
    c = torch.empty_like(s)
    back = torch.empty_like(s)
    nn, bc = [n] * 3, [0] * 3
    st = torch.cuda.current_stream()

    def step(ev=None):
        if ev:
            ev[0].record(st)
        lc.wavelet_md(lc.WAV_SYM8, False, nn, bc, nn, nn, 1, s, c, stream=st)
        if ev:
            ev[1].record(st)
        lc.wavelet_md(lc.WAV_SYM8, True, nn, bc, nn, nn, 1, c, back, stream=st)
        if ev:
            ev[2].record(st)

    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    with quiet_host():
        settle_gpu(torch, lambda i: step(), st)
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        for i in range(steps):
            step(evs[i])
        torch.cuda.synchronize()
    t_f = sum(e[0].elapsed_time(e[1]) for e in evs) / steps
    t_i = sum(e[1].elapsed_time(e[2]) for e in evs) / steps
    err = ((back.double() - s.double()).abs().max() / s.double().abs().max()).item()
    sz = 4 if dtype == torch.float32 else 8
    op_bytes = 2 * tot * sz
    return {'pairs': nn, 'values': tot, 'dwt_ms': round(t_f, 5), 'idwt_ms': round(t_i, 5),
            'GB_s': round(2 * op_bytes / ((t_f + t_i) * 1e-3) / 1e9, 1),
            'frac_hbm': round(2 * op_bytes / ((t_f + t_i) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            'algorithmic_bytes': 2 * op_bytes, 'round_trip_rel_err': err}


def run_hamiltonian(torch, lc, dtype, steps, warmup, n=64):
    """The local-Hamiltonian chain of one orbital (libconv.f90:238-295) on n^3
    pairs: IDWT grow, MF grow with the fused potential / density step, IMF
    shrink, free-BC Kinetic accumulate with energies, DWT shrink."""
    nn = [n] * 3
    ng, nf, nr = (2 * n) ** 3, (2 * (n + 7)) ** 3, (2 * (n + 7) + 15) ** 3
    g = torch.Generator(device='cuda').manual_seed(7)
    rnd = lambda k: torch.rand(k, dtype=torch.float64, device='cuda', generator=g).mul_(2).sub_(1).to(dtype)
    psig, pot = rnd(ng), rnd(nr)
    rho = torch.zeros(nr, dtype=dtype, device='cuda')
    hpsig = torch.empty_like(psig)
    work = torch.empty(lc.local_hamiltonian_work(nn), dtype=dtype, device='cuda')
    ekin = torch.zeros(3, dtype=torch.float64, device='cuda')
    a = [-0.5 / (H / 2) ** 2] * 3
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with quiet_host():
        settle_gpu(torch, lambda i: lc.local_hamiltonian(nn, a, psig, hpsig, pot, rho, work, ekin, stream=st), st)
        for _ in range(warmup):
            lc.local_hamiltonian(nn, a, psig, hpsig, pot, rho, work, ekin, stream=st)
        torch.cuda.synchronize()
        e0.record(st)
        for _ in range(steps):
            lc.local_hamiltonian(nn, a, psig, hpsig, pot, rho, work, ekin, stream=st)
        e1.record(st)
        torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    sz = 4 if dtype == torch.float32 else 8
    # per step: IDWT (ng + nf), MF (nf + nr) with pot read and rho r/w (3 nr),
    # IMF (nr + nf), Kinetic (nf in, hpsi read and written), DWT (nf + ng)
    byts = sz * (2 * ng + 7 * nf + 5 * nr)
    return {'pairs': nn, 'ms': round(ms, 5), 'GB_s': round(byts / (ms * 1e-3) / 1e9, 1),
            'frac_hbm': round(byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), 'algorithmic_bytes': byts,
            'fine_points_per_s_G': round(nf / (ms * 1e-3) / 1e9, 3)}


def run_poisson(torch, lc, dtype, steps, warmup, n=256, nord=16, geocode=3, op='div'):
    """op 'div': div_u_i with cc on an n^3 vector field, order 16, geocode 3
    (x, z periodic, y free): the configuration of BigDFT/Poisson_test2.rb.
    op 'rhopol': update_rhopol on the same grid (u, dlogeps and rhopol in,
    rhopol out, rhores2 summed on the device)."""
    nn = [n] * 3
    npts = n ** 3
    g = torch.Generator(device='cuda').manual_seed(9)
    h = [0.1, 0.1, 0.1]
    st = torch.cuda.current_stream()
    if op == 'div':
        u = torch.rand(3 * npts, dtype=torch.float64, device='cuda', generator=g).mul_(2).sub_(1).to(dtype)
        du = torch.empty(npts, dtype=dtype, device='cuda')
        cc = torch.empty(npts, dtype=dtype, device='cuda')

        def call():
            lc.div_u_i(geocode, nn, u, du, nord, h, cc, stream=st)
        per_point = 5  # u (3 components) in, du and cc out
    else:
        u = torch.rand(npts, dtype=torch.float64, device='cuda', generator=g).mul_(2).sub_(1).to(dtype)
        dl = torch.rand(3 * npts, dtype=torch.float64, device='cuda', generator=g).mul_(2).sub_(1).to(dtype)
        rp = torch.zeros(npts, dtype=dtype, device='cuda')
        r2 = torch.zeros(1, dtype=torch.float64, device='cuda')

        def call():
            # rhopol moves from step to step; the work of a step does not
            lc.update_rhopol(geocode, nn, u, nord, h, 0.5, dl, rp, r2, stream=st)
        per_point = 6  # u, dlogeps (3), rhopol in; rhopol out
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with quiet_host():
        settle_gpu(torch, lambda i: call(), st)
        for _ in range(warmup):
            call()
        torch.cuda.synchronize()
        e0.record(st)
        for _ in range(steps):
            call()
        e1.record(st)
        torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    sz = 4 if dtype == torch.float32 else 8
    byts = sz * per_point * npts
    return {'op': op, 'n': nn, 'nord': nord, 'geocode': geocode, 'ms': round(ms, 5),
            'GB_s': round(byts / (ms * 1e-3) / 1e9, 1),
            'frac_hbm': round(byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), 'algorithmic_bytes': byts,
            'Gpts_s': round(npts / (ms * 1e-3) / 1e9, 3)}


def batch_orbitals(count=512, seed=5):
    """Orbital sizes ~U{24..96} per dim and BC ~U{periodic, grow, shrink}."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return [([int(v) for v in rng.integers(24, 97, 3)], int(rng.choice([0, 1, -1])))
            for _ in range(count)]


def batch_items(torch, lc, orbs, dtype):
    """Per orbital: MF3D -> K3D (energies) on its grid, and one sym8 DWT of a
    (2 s)^3 array (the orbital's sizes in pairs)."""
    a = [-0.5 / H ** 2] * 3
    dots = torch.zeros((len(orbs), 3), dtype=torch.float64, device='cuda')
    mf, kin, dwt, pts = [], [], [], 0
    g = torch.Generator(device='cuda').manual_seed(55)
    for k, (s, bc) in enumerate(orbs):
        pts += s[0] * s[1] * s[2]
        if bc == 0:
            n_mf, xs, y1s, n_k, y2s = s, s, s, s, s
        elif bc == 1:
            y1s = [v + 15 for v in s]
            n_mf, xs, n_k, y2s = s, s, y1s, [v + 28 for v in y1s]
        else:
            y1s = [v + 28 for v in s]
            n_mf, xs, n_k, y2s = y1s, [v + 15 for v in y1s], s, s
        pr = lambda v: v[0] * v[1] * v[2]
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
    return mf + kin + dwt, pts, dots


def run_batch(torch, lc, dtype, reps, count=512):
    """512 orbitals: one graph replay of the batch plan vs one plain stream-API
    call per orbital and operator (same kernels, same stream)."""
    orbs = batch_orbitals(count)
    items, pts, dots = batch_items(torch, lc, orbs, dtype)
    prec = 4 if dtype == torch.float32 else 8
    plan = lc.Batch(items, prec=prec)
    # the stages' tiles and z chunking from the autotuner (lc_batch_tune,
    # BigDFT/GenericConvolution2.rb:744-837 in spirit): every variant of every
    # stage timed on the plan's own items, the fastest kept
    tuned_ms = plan.tune(repeat=3)
    variants = [plan.stage_variant(k) for k in range(len(tuned_ms))]
    byts = sum((it['x'].numel() + it['y'].numel()) * it['x'].element_size() for it in items)
    st = torch.cuda.current_stream()
    a = [-0.5 / H ** 2] * 3

    def singles():
        for it in items:
            bc = [it['bc']] * 3
            if it['kind'] == lc.BATCH_SEP:
                lc.conv_md(it['op'], it['n'], bc, it['nx'], it['ny'], 1, it['x'], it['y'], stream=st)
            elif it['kind'] == lc.BATCH_SUM:
                lc.conv_1ds(it['op'], it['n'], bc, it['nx'], it['ny'], 1, a, 0.0, 0.0, it['x'], it['y'],
                            dot=it['dot'], stream=st)
            else:
                lc.wavelet_md(it['op'], False, it['n'], bc, it['nx'], it['ny'], 1, it['x'], it['y'],
                              stream=st)

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with quiet_host():
            settle_gpu(torch, lambda i: fn(), st, chunk=1)
            fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            e0.record(st)
            for _ in range(reps):
                fn()
            e1.record(st)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
        return e0.elapsed_time(e1) / reps, (t1 - t0) * 1e3 / reps

    t_single, w_single = timed(singles)
    ref = [it['y'].clone() for it in items[::97]]
    ref_dots = dots.clone()
    t_graph, w_graph = timed(lambda: plan.run_graph(st))
    same = all(torch.equal(r, it['y']) for r, it in zip(ref, items[::97]))
    ddiff = ((dots - ref_dots).abs().max() / ref_dots.abs().max()).item()
    t_pers, w_pers = timed(lambda: plan.run_persistent_graph(st))
    pinfo = plan.persistent_info()
    same_p = all(torch.equal(r, it['y']) for r, it in zip(ref, items[::97]))
    info = plan.info()
    plan.close()
    return {'orbitals': count, 'nominal_points': pts, 'stages': info['stages'], 'blocks': info['blocks'],
            'graph_ms': round(t_graph, 4), 'per_orbital_ms': round(t_single, 4),
            'speedup': round(t_single / t_graph, 2),
            'persistent_graph_ms': round(t_pers, 4), 'persistent_speedup': round(t_single / t_pers, 2),
            'persistent_grid': pinfo['grid'], 'persistent_status': pinfo['status'],
            'persistent_outputs_equal_sampled': bool(same_p),
            'Gpts_s_graph': round(pts / (t_graph * 1e-3) / 1e9, 3),
            'algorithmic_bytes': byts, 'GB_s_graph': round(byts / (t_graph * 1e-3) / 1e9, 1),
            'frac_hbm_graph': round(byts / (t_graph * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            'tuned_stage_ms': [round(v, 4) for v in tuned_ms], 'tuned_stage_variants': variants,
            'Gpts_s_per_orbital': round(pts / (t_single * 1e-3) / 1e9, 3),
            'outputs_equal_sampled': bool(same), 'ekin_rel_diff': ddiff,
            'sizes': 'per dim U{24..96}, BC U{periodic, grow, shrink}, seed 5; each orbital: '
                     'MF3D -> K3D with energies, and a sym8 DWT of (2s)^3 values'}


# the sources that define the chain kernels' code: a counter summary taken
# from other sources describes other code
KERNEL_SOURCES = ['lc_common.hpp', 'lc_taps.h', 'lc_kernels.hpp', 'lc_fused.hpp', 'lc_body_conv.hpp',
                  'lc_body_sepq.hpp', 'lc_body_sum2.hpp', 'lc_fused_impl.hpp']


def kernel_source_digest():
    """SHA-256 (first 16 hex digits) of the chain kernels' sources and build flags."""
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES + ['../Makefile']:
        with open(os.path.join(ROOT, 'boast-lig_amd', 'csrc', name), 'rb') as f:
            h.update(name.encode() + b'\0' + f.read())
    return h.hexdigest()[:16]


def measured_traffic(kernel):
    """Bytes per launch of `kernel` from the committed counter summary, or None
    when the summary lacks the kernel or was taken from other kernel sources
    (its `source_digest` line differs from kernel_source_digest())."""
    import re
    if not kernel or not os.path.exists(TRAFFIC_FILE):
        return None
    digest, found = None, None
    for line in open(TRAFFIC_FILE):
        m = re.match(r'source_digest = (\w+)', line.strip())
        if m:
            digest = m.group(1)
        m = re.match(r'traffic\[(.*)\] = (\d+)', line.strip())
        if m and m.group(1) == kernel:
            found = int(m.group(2))
    return found if digest == kernel_source_digest() else None


def errs(a, b):
    import numpy as np
    d = a - b
    return dict(max_rel=float(np.abs(d).max() / np.abs(b).max()),
                rms_rel=float(np.sqrt((d * d).sum() / (b * b).sum())))


def oracle_chain(O, w, x):
    import numpy as np
    mf, kin = chain_args(w)
    y1 = np.zeros(w['ny1'] ** 3)
    O.s0s0_md(O.MD, mf['n'], mf['bc'], mf['nx'], mf['ny'], 1, x, y1)
    y2 = np.zeros(w['ny2'] ** 3)
    a = [-0.5 / H ** 2] * 3
    dot = O.s0s0_1ds(O.D2, kin['n'], kin['bc'], kin['nx'], kin['ny'], 1, a, 0.0, 0.0, y1, y2)
    return y1, y2, dot


def cpu_baseline(O, np, w, x):
    """The CPU baseline on the same workload and input: oracle/liblc_cpu.so,
    the reference's generated sweep structure restated in C++/OpenMP (three
    MagicFilter sweeps through work arrays, three Kinetic passes with the
    energies, lines unrolled along the contiguous dim, cache-staged strided
    lines), all host threads, work arrays allocated before timing.  One
    warm-up, then the min of 5 runs (BigDFT/MagicFilter_test.rb:69-79 keeps
    the min of repeated runs)."""
    mf, kin = chain_args(w)
    a = [-0.5 / H ** 2] * 3
    y1 = np.zeros(w['ny1'] ** 3)
    y2 = np.zeros(w['ny2'] ** 3)
    w1 = np.zeros(mf['ny'][0] * mf['nx'][1] * mf['nx'][2])
    w2 = np.zeros(mf['ny'][0] * mf['ny'][1] * mf['nx'][2])

    def chain():
        O.cpu_magicfilter3d(mf['n'], mf['bc'], mf['nx'], mf['ny'], x, y1, w1, w2)
        return O.cpu_kinetic3d(kin['n'], kin['bc'], kin['nx'], kin['ny'], a, y1, y2)

    chain()
    best = 1e30
    for _ in range(5):
        t0 = time.perf_counter()
        chain()
        best = min(best, time.perf_counter() - t0)
    model = ''
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                model = line.split(':', 1)[1].strip()
                break
    except OSError:
        pass
    return dict(value=w['pts'] / best / 1e9, unit='Gpts/s', cores=O.cpu_threads(), kind='port',
                sample='the full workload (%s), FP64, min of 5 after 1 warm-up; oracle/liblc_cpu.so '
                       '(the generated sweep structure in C++/OpenMP, x86-64-v3)' % w['desc'],
                seconds=round(best, 4), cpu=model, nproc=os.cpu_count(),
                omp_num_threads=os.environ.get('OMP_NUM_THREADS'))


def copy_ceiling(torch, mb=1024, reps=10):
    """Device copy of mb MiB buffers (rotated past the MALL): read + write
    bytes per second, the practical ceiling of a streaming kernel."""
    n = (mb << 20) // 8
    bufs = [(torch.ones(n, dtype=torch.float64, device='cuda'), torch.empty(n, dtype=torch.float64, device='cuda'))
            for _ in range(3)]
    st = torch.cuda.current_stream()
    for i in range(3):
        bufs[i][1].copy_(bufs[i][0])
    torch.cuda.synchronize()
    best = 1e30
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for i in range(reps):
            b = bufs[i % 3]
            b[1].copy_(b[0])
        e1.record(st)
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1) / reps)
    del bufs
    return round(2 * n * 8 / (best * 1e-3) / 1e9, 1)


def run_c1(torch, lc, reps=2000):
    """MagicFilter sweep along dim 0 of a 32^3 FP64 grid, periodic: kernel
    time per call (events around back-to-back calls) and the wall time of one
    call on the host (launch-latency bound)."""
    import numpy as np
    n = [32, 32, 32]
    g = torch.Generator(device='cuda').manual_seed(1)
    x = torch.rand(32 ** 3, dtype=torch.float64, device='cuda', generator=g) * 2 - 1
    y = torch.empty_like(x)
    st = torch.cuda.current_stream()
    for _ in range(20):
        lc.conv_1d(lc.FIL_MD, 0, n, 0, n, n, 1, x, y, stream=st)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with quiet_host():
        t0 = time.perf_counter()
        e0.record(st)
        for _ in range(reps):
            lc.conv_1d(lc.FIL_MD, 0, n, 0, n, n, 1, x, y, stream=st)
        e1.record(st)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / reps
    dev = e0.elapsed_time(e1) / reps
    byts = 2 * 32 ** 3 * 8
    out = {'workload': 'MF 1D sweep, dim 0, 32^3, periodic, FP64', 'device_us_per_call': round(dev * 1e3, 2),
           'host_us_per_call': round(wall * 1e6, 2), 'Gpts_s': round(32 ** 3 / (dev * 1e-3) / 1e9, 3),
           'GB_s': round(byts / (dev * 1e-3) / 1e9, 1), 'algorithmic_bytes': byts}
    # the same calls captured in a HIP graph and replayed: the sweep's own
    # latency, without the host's per-call cost (the sweep allocates nothing
    # and clears nothing, so it captures as it runs)
    per_graph = 100
    ref = y.clone()
    cs = torch.cuda.Stream()
    cs.wait_stream(st)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(cs):
        with torch.cuda.graph(graph, stream=cs):
            for _ in range(per_graph):
                lc.conv_1d(lc.FIL_MD, 0, n, 0, n, n, 1, x, y, stream=cs)
    st.wait_stream(cs)
    with quiet_host():
        for _ in range(3):
            y.zero_()
            graph.replay()
        torch.cuda.synchronize()
        same = bool(torch.equal(y, ref))
        e0.record(st)
        for _ in range(reps // per_graph):
            graph.replay()
        e1.record(st)
        torch.cuda.synchronize()
    gdev = e0.elapsed_time(e1) / ((reps // per_graph) * per_graph)
    out.update({'graph_device_us_per_call': round(gdev * 1e3, 2), 'graph_Gpts_s': round(32 ** 3 / (gdev * 1e-3) / 1e9, 3),
                'graph_GB_s': round(byts / (gdev * 1e-3) / 1e9, 1), 'graph_output_equal': same})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--workload', default='c4', choices=['c4', 'c2p', 'c2g'])
    ap.add_argument('--no-oracle', action='store_true', help='skip the FP64 oracle error check')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline')
    ap.add_argument('--fp64-only', action='store_true')
    ap.add_argument('--no-extra', action='store_true', help='skip the wavelet and batch lines')
    ap.add_argument('--copy-calib', action='store_true',
                    help='only run a 1 GiB device copy (calibrates the HBM byte counters)')
    ap.add_argument('--source-digest', action='store_true',
                    help='print the digest of the chain kernels\' sources (for the counter summary)')
    args = ap.parse_args()
    if args.source_digest:
        print(kernel_source_digest())
        return 0
    if args.copy_calib:
        import torch
        nb = 1 << 30
        src = torch.ones(nb // 8, dtype=torch.float64, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({'copy_calib_bytes_read': nb, 'copy_calib_bytes_written': nb, 'launches': 3}))
        return 0
    if args.gpus != 1:
        print('bench.py: measures one GPU; --gpus %d is not supported (orbitals would shard '
              'across GPUs without communication, a single grid by z slabs with halo planes)' % args.gpus)
        return 2
    import numpy as np
    import torch
    import libconv as lc
    assert torch.cuda.is_available(), 'bench.py needs a GPU'
    lc.lib()
    w = workload(args.workload)
    r64 = run_gpu(torch, lc, w, torch.float64, args.steps, args.warmup)
    r32 = None if args.fp64_only else run_gpu(torch, lc, w, torch.float32, args.steps, args.warmup)

    s = 8
    bx, by1, by2 = w['nx'] ** 3, w['ny1'] ** 3, w['ny2'] ** 3
    chain_bytes = s * (bx + 2 * by1 + by2)
    mf_bytes = s * (bx + by1)
    k_bytes = s * (by1 + by2)
    gpts = w['pts'] / (r64['ms'] * 1e-3) / 1e9
    # dominant operator of the step
    dom = 'kinetic3d' if r64['t_k'] >= r64['t_mf'] else 'magicfilter3d'
    dom_t = max(r64['t_k'], r64['t_mf'])
    dom_b = k_bytes if dom == 'kinetic3d' else mf_bytes
    achieved = dom_b / (dom_t * 1e-3) / 1e9
    mfa, kina = chain_args(w)
    if dom == 'kinetic3d':
        kname = lc.kernel_name(8, lc.FORM_SUM, lc.FIL_D2, kina['n'], kina['bc'], kina['nx'], kina['ny'])
    else:
        kname = lc.kernel_name(8, lc.FORM_SEP, lc.FIL_MD, mfa['n'], mfa['bc'], mfa['nx'], mfa['ny'])
    ceiling = copy_ceiling(torch)
    out = {
        'metric': 'Gpts/s & % of HBM roofline, 3D MagicFilter+Kinetic FP64 vs FP32 at 1e-6',
        'value': round(gpts, 4),
        'unit': 'Gpts/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(r64['ms'], 5),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic: U[-1,1] seeded + Gaussian bump (amplitude 1, sigma 10 points)',
        'config': {'workload': w['desc'], 'grid_out': [w['ny2']] * 3, 'h': H, 'bc': w['bc'],
                   'buffer_sets': r64['nsets']},
        'timing': {'settle': r64['settle'], 'ms_per_step_before_settle': round(r64['ms_cold'], 5),
                   'note': 'untimed settle phase (the same steps, %g ms of GPU time) before the warm-up '
                           'steps: after idle host work the part boosts, then throttles for ~10-30 ms; '
                           'ms_per_step_before_settle is the first timed-length run after start-up' % SETTLE_MS},
        'roofline': {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': measured_traffic(kname),
                     'kernel': kname or dom,
                     'avg_ms': round(dom_t, 5), 'algorithmic_bytes': dom_b,
                     'timed': ('HIP events around the operator call on its stream: the kernel, plus the '
                               '~5 us energy reduction kernel (sum_partials) for the Kinetic operator'),
                     'copy_ceiling_GBs': ceiling, 'frac_of_copy': round(achieved / ceiling, 4)},
        'chain': {'GB_s': round(chain_bytes / (r64['ms'] * 1e-3) / 1e9, 1),
                  'frac_hbm': round(chain_bytes / (r64['ms'] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                  'frac_of_copy': round(chain_bytes / (r64['ms'] * 1e-3) / 1e9 / ceiling, 4),
                  'algorithmic_bytes': chain_bytes, 'mf_ms': round(r64['t_mf'], 5),
                  'kin_ms': round(r64['t_k'], 5)},
    }
    if r32 is not None:
        g32 = w['pts'] / (r32['ms'] * 1e-3) / 1e9
        out['fp32'] = {'value': round(g32, 4), 'ms_per_step': round(r32['ms'], 5),
                       'ratio_vs_fp64': round(g32 / gpts, 3),
                       'chain_GB_s': round(chain_bytes / 2 / (r32['ms'] * 1e-3) / 1e9, 1),
                       'mf_ms': round(r32['t_mf'], 5), 'kin_ms': round(r32['t_k'], 5),
                       'ms_per_step_before_settle': round(r32['ms_cold'], 5)}
    if not args.no_extra:
        out['c1'] = run_c1(torch, lc)
        for wl in ('c2p', 'c2g'):
            wc = workload(wl)
            rc = run_gpu(torch, lc, wc, torch.float64, args.steps, args.warmup)
            cb = 8 * (wc['nx'] ** 3 + 2 * wc['ny1'] ** 3 + wc['ny2'] ** 3)
            line = {'workload': wc['desc'], 'ms': round(rc['ms'], 5),
                    'Gpts_s': round(wc['pts'] / (rc['ms'] * 1e-3) / 1e9, 3),
                    'GB_s': round(cb / (rc['ms'] * 1e-3) / 1e9, 1),
                    'frac_hbm': round(cb / (rc['ms'] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    'algorithmic_bytes': cb, 'mf_ms': round(rc['t_mf'], 5), 'kin_ms': round(rc['t_k'], 5)}
            if not args.no_oracle:
                import oracle as O
                y1r, y2r, dr = oracle_chain(O, wc, rc['x'])
                line['errors'] = {'y1': errs(rc['y1'], y1r), 'y2': errs(rc['y2'], y2r),
                                  'ekin_rel': float(abs(rc['ekin'].sum() - dr.sum()) / abs(dr.sum()))}
            out[wl] = line
        out['dwt3d'] = run_dwt(torch, lc, torch.float64, args.steps, args.warmup)
        out['dwt3d']['fp32'] = run_dwt(torch, lc, torch.float32, args.steps, args.warmup)
        out['hamiltonian'] = run_hamiltonian(torch, lc, torch.float64, args.steps, args.warmup)
        out['hamiltonian']['fp32'] = run_hamiltonian(torch, lc, torch.float32, args.steps, args.warmup)
        out['poisson_div'] = run_poisson(torch, lc, torch.float64, args.steps, args.warmup)
        out['poisson_div']['fp32'] = run_poisson(torch, lc, torch.float32, args.steps, args.warmup)
        out['poisson_rhopol'] = run_poisson(torch, lc, torch.float64, args.steps, args.warmup, op='rhopol')
        out['batch'] = run_batch(torch, lc, torch.float64, reps=3)
        out['batch']['fp32'] = run_batch(torch, lc, torch.float32, reps=3)
    if not args.no_oracle:
        import oracle as O
        y1r, y2r, dr = oracle_chain(O, w, r64['x'])
        out['errors_fp64'] = {'y1': errs(r64['y1'], y1r), 'y2': errs(r64['y2'], y2r),
                              'ekin_rel': float(abs(r64['ekin'].sum() - dr.sum()) / abs(dr.sum()))}
        if r32 is not None:
            # the FP32 run starts from the FP32-rounded input: compare with the
            # FP64 oracle on the same (rounded) input
            y1s, y2s, ds = oracle_chain(O, w, r32['x'])
            out['fp32']['errors'] = {'y1': errs(r32['y1'], y1s), 'y2': errs(r32['y2'], y2s),
                                     'ekin_rel': float(abs(r32['ekin'].sum() - ds.sum()) / abs(ds.sum()))}
            # and against the oracle on the FP64 input: adds the input's quantisation
            out['fp32']['errors_vs_fp64_input'] = {
                'y1': errs(r32['y1'], y1r), 'y2': errs(r32['y2'], y2r),
                'ekin_rel': float(abs(r32['ekin'].sum() - dr.sum()) / abs(dr.sum()))}
    if not args.no_cpu:
        import oracle as O
        out['cpu_baseline'] = cpu_baseline(O, np, w, r64['x'])
    print(json.dumps(out))
    return 0


if __name__ == '__main__':
    sys.exit(main())
