"""Benchmark of the TOP_GPU hot path on one MI355X.

Workload: the 128 x 256 x 128 cantilever (optCase 1, 4,194,304 voxels, 12,830,211 DOFs, three
multigrid levels).  One step = one elasticity PCG solve (tol 1e-3, cold start) with the
multigrid V-cycle preconditioner, on densities x = 0.12 + U(-0.05, 0.05) (seed 12345),
E = Emin + x^3 (1 - Emin).  Inputs are resident in HBM before the timed region.

The line printed on stdout is one JSON object.  `value` is the mixed-precision PCG solves/s;
the FP64 path is measured on the same GPU in the same run, and the relative compliance
difference between the two is reported.

    python bench.py [--steps K] [--warmup W] [--size 128 256 128] [--no-cpu]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

METRIC = 'PCG solves/sec + K·u HBM GB/s, mixed vs FP64 at 1e-6 rel compliance err'
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print('[bench]', *a, file=sys.stderr, flush=True)


def kbyu_alg_bytes(Nn, Ne, s_in, s_out, s_E):
    """Algorithmic bytes of one finest-level K*u: 3 Nn (s_in + s_out) + Ne s_E + 32 Ne."""
    return 3 * Nn * (s_in + s_out) + Ne * s_E + 32 * Ne


def densities(Ne, seed=12345):
    rng = np.random.default_rng(seed)
    return 0.12 + rng.uniform(-0.05, 0.05, Ne)


def cpu_baseline(nely, nelx, nelz, E, F, its_fp64, sample_iters):
    """Oracle (numpy restatement of the reference algorithm) on a bounded sample: full setup and
    assembly, then `sample_iters` PCG iterations; solves/s extrapolated with the GPU FP64 solve's
    iteration count."""
    from oracle import mesh as M
    from oracle import solver as S
    t0 = time.perf_counter()
    m = M.build_cuboid_model(nely, nelx, nelz, opt_case=1)
    sv = S.Solver(m)
    sv.assemble(E)
    t1 = time.perf_counter()
    log('cpu oracle setup+assembly %.1fs' % (t1 - t0))
    # the solve's fixed cost: normB, r = b, one V-cycle, one dot; then per iteration K*u + V-cycle
    ta = time.perf_counter()
    r = F.copy()
    z = sv.vcycle(r)
    tb = time.perf_counter()
    _, its, _ = S.pcg(sv.kbyu, sv.vcycle, F, 1e-30, sample_iters)
    tc = time.perf_counter()
    t_iter = ((tc - tb) - (tb - ta)) / sample_iters
    t_solve = (tb - ta) + its_fp64 * t_iter
    threads = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
    return {
        'value': 1.0 / t_solve, 'unit': 'solves/s', 'cores': threads, 'kind': 'port',
        'sample': ('numpy/scipy oracle (oracle/): setup + Galerkin assembly, then %d PCG iterations on the same '
                   '128x256x128 system; solve time = first V-cycle + %d iterations x %.2f s (GPU FP64 iteration '
                   'count); BLAS threaded, indexing single-threaded' % (sample_iters, its_fp64, t_iter)),
        'sec_per_iteration': t_iter, 'setup_s': t1 - t0,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--size', type=int, nargs=3, default=[128, 256, 128], metavar=('NELY', 'NELX', 'NELZ'))
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-iters', type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU only: the hot path has no multi-GPU mode in this build '
              '(requested --gpus %d)' % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from top_gpu_on_matlab_amd.model import cuboid_model
    from top_gpu_on_matlab_amd.solver import Solver

    nely, nelx, nelz = args.size
    Emin = 1e-6
    torch.cuda.init()
    t = time.perf_counter()
    m64 = cuboid_model(nely, nelx, nelz, 1, precision='fp64')
    mmx = cuboid_model(nely, nelx, nelz, 1, precision='mixed')
    Ne, Nn = m64.num_elements, m64.num_nodes
    log('models %.1fs: Ne=%d Nn=%d levels=%d' % (time.perf_counter() - t, Ne, Nn, m64.num_levels))
    x = densities(Ne)
    E = Emin + x ** 3 * (1 - Emin)
    F = m64.get_table(1, 'F')[:, 0].copy()
    dF = torch.from_numpy(F).cuda()
    dE = torch.from_numpy(E).cuda()
    solvers = {}
    for name, mm in (('fp64', m64), ('mixed', mmx)):
        t = time.perf_counter()
        s = Solver(mm)
        s.set_modulus(dE)
        s.assemble()
        torch.cuda.synchronize()
        solvers[name] = s
        log('%s solver ready (assembly %.2fs)' % (name, time.perf_counter() - t))

    res = {}
    for name in ('fp64', 'mixed'):
        s = solvers[name]
        for _ in range(args.warmup):
            s.pcg(dF)
        torch.cuda.synchronize()
        t = time.perf_counter()
        its_list = []
        for _ in range(args.steps):
            u, its, rn = s.pcg(dF)
            its_list.append(its)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        c = float(torch.dot(dF, u))
        res[name] = dict(solves_per_s=args.steps / dt, ms_per_solve=1e3 * dt / args.steps, pcg_its=its_list[-1],
                         resnorm=rn, compliance=c)
        log('%s: %.2f ms/solve, %d its, c=%.12e' % (name, 1e3 * dt / args.steps, its_list[-1], c))

    # K*u kernel durations (HIP events on the solver stream)
    ms64 = solvers['fp64'].time_kbyu(1, 'fp64', 50)
    ms32 = solvers['mixed'].time_kbyu(1, 'fp32', 50)
    b64 = kbyu_alg_bytes(Nn, Ne, 8, 8, 8)
    b32 = kbyu_alg_bytes(Nn, Ne, 4, 4, 4)
    gbs64, gbs32 = b64 / ms64 / 1e6, b32 / ms32 / 1e6
    tf64 = 1152 * Ne / ms64 / 1e9
    log('K*u fp64 %.1f us (%.0f GB/s alg, %.1f TFLOP/s); fp32 %.1f us (%.0f GB/s alg)'
        % (ms64 * 1e3, gbs64, tf64, ms32 * 1e3, gbs32))

    rel_err = abs(res['mixed']['compliance'] - res['fp64']['compliance']) / abs(res['fp64']['compliance'])
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(nely, nelx, nelz, E, F, res['fp64']['pcg_its'], args.cpu_iters)
        except Exception as e:  # the baseline is reported, never required for the GPU number
            log('cpu baseline failed: %r' % (e,))
    out = {
        'metric': METRIC,
        'value': res['mixed']['solves_per_s'],
        'unit': 'solves/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': res['mixed']['ms_per_solve'],
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32 V-cycle / f64 PCG',
        'data': 'synthetic',
        'config': {'workload': 'cantilever %dx%dx%d (nely x nelx x nelz), %d voxels, %d DOFs, one PCG solve '
                               '(tol 1e-3, cold start) per step' % (nely, nelx, nelz, Ne, 3 * Nn),
                   'levels': m64.num_levels, 'densities': '0.12+U(-0.05,0.05) seed 12345, SIMP p=3, Emin=1e-6'},
        'fp64': res['fp64'],
        'mixed': res['mixed'],
        'mixed_over_fp64': res['mixed']['solves_per_s'] / res['fp64']['solves_per_s'],
        'rel_compliance_err_mixed_vs_fp64': rel_err,
        'kbyu': {'fp64_us': ms64 * 1e3, 'fp64_GBs_alg': gbs64, 'fp64_TFLOPs': tf64,
                 'fp32_us': ms32 * 1e3, 'fp32_GBs_alg': gbs32},
        'roofline': {'kernel': 'finest-level K*u (k_kfine<double>)', 'bound': 'hbm', 'achieved': gbs64,
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': gbs64 / HBM_PEAK_GBS, 'traffic': None,
                     'fp64_flop_frac': tf64 / FP64_PEAK_TFLOPS},
        'cpu_baseline': cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
