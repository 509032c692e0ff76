"""Benchmark of the TOP_GPU hot path on one MI355X.

Workload: the adjoint-run cantilever, 128 x 256 x 128 voxels (nely x nelx x nelz; 4,194,304
elements, 12,830,211 DOFs, three multigrid levels), optCase 1, ft = 2 (PDE density filter),
rMin = 3, V0 = 0.12, x0 = 0.12 + U(-0.05, 0.05) (seed 12345), SIMP p = 3, Emin = 1e-6.

One step = one optimisation iteration of TOP_GPU.m (Galerkin assembly, one warm-started
elasticity PCG solve with the multigrid V-cycle, compliance + sensitivities, three PDE-filter
solves, OC update), driven by the FP64 solver.  In every step the same assembled system (same
moduli) is also solved by the mixed-precision solver (FP32 V-cycle inside FP64 PCG, warm-started
from its own previous solution), so both precisions are timed on identical systems and the
relative compliance difference |c_mixed - c_fp64| / c_fp64 measures precision, not trajectory
drift.  The metric is PCG solves/s: solves in the timed window divided by the summed wall time of
those solves (V-cycles included); `value` is the mixed-precision rate.  Inputs are resident in
HBM before the timed region.

    python bench.py [--steps K] [--warmup W] [--size NELY NELX NELZ] [--no-cpu]
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
# measured HBM bytes per launch of the FP64 PCG operator: the copy-calibrated FETCH_SIZE + WRITE_SIZE
# of two rocprofv3 --pmc runs of `bench.py --traffic-probe`, kept with their commands and raw counts in
# this summary, one `traffic_bytes_per_launch <workload> <kernel> <bytes>` line per measured workload
TRAFFIC_SUMMARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                               'kfine_walsh_fp64_c5_128x256x128_pmc_traffic.md')


def measured_traffic(shape, kernel='fp64_pcg_operator'):
    """HBM bytes per launch of `kernel` at `shape` from the committed counter summary, or None when that
    workload has not been measured."""
    want = '%dx%dx%d' % tuple(shape)
    try:
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                w = line.split()
                if len(w) == 4 and w[0] == 'traffic_bytes_per_launch' and w[1] == want and w[2] == kernel:
                    return float(w[3])
    except OSError:
        pass
    return None


# the mixed solver with its finest element products on the matrix cores in BF16 (reported beside FP32)
BF16_MODES = ('bf16', 'bf16x3')
BF16_MODES_TEXT = {
    'bf16': 'dense 24x24 on MFMA f32 32x32x16 bf16 (BF16 operands, FP32 accumulation), rest of the mixed V-cycle FP32',
    'bf16x3': 'dense 24x24 on MFMA f32 32x32x16 bf16 with each operand split into two BF16 parts (hi lo + lo hi + '
              'hi hi, FP32 accumulation), rest of the mixed V-cycle FP32',
}


def log(*a):
    print('[bench]', *a, file=sys.stderr, flush=True)


def kbyu_alg_bytes(Nn, Ne, s_in, s_out, s_E):
    """Algorithmic bytes of one finest-level K*u: 3 Nn (s_in + s_out) + Ne s_E + 32 Ne."""
    return 3 * Nn * (s_in + s_out) + Ne * s_E + 32 * Ne


class Run:
    """TOP_GPU.m's loop (TOP_GPU.m:176-292) split into single iterations.  The FP64 solver drives
    the design; the mixed solver solves each iteration's system too."""

    def __init__(self, shape, x0, V0=0.12, rMin=3.0, Emin=1e-6):
        import torch
        from top_gpu_on_matlab_amd import _lib
        from top_gpu_on_matlab_amd import top_gpu as G
        from top_gpu_on_matlab_amd.model import cuboid_model
        from top_gpu_on_matlab_amd.solver import Solver
        self.G, self.L, self.torch = G, _lib, torch
        self.V0, self.Emin = V0, Emin
        self.models = {p: cuboid_model(*shape, 1, precision=p) for p in ('fp64', 'mixed')}
        for p in BF16_MODES:  # the mixed solver with its finest element products as BF16 on MFMA
            self.models[p] = cuboid_model(*shape, 1, precision='mixed')
        self.solvers = {p: Solver(self.models[p]) for p in self.models}
        for p in BF16_MODES:
            self.solvers[p].set_tuning(kfine_mfma=p)
        self.model = self.models['fp64']
        self.filt = G.PDEFilter(self.model, rMin, tol=1e-4, max_it=800)
        self.Ne = self.model.num_elements
        self.F = torch.from_numpy(self.model.get_table(1, 'F')[:, 0].copy()).cuda()
        self.x = torch.as_tensor(x0).cuda().clone()
        self.xPhys = self.x.clone()
        E1 = torch.ones_like(self.x)
        self.U = {}
        for p, sv in self.solvers.items():
            sv.set_modulus(E1)
            sv.assemble()
            self.U[p], _, _ = sv.pcg(self.F)
        ce = G.unit_compliance(self.solvers['fp64'], self.U['fp64'])
        self.cSolid = G.reduce(_lib.TG_RED_DOT, E1, ce)
        self.ones = torch.ones_like(self.x)

    def solve(self, p, E):
        torch = self.torch
        sv = self.solvers[p]
        t0 = time.perf_counter()
        sv.set_modulus(E)
        sv.assemble()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        self.U[p], its, _ = sv.pcg(self.F, y0=self.U[p])
        t2 = time.perf_counter()
        c = float(torch.dot(self.F, self.U[p]))  # compliance F'U (equals sum E_e u_e'K0 u_e)
        return dict(pcg_s=t2 - t1, its=its, c=c, assembly_s=t1 - t0)

    def iterate(self):
        G, L, torch = self.G, self.L, self.torch
        t0 = time.perf_counter()
        E = G.simp(self.xPhys, self.Emin, 1.0, 3.0)
        r = {p: self.solve(p, E) for p in ('fp64', 'mixed') + BF16_MODES}
        t1 = time.perf_counter()
        ce = G.unit_compliance(self.solvers['fp64'], self.U['fp64'])
        dc = G.sensitivity(self.xPhys, ce, self.cSolid, self.Emin, 1.0, 3.0)
        dc = self.filt(dc)
        dv = self.filt(self.ones)
        fval = G.reduce(L.TG_RED_SUM, self.xPhys) / self.Ne - self.V0
        xnew, change = G.oc_update(self.x, dc, dv, fval, 0.2)
        self.x = xnew
        self.xPhys = self.filt(xnew)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        r['update_s'] = t2 - t1
        r['iter_s'] = t2 - t0 - sum(r[p]['pcg_s'] + r[p]['assembly_s'] for p in ('mixed',) + BF16_MODES)
        return r


def cpu_baseline(shape, E, F, its_ref, sample_iters):
    """CPU restatement of the reference algorithm (oracle/, numpy/scipy, FP64), on one core: BLAS is held
    to one thread and the oracle's sparse products and indexing are single-threaded anyway.

    Two samples, about 20-30 s of CPU work in all:
      * the benchmarked system itself (same shape, the first timed iteration's moduli): set-up and
        Galerkin assembly (reported, not part of solves/s), the first V-cycle and `sample_iters` PCG
        iterations; solves/s = 1 / (V-cycle + its_ref iterations), its_ref being the GPU FP64 solver's
        mean iteration count on the same systems;
      * a complete warm-start-free PCG solve of the 24 x 48 x 24 cantilever at its first optimisation
        iteration (uniform density 0.12, ft = 1), timed whole: no extrapolation.
    """
    from threadpoolctl import threadpool_limits
    from oracle import mesh as M
    from oracle import solver as S
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        m = M.build_cuboid_model(*shape, opt_case=1)
        sv = S.Solver(m)
        sv.assemble(E)
        t1 = time.perf_counter()
        log('cpu oracle setup+assembly %.1fs' % (t1 - t0))
        ta = time.perf_counter()
        sv.vcycle(F.copy())
        tb = time.perf_counter()
        S.pcg(sv.kbyu, sv.vcycle, F, 1e-30, sample_iters)
        tc = time.perf_counter()
        t_iter = ((tc - tb) - (tb - ta)) / sample_iters
        t_solve = (tb - ta) + its_ref * t_iter
        log('cpu oracle: V-cycle %.1fs, %.2f s per PCG iteration' % (tb - ta, t_iter))
        # complete solve of the small cantilever
        mc = M.build_cuboid_model(24, 48, 24, opt_case=1)
        svc = S.Solver(mc)
        Ec = np.full(mc.levels[0].numElements, 1e-6 + 0.12 ** 3 * (1 - 1e-6))
        svc.assemble(Ec)
        td = time.perf_counter()
        _, its_c, _ = svc.pcg(mc.F[:, 0], tol=1e-3, maxIT=800)
        te = time.perf_counter()
        log('cpu oracle: 24x48x24 solve %.2fs (%d PCG iterations)' % (te - td, its_c))
    return {
        'value': 1.0 / t_solve, 'unit': 'solves/s', 'cores': 1, 'kind': 'port',
        'sample': ('CPU restatement of the reference algorithm (oracle/, numpy/scipy, FP64, one thread): the '
                   '%dx%dx%d system of the first timed iteration, first V-cycle + %d PCG iterations timed, solve '
                   '= V-cycle + %.1f iterations (the GPU FP64 mean) x %.2f s/iteration; set-up and assembly '
                   '%.1f s not included' % (shape[0], shape[1], shape[2], sample_iters, its_ref, t_iter, t1 - t0)),
        'sec_per_pcg_iteration': t_iter, 'setup_s': t1 - t0,
        'full_solve_24x48x24': {'solves_per_s': 1.0 / (te - td), 'pcg_its': its_c, 'seconds': te - td},
    }


def traffic_probe(shape):
    """Finest K*u launches at `shape` (moduli 1e-6 + U(0,1)^3, seed 0; 3 + 3 timed applies and 3 PCG
    iterations, whose operator is the FP64 product with the fused p.Kp) and two 2 GiB device copies
    (the runtime blit and a plain elementwise load/store kernel): under rocprofv3 --pmc the copies,
    which move exactly 2 GiB in and 2 GiB out, give the factors that turn FETCH_SIZE and WRITE_SIZE
    into HBM bytes."""
    import torch
    from top_gpu_on_matlab_amd.model import cuboid_model
    from top_gpu_on_matlab_amd.solver import Solver
    m = cuboid_model(*shape, 1, precision='mixed')
    s = Solver(m)
    g = torch.Generator().manual_seed(0)
    s.set_modulus((1e-6 + torch.rand(m.num_elements, generator=g, dtype=torch.float64) ** 3).cuda())
    s.assemble()
    torch.cuda.synchronize()
    log('finest K*u: fp64 %.1f us, fp32 %.1f us' % (s.time_kbyu(1, 'fp64', 3) * 1e3, s.time_kbyu(1, 'fp32', 3) * 1e3))
    # three PCG iterations: the PCG operator itself (the FP64 product with the fused p.Kp)
    F = torch.from_numpy(m.get_table(1, 'F')[:, 0].copy()).cuda()
    s.pcg(F, max_it=3)
    torch.cuda.synchronize()
    del s, m
    torch.cuda.empty_cache()
    a = torch.empty(256 * 1024 * 1024, dtype=torch.float64, device='cuda')
    a.fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(2):
        b.copy_(a)
    for _ in range(2):
        torch.mul(a, 1.0, out=b)
    # read-only (a reduction: 2 GiB in, 8 bytes out) and write-only (a fill: 2 GiB out) kernels settle
    # each counter's scale on its own
    for _ in range(2):
        float(a.sum())
    for _ in range(2):
        b.fill_(2.0)
    torch.cuda.synchronize()
    log('copies done (2 GiB read + 2 GiB write each), reductions (2 GiB read), fills (2 GiB write)')


def device_copy_gbs(reps=5):
    """Read + write GB/s of a 2 GiB device-to-device copy (torch copy_, HIP events): the practical HBM
    ceiling that the published 8 TB/s peak is not."""
    import torch
    a = torch.empty(256 * 1024 * 1024, dtype=torch.float64, device='cuda')
    a.fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2 * a.numel() * 8 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return gbs


def kbyu_microbench(reps):
    """The finest-level K*u at every configuration's shape, as the scope table's micro-benchmark asks:
    moduli E = Emin + x^3 (1 - Emin) with x ~ U(0,1) (seed 2) on the solid elements, HIP events around
    `reps` launches after one warm-up launch, FP64 and the mixed solver's FP32 operator (the Walsh form,
    and the dense products on MFMA with BF16 and split-BF16 operands); bytes_alg per launch as in
    kbyu_alg_bytes.  The input vector is the solver's level-1 work vector: the kernels' cost does not
    depend on the values.  Beside them, the practical HBM ceilings: a 2 GiB device copy and a 2 GiB
    read-only reduction."""
    import tempfile
    import torch
    from top_gpu_on_matlab_amd.model import Model, cuboid_model
    from top_gpu_on_matlab_amd.solver import Solver
    from top_gpu_on_matlab_amd.synthetic import bottom_clamp_top_patch, value_noise_mask
    from top_gpu_on_matlab_amd.topvoxel import load_topvoxel_model, write_topvoxel
    out = {}
    a = torch.empty(256 * 1024 * 1024, dtype=torch.float64, device='cuda')
    a.uniform_(-1.0, 1.0)
    b = torch.empty_like(a)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    b.copy_(a)
    ev[0].record()
    for _ in range(10):
        b.copy_(a)
    ev[1].record()
    torch.cuda.synchronize()
    out['copy_2GiB'] = {'GBs': 2 * a.numel() * 8 * 10 / (ev[0].elapsed_time(ev[1]) * 1e-3) / 1e9}
    float(a.sum())
    ev[0].record()
    for _ in range(10):
        a.sum()
    ev[1].record()
    torch.cuda.synchronize()
    out['read_2GiB'] = {'GBs': a.numel() * 8 * 10 / (ev[0].elapsed_time(ev[1]) * 1e-3) / 1e9}
    del a, b
    torch.cuda.empty_cache()
    log('device copy %.0f GB/s (read + write), read-only reduction %.0f GB/s'
        % (out['copy_2GiB']['GBs'], out['read_2GiB']['GBs']))

    def models():
        yield 'C1 cantilever 24x48x24', cuboid_model(24, 48, 24, 1, precision='mixed')
        yield 'C2 MBB 10x60x10 super-element', cuboid_model(10, 60, 10, 2, super_element=True, precision='mixed')
        with tempfile.TemporaryDirectory() as td:
            mask = value_noise_mask(256, 128, 128, seed=0, octaves=4, period=32)
            path = td + '/value_noise.TopVoxel'
            write_topvoxel(path, mask, *bottom_clamp_top_patch(mask))
            yield 'C3 value noise 256x128x128', load_topvoxel_model(path, precision='mixed')
        yield 'C4 cantilever 96x192x96', cuboid_model(96, 192, 96, 1, precision='mixed')
        yield 'C5 cantilever 128x256x128', cuboid_model(128, 256, 128, 1, precision='mixed')

    for name, m in models():
        s = Solver(m)
        g = torch.Generator().manual_seed(2)
        nE = m.num_elements * (64 if m.super_element else 1)
        x = torch.rand(nE, generator=g, dtype=torch.float64)
        s.set_modulus((1e-6 + x ** 3 * (1 - 1e-6)).cuda())
        s.assemble()
        torch.cuda.synchronize()
        Ne, Nn = m.num_elements, m.num_nodes
        row = {'elements': Ne, 'nodes': Nn}
        variants = [('fp64', 'fp64', None, 8)] + ([] if m.super_element else
                                                  [('fp32', 'fp32', 'walsh', 4), ('bf16', 'fp32', 'bf16', 4),
                                                   ('bf16x3', 'fp32', 'bf16x3', 4)])
        for key, prec, mode, sz in variants:
            if mode:
                s.set_tuning(kfine_mfma=mode)
            us = s.time_kbyu(1, prec, reps) * 1e3
            ba = kbyu_alg_bytes(Nn, Ne, sz, sz, sz)
            row[key] = {'us': us, 'bytes_alg': ba, 'GBs_alg': ba / us / 1e3}
        if not m.super_element:
            s.set_tuning(kfine_mfma='walsh')
        out[name] = row
        log('%s: %s' % (name, ', '.join('%s %.1f us (%.0f GB/s alg)' % (k, v['us'], v['GBs_alg'])
                                         for k, v in row.items() if isinstance(v, dict))))
        del s, m
        torch.cuda.empty_cache()
    return out


def mex_timing(shape, reps):
    """The six MEX entry points on the connectivity and transfer maps of the `shape` cantilever (the
    super-element assembly on the 10x60x10 MBB model, its own configuration), FP64, inputs resident on
    the device; HIP events around `reps` calls after one warm-up call.  The scatter entry points reuse
    their cached incidence after the first call (the content hash of the index array is recomputed
    every call and included)."""
    import torch
    from top_gpu_on_matlab_amd import mex
    from top_gpu_on_matlab_amd.model import cuboid_model
    from top_gpu_on_matlab_amd.mex import GpuArray

    def ga(a):
        a = np.asarray(a)
        return GpuArray(torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order='F'))).cuda(), a.shape)

    def zeros(dims, dtype=torch.float64):
        return GpuArray(torch.zeros(int(np.prod(dims)), dtype=dtype, device='cuda'), dims)

    rng = np.random.default_rng(1)
    m = cuboid_model(*shape, 1)
    i1, i2 = m.level_info(1), m.level_info(2)
    Ne, Nn, Nc, Nnc = i1['numElements'], i1['numNodes'], i2['numElements'], i2['numNodes']
    eNod, eNod2 = ga(m.get_map(1, 'eNodMat')), ga(m.get_map(2, 'eNodMat'))
    up2 = ga(m.get_map(2, 'elementUpwardMap'))
    tm2 = m.get_map(2, 'transferMat')
    nInt = i2['intermediateNumNodes']
    u = ga(rng.uniform(-1, 1, (Nn, 3)))
    uMat = zeros((Ne, 24))
    E = ga(1e-6 + rng.random(Ne) ** 3)
    Ke = m.get_table(1, 'Ke')
    muk = ga(m.get_map(1, 'mapUniqueKes'))
    uf, ux = ga(m.get_table(1, 'uniqueKesFree')), ga(m.get_table(1, 'uniqueKesFixed'))
    Pch = ga(rng.standard_normal((24, 24, up2.dims[1])))
    cases = {
        'Gathering_inplace': (mex.Gathering_inplace, (uMat, u, eNod),
                              'finest: uMat [%d x 24] <- u [%d x 3] via eNodMat' % (Ne, Nn)),
        'Scattering_inplace': (mex.Scattering_inplace, (zeros((Nn, 3)), eNod, ga(rng.standard_normal((Ne, 24)))),
                               'finest: Y [%d x 3] += uMat [%d x 24] via eNodMat' % (Nn, Ne)),
        'scatter_accum3_inplace (restriction)': (
            mex.scatter_accum3_inplace, (ga(m.get_map(2, 'eNodMat').reshape(-1, order='F')),
                                         ga(rng.standard_normal((8 * Nc, 3))), zeros((Nnc, 3))),
            'L1->L2: keys = level-2 eNodMat (N = %d) into M = %d coarse nodes' % (8 * Nc, Nnc)),
        'scatter_accum3_inplace (prolongation)': (
            mex.scatter_accum3_inplace, (ga(tm2.reshape(-1, order='F')), ga(rng.standard_normal((tm2.size, 3))),
                                         zeros((nInt, 3))),
            'L2->L1: keys = level-2 transferMat (N = %d) into M = %d intermediate nodes' % (tm2.size, nInt)),
        'assembleKs_level2_inplace': (
            mex.assembleKs_level2_inplace, (zeros((24, 24, Nc)), up2, E, ga(Ke.reshape(-1, order='F')), Pch, muk, uf, ux),
            'Ks [24 x 24 x %d], 64 children each, %d boundary element matrices' % (Nc, uf.dims[1])),
    }
    if m.num_levels >= 3:
        i3 = m.level_info(3)
        up3 = ga(m.get_map(3, 'elementUpwardMap'))
        cases['assembleKs_higherLevel_inplace'] = (
            mex.assembleKs_higherLevel_inplace, (zeros((24, 24, i3['numElements'])), up3,
                                                 ga(rng.standard_normal((24, 24, Nc))),
                                                 ga(rng.standard_normal((24, 24, up3.dims[1])))),
            'Ks [24 x 24 x %d] from KsPrev [24 x 24 x %d]' % (i3['numElements'], Nc))
    ms = cuboid_model(10, 60, 10, 2, super_element=True)
    s1, s2 = ms.level_info(1), ms.level_info(2)
    sup = ga(ms.get_map(2, 'elementUpwardMap'))
    cases['assembleKs_level2_superEle_inplace'] = (
        mex.assembleKs_level2_superEle_inplace,
        (zeros((24, 24, s2['numElements'])), sup, ga(rng.standard_normal((576, s1['numElements']))),
         ga(rng.standard_normal((24, 24, sup.dims[1]))), ga(ms.get_map(1, 'mapUniqueKes')),
         ga(rng.random((64, s1['numElements']))), ga(ms.get_table(1, 'uniqueKesFree')),
         ga(ms.get_table(1, 'uniqueKesFixed'))),
        '10x60x10 MBB super-element: Ks [24 x 24 x %d] from %d fine elements' % (s2['numElements'], s1['numElements']))
    out = {}
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for name, (fn, args, what) in cases.items():
        fn(*args)  # warm-up (the scatter entries build their incidence here)
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(reps):
            fn(*args)
        ev1.record()
        torch.cuda.synchronize()
        out[name] = {'ms_per_call': ev0.elapsed_time(ev1) / reps, 'calls': reps, 'shape': what}
        log('%-40s %8.3f ms  (%s)' % (name, out[name]['ms_per_call'], what))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--size', type=int, nargs=3, default=[128, 256, 128], metavar=('NELY', 'NELX', 'NELZ'))
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-iters', type=int, default=3)
    ap.add_argument('--traffic-probe', action='store_true',
                    help='only run 3 + 3 finest K*u launches (FP64, FP32) and the 2 GiB copies that calibrate '
                         'the FETCH_SIZE / WRITE_SIZE counters; for rocprofv3 --pmc runs')
    ap.add_argument('--kfine', choices=('walsh', 'mfma32'), default='walsh',
                    help='mixed solver: finest element products in the Walsh form on the VALU (default) or dense on '
                         'MFMA with FP32 operands (for comparison)')
    ap.add_argument('--coarse', choices=('step', 'persistent', 'streaming'), default=None,
                    help='coarsest solve: launched step by step, as one persistent launch, or persistent with '
                         "non-temporal factor loads (default: the library's default)")
    ap.add_argument('--kbyu', type=int, default=0, metavar='REPS',
                    help='only time the finest K*u at every configuration shape (REPS launches each) and the '
                         'device copy / read bandwidths')
    ap.add_argument('--mex', type=int, default=0, metavar='REPS',
                    help='only time the six MEX entry points (REPS calls each) at the --size shapes')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU only: the hot path has no multi-GPU mode in this build '
              '(requested --gpus %d)' % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.traffic_probe:
        traffic_probe(tuple(args.size))
        return
    if args.mex:
        print(json.dumps({'mex_ms_per_call': mex_timing(tuple(args.size), args.mex)}), flush=True)
        return
    if args.kbyu:
        print(json.dumps({'kbyu_microbench': kbyu_microbench(args.kbyu)}), flush=True)
        return
    import torch
    shape = tuple(args.size)
    rng = np.random.default_rng(12345)
    Ne_full = shape[0] * shape[1] * shape[2]
    x0 = 0.12 + rng.uniform(-0.05, 0.05, Ne_full)
    t = time.perf_counter()
    run = Run(shape, x0)
    run.solvers['mixed'].set_tuning(kfine_mfma=args.kfine)
    if args.coarse:
        for sv in run.solvers.values():
            sv.set_tuning(coarse_persistent=args.coarse != 'step', coarse_streaming=args.coarse == 'streaming')
    torch.cuda.synchronize()
    log('setup + solid solves %.1fs (Ne=%d)' % (time.perf_counter() - t, run.Ne))
    hist = []
    for k in range(args.warmup + args.steps):
        if k == args.warmup:
            # per-launch HIP-event timing, inside the timed steps, of the roofline kernel: the FP64
            # finest K*u that is the PCG operator, as the mixed solver launches it (the events cost
            # ~7 us per launch, ~1% of its PCG time, which `value` includes); the FP64 solver runs bare
            run.solvers['mixed'].time_kfine(True, only_operator=True)
        r = run.iterate()
        hist.append(r)
        log('it %d: fp64 PCG %.1f ms (%d its) c=%.12e | mixed PCG %.1f ms (%d its) c=%.12e | rel %.1e | %s | '
            'assembly %.0f ms, FP64 iteration %.0f ms'
            % (k + 1, 1e3 * r['fp64']['pcg_s'], r['fp64']['its'], r['fp64']['c'], 1e3 * r['mixed']['pcg_s'],
               r['mixed']['its'], r['mixed']['c'], abs(r['mixed']['c'] - r['fp64']['c']) / abs(r['fp64']['c']),
               ' | '.join('%s PCG %.1f ms (%d its) rel %.1e' % (p, 1e3 * r[p]['pcg_s'], r[p]['its'],
                                                                abs(r[p]['c'] - r['fp64']['c']) / abs(r['fp64']['c']))
                          for p in BF16_MODES),
               1e3 * r['fp64']['assembly_s'], 1e3 * r['iter_s']))
    timed = hist[args.warmup:]
    res = {}
    for p in ('fp64', 'mixed') + BF16_MODES:
        pcg = sum(h[p]['pcg_s'] for h in timed)
        res[p] = dict(solves_per_s=len(timed) / pcg, ms_per_solve=1e3 * pcg / len(timed),
                      pcg_its=[h[p]['its'] for h in timed], assembly_ms=1e3 * np.mean([h[p]['assembly_s'] for h in timed]),
                      compliance=[h[p]['c'] for h in timed])
    res['fp64']['iteration_ms'] = 1e3 * np.mean([h['iter_s'] for h in timed])
    errs = [abs(h['mixed']['c'] - h['fp64']['c']) / abs(h['fp64']['c']) for h in timed]
    errs16 = {p: [abs(h[p]['c'] - h['fp64']['c']) / abs(h['fp64']['c']) for h in timed] for p in BF16_MODES}
    # finest K*u: average launch duration over the timed steps (HIP events around every launch on the
    # solver stream); FP64 = the PCG operator of the FP64 solver, FP32 = the mixed V-cycle's residual
    # and smoothing applies
    Ne, Nn = run.Ne, run.model.num_nodes
    kst = {}
    n, ms = run.solvers['mixed'].kfine_timing('fp64', 'ku')
    kst['fp64_ku'] = dict(launches=n, avg_us=1e3 * ms / max(n, 1))
    b64, b32 = kbyu_alg_bytes(Nn, Ne, 8, 8, 8), kbyu_alg_bytes(Nn, Ne, 4, 4, 4)
    us64 = kst['fp64_ku']['avg_us']
    gbs64 = b64 / us64 / 1e3
    tf64 = 1152 * Ne / us64 / 1e6
    for k, v in kst.items():
        v['GBs_alg'] = (b64 if k.startswith('fp64') else b32) / v['avg_us'] / 1e3
    log('finest K*u in the timed steps: ' + ', '.join('%s %.1f us x%d (%.0f GB/s alg)'
                                                      % (k, v['avg_us'], v['launches'], v['GBs_alg']) for k, v in kst.items()))
    copy_gbs = device_copy_gbs()
    log('device copy (the practical HBM ceiling): %.0f GB/s; the FP64 PCG operator reaches %.2f of it'
        % (copy_gbs, gbs64 / copy_gbs))
    cpu = None
    if not args.no_cpu:
        try:
            E = run.xPhys.cpu().numpy() ** 3 * (1 - 1e-6) + 1e-6
            cpu = cpu_baseline(shape, E, run.F.cpu().numpy(), float(np.mean(res['fp64']['pcg_its'])),
                               args.cpu_iters)
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
        'config': {'workload': 'cantilever %dx%dx%d (nely x nelx x nelz), %d elements, %d DOFs, 3 MG levels; '
                               'one step = one FP64-driven optimisation iteration (ft=2, rMin=3, V0=0.12, '
                               'x0=0.12+U(-0.05,0.05) seed 12345, Emin=1e-6) whose system is solved by both '
                               'precisions; value = mixed PCG solves / summed PCG time'
                               % (shape[0], shape[1], shape[2], Ne, 3 * Nn)},
        'fp64': res['fp64'],
        'mixed': res['mixed'],
        'mixed_over_fp64': res['mixed']['solves_per_s'] / res['fp64']['solves_per_s'],
        'rel_compliance_err_mixed_vs_fp64': {'max': max(errs), 'per_iteration': errs},
        **{'mixed_' + p: dict(res[p], element_products=BF16_MODES_TEXT[p],
                              over_fp64=res[p]['solves_per_s'] / res['fp64']['solves_per_s'],
                              rel_compliance_err_vs_fp64={'max': max(errs16[p]), 'per_iteration': errs16[p]})
           for p in BF16_MODES},
        'kbyu': kst,
        'roofline': {'kernel': 'finest-level K*u, FP64, PCG operator with the fused p.Kp (k_kfine_walsh<double,0,true,true,false,0>), '
                               'launches of the mixed solver in the timed steps', 'bound': 'hbm',
                     'achieved': gbs64, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': gbs64 / HBM_PEAK_GBS,
                     'traffic': measured_traffic(shape), 'bytes_alg': b64, 'avg_launch_us': us64, 'launches': kst['fp64_ku']['launches'],
                     # nominal, not a utilisation: the dense 24x24 element product's 1,152 flop per element at
                     # this launch time; the Walsh form executes far fewer
                     'dense_equivalent_fp64_flop_frac': tf64 / FP64_PEAK_TFLOPS,
                     'copy_GBs': copy_gbs, 'frac_of_copy': gbs64 / copy_gbs},
        'cpu_baseline': cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
