"""The drop-in boundary driven from C, with no Python or torch in the process.  Each test compiles a C99
program against include/topgpu.h (gcc, -Wall -Wextra -Werror), links it to libtopgpu.so and runs it:

- MEX layer: device buffers from hipMalloc wrapped as tg_array descriptors (as a MEX shim does with
  mxGPUArray) through Scattering_inplace and Gathering_inplace on a random 700-element connectivity.
  The scatter must match a host scatter to 1e-13, skipping an out-of-range node id as the reference
  does; the gather must match a host gather exactly; a refused class must give the reference's
  identifier.
- Solver layer: a 16 x 32 x 16 cantilever built, assembled and solved by PCG to 1e-8 through the
  model and solver entry points; the solution's true residual, from the library's K*u, within 1.5e-8."""
import os
import shutil
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, 'top_gpu_on_matlab_amd')
ROCM = os.environ.get('ROCM_PATH', '/opt/rocm')

MEX_SRC = r'''
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <hip/hip_runtime_api.h>
#include "topgpu.h"

#define NN 1000
#define NE 700

static unsigned lcg_state = 12345u;
static unsigned lcg(void) { lcg_state = lcg_state * 1664525u + 1013904223u; return lcg_state >> 8; }
static double val(void) { return (double)(lcg() % 20001u) / 10000.0 - 1.0; }

static double u[NN * 3], v[NE * 24], m[NE * 24], y0[NN * 3], y[NN * 3], yref[NN * 3];
static int32_t en[NE * 8];

#define HIPCHECK(x) do { if ((x) != hipSuccess) { printf("hip error at %s\n", #x); return 1; } } while (0)

int main(void) {
    void *dU, *dM, *dE, *dY;
    tg_array g[3], s[3];
    int e, n, c, rc;
    double err = 0, mx = 0;
    for (n = 0; n < NN * 3; ++n) { u[n] = val(); y0[n] = val(); }
    for (e = 0; e < NE * 24; ++e) v[e] = val();
    for (e = 0; e < NE * 8; ++e) en[e] = 1 + (int32_t)(lcg() % NN);   /* column-major [Ne x 8], 1-based */
    en[3 + NE * 2] = NN + 5;                                            /* out of range: the scatter skips it */
    HIPCHECK(hipMalloc(&dU, sizeof u));
    HIPCHECK(hipMalloc(&dM, sizeof m));
    HIPCHECK(hipMalloc(&dE, sizeof en));
    HIPCHECK(hipMalloc(&dY, sizeof y));
    HIPCHECK(hipMemcpy(dU, u, sizeof u, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dE, en, sizeof en, hipMemcpyHostToDevice));
    en[3 + NE * 2] = 1;                                                 /* the gather gets a valid copy */
    HIPCHECK(hipMemcpy(dM, v, sizeof v, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dY, y0, sizeof y0, hipMemcpyHostToDevice));
    /* Scattering_inplace(Y [Nn x 3], eNodMat [Ne x 8], uMat [Ne x 24]): Y(eNod(e,j), c) += uMat(e, 3j + c) */
    memset(s, 0, sizeof s);
    s[0].data = dY; s[0].cls = TG_DOUBLE; s[0].ndim = 2; s[0].dims[0] = NN; s[0].dims[1] = 3;
    s[1].data = dE; s[1].cls = TG_INT32; s[1].ndim = 2; s[1].dims[0] = NE; s[1].dims[1] = 8;
    s[2].data = dM; s[2].cls = TG_DOUBLE; s[2].ndim = 2; s[2].dims[0] = NE; s[2].dims[1] = 24;
    rc = Scattering_inplace(0, NULL, 3, s, NULL);
    if (rc != TG_OK) { printf("Scattering_inplace: %s %s\n", tg_last_error_id(), tg_last_error_message()); return 1; }
    HIPCHECK(hipMemcpy(y, dY, sizeof y, hipMemcpyDeviceToHost));
    memcpy(yref, y0, sizeof y0);
    for (e = 0; e < NE; ++e)
        for (n = 0; n < 8; ++n) {
            if (e == 3 && n == 2) continue;
            for (c = 0; c < 3; ++c) yref[(en[e + NE * n] - 1) + NN * c] += v[e + NE * (3 * n + c)];
        }
    for (n = 0; n < NN * 3; ++n) {
        err = fmax(err, fabs(y[n] - yref[n]));
        mx = fmax(mx, fabs(yref[n]));
    }
    if (err > 1e-13 * mx) { printf("scatter error %g\n", err / mx); return 1; }
    /* Gathering_inplace(uMat [Ne x 24], u [Nn x 3], eNodMat [Ne x 8]): uMat(e, 3n + c) = u(eNod(e,n), c) */
    HIPCHECK(hipMemcpy(dE, en, sizeof en, hipMemcpyHostToDevice));
    memset(g, 0, sizeof g);
    g[0] = s[2];
    g[1].data = dU; g[1].cls = TG_DOUBLE; g[1].ndim = 2; g[1].dims[0] = NN; g[1].dims[1] = 3;
    g[2] = s[1];
    rc = Gathering_inplace(0, NULL, 3, g, NULL);
    if (rc != TG_OK) { printf("Gathering_inplace: %s %s\n", tg_last_error_id(), tg_last_error_message()); return 1; }
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipMemcpy(m, dM, sizeof m, hipMemcpyDeviceToHost));
    for (e = 0; e < NE; ++e)
        for (n = 0; n < 8; ++n)
            for (c = 0; c < 3; ++c)
                if (m[e + NE * (3 * n + c)] != u[(en[e + NE * n] - 1) + NN * c]) { printf("gather %d %d %d\n", e, n, c); return 1; }
    /* a class the entry point refuses, with the reference's identifier */
    g[1].cls = TG_INT32;
    rc = Gathering_inplace(0, NULL, 3, g, NULL);
    if (rc == TG_OK || strcmp(tg_last_error_id(), "Gathering_inplace:Type") != 0) { printf("type check: %s\n", tg_last_error_id()); return 1; }
    hipFree(dU); hipFree(dM); hipFree(dE); hipFree(dY);
    printf("scatter rel err %.1e; gather exact; ok\n", err / mx);
    return 0;
}
'''

SOLVER_SRC = r'''
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <hip/hip_runtime_api.h>
#include "topgpu.h"

#define NY 16
#define NX 32
#define NZ 16

#define HIPCHECK(x) do { if ((x) != hipSuccess) { printf("hip error at %s\n", #x); return 1; } } while (0)
#define TGCHECK(x) do { if ((x) != TG_OK) { printf("%s: %s %s\n", #x, tg_last_error_id(), tg_last_error_message()); return 1; } } while (0)

int main(void) {
    tg_params p;
    tg_model *m = NULL;
    tg_solver *s = NULL;
    tg_level_info li;
    static uint8_t vox[NY * NX * NZ];
    double *F, *E, *ky, *yh, *dE, *dB, *dY, *dKY, num = 0, den = 0, c = 0, res = 0;
    int64_t nd, ne, e, i;
    int32_t its = 0, levels = 0;
    tg_params_default(&p);
    p.cell_size = 1.0 / NX;
    memset(vox, 1, sizeof vox);
    TGCHECK(tg_model_create(vox, NY, NX, NZ, &p, &m));
    TGCHECK(tg_model_builtin_case(m, 1));            /* cantilever */
    TGCHECK(tg_model_finalize(m));
    TGCHECK(tg_model_level_info(m, 1, &li));
    nd = 3 * li.numNodes;
    ne = li.numElements;
    F = malloc((size_t)tg_model_table_size(m, 1, TG_TAB_LOADS) * sizeof(double));
    E = malloc((size_t)ne * sizeof(double));
    ky = malloc((size_t)nd * sizeof(double));
    yh = malloc((size_t)nd * sizeof(double));
    TGCHECK(tg_model_get_table(m, 1, TG_TAB_LOADS, F));   /* first load case: F[0 .. nd) */
    for (e = 0; e < ne; ++e) {                       /* a smooth density field, SIMP p = 3 */
        const double ey = (double)(e % NY), ex = (double)((e / NY) % NX), ez = (double)(e / (NY * NX));
        const double x = 0.12 + 0.05 * sin(ex / 7.0) * cos(ey / 5.0) * sin(ez / 6.0 + 1.0);
        E[e] = 1e-6 + x * x * x * (1.0 - 1e-6);
    }
    HIPCHECK(hipMalloc((void **)&dE, (size_t)ne * sizeof(double)));
    HIPCHECK(hipMalloc((void **)&dB, (size_t)nd * sizeof(double)));
    HIPCHECK(hipMalloc((void **)&dY, (size_t)nd * sizeof(double)));
    HIPCHECK(hipMalloc((void **)&dKY, (size_t)nd * sizeof(double)));
    HIPCHECK(hipMemcpy(dE, E, (size_t)ne * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dB, F, (size_t)nd * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemset(dY, 0, (size_t)nd * sizeof(double)));
    TGCHECK(tg_solver_create(m, NULL, &s));          /* NULL: the default stream */
    TGCHECK(tg_solver_set_modulus(s, dE));
    TGCHECK(tg_solver_assemble(s));
    TGCHECK(tg_solver_pcg(s, dB, dY, 1e-8, 400, 0, &its, &res));
    TGCHECK(tg_solver_kbyu(s, 1, dY, dKY));
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipMemcpy(ky, dKY, (size_t)nd * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(yh, dY, (size_t)nd * sizeof(double), hipMemcpyDeviceToHost));
    for (i = 0; i < nd; ++i) {
        num += (ky[i] - F[i]) * (ky[i] - F[i]);
        den += F[i] * F[i];
        c += F[i] * yh[i];
    }
    levels = tg_model_num_levels(m);
    tg_solver_destroy(s);
    tg_model_destroy(m);
    HIPCHECK(hipFree(dE));
    HIPCHECK(hipFree(dB));
    HIPCHECK(hipFree(dY));
    HIPCHECK(hipFree(dKY));
    free(F);
    free(E);
    free(ky);
    free(yh);
    printf("levels %d, PCG %d its, reported residual %.2e, true residual %.2e, compliance %.6e\n", (int)levels,
           (int)its, res, sqrt(num / den), c);
    if (!(levels == 3 && its > 0 && its < 400 && res < 1e-8 && sqrt(num / den) < 1.5e-8 && c > 0)) return 1;
    puts("ok");
    return 0;
}
'''


def _compile_and_run(tmp_path, name, source):
    gcc = shutil.which('gcc')
    if gcc is None:
        pytest.skip('no C compiler')
    src = tmp_path / (name + '.c')
    src.write_text(source)
    exe = tmp_path / name
    cmd = [gcc, '-std=c99', '-Wall', '-Wextra', '-Werror', '-D__HIP_PLATFORM_AMD__', '-I' + os.path.join(ROCM, 'include'),
           '-I' + os.path.join(ROOT, 'include'), str(src), '-o', str(exe), '-L' + PKG, '-ltopgpu',
           '-Wl,-rpath,' + PKG, '-L' + os.path.join(ROCM, 'lib'), '-lamdhip64', '-Wl,-rpath,' + os.path.join(ROCM, 'lib'),
           '-lm']
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print('\n[c program %s]' % name, flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout.strip(), flush=True)
    assert r.returncode == 0 and r.stdout.strip().endswith('ok'), (r.stdout, r.stderr)


def test_mex_entry_points_from_a_c_program(tmp_path):
    _compile_and_run(tmp_path, 'mex_c', MEX_SRC)


def test_solver_entry_points_from_a_c_program(tmp_path):
    _compile_and_run(tmp_path, 'solver_c', SOLVER_SRC)
