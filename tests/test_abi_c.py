"""The C ABI as a C caller sees it: include/topgpu.h compiles as C99 with every warning an error, and a
C program linked against libtopgpu.so (the way a MEX shim or another language's FFI binds it) runs the
entry points that need no device: the version string, default parameters, a model's hierarchy and
boundary conditions, and a MEX entry's argument check with the reference's error identifier."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, 'top_gpu_on_matlab_amd')

C_SRC = r'''
#include <stdio.h>
#include <string.h>
#include "topgpu.h"

int main(void) {
    const char *v = tg_version();
    tg_params p;
    uint8_t vox[16 * 8 * 8];
    tg_model *m = NULL;
    tg_level_info li;
    tg_array in[2];
    int32_t levels;
    int rc;
    if (!v || !strstr(v, "gfx950")) { printf("version: %s\n", v ? v : "(null)"); return 1; }
    tg_params_default(&p);
    memset(vox, 1, sizeof vox);
    if (tg_model_create(vox, 8, 16, 8, &p, &m) != TG_OK || tg_model_builtin_case(m, 1) != TG_OK ||
        tg_model_finalize(m) != TG_OK) {
        printf("model: %s %s\n", tg_last_error_id(), tg_last_error_message());
        return 1;
    }
    levels = tg_model_num_levels(m);
    if (tg_model_level_info(m, 1, &li) != TG_OK) { printf("level info: %s\n", tg_last_error_message()); return 1; }
    printf("levels %d finest %d x %d x %d nodes %lld\n", (int)levels, (int)li.resY, (int)li.resX, (int)li.resZ,
           (long long)li.numNodes);
    if (levels != 3 || li.resY != 8 || li.resX != 16 || li.resZ != 8 || li.numNodes != 9LL * 17 * 9) return 1;
    /* Gathering_inplace with two inputs: the reference's usage error, before any device work */
    memset(in, 0, sizeof in);
    in[0].cls = TG_DOUBLE; in[0].ndim = 2; in[0].dims[0] = 4; in[0].dims[1] = 24;
    in[1].cls = TG_DOUBLE; in[1].ndim = 2; in[1].dims[0] = 10; in[1].dims[1] = 3;
    rc = Gathering_inplace(0, NULL, 2, in, NULL);
    if (rc == TG_OK || strcmp(tg_last_error_id(), "Gathering_inplace:Args") != 0) {
        printf("Gathering_inplace: %d %s\n", rc, tg_last_error_id());
        return 1;
    }
    tg_model_destroy(m);
    puts("ok");
    return 0;
}
'''


def test_header_compiles_as_c99_and_host_entry_points_run(tmp_path):
    gcc = shutil.which('gcc')
    if gcc is None:
        pytest.skip('no C compiler')
    assert os.path.exists(os.path.join(PKG, 'libtopgpu.so')), 'build libtopgpu.so first'
    src = tmp_path / 'abi.c'
    src.write_text(C_SRC)
    exe = tmp_path / 'abi'
    cmd = [gcc, '-std=c99', '-Wall', '-Wextra', '-Werror', '-pedantic', '-I' + os.path.join(ROOT, 'include'),
           str(src), '-o', str(exe), '-L' + PKG, '-ltopgpu', '-Wl,-rpath,' + PKG]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith('ok'), (r.stdout, r.stderr)
