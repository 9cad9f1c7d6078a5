"""The drop-in boundary from C: a C caller written against include/*.h (the
reference's vie_iterative.h / vie_fft.h / dielectric_fft.h names) compiles
with gcc, links libvie_hip.so and runs the host-side API (no device): the
reference's constants, backend names, a rejected plan, and a callback operator
of the caller's own refused by vie_gmres (the host program keeps its own)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

C_SOURCE = r"""
#include <stdio.h>
#include <string.h>
#include "vie_hip.h"

static int stencil(void *ctx, const complex double *x, complex double *y)
{
  size_t i, n = *(size_t *)ctx;
  for(i = 0; i < n; i++) {
    y[i] = (3.0 + 0.2*I)*x[i];
    if(i > 0) y[i] += (0.3 - 0.1*I)*x[i-1];
    if(i + 1 < n) y[i] += (-0.5 + 0.3*I)*x[i+1];
  }
  return 1;
}

int main(void)
{
  size_t n = 40, i;
  complex double b[40], x[40], e[40];
  vie_linear_operator op = {40, stencil, &n, NULL, NULL};
  vie_gmres_options o;
  vie_gmres_report r;
  dielectric_fft_system *s = NULL;
  vie_fft_plan *p = NULL;
  size_t shape[3] = {1, 1, 1}, xyz[3] = {0, 0, 0};
  for(i = 0; i < n; i++) { e[i] = sin((double)i) + I*cos(0.4*i); x[i] = 0; }
  stencil(&n, e, b);
  vie_gmres_default_options(&o);
  o.restart = 3;
  /* a callback operator of the caller's own is the host program's to solve */
  if(vie_gmres(&op, b, x, &o, &r) || r.status != VIE_ITERATIVE_INVALID_INPUT) return 1;
  for(i = 0; i < n; i++) if(x[i] != 0) return 2;
  if(dielectric_gmres_settings.restart != 30) return 3;
  if(strcmp(dielectric_fft_backend(s), "none")) return 4;
  if(vie_fft_backend_set("fftw") || !vie_fft_backend_set("hip")) return 5;
  if(vie_fft_create(&p, shape, 1, xyz, .25, 5.0, 3.0, 1000000) || p) return 6;   /* k h > 1 */
  printf("C ABI ok: foreign operator refused (%s)\n", vie_iterative_status_name(r.status));
  return 0;
}
"""


def test_c_caller_compiles_links_and_runs(tmp_path):
    src = tmp_path / "caller.c"
    src.write_text(C_SOURCE)
    exe = tmp_path / "caller"
    lib = os.path.join(ROOT, "nec2c_amd", "lib")
    r = subprocess.run(["gcc", "-std=c11", "-O1", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L" + lib, "-lvie_hip", "-Wl,-rpath," + lib, "-lm"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert "C ABI ok" in r.stdout
