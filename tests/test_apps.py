"""The apps' C plug-in headers (fivonagi_amd/apps/<app>/parameters.h, init.h)
against the Python workloads (fivonagi_amd/workloads.py) the GPU tests and
the oracle comparisons use: a C program built here includes every header and
writes what it produces; parameters must agree field by field and initial
states bit for bit.  Also: the hifu-beam C host program builds and refuses a
bad command line before any GPU call."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

import fivonagi_amd as fa
from fivonagi_amd import build as fb
from fivonagi_amd import workloads as wl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APPS = os.path.join(ROOT, "fivonagi_amd", "apps")

PROGRAM = r'''
#include <stdio.h>
#include <stdlib.h>
#include "fivonagi.h"
#include "common/parameters.h"
#include "common/init.h"
#include "edge-drive/parameters.h"
#include "edge-drive/init.h"
#include "hifu-beam/parameters.h"
#include "hifu-beam/init.h"
#include "taylor-angle-xy/parameters.h"
#include "taylor-angle-xy/init.h"

static void dump(const char *dir, const char *name, const fivonagi_params *p, const float *q) {
  char path[512];
  snprintf(path, sizeof path, "%s/%s.params", dir, name);
  FILE *f = fopen(path, "wb");
  fwrite(p, sizeof(*p), 1, f);
  fclose(f);
  if (!q) return;
  snprintf(path, sizeof path, "%s/%s.q", dir, name);
  f = fopen(path, "wb");
  fwrite(q, sizeof(float), 3 * (size_t)p->nx * p->ny, f);
  fclose(f);
}

int main(int argc, char **argv) {
  const char *dir = argv[1];
  fivonagi_params p;
  float *q = malloc(sizeof(float) * 3 * 512 * 1024);
  size_t n;
  fivonagi_pulse_parameters(&p, 400, 400);
  n = (size_t)p.nx * p.ny;
  fivonagi_pulse_init(&p, 0.01, 10.0, q, q + n, q + 2 * n);
  dump(dir, "pulse", &p, q);
  static double amp[1024];
  fivonagi_edge_drive_parameters(&p, 256, 128, 0.02, 30.0, 20.0, amp);
  n = (size_t)p.nx * p.ny;
  fivonagi_edge_drive_init(&p, q, q + n, q + 2 * n);
  dump(dir, "edge", &p, q);
  FILE *f = fopen("/dev/null", "w");
  char path[512];
  snprintf(path, sizeof path, "%s/edge.amp", dir);
  f = fopen(path, "wb");
  fwrite(amp, sizeof(double), 128, f);
  fclose(f);
  fivonagi_hifu_parameters(&p, 1, 30.0, 0.9, 0, 0, 1);
  n = (size_t)p.nx * p.ny;
  fivonagi_hifu_init(&p, q, q + n, q + 2 * n);
  dump(dir, "hifu", &p, q);
  fivonagi_collect_opts o;
  fivonagi_hifu_collect_options(&p, &o);
  snprintf(path, sizeof path, "%s/hifu.collect", dir);
  f = fopen(path, "wb");
  fwrite(&o, sizeof o, 1, f);
  fclose(f);
  fivonagi_taylor_parameters(&p, 2, 3.14159265358979323846 / 8, 0.7, 1);
  n = (size_t)p.nx * p.ny;
  fivonagi_taylor_init(&p, q, q + n, q + 2 * n);
  dump(dir, "taylor", &p, q);
  for (int fp64 = 0; fp64 < 2; ++fp64) {
    fivonagi_taylor_collect_options(&p, fp64, &o);
    snprintf(path, sizeof path, "%s/taylor%d.collect", dir, fp64);
    f = fopen(path, "wb");
    fwrite(&o, sizeof o, 1, f);
    fclose(f);
  }
  return 0;
}
'''


@pytest.fixture(scope="module")
def dumped(tmp_path_factory):
    d = tmp_path_factory.mktemp("apps")
    src = d / "dump.c"
    src.write_text(PROGRAM)
    exe = d / "dump"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-I", APPS, str(src), "-o", str(exe), "-lm"], check=True)
    subprocess.run([str(exe), str(d)], check=True)
    return d


def params(d, name):
    return fa.Params.from_buffer_copy((d / f"{name}.params").read_bytes())


def check(p, prob, skip=()):
    """Every field of the C header's fivonagi_params against the Problem's."""
    ref = prob.cparams()
    for f, _ in fa.Params._fields_:
        if f in ("edge_amp", "edge_phase") or f in skip:
            continue
        a, b = getattr(p, f), getattr(ref, f)
        assert a == b, (f, a, b)


def state(d, name, prob):
    q = np.fromfile(d / f"{name}.q", np.float32).reshape(3, prob.ny, prob.nx)
    np.testing.assert_array_equal(q, prob.q0)


def test_pulse_app(dumped):
    prob = wl.gauss_pulse(400, 400)
    check(params(dumped, "pulse"), prob)
    state(dumped, "pulse", prob)


def test_edge_drive_app(dumped):
    prob = wl.edge_drive(256, 128, halfwidth_cells=30.0)
    check(params(dumped, "edge"), prob)
    state(dumped, "edge", prob)
    amp = np.fromfile(dumped / "edge.amp", np.float64)
    np.testing.assert_array_equal(amp, prob.edge_amp)


def test_hifu_beam_app(dumped):
    prob = wl.hifu(1)
    check(params(dumped, "hifu"), prob)
    state(dumped, "hifu", prob)
    o = fa.CollectOpts.from_buffer_copy((dumped / "hifu.collect").read_bytes())
    ref = wl.hifu_collect(prob)
    assert (o.jcut, list(o.mic), o.eta) == (ref["jcut"], list(ref["mic"]), ref["eta"])


def test_taylor_app(dumped):
    prob = wl.taylor(2, theta=math.pi / 8, moving=True, cflw=0.7)
    check(params(dumped, "taylor"), prob)
    state(dumped, "taylor", prob)
    for fp64 in (0, 1):
        o = fa.CollectOpts.from_buffer_copy((dumped / f"taylor{fp64}.collect").read_bytes())
        ref = wl.taylor_collect(prob, fp64=bool(fp64))
        assert (o.eta, o.cut_half, o.tan_theta) == (ref["eta"], ref["cut_half"],
                                                     ref["tan_theta"])


def test_hifu_driver_builds_and_checks_its_arguments():
    assert os.path.exists(fb.DRIVER), "build the library first (python -m fivonagi_amd.build)"
    r = subprocess.run([fb.DRIVER], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


def test_every_registered_app_has_its_headers():
    reg = open(os.path.join(APPS, "apps.h")).read()
    for app in ("common", "edge-drive", "hifu-beam", "taylor-angle-xy"):
        assert f'#include "{app}/boundary.h"' in reg
        for h in ("boundary.h", "parameters.h", "init.h"):
            assert os.path.exists(os.path.join(APPS, app, h)), (app, h)
