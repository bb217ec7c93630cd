"""The C ABI library loads and exports every entry point include/*.h
declares, and the ctypes mirror has the header's struct layouts.
No device work is done here."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from sharpclaw_amd import _abi
from sharpclaw_amd._lib import LIB_PATH, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "sharpclaw_amd.h")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sc_\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = declared()
    assert len(names) >= 18, names
    L = lib()
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (sc_\w+)", out))
    assert set(names) <= exported, set(names) - exported


def test_version_and_no_device_error():
    L = lib()
    assert L.sc_version() >= 1
    p = _abi.make_params("euler", (16,))
    h = C.c_void_p()
    import sharpclaw_amd
    rc = L.sc_create(C.byref(p), C.byref(h))
    if rc == 0:  # a GPU is present: clean up, nothing else to check here
        L.sc_destroy(h)
        pytest.skip("device present")
    assert rc == sharpclaw_amd._abi.SC_ERR_HIP
    assert b"no HIP device" in L.sc_last_error()


def test_invalid_options_rejected_before_device_work():
    """Option checks (the reference's stop paths) run before any HIP call."""
    L = lib()
    cases = [
        dict(physics="advection", nx=(10,), mthbc=(3, 3, 1, 1)),   # wall needs component 2
        dict(physics="euler", nx=(10,), mthbc=(2, 1, 1, 1)),       # periodic on one side only
        dict(physics="euler", nx=(10,), time_integrator=7),
        dict(physics="advection", nx=(10, 10), maux=2, char_decomp=1),  # 2D advection: weno5 only
        dict(physics="euler", nx=(10, 10), lim_type=1, char_decomp=2, mthlim=(1, 1, 1, 0)),  # 4 fields
        dict(physics="euler", nx=(10,), lim_type=1, char_decomp=0),  # tvd2: undefined first slope
        dict(physics="euler", nx=(10,), lim_type=1, char_decomp=2, mthlim=(9, 1, 1, 1)),
        dict(physics="euler", nx=(10,), lim_type=0),                 # q2qlqr_poly
    ]
    for c in cases:
        p = _abi.make_params(c.pop("physics"), c.pop("nx"), **c)
        h = C.c_void_p()
        rc = L.sc_create(C.byref(p), C.byref(h))
        assert rc in (_abi.SC_ERR_ARG, _abi.SC_ERR_UNSUPPORTED), (c, rc)


def test_batch_options_rejected_before_device_work():
    L = lib()
    p2 = _abi.make_params("euler", (8, 8))
    mx = (C.c_int * 2)(16, 32)
    lo = (C.c_double * 2)(0.0, 0.0)
    hi = (C.c_double * 2)(1.0, 1.0)
    h = C.c_void_p()
    assert L.sc_batch_create(C.byref(p2), 2, mx, lo, hi, C.byref(h)) == _abi.SC_ERR_UNSUPPORTED
    p1 = _abi.make_params("shallow_water", 16)
    bad = (C.c_int * 2)(16, 2)
    assert L.sc_batch_create(C.byref(p1), 2, bad, lo, hi, C.byref(h)) == _abi.SC_ERR_ARG
    assert b"mx" in L.sc_last_error()
    assert L.sc_batch_create(C.byref(p1), 0, mx, lo, hi, C.byref(h)) == _abi.SC_ERR_ARG
    assert not h.value


def test_server_arguments_checked_before_device_work():
    L = lib()
    h = C.c_void_p()
    assert L.sc_server_create(0, 1 << 20, 0, C.byref(h)) == _abi.SC_ERR_ARG
    assert L.sc_server_create(0, 16, 4, C.byref(h)) == _abi.SC_ERR_ARG
    assert L.sc_server_create(0, 1 << 20, 4, None) == _abi.SC_ERR_ARG
    assert not h.value


def test_struct_layouts_match_header():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "sharpclaw_amd.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(sc_params), offsetof(sc_params, xlower),
         offsetof(sc_params, physics), offsetof(sc_params, device), sizeof(sc_run_ctl),
         sizeof(sc_run_stats));
  printf("%zu %zu\n", offsetof(sc_run_stats, stage_evals), offsetof(sc_run_stats, gpu_ms));
  printf("%zu %zu %zu %zu\n", sizeof(sc_batch_result), offsetof(sc_batch_result, buffer),
         offsetof(sc_batch_result, stage_evals), offsetof(sc_batch_result, done_ms));
  printf("%zu %zu %zu %zu %zu\n", sizeof(sc_request_stats), offsetof(sc_request_stats, submit_ms),
         offsetof(sc_request_stats, rounds), sizeof(sc_server_stats), offsetof(sc_server_stats, clock_ms));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "t.c"), "w") as f:
            f.write(src)
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", os.path.join(d, "t"),
                               os.path.join(d, "t.c")])
        out = subprocess.check_output([os.path.join(d, "t")], text=True).split()
    P, S = _abi.ScParams, _abi.ScRunStats
    want = [C.sizeof(P), P.xlower.offset, P.physics.offset, P.device.offset,
            C.sizeof(_abi.ScRunCtl), C.sizeof(S), S.stage_evals.offset, S.gpu_ms.offset]
    B = _abi.ScBatchResult
    want += [C.sizeof(B), B.buffer.offset, B.stage_evals.offset, B.done_ms.offset]
    R, V = _abi.ScRequestStats, _abi.ScServerStats
    want += [C.sizeof(R), R.submit_ms.offset, R.rounds.offset, C.sizeof(V), V.clock_ms.offset]
    assert [int(x) for x in out] == want
