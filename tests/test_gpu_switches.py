"""GPU tests of the library's environment switches (each is read once per
process, so each setting runs in a child process of its own, one at a
time): the alternate code paths they select still meet the same bounds as
the defaults -- the mixed LU-IR's FP64 backward error <= 1e-14 on a matrix
tall enough for multi-group tournament leaves, and the FP64 LU / LLT
drop-ins' backward error and pivots."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
import numpy as np, torch
sys.path.insert(0, sys.argv[1])
import eigen_amd as ea
from eigen_amd import mixed
from oracle import pyoracle as po
from tests import inputs
torch.cuda.set_device(0)
out = {}
n = 3000
a = inputs.normal(77, n, n)
b = inputs.normal(78, n)
A = ea.colmajor(torch.from_numpy(a).cuda())
B = torch.from_numpy(b).cuda()
for prec in ("fp16x3", "fp32"):
    x, rep = mixed.gesv_ir(A, B, precision=prec)
    out[prec] = [po.backward_error(a, x.cpu().numpy(), b), rep["converged"]]
m = 700
am = inputs.normal(79, m, m)
lu = ea.PartialPivLU(ea.colmajor(torch.from_numpy(am).cuda()))
t = lu.transpositionsP().cpu().numpy()
_, t_r, _, _, margin = po.lu(am, margins=True)
d = np.nonzero(t != t_r)[0]
out["lu_pivots_ok"] = bool(len(d) == 0 or margin[d[0]] < 1e-12)
bm = inputs.normal(80, m)
out["lu"] = po.backward_error(am, lu.solve(torch.from_numpy(bm).cuda()).cpu().numpy(), bm)
s = inputs.spd(81, m)
f = ea.LLT(ea.colmajor(torch.from_numpy(s).cuda()))
out["llt"] = po.backward_error(s, f.solve(torch.from_numpy(bm).cuda()).cpu().numpy(), bm)
print("RESULT " + json.dumps(out), flush=True)
"""


@pytest.mark.parametrize("env", [
    {"EAMD_LOOKAHEAD": "0"},
    {"EAMD_TOURNAMENT": "0"},
    {"EAMD_FUSE_PANEL": "0"},
    {"EAMD_TSL_WIDE": "0"},
    {"EAMD_LU_NB": "32", "EAMD_LLT_NB": "64"},
    {"EAMD_LU_NB": "256", "EAMD_LLT_NB": "256"},
], ids=lambda e: ",".join(f"{k}={v}" for k, v in e.items()))
def test_switch(gpu, env):
    print("switch", env, flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], env={**os.environ, **env},
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("RESULT ")][-1]
    res = json.loads(line[len("RESULT "):])
    for prec in ("fp16x3", "fp32"):
        be, conv = res[prec]
        assert conv == 1 and be <= 1e-14, (prec, res)
    assert res["lu_pivots_ok"], res
    assert res["lu"] <= 1e-14 and res["llt"] <= 1e-14, res


GIVE_UP_CHILD = r"""
import json, sys
import torch
sys.path.insert(0, sys.argv[1])
import eigen_amd as ea
from tests import inputs
torch.cuda.set_device(0)
n = 4096
lu = ea.PartialPivLU(ea.colmajor(torch.from_numpy(inputs.normal(91, n, n)).cuda()))
out = {}
try:
    lu.solve(torch.from_numpy(inputs.normal(92, n)).cuda())
    out["raised"] = False
except ea.EamdError as e:
    out["raised"], out["msg"] = True, str(e)
torch.cuda.synchronize()
print("RESULT " + json.dumps(out), flush=True)
"""


def test_trsv_wait_gives_up_and_reports(gpu):
    """The triangular-solve chain's bounded wait, forced from the host:
    with EAMD_TRSV_SPIN=0 every workgroup whose predecessor has not yet
    published gives up at its first poll (its kernel returns normally and
    the others see the error word), and the solve reports the time-out as
    an error instead of hanging or returning a partial result."""
    print("trsv give-up: EAMD_TRSV_SPIN=0", flush=True)
    r = subprocess.run([sys.executable, "-c", GIVE_UP_CHILD, ROOT],
                       env={**os.environ, "EAMD_TRSV_SPIN": "0"},
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("RESULT ")][-1]
    res = json.loads(line[len("RESULT "):])
    assert res["raised"] and "timed out" in res["msg"], res


LEAF_GIVE_UP_CHILD = r"""
import json, sys
import torch
sys.path.insert(0, sys.argv[1])
import eigen_amd as ea
from tests import inputs
torch.cuda.set_device(0)
n = 4096
A = ea.colmajor(torch.from_numpy(inputs.normal(93, n, n)).cuda())
out = {}
try:
    ea.PartialPivLU(A)
    out["raised"] = False
except ea.EamdError as e:
    out["raised"], out["msg"] = True, str(e)
torch.cuda.synchronize()
print("RESULT " + json.dumps(out), flush=True)
"""


def test_leaf_exchange_gives_up_and_reports(gpu):
    """The FP64 LU leaf's grid exchange, forced from the host: with
    EAMD_LEAF_SPIN=0 a workgroup gives up at its first poll that finds
    another workgroup's words missing (every kernel returns normally; the
    aborted leaves leave empty permutation lists, so the interchanges after
    them do nothing), and the factorization reports the time-out as an
    error."""
    print("leaf give-up: EAMD_LEAF_SPIN=0", flush=True)
    r = subprocess.run([sys.executable, "-c", LEAF_GIVE_UP_CHILD, ROOT],
                       env={**os.environ, "EAMD_LEAF_SPIN": "0"},
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("RESULT ")][-1]
    res = json.loads(line[len("RESULT "):])
    assert res["raised"] and "timed out" in res["msg"], res
