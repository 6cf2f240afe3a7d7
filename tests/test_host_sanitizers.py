"""The planner and the run-time kernel emitter (host C++, csrc/plan.cpp and
csrc/jit.cpp) built with g++ -fsanitize=address,undefined and run over a
few hundred op lists: templates of the reference (generate_circuit types,
TFIM / XXZ shared parameters), random gate mixes with fSim, fixed gates and
wide shared generators, with and without derivative ops, 2-22 qubits,
random tile options.  Any memory error, leak or undefined behaviour aborts
the driver (no GPU involved)."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from pyramaterised_amd import _lib as L
from tests.helpers import build_pkg, random_spec

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pyramaterised_amd", "csrc")


def _cases():
    import pyramaterised_amd as pyqc
    from pyramaterised_amd import templates as T
    rng = np.random.default_rng(2024)
    out = []

    def add(qc, n):
        for gens in (False, True):
            try:
                ops = qc.ops(with_generators=gens)
            except NotImplementedError:
                continue
            o = L.PlanOpts()
            o.tile_bits = int(rng.choice([0, 0, 6, 10, 11, 12, 13]))
            o.low_bits = int(rng.integers(0, 5))
            o.dtype = int(rng.integers(0, 2))
            o.flags = int(rng.choice([0, L.PQS_PLAN_ROLLING, L.PQS_PLAN_ROLLING | L.PQS_PLAN_NO_EXCHANGE,
                                      L.PQS_PLAN_GREEDY_SWEEPS, L.PQS_PLAN_NO_DIRECT | L.PQS_PLAN_ROLLING]))
            out.append((n, ops, o))

    for kind in ("NPQC", "TFIM", "TFIM_modified", "XXZ", "Circuit_1", "Circuit_2", "Circuit_9", "qg_circuit",
                 "generic_HE", "clifford", "y_CPHASE", "double_y_CPHASE", "fermionic", "zfsim", "fsim",
                 "fixed_fsim"):
        for n in (4, 6, 9, 12, 14):
            add(T.generate_circuit(kind, n, 2, shuffle=False), n)
    for n in (2, 3, 5, 8, 11, 13, 16, 18, 22):
        for _ in range(4):
            qc = pyqc.PQC(n)
            for g in build_pkg(random_spec(n, int(rng.integers(1, 4)), rng), n, pyqc):
                qc.add_layer([g])
            add(qc, n)
    return out


def test_planner_and_emitter_under_sanitizers(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None or not os.path.exists("/opt/rocm/include/hip/hiprtc.h"):
        pytest.skip("g++ or the ROCm headers are not available")
    cases = _cases()
    assert len(cases) > 150
    data = tmp_path / "ops.bin"
    with open(data, "wb") as f:
        for n, ops, o in cases:
            arr = (L.Op * max(1, len(ops)))(*ops)
            f.write(np.array([n, len(ops)], dtype=np.int32).tobytes())
            f.write(bytes(o))
            f.write(bytes(arr)[:len(ops) * L.C.sizeof(L.Op)])
    exe = tmp_path / "plan_sanitize"
    cmd = [gxx, "-std=c++20", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"),
           "-I" + CSRC, os.path.join(ROOT, "tests", "native", "plan_sanitize.cpp"), os.path.join(CSRC, "plan.cpp"),
           os.path.join(CSRC, "jit.cpp"), "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib", "-o", str(exe)]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe), str(data)], capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    words = r.stdout.split()
    counts = dict(zip(words[0::2], (int(w) for w in words[1::2])))
    assert counts["cases"] == len(cases)
    assert counts["planned"] > 0.8 * len(cases) and counts["sources"] > 0
