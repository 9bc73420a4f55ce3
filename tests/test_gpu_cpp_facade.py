"""The C++ drop-in facade (include/viennals_gpu/lsAdvect.hpp) used like the
reference API by tests/cpp/advect_trench.cpp; its result equals the oracle's
on the same initial band."""
import os
import subprocess
import sys

import numpy as np
import pytest

import cases
import pyoracle as po

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def read_runs(path, D=3):
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(8), np.int64)[0])
        idx = np.frombuffer(f.read(4 * D * n), np.int32).reshape(n, D)
        val = np.frombuffer(f.read(8 * n), np.float64)
    return idx, val


def test_cpp_facade_trench(tmp_path):
    print("[gpu-case] C++ facade: trench WENO5/RK3 T=2", flush=True)
    exe = os.path.join(ROOT, "tests", "cpp", "advect_trench")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    prefix = str(tmp_path / "trench")
    out = subprocess.run([exe, prefix, "2.0"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    i0, v0 = read_runs(prefix + "_initial.bin")
    i1, v1 = read_runs(prefix + "_final.bin")
    o = po.Domain(3, [-32, -32, -32], [32, 32, 32], [0, 0, 1], 1.0)
    o.set(i0, v0, 2)
    assert o.check()[0] == 0
    p = po.make_params(spatial=11, temporal=2, cfl=0.4, time=2.0, vel=cases.VEL_ETCH_DIRECTIONAL)
    st, r = po.advect(o, p)
    io, vo = o.get()
    assert cases.bit_equal(io, vo, i1, v1), cases.first_diff(io, vo, i1, v1)
    assert f"steps {r['steps']}" in out.stdout


def test_cpp_facade_user_velocity_field(tmp_path):
    """A VelocityField subclass written as reference code writes it runs
    through the host-callback path and gives the built-in functor's result."""
    print("[gpu-case] C++ facade: trench with a user VelocityField subclass", flush=True)
    exe = os.path.join(ROOT, "tests", "cpp", "advect_trench")
    outs = {}
    for mode in ("builtin", "host"):
        prefix = str(tmp_path / mode)
        out = subprocess.run([exe, prefix, "1.0", mode], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        outs[mode] = (read_runs(prefix + "_final.bin"), out.stdout)
    (ib, vb), sb = outs["builtin"]
    (ih, vh), sh = outs["host"]
    assert sb == sh
    assert np.array_equal(ib, ih) and np.array_equal(vb, vh)


def test_cpp_facade_velocity_update_callback(tmp_path):
    """Advect::setVelocityUpdateCallback through the C++ facade: a lambda
    that lowers the built-in functor's constant between the RK stages, equal
    to the oracle changing its velocity in the same calls."""
    print("[gpu-case] C++ facade: trench with a velocity update callback", flush=True)
    exe = os.path.join(ROOT, "tests", "cpp", "advect_trench")
    prefix = str(tmp_path / "callback")
    out = subprocess.run([exe, prefix, "1.0", "callback"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    i0, v0 = read_runs(prefix + "_initial.bin")
    i1, v1 = read_runs(prefix + "_final.bin")
    o = po.Domain(3, [-32, -32, -32], [32, 32, 32], [0, 0, 1], 1.0)
    o.set(i0, v0, 2)
    calls = []

    def cb(view, p):
        calls.append(1)
        p.velB -= 0.01
    p = po.make_params(spatial=11, temporal=2, cfl=0.4, time=1.0, vel=cases.VEL_ETCH_DIRECTIONAL, velocity_update=cb)
    st, r = po.advect(o, p)
    assert st == 0
    io, vo = o.get()
    assert cases.bit_equal(io, vo, i1, v1), cases.first_diff(io, vo, i1, v1)
    assert f"steps {r['steps']}" in out.stdout and f"callbacks {len(calls)}" in out.stdout
    assert len(calls) == 2 * r["steps"]
