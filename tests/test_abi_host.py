"""
CPU tests of the boundary and host logic: the C-ABI library loads and exports
every entry point include/dadi_hip.h declares, the struct layouts agree, and
the host-side schedule reproduces dadi's Python time loops exactly.  No GPU
calls are made.
"""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "dadi_hip.h")).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(dh_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    from dadi_amd import _lib
    L = _lib.load()
    declared = _declared()
    assert len(declared) >= 12
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (dh_\w+)", out))
    for name in declared:
        assert name in exported, name
        assert hasattr(L, name)
    assert set(_lib.EXPORTED) <= set(declared)


def test_struct_layout_and_errors_without_gpu():
    from dadi_amd import _lib
    L = _lib.load()
    assert L.dh_sizeof_segment() == _lib.SEG_DTYPE.itemsize == 336
    assert L.dh_sizeof_instance() == _lib.INST_DTYPE.itemsize
    assert L.dh_sizeof_sweep_item() == _lib.ITEM_DTYPE.itemsize == 112
    assert L.dh_abi_version() == _lib.ABI_VERSION == 6
    # calls before dh_init fail cleanly with a message (no GPU touched)
    cfg = _lib.DhLaunch(1, 1, 0, 0, 0)
    rc = L.dh_integrate_nd(None, 1, 2, 0, 10, None, cfg, None)
    assert rc == -1
    assert "dh_init" in _lib.last_error()
    assert L.dh_lds_bytes(2, 100, 1, 0, 0, 0, 0) == 8 * 100 * (4 + 6 + 64)
    assert L.dh_lds_bytes(2, 100, 2, 8, 0, 0, 0) == 8 * (100 * 10 + 2 * 64 * (13 + 9))
    assert L.dh_lds_bytes(2, 100, 2, 8, 1, 0, 0) == 8 * (100 * 10 + 2 * 64 * (26 + 9))
    # gamma columns in a global workspace: tables and staging buffers only
    assert L.dh_lds_bytes(2, 100, 2, 0, 0, 1, 0) == 8 * (100 * 10 + 2 * 64 * 9)
    # the delta trick adds V at the interval midpoints per axis
    assert L.dh_lds_bytes(2, 100, 1, 0, 0, 0, 1) == 8 * 100 * (4 + 8 + 64)
    assert ctypes.sizeof(_lib.DhLaunch) == 48
    assert L.dh_nd_groups(2, 100, 2, 0, 0, 1, 0, 10) == -1   # before dh_init


def test_const_steps_match_python_loop():
    from dadi_amd import schedule as S
    rng = np.random.default_rng(0)
    for _ in range(3000):
        t0 = float(rng.uniform(0, 1) * rng.integers(0, 2))
        T = t0 + float(rng.uniform(1e-4, 2))
        dt = 1e-3 / float(rng.uniform(0.025, 30))
        nf, tails = S.const_steps(t0, T, dt)
        t, seq = t0, []
        while t < T:
            d = min(dt, T - t)
            seq.append(d)
            t += d
        assert seq[:nf] == [dt] * nf and seq[nf:] == tails


def test_schedule_matches_oracle_driver(oracle):
    """Injected values, dt and parameters of the time-varying schedule equal
    what the oracle's restated loop uses, step by step."""
    from dadi_amd import schedule as S
    from dadi_amd.Misc import ensure_1arg_func as f
    xx = oracle.default_grid(25)
    T = 0.05
    nu_f = [f(lambda t: 0.5 + 10 * t), f(2.0)]
    m_f = [[f(lambda t: 1 + t)], [f(0.3)]]
    g_f = [f(-1.0), f(lambda t: 3 * t)]
    h_f = [f(0.5), f(0.2)]
    seg = S.tv_segments(2, xx, 0.0, T, nu_f, m_f, g_f, h_f, f(1.3), None)
    # replay the oracle's loop
    t = 0.0
    steps = []
    nu = [fn(t) for fn in nu_f]
    mm = [[fn(t) for fn in fs] for fs in m_f]
    ga = [fn(t) for fn in g_f]
    hh = [fn(t) for fn in h_f]
    while t < T:
        dt = min(oracle.compute_dt(nu[k], mm[k], ga[k], hh[k]) for k in range(2))
        this = min(dt, T - t)
        nt = t + this
        nu = [fn(nt) for fn in nu_f]
        mm = [[fn(nt) for fn in fs] for fs in m_f]
        ga = [fn(nt) for fn in g_f]
        hh = [fn(nt) for fn in h_f]
        steps.append((this, oracle.inject_value(2, this, xx, 1.3), nu[0], mm[0][0], ga[1]))
        t = nt
    expanded = np.repeat(seg, seg["nsteps"])
    assert len(expanded) == len(steps)
    for rec, (d, inj, nu0, m0, g1) in zip(expanded, steps):
        assert rec["dt"] == d and rec["inject"] == inj
        assert rec["nu"][0] == nu0 and rec["mig"][0][0] == m0 and rec["gamma"][1] == g1


def test_validation_errors_raise_before_launch():
    import dadi_amd
    from dadi_amd import Integration as I
    xx = dadi_amd.Numerics.default_grid(10)
    phi = np.ones((10, 10))
    with pytest.raises(ValueError):
        I.two_pops(phi, xx, 0.1, nu1=-1.0)
    with pytest.raises(ValueError):
        I.two_pops(phi, xx, 0.1, m12=1.0, frozen1=True)
    with pytest.raises(ValueError):
        I.two_pops(phi, xx, 0.1, initial_t=0.2)
    with pytest.raises(ValueError):
        I.two_pops(phi, xx, 0.1, nu2=0.0)
    with pytest.raises(ValueError):
        I.one_pop(np.ones(10), xx, 0.1, nu=lambda t: -1.0)
    # nothing to do: returned unchanged without touching the device
    assert np.array_equal(I.two_pops(phi, xx, 0.0), phi)
    assert np.array_equal(I.one_pop(np.ones(10), xx, 0.3, frozen=True), np.ones(10))


def test_disabled_backend_raises_loudly():
    import dadi_amd
    from dadi_amd import Integration as I, _lib
    xx = dadi_amd.Numerics.default_grid(10)
    prev = I.cuda_enabled
    try:
        dadi_amd.cuda_enabled(False)
        assert dadi_amd.cuda_enabled() is False
        with pytest.raises(_lib.BackendUnavailable):
            I.two_pops(np.ones((10, 10)), xx, 0.1)
    finally:
        I.cuda_enabled = prev


def test_no_gpu_means_no_silent_fallback():
    """Without a GPU the integrators raise instead of computing on the CPU."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    import dadi_amd
    from dadi_amd import _lib
    xx = dadi_amd.Numerics.default_grid(10)
    with pytest.raises(_lib.BackendUnavailable):
        dadi_amd.Integration.two_pops(np.ones((10, 10)), xx, 0.1)
    assert dadi_amd.cuda_enabled(True) is False
    dadi_amd.Integration.cuda_enabled = True


def test_product_package_never_imports_oracle():
    pkg = os.path.join(ROOT, "dadi_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".h", ".c", ".cpp")):
                text = open(os.path.join(dirpath, fn)).read()
                assert "oracle" not in text.replace("the oracle", ""), fn


def test_one_pop_x_argument_errors():
    """dadi/Integration.py:1285-1318: errors come before any GPU work."""
    import dadi_amd
    xx = np.linspace(0, 1, 20)
    phi = np.ones(20)
    with pytest.raises(NotImplementedError):
        dadi_amd.Integration.one_pop_X(phi, xx, 0.1, nu=lambda t: 1.0)
    with pytest.raises(ValueError):
        dadi_amd.Integration.one_pop_X(phi, xx, 0.1, beta=0)
    with pytest.raises(ValueError):
        dadi_amd.Integration.one_pop_X(phi, xx, 0.1, alpha=-1)
    with pytest.raises(ValueError):
        dadi_amd.Integration.one_pop_X(phi, xx, 0.1, initial_t=0.2)
    out = dadi_amd.Integration.one_pop_X(phi, xx, 0.1, frozen=True)
    assert np.array_equal(out, phi) and out is not phi


def test_import_raises_hw_queues_before_gpu_use():
    """Importing the package raises GPU_MAX_HW_QUEUES (4 by default) to 16 so
    that the serving loop's per-class streams get hardware queues of their
    own; a larger value is kept, DADI_AMD_HW_QUEUES=0 leaves it alone and
    other targets are capped at 32."""
    code = "import os, dadi_amd; print(os.environ.get('GPU_MAX_HW_QUEUES'))"
    cases = [({"GPU_MAX_HW_QUEUES": "4"}, "16"), ({}, "16"),
             ({"GPU_MAX_HW_QUEUES": "24"}, "24"),
             ({"GPU_MAX_HW_QUEUES": "4", "DADI_AMD_HW_QUEUES": "0"}, "4"),
             ({"DADI_AMD_HW_QUEUES": "64"}, "32")]
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for extra, want in cases:
        env = {k: v for k, v in os.environ.items()
               if k not in ("GPU_MAX_HW_QUEUES", "DADI_AMD_HW_QUEUES")}
        env.update(extra)
        out = subprocess.run([sys.executable, "-c", code], cwd=root, env=env,
                             capture_output=True, text=True, check=True).stdout.split()
        assert out[-1] == want, (extra, out)
