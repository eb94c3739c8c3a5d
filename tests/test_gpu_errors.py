"""GPU: the error contract of the C ABI (include/rrtmg_hip.h) and its
workspaces.  The reference stops the model on bad input (stop / FATAL); the
library returns negative codes with a message instead and leaves its outputs
untouched, and the library keeps working after a failed call."""
import os
import subprocess
import sys

import numpy as np
import pytest

from mima_amd.columns import make_columns, subset

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def rr():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from mima_amd import rrtmg
    rrtmg.rrtmg_sw_ini()
    rrtmg.rrtmg_lw_ini()
    yield rrtmg
    rrtmg.set_precision("fp64")
    rrtmg.set_checks(False)


def _cols(n=96, seed=3):
    full = make_columns(32, 16, 40, seed=seed)
    return subset(full, np.arange(n) * (full["ncol"] // n))


def test_device_checks_reject_non_finite_and_bad_plev(rr):
    """With checks on, device-pointer calls refuse what host-pointer calls
    refuse, and name the array; outputs stay untouched."""
    import torch
    from mima_amd.rrtmg import RRTMGError
    c = _cols()
    d = rr.to_device(c)
    rr.set_precision("fp64")
    assert rr.set_checks(True) is False
    try:
        good = rr.sw_columns(d)
        for side, key, where in (("sw", "tlay", (3, 2)), ("lw", "o3vmr", (7, 5)), ("sw", "coszen", (4,)),
                                 ("lw", "emis", (2, 1))):
            bad = dict(d)
            bad[key] = d[key].clone()
            bad[key][where] = float("nan")
            with pytest.raises(RRTMGError) as e:
                (rr.sw_columns if side == "sw" else rr.lw_columns)(bad)
            assert e.value.code == -22 and key in str(e.value), (side, key, str(e.value))
        bad = dict(d)
        bad["plev"] = d["plev"].clone()
        bad["plev"][6, 9] = bad["plev"][5, 9] + 1.0
        with pytest.raises(RRTMGError) as e:
            rr.lw_columns(bad)
        assert e.value.code == -22 and "plev" in str(e.value)
        bad["plev"] = d["plev"].clone()
        bad["plev"][-1, 0] = 0.0
        with pytest.raises(RRTMGError) as e:
            rr.sw_columns(bad)
        assert e.value.code == -22
        cot = torch.zeros_like(d["tlay"])
        cot[0, 0] = float("inf")
        with pytest.raises(RRTMGError) as e:
            rr.sw_vjp_columns(d, swhr_bar=cot)
        assert e.value.code == -22 and "swhr_bar" in str(e.value)
        again = rr.sw_columns(d)
        for k in good:
            assert torch.equal(good[k], again[k]), k
    finally:
        rr.set_checks(False)


def test_lw_idrv_and_cloudy_requests_are_refused(rr):
    from mima_amd.rrtmg import LW_KEYS, RRTMGError
    c = _cols(16)
    args = [c[k] for k in LW_KEYS]
    for icld, idrv in ((0, 1), (2, 0)):
        with pytest.raises(RRTMGError) as e:
            rr.rrtmg_lw(c["ncol"], c["nlay"], icld, idrv, *args)
        assert e.value.code == -95


def test_enoinit_before_init_in_a_fresh_process():
    """Every entry point refuses to run before its side's init (ENOINIT),
    in a process that has not initialised the library."""
    child = (
        "import sys, ctypes\n"
        "sys.path.insert(0, %r)\n"
        "import numpy as np\n"
        "from mima_amd import rrtmg\n"
        "from mima_amd.columns import make_columns\n"
        "from mima_amd.rrtmg import RRTMGError\n"
        "c = make_columns(4, 4, 20, seed=1)\n"
        "codes = []\n"
        "for f in (rrtmg.sw_columns, rrtmg.lw_columns, rrtmg.sw_vjp_columns, rrtmg.lw_vjp_columns):\n"
        "    try:\n"
        "        f(c)\n"
        "        codes.append(0)\n"
        "    except RRTMGError as e:\n"
        "        codes.append(e.code)\n"
        "print('CODES', codes)\n"
        "rrtmg.rrtmg_sw_ini()\n"
        "print('AFTER', np.isfinite(rrtmg.sw_columns(c)['swdflx']).all())\n"
    ) % ROOT
    r = subprocess.run([sys.executable, "-c", child], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "CODES [-2, -2, -2, -2]" in r.stdout, r.stdout
    assert "AFTER True" in r.stdout


def _raw_stream():
    """A new HIP stream from the runtime torch uses (torch.cuda.Stream()
    hands out streams from a pool, which earlier calls may have used)."""
    import ctypes
    import torch  # noqa: F401  (loads the HIP runtime first)
    hip = ctypes.CDLL("libamdhip64.so")
    h = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(h)) == 0
    return hip, h


def test_injected_allocation_failure_gives_enomem(rr):
    """A workspace that cannot be allocated fails the call with ENOMEM
    (injected with RRTMG_HIP_FAIL_ALLOC on a new stream, which has no
    workspace yet); the next call on that stream allocates and works."""
    import torch
    from mima_amd.rrtmg import RRTMGError
    c = _cols(64, seed=5)
    d = rr.to_device(c)
    rr.set_precision("mixed")
    ref = rr.lw_columns(d)
    torch.cuda.synchronize()
    hip, st = _raw_stream()
    try:
        os.environ["RRTMG_HIP_FAIL_ALLOC"] = "1"
        try:
            with pytest.raises(RRTMGError) as e:
                rr.lw_columns(d, stream=st.value)
            assert e.value.code == -12 and "allocation" in str(e.value)
        finally:
            del os.environ["RRTMG_HIP_FAIL_ALLOC"]
        ok = rr.lw_columns(d, stream=st.value)
        assert hip.hipStreamSynchronize(st) == 0
        for k in ok:
            assert torch.equal(ok[k], ref[k]), k
    finally:
        torch.cuda.synchronize()
        hip.hipStreamDestroy(st)


def test_calls_on_two_streams_overlap_safely(rr):
    """Workspaces are per stream: forward calls issued on two streams without
    synchronisation between them give exactly the results of the same calls
    made one after the other."""
    import torch
    rr.set_precision("mixed")
    a = rr.to_device(_cols(480, seed=6))
    b = rr.to_device(_cols(480, seed=7))
    ref = [rr.sw_columns(a), rr.lw_columns(b), rr.sw_columns(b), rr.lw_columns(a)]
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    outs = []
    for _ in range(3):
        with torch.cuda.stream(s1):   # outputs allocated on the stream that writes them
            o1 = rr.sw_columns(a)
            o4 = rr.lw_columns(a)
        with torch.cuda.stream(s2):
            o2 = rr.lw_columns(b)
            o3 = rr.sw_columns(b)
        outs.append([o1, o2, o3, o4])
    torch.cuda.synchronize()
    for rnd in outs:
        for r, o in zip(ref, rnd):
            for k in r:
                assert torch.equal(r[k], o[k]), k
