"""rrtmg_hip_sw_noalloc, the SW forward that allocates nothing, called through
the C ABI: its outputs must be bitwise those of rrtmg_hip_sw with device
pointers (the same kernels on a workspace carved from the library's fixed
device array), on two calls in a row with the second smaller, in FP64 and
FP32-mixed; at the largest guaranteed size; with no growth of device memory
on a fresh stream; and a call beyond the array's capacity is refused with
RRTMG_HIP_EINVAL, leaving the outputs untouched."""
import ctypes

import pytest

from mima_amd.columns import make_columns

pytestmark = pytest.mark.gpu

EINVAL = -22
MAX_CELLS = 262144  # RRTMG_HIP_NOALLOC_MAX_CELLS


@pytest.fixture(scope="module")
def rr():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from mima_amd import rrtmg
    rrtmg.rrtmg_sw_ini()
    yield rrtmg
    rrtmg.set_precision("fp64")


def _outs(d):
    import torch
    n, L = d["ncol"], d["nlay"]
    return [torch.full((m, n), float("nan"), dtype=torch.float64, device="cuda") for m in (L + 1, L + 1, L, L + 1,
                                                                                           L + 1, L)]


def _noalloc(rr, d, stream=None, outs=None):
    """rrtmg_hip_sw_noalloc on device columns d; returns (rc, outputs)."""
    import torch
    n, L = d["ncol"], d["nlay"]
    outs = _outs(d) if outs is None else outs
    ptr = lambda t: ctypes.cast(ctypes.c_void_p(t.data_ptr()), ctypes.POINTER(ctypes.c_double))  # noqa: E731
    st = torch.cuda.current_stream() if stream is None else stream
    rc = rr.lib().rrtmg_hip_sw_noalloc(
        int(n), int(L), 0, 0, *[ptr(d[k]) for k in rr.SW_KEYS], ctypes.c_double(d["adjes"]), int(d["dyofyr"]),
        ctypes.c_double(d["scon"]), *[ptr(o) for o in outs], ctypes.c_void_p(st.cuda_stream))
    return rc, dict(zip(("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc"), outs))


def _same(rr, d, got):
    import torch
    ref = rr.sw_columns(d)
    torch.cuda.synchronize()
    for k, v in ref.items():
        assert torch.equal(got[k], v), k


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_matches_sw_on_two_calls_second_smaller(rr, prec):
    import torch
    rr.set_precision(prec)
    big = rr.to_device(make_columns(64, 32, 40, seed=21))
    small = rr.to_device(make_columns(12, 7, 26, seed=22))
    print("%s: noalloc %d x %d then %d x %d" % (prec, big["ncol"], big["nlay"], small["ncol"], small["nlay"]),
          flush=True)
    rc, a = _noalloc(rr, big)
    assert rc == 0, rr.lib().rrtmg_hip_last_error()
    rc, b = _noalloc(rr, small)
    assert rc == 0, rr.lib().rrtmg_hip_last_error()
    torch.cuda.synchronize()
    _same(rr, big, a)
    _same(rr, small, b)


def test_noalloc_largest_guaranteed_size(rr):
    """ncol x (nlay + 1) = RRTMG_HIP_NOALLOC_MAX_CELLS in FP64 (the largest
    workspace per cell); the band kernel's scratch rows then go through
    several launches."""
    import torch
    rr.set_precision("fp64")
    d = rr.to_device(make_columns(64, 64, 63, seed=23))
    assert d["ncol"] * (d["nlay"] + 1) == MAX_CELLS
    print("fp64: noalloc %d x %d" % (d["ncol"], d["nlay"]), flush=True)
    rc, got = _noalloc(rr, d)
    assert rc == 0, rr.lib().rrtmg_hip_last_error()
    torch.cuda.synchronize()
    _same(rr, d, got)


def test_noalloc_allocates_no_device_memory(rr):
    """On a stream the library has never seen, rrtmg_hip_sw would allocate a
    workspace of about 430 MB for these columns (its scratch rows in one
    launch); the no-allocation call on such a stream leaves the device's free
    memory unchanged (outputs allocated beforehand)."""
    import torch
    rr.set_precision("fp64")
    d = rr.to_device(make_columns(64, 48, 60, seed=24))
    outs = _outs(d)
    st = torch.cuda.Stream()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    rc, got = _noalloc(rr, d, stream=st, outs=outs)
    assert rc == 0, rr.lib().rrtmg_hip_last_error()
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    print("free device memory before %d, after %d bytes" % (free0, free1))
    assert free0 - free1 < (4 << 20), (free0, free1)
    _same(rr, d, got)


def test_noalloc_rejects_size_beyond_its_array(rr):
    import torch
    rr.set_precision("fp64")
    d = rr.to_device(make_columns(128, 64, 60, seed=25))
    assert d["ncol"] * (d["nlay"] + 1) > MAX_CELLS
    rc, got = _noalloc(rr, d)
    torch.cuda.synchronize()
    msg = rr.lib().rrtmg_hip_last_error().decode()
    print(msg)
    assert rc == EINVAL and "no-allocation" in msg
    for v in got.values():
        assert bool(torch.isnan(v).all())
