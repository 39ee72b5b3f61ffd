"""gd_direct_sum_fp64_noalloc (partials in a static device array) against
gd_direct_sum_fp64 (library-allocated partials), through the C-ABI: two
calls in a row, the second smaller, each equal to the allocating entry
point's result and within 1e-12 of the oracle's FP64 sum; and sizes beyond
the static array are refused before any launch."""
import numpy as np
import pytest

from cudagravitydemo_amd import GravityError, ics, load_library
from cudagravitydemo_amd._lib import check
from cudagravitydemo_amd.engine import DeviceBuffer, from_device, to_device

pytestmark = pytest.mark.gpu


def _judge(lib, fn, pos_d, lo_d, n, sample, eps2):
    ns = sample.shape[0]
    s_d = to_device(lib, sample)
    out = DeviceBuffer(lib, 24 * ns)
    try:
        fn(pos_d.ptr, lo_d.ptr if lo_d else None, n, s_d.ptr, ns, eps2, 1.0, out.ptr)
        check(lib)
        return from_device(lib, out.ptr, np.float64, (ns, 3))
    finally:
        s_d.close()
        out.close()


@pytest.mark.parametrize("hilo", [False, True])
def test_noalloc_judge_matches_allocating_judge(hilo):
    lib = load_library()
    lib.gd_set_exit_on_error(0)
    cases = [(200_000, 4096, 11), (30_001, 700, 12)]   # the second call is smaller
    for n, ns, seed in cases:
        print(f"case noalloc judge n={n} ns={ns} hilo={hilo}", flush=True)
        pos, _ = ics.plummer(n, seed=seed)
        lo = (np.random.default_rng(seed).standard_normal((n, 3)) * 1e-8).astype(np.float32) \
            if hilo else None
        sample = ics.subsample(n, ns, seed=seed + 1)
        pos_d = to_device(lib, pos)
        lo_d = to_device(lib, lo) if hilo else None
        try:
            eps2 = 0.01 ** 2
            got = _judge(lib, lib.gd_direct_sum_fp64_noalloc, pos_d, lo_d, n, sample, eps2)
            want = _judge(lib, lib.gd_direct_sum_fp64, pos_d, lo_d, n, sample, eps2)
            rel = np.abs(got - want).max() / np.abs(want).max()
            assert rel <= 1e-13, rel
            # the oracle's FP64 sum over the same bodies (hi + lo folded in)
            chk = np.arange(0, ns, max(1, ns // 64))
            p64 = pos.copy()
            ref_pos = p64.astype(np.float64)
            if hilo:
                ref_pos[:, :3] += lo
            d = ref_pos[sample[chk], None, :3] - ref_pos[None, :, :3]
            r2 = (d * d).sum(-1) + eps2
            f = -ref_pos[None, :, 3] / (r2 * np.sqrt(r2))
            f[np.arange(chk.size), sample[chk]] = 0.0
            exact = (f[..., None] * d).sum(1)
            err = np.abs(got[chk] - exact).max() / np.abs(exact).max()
            assert err <= 1e-12, err
        finally:
            pos_d.close()
            if lo_d:
                lo_d.close()


def test_noalloc_judge_refuses_too_many_samples():
    lib = load_library()
    lib.gd_set_exit_on_error(0)
    n, ns = 1024, (1 << 19) + 1
    pos, _ = ics.uniform_cube(n, seed=3)
    pos_d = to_device(lib, pos)
    s_d = to_device(lib, np.zeros(ns, np.int32))
    out = DeviceBuffer(lib, 24 * ns)
    try:
        lib.gd_direct_sum_fp64_noalloc(pos_d.ptr, None, n, s_d.ptr, ns, 1e-4, 1.0, out.ptr)
        with pytest.raises(GravityError, match="exceed the static scratch"):
            check(lib)
    finally:
        for b in (pos_d, s_d, out):
            b.close()
