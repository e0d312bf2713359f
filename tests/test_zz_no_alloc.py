"""tn_svd_f64_noalloc (the SVD entry point that allocates nothing) against
tn_svd_f64 through the C ABI: identical bits on two calls in a row, the
second smaller; the library's memory pool is untouched; a shape beyond the
stated maximum is rejected with a clear error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _stats(tn):
    v = [ctypes.c_size_t() for _ in range(4)]
    tn.call("tn_mem_stats", *[ctypes.byref(x) for x in v])
    return [x.value for x in v]


@pytest.mark.parametrize("with_u", [True, False])
def test_svd_noalloc_matches_allocating_entry(tn, with_u):
    rng = np.random.default_rng(77)
    for M, N in ((128, 64), (40, 24)):          # the second call is smaller
        print("svd noalloc %dx%d with_u=%s" % (M, N, with_u), flush=True)
        a = rng.uniform(-1, 1, (M, N))
        k = min(M, N)
        dA = tn.DeviceArray.from_host(a)
        outs = []
        for fn in ("tn_svd_f64", "tn_svd_f64_noalloc"):
            dU = tn.DeviceArray((M, k), zero=True) if with_u else None
            dS, dV = tn.DeviceArray((k,), zero=True), tn.DeviceArray((k, N), zero=True)
            tn.sync()
            before = _stats(tn)
            tn.call(fn, tn._p(dA), M, N, tn._p(dU) if with_u else None, tn._p(dS), tn._p(dV), None)
            after = _stats(tn)
            if fn == "tn_svd_f64_noalloc":
                assert after[:3] == before[:3]     # in use, cached, peak: no pool allocation
            outs.append([x.to_host() for x in ((dU, dS, dV) if with_u else (dS, dV))])
        for x, y in zip(*outs):
            assert x.tobytes() == y.tobytes()
        s = outs[1][1 if with_u else 0]
        assert np.allclose(s, np.linalg.svd(a, compute_uv=False), rtol=0, atol=1e-12 * s[0])


def test_svd_noalloc_rejects_shapes_beyond_its_maximum(tn):
    dA = tn.DeviceArray((130 * 64,), zero=True)
    dS, dV = tn.DeviceArray((64,)), tn.DeviceArray((64 * 64,))
    with pytest.raises(tn.TNError, match="exceeds the maximum"):
        tn.call("tn_svd_f64_noalloc", tn._p(dA), 130, 64, None, tn._p(dS), tn._p(dV), None)
    with pytest.raises(tn.TNError, match="exceeds the maximum"):
        tn.call("tn_svd_f64_noalloc", tn._p(dA), 65, 65, None, tn._p(dS), tn._p(dV), None)
