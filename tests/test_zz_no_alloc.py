"""gpf_reduce_noalloc, the no-allocation variant of gpf_reduce (partials in the library's
static device array, result in caller device memory, no synchronisation), against
gpf_reduce on two calls in a row, the second with a smaller input: the results must be
bit-identical, since both run the same kernels over the same fixed partition."""
import numpy as np
import pytest

import gopf_amd as G

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

KEYS = ("sum", "sum_sq", "min", "max", "nonfinite")


def _noalloc(d, n, prec):
    out = torch.full((5,), -7.0, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    G.ReduceNoAlloc(d.data_ptr(), n, prec, 0, stream, out.data_ptr())
    torch.cuda.synchronize()
    return dict(zip(KEYS, out.cpu().numpy().tolist()))


@pytest.mark.parametrize("prec,dtype", [("fp64", np.float64), ("fp32", np.float32)])
def test_noalloc_matches_gpf_reduce_on_two_calls(prec, dtype):
    rng = np.random.default_rng(11)
    for n in (3_000_001, 70_000):      # the second call is smaller: stale partials must not leak in
        print(f"gpf_reduce_noalloc {prec} n={n}", flush=True)
        x = (rng.standard_normal(n) * 2.0 - 0.25).astype(dtype)
        x[[5, n - 2]] = [np.nan, np.inf]
        d = torch.from_numpy(x).to("cuda")
        ref = G.Reduce(d.data_ptr(), n, prec)
        got = _noalloc(d, n, prec)
        for k in KEYS:
            assert got[k] == ref[k], (k, got[k], ref[k])
        assert int(got["nonfinite"]) == 2


def test_noalloc_refuses_more_than_its_maximum():
    d = torch.zeros(4, dtype=torch.float64, device="cuda")
    out = torch.zeros(5, dtype=torch.float64, device="cuda")
    with pytest.raises(G.GpfError) as ei:
        G.ReduceNoAlloc(d.data_ptr(), (1 << 28) + 1, "fp64", 0, 0, out.data_ptr())
    assert ei.value.code == G.GPF_EINVAL and "maximum" in str(ei.value)


@pytest.mark.parametrize("n", [-1, (1 << 28) + 1])
def test_noalloc_argument_checks_precede_the_device(n):
    """Sizes the static array cannot serve are refused before any device call."""
    import ctypes
    L = G.lib()
    buf = (ctypes.c_double * 8)()
    assert L.gpf_reduce_noalloc(buf, n, G.GPF_FP64, 0, None, buf) == G.GPF_EINVAL
