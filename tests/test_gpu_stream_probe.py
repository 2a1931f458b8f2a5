"""The streaming probe (femtc_stream_probe) that measures the practical HBM
bandwidth computes what it says: dst = sum of the source arrays, and with no
destination one partial sum per workgroup whose total is the sum of all."""
import numpy as np
import pytest

from femtc_amd import _lib

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _probe(nr, nw, src, dst, n, sink, wg=0):
    return _lib.lib().femtc_stream_probe(
        nr, nw, src.data_ptr(), 0 if dst is None else dst.data_ptr(), n,
        0 if sink is None else sink.data_ptr(),
        0 if sink is None else sink.numel(), wg,
        torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("nr,wg", [(1, 0), (5, 1), (8, 8)])
@pytest.mark.parametrize("n", [2, 1000, 1 << 20])
def test_stream_probe_sums(nr, wg, n):
    g = torch.Generator(device="cpu").manual_seed(nr * 7 + n)
    srch = torch.randint(-8, 8, (nr * n,), generator=g).double()
    src = srch.cuda()
    dst = torch.empty(n, device="cuda", dtype=torch.float64)
    _lib.check(_probe(nr, 1, src, dst, n, None, wg), "stream_probe")
    torch.cuda.synchronize()
    assert torch.equal(dst.cpu(), srch.view(nr, n).sum(0))
    sink = torch.zeros(1 << 14, device="cuda", dtype=torch.float64)
    _lib.check(_probe(nr, 0, src, None, n, sink, wg), "stream_probe")
    torch.cuda.synchronize()
    assert float(sink.sum()) == float(srch.sum())


def test_stream_probe_rejects_bad_arguments():
    src = torch.zeros(16, device="cuda", dtype=torch.float64)
    for args in ((0, 1, src, src, 8, None), (9, 1, src, src, 8, None),
                 (1, 1, src, src, 7, None), (1, 0, src, None, 8, None),
                 (1, 2, src, src, 8, None), (1, 1, src, src, 8, None, -1),
                 (1, 1, src, src, 8, None, 65)):
        assert _probe(*args) == _lib.ERR_INVALID_ARGUMENT
