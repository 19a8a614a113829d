"""tx_frontend_stats_noalloc: Equalize's statistics with the library's fixed
scratch array instead of a caller's work buffer.  Through the C ABI its
offset and gain are bit-identical to tx_frontend_stats's, on two calls in a
row where the second has fewer channels and samples (the scratch keeps the
first call's sums, so the second shows that they are cleared), and sizes
past the fixed maximum are refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from transientx_amd import _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _stats_work(L, x, td, fd, stream):
    import torch

    ns, nch = x.shape
    nco = nch // fd
    off = torch.empty(nco, dtype=torch.float32, device=x.device)
    gain = torch.empty_like(off)
    nb = L.tx_frontend_stats_work_bytes(nch, fd)
    work = torch.empty(nb, dtype=torch.uint8, device=x.device)
    _lib.check(L.tx_frontend_stats(stream, x.data_ptr(), ns, nch, td, fd, off.data_ptr(), gain.data_ptr(),
                                   work.data_ptr(), nb))
    return off, gain


def _stats_noalloc(L, x, td, fd, stream):
    import torch

    ns, nch = x.shape
    nco = nch // fd
    off = torch.empty(nco, dtype=torch.float32, device=x.device)
    gain = torch.empty_like(off)
    _lib.check(L.tx_frontend_stats_noalloc(stream, x.data_ptr(), ns, nch, td, fd, off.data_ptr(),
                                           gain.data_ptr()))
    return off, gain


def test_noalloc_stats_match_work_buffer_entry():
    import torch

    L = _lib.lib()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(77)
    cases = [(4099, 4096, 1, 1), (1001, 130, 2, 2)]  # the second call is smaller
    for ns, nch, td, fd in cases:
        say("case noalloc stats ns", ns, "nchans", nch, "td", td, "fd", fd)
        x = rng.normal(128, 20, size=(ns, nch)).clip(0, 255).round().astype(np.uint8)
        x[:, 1] = 200  # a constant channel: gain 1
        g = torch.from_numpy(x).cuda()
        o1, g1 = _stats_noalloc(L, g, td, fd, stream)
        o2, g2 = _stats_work(L, g, td, fd, stream)
        torch.cuda.synchronize()
        a1, b1 = o1.cpu().numpy(), g1.cpu().numpy()
        a2, b2 = o2.cpu().numpy(), g2.cpu().numpy()
        assert np.array_equal(a1.view(np.uint32), a2.view(np.uint32)), (ns, nch)
        assert np.array_equal(b1.view(np.uint32), b2.view(np.uint32)), (ns, nch)
        assert np.isfinite(a1).all() and (b1 > 0).all()


def test_noalloc_stats_refuses_more_than_its_scratch():
    import torch

    L = _lib.lib()
    dummy = torch.zeros(16, dtype=torch.uint8, device="cuda")
    p = dummy.data_ptr()
    # 65537 output channels: refused from the arguments alone, nothing launched
    rc = L.tx_frontend_stats_noalloc(None, p, 1, 65537, 1, 1, p, p)
    assert rc == _lib.TX_EINVAL
    assert b"65536" in L.tx_last_error()
