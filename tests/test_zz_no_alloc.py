"""GPU: mitos_cic_overdensity_noalloc (no allocation, fixed device scratch)
against mitos_cic_overdensity, on two calls in a row (the second smaller),
on the null stream and on a caller stream; sizes beyond the fixed scratch
are rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _positions(rng, n, L):
    import torch
    return [torch.from_numpy(rng.uniform(-0.1 * L, 1.1 * L, n).astype(np.float32)).cuda()
            for _ in range(3)]


def test_noalloc_cic_matches_allocating_entry_point(gpu):
    import torch
    lib = gpu.lib()
    rng = np.random.default_rng(77)
    stream = torch.cuda.Stream()
    for N, L, n in ((128, 300.0, 3_000_000), (32, 50.0, 40_000)):
        print("case noalloc cic", N, L, n, flush=True)
        X, Y, Z = _positions(rng, n, L)
        ref = torch.empty((N, N, N), dtype=torch.float64, device="cuda")
        got = torch.empty_like(ref)
        got2 = torch.empty_like(ref)
        torch.cuda.synchronize()
        assert lib.mitos_cic_overdensity(N, L, n, X.data_ptr(), Y.data_ptr(), Z.data_ptr(),
                                         ref.data_ptr()) == 0
        assert lib.mitos_cic_overdensity_noalloc(N, L, n, X.data_ptr(), Y.data_ptr(),
                                                 Z.data_ptr(), got.data_ptr(), None) == 0
        assert torch.equal(got, ref)
        assert lib.mitos_cic_overdensity_noalloc(N, L, n, X.data_ptr(), Y.data_ptr(),
                                                 Z.data_ptr(), got2.data_ptr(),
                                                 stream.cuda_stream) == 0
        stream.synchronize()
        assert torch.equal(got2, ref)
    X, Y, Z = _positions(rng, 1000, 10.0)
    out = torch.empty((160, 160, 160), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert lib.mitos_cic_overdensity_noalloc(160, 10.0, 1000, X.data_ptr(), Y.data_ptr(),
                                             Z.data_ptr(), out.data_ptr(), None) == 8
