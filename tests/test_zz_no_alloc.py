"""GPU test of the allocation-free entry point gm_norm2_noalloc against gm_norm2
(same reduction, partials in the library's static device array instead of a
caller workspace): two calls in a row, the second on a smaller vector, each
equal to gm_norm2's result on the same input."""
import pytest


@pytest.mark.gpu
def test_norm2_noalloc_matches_norm2_on_two_calls():
    import torch
    from geowarp_amd import _lib as L
    lib = L.lib()
    f = lib.gm_norm2_noalloc
    dev = torch.device('cuda', 0)
    stream = L.c_void_p(torch.cuda.current_stream().cuda_stream)
    g = torch.Generator(device='cpu').manual_seed(11)
    results = []
    for n in (3_000_001, 40_000):  # the second call is smaller: the 1024 partials of the first stay behind
        v = torch.randn(n, generator=g, dtype=torch.float64).to(dev)
        part = torch.zeros(2048, dtype=torch.float64, device=dev)
        ref = torch.zeros(1, dtype=torch.float64, device=dev)
        out = torch.full((1,), float('nan'), dtype=torch.float64, device=dev)
        L.check(lib.gm_norm2(n, L.ptr(v), L.ptr(part), L.ptr(ref), stream), 'gm_norm2')
        L.check(f(n, L.ptr(v), L.ptr(out), stream), 'gm_norm2_noalloc')
        torch.cuda.synchronize()
        a, b = float(out.item()), float(ref.item())
        exact = float((v.double() ** 2).sum().item())
        print('n=%d noalloc=%r norm2=%r bitwise=%s torch=%r' % (n, a, b, a == b, exact))
        assert abs(a - b) <= 1e-14 * abs(b)
        assert abs(a - exact) <= 1e-12 * abs(exact)
        results.append(a)
    assert results[0] != results[1]


@pytest.mark.gpu
def test_norm2_noalloc_rejects_bad_sizes():
    import torch
    from geowarp_amd import _lib as L
    lib = L.lib()
    f = lib.gm_norm2_noalloc
    stream = L.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = torch.zeros(1, dtype=torch.float64, device='cuda:0')
    assert f(-1, None, L.ptr(out), stream) != 0
    assert f((1 << 40) + 1, L.ptr(out), L.ptr(out), stream) != 0
    assert 'n =' in lib.gm_last_error().decode()
    # an empty vector has norm 0
    assert f(0, None, L.ptr(out), stream) == 0
    torch.cuda.synchronize()
    assert float(out.item()) == 0.0
