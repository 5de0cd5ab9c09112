"""The allocation-free pursuit (ksvd_omp_noalloc) against ksvd_omp.

Two calls in a row through the C ABI on caller-owned device memory, the
second with a smaller input (fewer patches, a shorter patch and fewer atoms),
each checked bit for bit against the codes ksvd_omp gives on the same patches
and dictionary: both run the same Gram and pursuit kernels.  Sizes beyond the
scratch maximum are refused before anything is enqueued.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def context_codes(kl, seed, W, patch, slide, atoms, k):
    clean, noisy = kl.synth_image(seed, W, W, 20.0)
    ks = kl.KSvd(patch=patch, slide=slide, atoms=atoms, omp_iter=k)
    ks.set_picture(noisy)
    ks.extract()
    ks.init_dictionary()
    ks.omp()
    out = ks.patches(), ks.dictionary(), ks.codes()
    ks.close()
    return out


def noalloc_codes(kl, Y, D, k, stream):
    P, dim = Y.shape
    K = D.shape[0]
    dev = torch.device("cuda:0")
    Yd = torch.from_numpy(Y).to(dev)
    Dd = torch.from_numpy(D).to(dev)
    idx = torch.full((P, k), -7, dtype=torch.int32, device=dev)
    coef = torch.full((P, k), 7.0, dtype=torch.float32, device=dev)
    cnt = torch.full((P,), -7, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    s = stream.cuda_stream if stream is not None else 0
    kl.omp_noalloc(Yd.data_ptr(), dim, P, Dd.data_ptr(), K, k, idx.data_ptr(), coef.data_ptr(), cnt.data_ptr(), s)
    torch.cuda.synchronize()
    return idx.cpu().numpy(), coef.cpu().numpy(), cnt.cpu().numpy()


def test_noalloc_matches_ksvd_omp_twice(kl):
    assert torch.cuda.is_available()
    # first call: the 512 x 512 workload's shape (12 x 12 patches, 256 atoms)
    Y1, D1, c1 = context_codes(kl, 2, 512, (12, 12), (3, 3), 256, 5)
    g1 = noalloc_codes(kl, Y1, D1, 5, None)
    # second call, smaller: 8 x 8 patches of a 128 x 128 picture, 200 atoms,
    # OMP limit 4, on a stream of its own
    Y2, D2, c2 = context_codes(kl, 9, 128, (8, 8), (2, 2), 200, 4)
    assert Y2.shape[0] < Y1.shape[0]
    g2 = noalloc_codes(kl, Y2, D2, 4, torch.cuda.Stream())
    for name, g, c in (("first", g1, c1), ("second", g2, c2)):
        for a, b in zip(g, c):
            np.testing.assert_array_equal(a, b, err_msg=f"{name} call")
    print(f"\nallocation-free OMP: {len(c1[2])} then {len(c2[2])} patches, codes equal to ksvd_omp's bit for bit")


def test_noalloc_refuses_sizes_beyond_the_maximum(kl):
    dev = torch.device("cuda:0")
    buf = torch.zeros(1024, device=dev)
    p = buf.data_ptr()
    for dim, K, k in ((145, 64, 5), (64, 257, 5), (64, 64, 9)):
        with pytest.raises(kl.KsvdError) as e:
            kl.omp_noalloc(p, dim, 4, p, K, k, p, p, p, 0)
        assert e.value.code == kl.KSVD_ERR_UNSUPPORTED
        assert "maximum" in str(e.value)
