"""tc_mlp_train_step_noalloc (the training step that allocates nothing) against
tc_mlp_train_step through the C ABI: two trainers from the same seed take the
same steps, the first on 70000 samples (the gradient kernel's full grid of 512
blocks, so every record of the static scratch is written), the second on 3000
(47 blocks: the scratch keeps the first call's other records, which must not
leak into the result). Both entry points run the same kernels with the same
block partition and a fixed summation order, so losses and weights must agree
bit for bit."""
import numpy as np
import pytest

import turbchannel_amd as tca

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to("cuda")
    torch.cuda.synchronize()
    return t


def test_no_alloc_train_step_matches_train_step():
    rng = np.random.default_rng(4)
    n0 = 70000
    X = rng.normal(size=(n0, 10)).astype(np.float32)
    T = (0.3 * rng.normal(size=(n0, 6))).astype(np.float32)
    x, t = dev(X), dev(T)
    ref, nal = tca.Mlp(seed=11), tca.Mlp(seed=11)
    for n in (n0, 3000):
        print(f"no-alloc training step, n = {n}", flush=True)
        la = ref.train_step(x.data_ptr(), t.data_ptr(), n, lr=1e-3)
        lb = nal.train_step(x.data_ptr(), t.data_ptr(), n, lr=1e-3, no_alloc=True)
        assert la == lb, (n, la, lb)
        assert np.array_equal(ref.get_params(), nal.get_params()), n
    assert nal.steps() == 2
    ref.close()
    nal.close()


def test_no_alloc_train_step_rejects_bad_arguments():
    net = tca.Mlp(seed=1)
    x = dev(np.zeros((64, 10)))
    t = dev(np.zeros((64, 6)))
    with pytest.raises(tca.TCError):
        net.train_step(x.data_ptr(), t.data_ptr(), 0, no_alloc=True)
    with pytest.raises(tca.TCError):
        net.train_step(x.data_ptr(), t.data_ptr(), 2**31, no_alloc=True)
    with pytest.raises(tca.TCError):
        net.train_step(x.data_ptr(), t.data_ptr(), 64, lr=-1.0, no_alloc=True)
    net.close()
