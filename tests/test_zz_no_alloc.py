"""engine_mlp_grad_noalloc: the surrogate's loss gradient on device buffers,
with its scratch in a static per-device array (no allocation in the call),
against engine_mlp_grad on the same samples.

Two calls in a row, the second with fewer samples, so that the scratch the
first call left behind must not leak into the second.  Up to 512 tiles both
entry points launch the same grid and sum every gradient in the same fixed
order, so the gradients and the loss agree bit for bit.  Sizes beyond the
scratch limit are refused, and a pair id outside the network turns the loss
into NaN (the kernels read it as pair 0).
"""
import numpy as np
import pytest

import mdcore as md
import surrogate as sg
import systems

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def net():
    sysm = systems.mixture(3000, seed=4, nspecies=3)
    eng = systems.build_engine(sysm, md.engine_flag_gpu)
    dom = sg.pair_domains(eng, sysm)
    eng.mlp_init(*dom, seed=3)
    yield sysm, eng, dom
    eng.close()


def samples(eng, sysm, dom, n, seed):
    npair = len(dom[0])
    rng = np.random.default_rng(seed)
    p = np.repeat(rng.integers(0, npair, n // 16), 16).astype(np.int32)
    r = (dom[0][p] + (dom[1][p] - dom[0][p]) * rng.random(n)).astype(np.float32)
    pots, _, _ = sg.pair_potentials(eng, sysm)
    E = np.zeros(n)
    F = np.zeros(n)
    for k in range(npair):
        m = p == k
        E[m], F[m] = pots[k].eval_r(r[m].astype(np.float64))
    return r, p, E.astype(np.float32), F.astype(np.float32)


def noalloc(eng, r, p, E, F, wE, wF):
    nparams = len(eng.mlp_params())
    dev = [torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0") for a in (r, p, E, F)]
    grad = torch.full((nparams,), 7.0, dtype=torch.float32, device="cuda:0")
    loss = torch.full((1,), 7.0, dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    eng.mlp_grad_noalloc(dev[0].data_ptr(), dev[1].data_ptr(), dev[2].data_ptr(), dev[3].data_ptr(), len(r),
                         grad.data_ptr(), loss.data_ptr(), wE, wF)
    torch.cuda.synchronize()
    return grad.cpu().numpy(), float(loss.cpu().numpy()[0])


def test_noalloc_gradient_equals_engine_mlp_grad(net):
    sysm, eng, dom = net
    for n, seed in ((8192, 21), (1024, 22)):
        r, p, E, F = samples(eng, sysm, dom, n, seed)
        g_ref, loss_ref = eng.mlp_grad(r, p, E, F, 0.5, 2.0)
        g, loss = noalloc(eng, r, p, E, F, 0.5, 2.0)
        print(f"n={n}: loss {loss:.6e} (engine_mlp_grad {loss_ref:.6e}), "
              f"max |dg| {np.abs(g - g_ref).max():.3e}", flush=True)
        assert np.all(np.isfinite(g))
        assert np.array_equal(g, g_ref)
        assert np.float32(loss) == np.float32(loss_ref)


def test_noalloc_limits(net):
    sysm, eng, dom = net
    nparams = len(eng.mlp_params())
    buf = torch.zeros(16 * 4097, dtype=torch.float32, device="cuda:0")
    ibuf = torch.zeros(16 * 4097, dtype=torch.int32, device="cuda:0")
    grad = torch.zeros(nparams, dtype=torch.float32, device="cuda:0")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    a = [buf.data_ptr(), ibuf.data_ptr(), buf.data_ptr(), buf.data_ptr()]
    for n in (16 * 4097, 24, 0):
        with pytest.raises(md.MdcoreError):
            eng.mlp_grad_noalloc(*a, n, grad.data_ptr(), loss.data_ptr())
    with pytest.raises(md.MdcoreError):
        eng.mlp_grad_noalloc(0, *a[1:], 1024, grad.data_ptr(), loss.data_ptr())


def test_noalloc_bad_pair_gives_nan_loss(net):
    sysm, eng, dom = net
    r, p, E, F = samples(eng, sysm, dom, 512, 23)
    p[16:32] = 1000
    _, loss = noalloc(eng, r, p, E, F, 1.0, 1.0)
    assert np.isnan(loss)
    # the next call on good samples is not affected
    r, p, E, F = samples(eng, sysm, dom, 512, 24)
    g_ref, loss_ref = eng.mlp_grad(r, p, E, F)
    g, loss = noalloc(eng, r, p, E, F, 1.0, 1.0)
    assert np.array_equal(g, g_ref) and np.float32(loss) == np.float32(loss_ref)
