"""GPU checks of the differentiable direct-lighting render (skw_dl_render_replay / skw_dl_adjoint,
through the C ABI), wrapped here, and only here, as a torch.autograd.Function.

Bars: the replayed estimator at equal parameters is skw_dl_render's image bit for bit (after the
float conversion); with split parameters it matches the oracle within 1e-5 relative; the adjoint
gradient matches the oracle's FP64 forward-mode gradient within 1e-6 relative (of the largest
entry) and central finite differences of the GPU estimator (step 1e-4 relative) within 1e-3
relative."""
import numpy as np
import pytest

import oracle
import skwarka_amd as sk
from skwarka_amd import workloads as wl

pytestmark = pytest.mark.gpu

W, H, SPP, NLS, NBS, SEED = 48, 40, 2, 2, 2, 13


def _setup(torch):
    s = wl.scene_room_small()
    scene = sk.Scene(s["meshes"])
    gdl = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"],
                     prim_group=s["prim_group"], n_groups=s["n_groups"])
    cam, cam15 = sk.camera(*s["camera"], W, H)
    return s, scene, gdl, cam, cam15


def _oracle(s):
    acc = oracle.Accel(s["meshes"])
    return oracle.DLScene(acc, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"],
                          prim_group=s["prim_group"], n_groups=s["n_groups"])


def _replay(torch, gdl, cam, te, ts):
    img = torch.empty((W * H, 3), dtype=torch.float64, device="cuda")
    gdl.render_replay(te, ts, cam, W, H, SPP, img.data_ptr(), NLS, NBS, SEED)
    torch.cuda.synchronize()
    return img


def _function(torch, gdl, cam):
    class DirectLighting(torch.autograd.Function):
        """image = F(theta; theta); backward = skw_dl_adjoint replaying the same paths."""

        @staticmethod
        def forward(ctx, theta):
            th = theta.detach().cpu().numpy()
            ctx.theta = th
            return _replay(torch, gdl, cam, th, th)

        @staticmethod
        def backward(ctx, grad_img):
            adj = grad_img.contiguous().to(torch.float64)
            grad = torch.empty(gdl.n_params, dtype=torch.float64, device="cuda")
            gdl.adjoint(ctx.theta, cam, W, H, SPP, adj.data_ptr(), grad.data_ptr(), NLS, NBS, SEED)
            torch.cuda.synchronize()
            return grad

    return DirectLighting


def test_replay_at_equal_parameters_is_the_render(torch_cuda):
    print("case: replay(theta0, theta0) vs skw_dl_render, small room", flush=True)
    torch = torch_cuda
    s, scene, gdl, cam, _ = _setup(torch)
    img = _replay(torch, gdl, cam, s["theta0"], s["theta0"]).cpu().numpy()
    rgb = torch.empty((W * H, 3), dtype=torch.float32, device="cuda")
    gdl.render(cam, W, H, SPP, rgb.data_ptr(), NLS, NBS, SEED)
    torch.cuda.synchronize()
    assert np.array_equal(img.astype(np.float32), rgb.cpu().numpy())
    assert img.max() > 0


def test_replay_with_split_parameters_matches_oracle(torch_cuda):
    print("case: replay(theta_target, theta0) vs oracle", flush=True)
    torch = torch_cuda
    s, scene, gdl, cam, cam15 = _setup(torch)
    te = wl.perturbed_theta(s["theta0"])
    got = _replay(torch, gdl, cam, te, s["theta0"]).cpu().numpy()
    ref = _oracle(s).render_replay(te, s["theta0"], cam15, W, H, SPP, NLS, NBS, SEED)
    err = np.abs(got - ref)
    assert np.all(err <= 1e-5 * np.abs(ref) + 1e-12), err.max()
    assert not np.array_equal(got, _replay(torch, gdl, cam, s["theta0"], s["theta0"]).cpu().numpy())


def test_adjoint_matches_oracle_and_finite_differences(torch_cuda):
    print("case: adjoint vs FP64 oracle gradient and central FD, small room", flush=True)
    torch = torch_cuda
    s, scene, gdl, cam, cam15 = _setup(torch)
    th0 = s["theta0"]
    target = torch.from_numpy(_oracle(s).render_replay(wl.perturbed_theta(th0), wl.perturbed_theta(th0), cam15, W,
                                                       H, SPP, NLS, NBS, SEED + 1)).cuda()
    fn = _function(torch, gdl, cam)
    theta = torch.tensor(th0, dtype=torch.float64, device="cuda", requires_grad=True)
    img = fn.apply(theta)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    g = theta.grad.cpu().numpy()
    # FP64 CPU reference gradient (forward-mode dual numbers over the same paths)
    adj = (2.0 * (img.detach() - target)).cpu().numpy()
    g_ref, _ = _oracle(s).gradient(th0, adj, cam15, W, H, SPP, NLS, NBS, SEED)
    scale = np.max(np.abs(g_ref))
    assert scale > 0
    assert np.all(np.abs(g - g_ref) <= 1e-6 * scale), np.max(np.abs(g - g_ref)) / scale
    # central differences of the GPU estimator, sampling held at theta0, step 1e-4 relative to
    # the parameter (roughness enters as an exponent 1/r, so an absolute step would be coarse
    # for r near 0.01)

    def loss_at(te):
        return float(((_replay(torch, gdl, cam, te, th0) - target) ** 2).sum().item())

    fd = np.zeros_like(g)
    for j in range(len(th0)):
        h = 1e-4 * max(abs(th0[j]), 1e-3)
        tp, tm = th0.copy(), th0.copy()
        tp[j] += h
        tm[j] -= h
        fd[j] = (loss_at(tp) - loss_at(tm)) / (tp[j] - tm[j])
    fscale = np.max(np.abs(fd))
    assert np.all(np.abs(g - fd) <= 1e-3 * np.abs(fd) + 1e-6 * fscale), np.max(np.abs(g - fd))
    assert np.count_nonzero(g[:64]) > 5 and np.count_nonzero(g[64:128]) > 2 and np.all(g[128:] != 0)


def test_adjoint_repeats_and_rejects_bad_theta(torch_cuda):
    print("case: adjoint twice, bad theta", flush=True)
    torch = torch_cuda
    s, scene, gdl, cam, _ = _setup(torch)
    adj = torch.ones((W * H, 3), dtype=torch.float64, device="cuda")
    g1 = torch.empty(gdl.n_params, dtype=torch.float64, device="cuda")
    g2 = torch.empty(gdl.n_params, dtype=torch.float64, device="cuda")
    gdl.adjoint(s["theta0"], cam, W, H, SPP, adj.data_ptr(), g1.data_ptr(), NLS, NBS, SEED)
    gdl.adjoint(s["theta0"], cam, W, H, SPP, adj.data_ptr(), g2.data_ptr(), NLS, NBS, SEED)
    torch.cuda.synchronize()
    a, b = g1.cpu().numpy(), g2.cpu().numpy()
    assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(a).max())
    bad = s["theta0"].copy()
    bad[3] = np.nan
    with pytest.raises(sk.SkwError):
        gdl.adjoint(bad, cam, W, H, SPP, adj.data_ptr(), g1.data_ptr(), NLS, NBS, SEED)
    with pytest.raises(ValueError):
        gdl.adjoint(s["theta0"][:10], cam, W, H, SPP, adj.data_ptr(), g1.data_ptr(), NLS, NBS, SEED)


def test_tape_adjoint_equals_recomputed_adjoint(torch_cuda):
    print("case: stored-derivative (tape) adjoint vs recomputed adjoint, autograd through the tape", flush=True)
    torch = torch_cuda
    s, scene, gdl, cam, cam15 = _setup(torch)
    th0 = s["theta0"]
    img = torch.empty((W * H, 3), dtype=torch.float64, device="cuda")
    tape = gdl.render_tape(th0, cam, W, H, SPP, img.data_ptr(), 1 << 30, NLS, NBS, SEED)
    torch.cuda.synchronize()
    assert torch.equal(img, _replay(torch, gdl, cam, th0, th0))
    rng = np.random.default_rng(7)
    adj = torch.from_numpy(rng.normal(size=(W * H, 3))).cuda()
    g_tape = torch.empty(gdl.n_params, dtype=torch.float64, device="cuda")
    g_rep = torch.empty_like(g_tape)
    for _ in range(2):  # the tape serves any number of adjoints
        tape.adjoint(adj.data_ptr(), g_tape.data_ptr())
        gdl.adjoint(th0, cam, W, H, SPP, adj.data_ptr(), g_rep.data_ptr(), NLS, NBS, SEED)
        torch.cuda.synchronize()
        a, b = g_tape.cpu().numpy(), g_rep.cpu().numpy()
        assert np.all(np.abs(a - b) <= 1e-10 * np.abs(b).max()), np.abs(a - b).max()
    tape.close()
    # a tape larger than the budget is refused, so the caller recomputes instead
    with pytest.raises(sk.SkwError, match="max_tape_bytes"):
        gdl.render_tape(th0, cam, W, H, SPP, img.data_ptr(), 1024, NLS, NBS, SEED)

    class TapedDirectLighting(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta):
            out = torch.empty((W * H, 3), dtype=torch.float64, device="cuda")
            ctx.tape = gdl.render_tape(theta.detach().cpu().numpy(), cam, W, H, SPP, out.data_ptr(), 1 << 30, NLS,
                                       NBS, SEED)
            return out

        @staticmethod
        def backward(ctx, grad_img):
            grad = torch.empty(gdl.n_params, dtype=torch.float64, device="cuda")
            ctx.tape.adjoint(grad_img.contiguous().data_ptr(), grad.data_ptr())
            torch.cuda.synchronize()
            ctx.tape.close()
            return grad

    theta = torch.tensor(th0, dtype=torch.float64, device="cuda", requires_grad=True)
    loss = (TapedDirectLighting.apply(theta) * adj).sum()
    loss.backward()
    assert np.all(np.abs(theta.grad.cpu().numpy() - b) <= 1e-10 * np.abs(b).max())
