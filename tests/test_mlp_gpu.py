"""bf16 MFMA product of the learned plane proposal against an fp32 reference."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ref(A, B, ta, tb):
    a = A.float().T if ta else A.float()
    b = B.float().T if tb else B.float()
    return a @ b


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(100, 70, 40), (512, 256, 256), (33, 256, 8), (256, 4, 1000)])
def test_gemm_bf16_matches_fp32_reference(ctx, ta, tb, M, N, K):
    import torch

    print("gemm", ta, tb, M, N, K, flush=True)
    g = torch.Generator().manual_seed(M * 7 + N * 3 + K + 11 * ta + 5 * tb)
    A = (torch.randn((K, M) if ta else (M, K), generator=g)).to(torch.bfloat16).cuda()
    B = (torch.randn((N, K) if tb else (K, N), generator=g)).to(torch.bfloat16).cuda()
    bias = torch.randn(N, generator=g).float().cuda()
    C = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    ctx.gemm_bf16_d(ta, tb, M, N, K, A, A.shape[1], B, B.shape[1], C, N, d_bias=bias)
    torch.cuda.synchronize()
    ref = _ref(A.cpu(), B.cpu(), ta, tb) + bias.cpu()
    tol = 1e-3 * (float(ref.abs().max()) + 1.0)
    assert float((C.cpu() - ref).abs().max()) <= tol
    # relu and accumulation into the fp32 output
    ctx.gemm_bf16_d(ta, tb, M, N, K, A, A.shape[1], B, B.shape[1], C, N, relu=True, acc=True)
    torch.cuda.synchronize()
    ref2 = torch.clamp(ref + _ref(A.cpu(), B.cpu(), ta, tb), min=0)
    assert float((C.cpu() - ref2).abs().max()) <= 2 * tol


def _torch_model(params, feat, planes, B, E, P):
    """fp32 reference of the model on the same parameters."""
    import torch

    from visacd_amd.native import mlp_unpack

    w = {k: torch.tensor(v, requires_grad=True) for k, v in mlp_unpack(params.copy()).items()}
    x = torch.cat([feat.reshape(B * E, 7), torch.zeros(B * E, 1)], 1)
    h = torch.relu(x @ w["W1"].T + w["b1"])
    h = torch.relu(h @ w["W2"].T + w["b2"])
    h = torch.relu(h @ w["W3"].T + w["b3"])
    z = h.reshape(B, E, 256).max(1).values
    u = z @ w["Wz"].T + w["c1"]
    q = planes.reshape(B, P, 4)
    g1 = torch.relu(u[:, None, :] + q @ w["Wp"].T)
    g2 = torch.relu(g1 @ w["V2"].T + w["c2"])
    out = g2 @ w["v3"] + w["c3"]
    return out.reshape(B * P), w


def _batch(seed, B=3, E=64, P=96):
    import torch

    g = torch.Generator().manual_seed(seed)
    feat = torch.rand((B, E, 7), generator=g) * 2 - 1
    feat[..., 6] = torch.linalg.norm(feat[..., 3:6] - feat[..., 0:3], dim=-1)
    n = torch.randn((B, P, 3), generator=g)
    n = n / n.norm(dim=-1, keepdim=True)
    planes = torch.cat([n, torch.rand((B, P, 1), generator=g) - 0.5], -1)
    target = torch.rand((B, P), generator=g) * 0.3
    return feat, planes, target


def test_mlp_forward_matches_fp32_reference(ctx):
    import torch

    from visacd_amd import native

    print("mlp forward", flush=True)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    m = native.Mlp(ctx, seed=3)
    B, E, P = 3, 64, 96
    feat, planes, _ = _batch(1, B, E, P)
    out = torch.empty(B * P, device="cuda")
    m.forward_d(B, E, feat.cuda(), P, planes.cuda(), out)
    torch.cuda.synchronize()
    ref, _ = _torch_model(m.get(), feat, planes, B, E, P)
    ref = ref.detach()
    scale = float(ref.abs().max()) + 1e-3
    assert float((out.cpu() - ref).abs().max()) <= 0.05 * scale
    m.close()


def test_fused_inference_head_matches_the_training_forward(ctx):
    """The fused inference head (g1 and g2 kept in registers) gives the same
    predictions as the unfused forward of a training step, up to fp32
    summation order: the step's loss (lr 0) equals the loss of the fused
    outputs.  P is not a multiple of 32, so row tiles span parts."""
    import torch

    from visacd_amd import native

    print("mlp fused head", flush=True)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    m = native.Mlp(ctx, seed=5)
    B, E, P = 4, 64, 200
    feat, planes, target = _batch(2, B, E, P)
    out = torch.empty(B * P, device="cuda")
    m.forward_d(B, E, feat.cuda(), P, planes.cuda(), out)
    torch.cuda.synchronize()
    loss_fused = float(((out.cpu() - target.reshape(-1)) ** 2).mean())
    loss_step = m.train_step_d(B, E, feat.cuda(), P, planes.cuda(), target.reshape(-1).cuda(), 0.0)
    assert abs(loss_fused - loss_step) <= 1e-4 * loss_step, (loss_fused, loss_step)
    # and the same values again (deterministic)
    out2 = torch.empty_like(out)
    m.forward_d(B, E, feat.cuda(), P, planes.cuda(), out2)
    torch.cuda.synchronize()
    assert torch.equal(out, out2)
    m.close()


def test_mlp_first_adam_step_follows_reference_gradient(ctx):
    """Adam's first step moves each weight by -lr * sign(g) (|g| >> eps): the
    update's signs must agree with the fp32 reference gradient."""
    import torch

    from visacd_amd import native

    print("mlp gradient", flush=True)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    m = native.Mlp(ctx, seed=5)
    B, E, P = 3, 64, 96
    feat, planes, target = _batch(2, B, E, P)
    w0 = m.get()
    out, w = _torch_model(w0, feat, planes, B, E, P)
    loss_ref = ((out - target.reshape(-1)) ** 2).mean()
    loss_ref.backward()
    lr = 1e-4
    loss = m.train_step_d(B, E, feat.cuda(), P, planes.cuda(), target.cuda(), lr)
    assert abs(loss - float(loss_ref)) <= 0.05 * float(loss_ref) + 1e-6
    dw = native.mlp_unpack(m.get() - w0)
    for name in ("v3", "V2", "c2", "Wp", "c1", "Wz", "W3", "W2", "W1", "b1"):
        gref = w[name].grad.numpy()
        big = np.abs(gref) > 0.05 * np.abs(gref).max()
        agree = np.mean(np.sign(dw[name][big]) == -np.sign(gref[big]))
        print(name, int(big.sum()), agree, flush=True)
        assert agree >= 0.9, name
    m.close()


def test_mlp_training_lowers_the_loss(ctx):
    import torch

    from visacd_amd import native

    print("mlp training", flush=True)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    m = native.Mlp(ctx, seed=7)
    B, E, P = 4, 64, 128
    feat, planes, target = _batch(3, B, E, P)
    feat, planes, target = feat.cuda(), planes.cuda(), target.cuda()
    losses = [m.train_step_d(B, E, feat, P, planes, target, 1e-3) for _ in range(60)]
    assert losses[-1] < 0.5 * losses[0]
    m.close()
