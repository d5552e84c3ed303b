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


@pytest.mark.parametrize("B,E,P", [(3, 64, 96), (1, 512, 16384)])
def test_mlp_forward_matches_fp32_reference(ctx, B, E, P):
    """bf16 operands, fp32 accumulation: within 2 % of the largest output of
    the fp32 model on the same parameters, at a small batch and at the size
    the proposal runs (512 edges, 16,384 planes), where the top-256 ranking
    it feeds must mostly agree with the fp32 model's."""
    import torch

    from visacd_amd import native

    print("mlp forward", B, E, P, flush=True)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    m = native.Mlp(ctx, seed=3)
    feat, planes, _ = _batch(1, B, E, P)
    out = torch.empty(B * P, device="cuda")
    m.forward_d(B, E, feat.cuda(), P, planes.cuda(), out)
    torch.cuda.synchronize()
    with torch.no_grad():
        ref, _ = _torch_model(m.get(), feat, planes, B, E, P)
    ref = ref.detach()
    scale = float(ref.abs().max()) + 1e-3
    err = float((out.cpu() - ref).abs().max())
    print("  max error", err, "scale", scale, flush=True)
    assert err <= 0.02 * scale
    if P >= 4096:
        top = set(torch.topk(out.cpu(), 256).indices.tolist())
        top_ref = set(torch.topk(ref, 256).indices.tolist())
        print("  top-256 overlap", len(top & top_ref), flush=True)
        assert len(top & top_ref) >= 0.8 * 256
    m.close()


@pytest.mark.parametrize("P", [200, 20])
def test_fused_inference_head_matches_the_training_forward(ctx, P):
    """The fused inference head (g1 and g2 kept in registers) gives the
    predictions of the unfused forward of a training step: the step's loss
    (lr 0) equals the loss of the fused outputs within 5e-4.  The two differ
    by the fused head's bf16 splits of q, Wp and u (about 1e-5 of g1 before
    its bf16 rounding, which flips that rounding for a few values) and fp32
    summation order.  P is not a multiple of 32, so row tiles span parts (two
    parts at P = 200; several at P = 20, the small-part variant)."""
    import torch

    from visacd_amd import native

    print("mlp fused head", P, flush=True)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    m = native.Mlp(ctx, seed=5)
    B, E = 4, 64
    feat, planes, target = _batch(2, B, E, P)
    out = torch.empty(B * P, device="cuda")
    m.forward_d(B, E, feat.cuda(), P, planes.cuda(), out)
    torch.cuda.synchronize()
    loss_fused = float(((out.cpu() - target.reshape(-1)) ** 2).mean())
    loss_step = m.train_step_d(B, E, feat.cuda(), P, planes.cuda(), target.reshape(-1).cuda(), 0.0)
    assert abs(loss_fused - loss_step) <= 5e-4 * loss_step, (loss_fused, loss_step)
    # and the same values again (deterministic)
    out2 = torch.empty_like(out)
    m.forward_d(B, E, feat.cuda(), P, planes.cuda(), out2)
    torch.cuda.synchronize()
    assert torch.equal(out, out2)
    m.close()



def test_fused_forward_calls_in_a_row_are_independent(ctx):
    """The fused forward's pooled maxima live in a buffer that each call's
    projection clears for the next: calls in a row with other inputs, a
    smaller batch, then a larger one (the buffer regrows) and edge tiles
    that straddle parts (48 edges per part) give bit for bit the outputs of
    a fresh model on each input."""
    import torch

    from visacd_amd import native

    print("mlp fused calls in a row", flush=True)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    runs = [(4, 48, 40, 11), (2, 64, 40, 12), (6, 48, 33, 13), (2, 64, 40, 12), (5, 512, 64, 14)]

    def run(m, B, E, P, seed):
        feat, planes, _ = _batch(seed, B, E, P)
        out = torch.empty(B * P, device="cuda")
        m.forward_d(B, E, feat.cuda(), P, planes.cuda(), out)
        torch.cuda.synchronize()
        return out.cpu()

    m = native.Mlp(ctx, seed=7)
    outs = [run(m, *r) for r in runs]
    m.close()
    for r, o in zip(runs, outs):
        fresh = native.Mlp(ctx, seed=7)
        ref = run(fresh, *r)
        fresh.close()
        assert torch.equal(o, ref), r

def test_mlp_gradient_matches_fp32_reference_and_adam_is_exact(ctx):
    """The training kernels' gradient (read back with vacd_mlp_get_grads)
    agrees with the fp32 model's autograd gradient on the same batch and
    parameters (bf16 operands: relative L2 error per tensor within 3 % for
    the head and 10 % for the encoder),
    and two Adam steps move the fp32 master weights exactly as Adam with
    betas 0.9/0.999, eps 1e-8 and bias correction does on those gradients."""
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
    lr = np.float32(1e-3)
    loss = m.train_step_d(B, E, feat.cuda(), P, planes.cuda(), target.cuda(), float(lr))
    assert abs(loss - float(loss_ref.detach())) <= 0.01 * float(loss_ref.detach())
    g1 = m.grads()
    gd = native.mlp_unpack(g1.copy())
    for name in ("v3", "c3", "V2", "c2", "Wp", "c1", "Wz", "W3", "b3", "W2", "b2", "W1", "b1"):
        gref = w[name].grad.numpy().ravel()
        gg = gd[name].ravel()
        if name == "W1":  # the padded 8th input column has no reference counterpart
            gref, gg = gref.reshape(256, 8)[:, :7].ravel(), gg.reshape(256, 8)[:, :7].ravel()
        nr = np.linalg.norm(gref)
        rel = np.linalg.norm(gg - gref) / nr
        cos = float(gg @ gref) / (np.linalg.norm(gg) * nr)
        print(" ", name, "rel", rel, "cos", cos, flush=True)
        # bf16 activations: the head within 3 %; the encoder behind the
        # max-pool (whose winners bf16 rounding can move) within 10 %
        head = name in ("v3", "c3", "V2", "c2", "Wp", "c1", "Wz")
        assert rel <= (0.03 if head else 0.10) and cos >= (0.999 if head else 0.995), name
    f = np.float32
    b1, b2, eps = f(0.9), f(0.999), f(1e-8)
    w1 = m.get()
    m1 = (f(1) - b1) * g1
    v1 = (f(1) - b2) * g1 * g1
    c1, c2 = f(1) - b1, f(1) - b2
    exp1 = w0 - lr * (m1 / c1) / (np.sqrt(v1 / c2) + eps)
    np.testing.assert_allclose(w1, exp1, rtol=0, atol=1e-7)
    m.train_step_d(B, E, feat.cuda(), P, planes.cuda(), target.cuda(), float(lr))
    g2 = m.grads()
    m2 = b1 * m1 + (f(1) - b1) * g2
    v2 = b2 * v1 + (f(1) - b2) * g2 * g2
    c1, c2 = f(1) - b1 * b1, f(1) - b2 * b2
    exp2 = w1 - lr * (m2 / c1) / (np.sqrt(v2 / c2) + eps)
    np.testing.assert_allclose(m.get(), exp2, rtol=0, atol=1e-7)
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
