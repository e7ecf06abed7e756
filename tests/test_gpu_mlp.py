"""The radiance-cache network (solhip_mlp_*) against a plain PyTorch fp32
reference of the same op (CPU, float32).

Forward: the reference uses the same BF16-rounded inputs and weights and
rounds each hidden activation to BF16 as the kernel does; outputs agree to
2e-2 of their scale (BF16 activation rounding, different summation order).
Gradient: torch autograd of the relative-L2 loss (denominator held
constant) on the BF16-rounded weights; per layer the cosine similarity with
the kernel's gradient exceeds 0.999 and the norms agree to 3 % (the kernel
back-propagates BF16 gradients).  Adam: one step from the kernel's gradient
equals the update formula in fp32 to 1e-6.  Training: the loss on a smooth
target falls by 10x within 300 steps."""
import numpy as np
import pytest
import torch

import solhip

pytestmark = pytest.mark.gpu

P = solhip.api.MLP_PARAMS


def bf16_round(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(torch.bfloat16).to(torch.float32)


def unpack(w):
    """master parameter vector -> [W0..W3 (64x64), W4 (64x3)] (k = input rows)"""
    Ws = [torch.from_numpy(w[l * 4096:(l + 1) * 4096].reshape(64, 64).copy()) for l in range(4)]
    Ws.append(torch.from_numpy(w[16384:].reshape(64, 16)[:, :3].copy()))
    return Ws


def pack_grad(gs):
    out = np.zeros(P, np.float32)
    for l in range(4):
        out[l * 4096:(l + 1) * 4096] = gs[l].reshape(-1)
    g4 = np.zeros((64, 16), np.float32)
    g4[:, :3] = gs[4]
    out[16384:] = g4.reshape(-1)
    return out


def ref_forward(x, w, round_acts=True):
    h = bf16_round(x)
    Ws = [bf16_round(W.numpy()) for W in unpack(w)]
    for l in range(4):
        h = torch.relu(h @ Ws[l])
        if round_acts:
            h = h.to(torch.bfloat16).to(torch.float32)
    return h @ Ws[4]


def _inputs(ctx, n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, size=(n, 64)).astype(np.float32)
    xb = torch.from_numpy(x).to(torch.bfloat16)
    xbuf = ctx.upload(xb.view(torch.int16).numpy())
    return x, xbuf


def test_forward_matches_torch(ctx):
    m = solhip.Mlp(ctx, seed=7)
    for n in (1, 31, 1000, 4099):
        x, xbuf = _inputs(ctx, n, n)
        ybuf = ctx.buffer(n * 16)
        print("mlp forward n=%d" % n, flush=True)
        m.forward(xbuf, n, ybuf)
        y = ybuf.download(np.float32).reshape(n, 4)[:, :3]
        ref = ref_forward(x, m.params()).numpy()
        scale = max(1e-3, np.abs(ref).max())
        assert np.abs(y - ref).max() <= 2e-2 * scale, (n, np.abs(y - ref).max(), scale)
    m.close()


def test_gradient_matches_torch(ctx):
    m = solhip.Mlp(ctx, seed=3)
    n = 3000
    x, xbuf = _inputs(ctx, n, 11)
    rng = np.random.default_rng(5)
    t = np.zeros((n, 4), np.float32)
    t[:, :3] = rng.uniform(0.0, 2.0, size=(n, 3))
    tbuf = ctx.upload(t)
    print("mlp gradient", flush=True)
    g, loss = m.gradient(xbuf, tbuf, n)
    w = m.params()
    Ws = [bf16_round(W.numpy()).requires_grad_(True) for W in unpack(w)]
    h = bf16_round(x)
    for l in range(4):
        h = torch.relu(h @ Ws[l])
    p = h @ Ws[4]
    tt = torch.from_numpy(t[:, :3])
    ref_loss = (((p - tt) ** 2) / (p.detach() ** 2 + 0.01)).sum() / n
    ref_loss.backward()
    assert abs(loss - ref_loss.item()) <= 2e-2 * ref_loss.item()
    ref = pack_grad([W.grad.numpy() for W in Ws])
    for l in range(5):
        sl = slice(l * 4096, (l + 1) * 4096) if l < 4 else slice(16384, P)
        a, b = g[sl].astype(np.float64), ref[sl].astype(np.float64)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        assert cos > 0.999, (l, cos)
        assert abs(np.linalg.norm(a) / np.linalg.norm(b) - 1.0) < 0.03, l
    # padded output columns get no gradient
    assert (g[16384:].reshape(64, 16)[:, 3:] == 0).all()
    m.close()


def test_adam_step_and_training(ctx):
    lr = 5e-3
    m = solhip.Mlp(ctx, seed=9, lr=lr)
    n = 4096
    rng = np.random.default_rng(2)
    x = rng.uniform(-1.0, 1.0, size=(n, 64)).astype(np.float32)
    t = np.zeros((n, 4), np.float32)
    t[:, 0] = 0.5 + 0.4 * np.sin(2.0 * x[:, 0]) * np.cos(x[:, 1])
    t[:, 1] = 0.5 + 0.3 * x[:, 2] * x[:, 3]
    t[:, 2] = 1.0 + 0.2 * np.tanh(x[:, 4] + x[:, 5])
    xbuf = ctx.upload(torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy())
    tbuf = ctx.upload(t)
    w0 = m.params()
    g, _ = m.gradient(xbuf, tbuf, n)
    m.train(xbuf, tbuf, n, 1)
    w1 = m.params()
    # first Adam step: m = 0.1 g, v = 0.01 g^2, bias corrections 0.1 / 0.01
    mh = (0.1 * g) / np.float32(1 - 0.9)
    vh = (0.01 * g * g) / np.float32(1 - 0.99)
    exp = w0 - lr * mh / (np.sqrt(vh) + 1e-8)
    exp[16384:].reshape(64, 16)[:, 3:] = 0.0
    assert np.abs(w1 - exp).max() <= 1e-6 + 1e-6 * np.abs(exp).max()
    print("mlp training", flush=True)
    first = m.train(xbuf, tbuf, n, 1)
    last = m.train(xbuf, tbuf, n, 300)
    assert last < 0.1 * first, (first, last)
    m.close()


def _nrc_frames(ctx, gs, sd, W, H, frames, nrc, spp=1, depth=6, sky=0, rr=0):
    s = solhip.spec_default(solhip.PROG_PATHTRACE)
    s.enable_sky, s.spp_per_frame, s.max_depth, s.nee, s.nrc, s.rr_start = sky, spp, depth, 1, nrc, rr
    prog = ctx.program(solhip.PROG_PATHTRACE, s)
    acc = ctx.upload(np.zeros((H, W, 4), np.float32))
    img = solhip.Images(acc.h, 0, 0, 0, 0, 0)
    out = []
    for f in range(frames):
        ctx.trace(prog, gs.tlas, gs.tables, sd.uniforms(W, H, f), 0, img, W, H)
        out.append(acc.download(np.float32).reshape(H, W, 4)[..., :3].copy())
    ctx.program_destroy(prog)
    return out


def test_radiance_cache_learns_the_cornell_box(ctx):
    """the cache-terminated estimator (biased) gets close to the converged
    plain image: relative RMSE of its 64-frame running mean below 0.2, and
    far below the error of an untrained cache's first frames"""
    from solhip import scenes
    sd = scenes.cornell_box()
    gs = solhip.GpuScene(ctx, sd)
    W = H = 64
    print("reference (plain, 2048 spp)", flush=True)
    ref = _nrc_frames(ctx, gs, sd, W, H, 8, 0, spp=256)[-1]
    print("cache frames", flush=True)
    imgs = _nrc_frames(ctx, gs, sd, W, H, 64, 1)
    for im in imgs:
        assert np.isfinite(im).all() and (im >= 0).all()
    norm = np.sqrt((ref ** 2).mean())
    rel = [np.sqrt(((im - ref) ** 2).mean()) / norm for im in imgs]
    print("relative RMSE by frame:", [round(r, 4) for r in rel[::8]], flush=True)
    assert rel[-1] < 0.2, rel[-1]
    assert rel[-1] < 0.7 * rel[1], (rel[1], rel[-1])


def test_radiance_cache_on_the_instanced_scene(ctx, oracle):
    """the cache on the batched instanced workload (its generator at 200
    instances of the five BLAS kinds; gradient sky, depth 8, roulette from
    bounce 3, next-event estimation, paths ended at their first bounce vertex
    into the cache): its running mean approaches the converged plain image
    and far below an untrained cache's first frames; a plain program on the
    same scene afterwards still renders the oracle's image bit for bit"""
    from solhip import scenes
    sd = scenes.instanced(200, seed=2, tube=(50, 100), spread=8.0)
    gs = solhip.GpuScene(ctx, sd)
    W, H = 64, 48
    print("reference (plain, 1024 spp)", flush=True)
    ref = _nrc_frames(ctx, gs, sd, W, H, 4, 0, spp=256, depth=8, sky=2, rr=3)[-1]
    print("cache frames", flush=True)
    imgs = _nrc_frames(ctx, gs, sd, W, H, 64, 1, depth=8, sky=2, rr=3)
    for im in imgs:
        assert np.isfinite(im).all() and (im >= 0).all()
    norm = np.sqrt((ref ** 2).mean())
    rel = [np.sqrt(((im - ref) ** 2).mean()) / norm for im in imgs]
    print("relative RMSE by frame:", [round(r, 4) for r in rel[::8]], flush=True)
    assert rel[-1] < 0.25, rel[-1]
    assert rel[-1] < 0.7 * rel[1], (rel[1], rel[-1])
    # nrc = 0 after the cache ran: the plain estimator, bit-identical with the oracle
    s = solhip.spec_default(solhip.PROG_PATHTRACE)
    s.enable_sky, s.spp_per_frame, s.max_depth, s.nee, s.rr_start = 2, 2, 8, 1, 3
    prog = ctx.program(solhip.PROG_PATHTRACE, s)
    acc = ctx.upload(np.zeros((H, W, 4), np.float32))
    u = sd.uniforms(W, H, 5)
    ctx.trace(prog, gs.tlas, gs.tables, u, 5, solhip.Images(acc.h, 0, 0, 0, 0, 0), W, H)
    got = acc.download(np.float32).reshape(H, W, 4)
    want = np.zeros((H, W, 4), np.float32)
    oracle.Scene(sd).render_pathtrace(u, oracle.PTConfig(2, 2, 8, 3, 1, 0), 5, want, None, W, H)
    assert np.count_nonzero(got.view(np.uint32) != want.view(np.uint32)) == 0
    ctx.program_destroy(prog)
