"""GPU: the BF16 MFMA GEMM against FP32 products of the same BF16-rounded
inputs (pins the MFMA lane layout, strides, transposes, epilogues), and the
FNO forward / hand-written backward against the FP64 torch.fft oracle."""

import math

import numpy as np
import pytest
import torch

from oracle import fno_oracle
from tests._helpers import rel_l2

pytestmark = pytest.mark.gpu


def _bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


@pytest.mark.parametrize("M,N,K,batch", [(128, 128, 64, 1), (77, 45, 100, 3), (1, 300, 513, 2),
                                         (256, 32, 32, 4)])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_gemm_matches_fp32_product_of_bf16_inputs(M, N, K, batch, ta, tb):
    from diffnano_amd.fno import gemm
    g = torch.Generator(device="cuda").manual_seed(M * 7 + N + K)
    A = torch.randn(batch, K, M, device="cuda", generator=g) if ta else \
        torch.randn(batch, M, K, device="cuda", generator=g)
    B = torch.randn(batch, N, K, device="cuda", generator=g) if tb else \
        torch.randn(batch, K, N, device="cuda", generator=g)
    Cin = torch.randn(batch, M, N, device="cuda", generator=g)
    bias = torch.randn(N, device="cuda", generator=g)
    C = Cin.clone()
    sA = (1, M, M * K) if ta else (K, 1, M * K)
    sB = (1, K, N * K) if tb else (N, 1, N * K)
    gemm(M, N, K, A, sA, B, sB, C, (N, 1, M * N), batch=batch, alpha=0.5, beta=-1.25, Cin=C,
         bias=bias, bias_mode=1)
    Am = _bf(A).transpose(1, 2) if ta else _bf(A)
    Bm = _bf(B).transpose(1, 2) if tb else _bf(B)
    ref = 0.5 * torch.bmm(Am.contiguous(), Bm.contiguous()) - 1.25 * Cin + bias
    torch.cuda.synchronize()
    assert rel_l2(C.cpu().numpy(), ref.cpu().numpy()) < 2e-6


def test_gemm_bf16_operands_gelu_and_pre():
    from diffnano_amd.fno import gemm
    torch.manual_seed(0)
    A = torch.randn(96, 40, device="cuda").to(torch.bfloat16)
    B = torch.randn(40, 200, device="cuda").to(torch.bfloat16)
    bias = torch.randn(96, device="cuda")
    C = torch.empty(96, 200, device="cuda")
    pre = torch.empty_like(C)
    gemm(96, 200, 40, A, (40, 1, 0), B, (200, 1, 0), C, (200, 1, 0), bias=bias, bias_mode=2,
         act=1, pre=pre)
    z = A.float() @ B.float() + bias[:, None]
    assert rel_l2(pre.cpu().numpy(), z.cpu().numpy()) < 2e-6
    gel = torch.nn.functional.gelu(z)
    assert rel_l2(C.cpu().numpy(), gel.cpu().numpy()) < 2e-6


@pytest.mark.parametrize("M,K", [(32, 9), (32, 32), (128, 32), (6, 128), (128, 6), (32, 128)])
@pytest.mark.parametrize("N", [4096, 1000])
@pytest.mark.parametrize("xbf", [False, True])
def test_pointwise_kernels(M, K, N, xbf, lib_opts):
    """The FNO's channel-mixing shapes (K <= 32 on the matrix-core kernel,
    K = 128 on the vector-ALU one): every epilogue (alpha, beta Cin, row
    bias, pre, GELU), a ragged cell count, a weight read through a
    transposed view, against FP32 products of the same BF16-rounded inputs
    and against the vector-ALU kernel."""
    from diffnano_amd.fno import gemm
    g = torch.Generator(device="cuda").manual_seed(M * 131 + K * 7 + N + int(xbf))
    Wt = torch.randn(K, M, device="cuda", generator=g)          # used as W = Wt^T
    X = torch.randn(K, N, device="cuda", generator=g)
    if xbf:
        X = X.to(torch.bfloat16)
    Cin = torch.randn(M, N, device="cuda", generator=g)
    bias = torch.randn(M, device="cuda", generator=g)
    outs = []
    for valu in (False, True):
        if valu:
            lib_opts.set("gemm_pw_valu", 1)
        C = torch.full((M, N), float("nan"), device="cuda")
        pre = torch.full_like(C, float("nan"))
        gemm(M, N, K, Wt, (1, M, 0), X, (N, 1, 0), C, (N, 1, 0), alpha=0.75, beta=-0.5, Cin=Cin,
             bias=bias, bias_mode=2, act=1, pre=pre)
        torch.cuda.synchronize()
        outs.append((C, pre))
    z = 0.75 * (_bf(Wt).t().contiguous() @ _bf(X.float())) - 0.5 * Cin + bias[:, None]
    (C, pre), (C2, pre2) = outs
    assert rel_l2(pre.cpu().numpy(), z.cpu().numpy()) < 2e-6
    assert rel_l2(C.cpu().numpy(), torch.nn.functional.gelu(z).cpu().numpy()) < 2e-6
    assert rel_l2(C.cpu().numpy(), C2.cpu().numpy()) < 1e-6
    assert rel_l2(pre.cpu().numpy(), pre2.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("abf,bbf", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("M,N,K,batch", [(128, 128, 64, 1), (200, 72, 640, 2), (9, 300, 128, 3)])
def test_gemm_nt_fast_path_matches_generic_and_fp32(M, N, K, batch, abf, bbf, lib_opts):
    """Both operands contiguous along K (the fast path), BF16 or FP32 in
    memory, ragged M/N, several K steps and batches."""
    from diffnano_amd.fno import gemm
    g = torch.Generator(device="cuda").manual_seed(M + N + K)
    A = torch.randn(batch, M, K, device="cuda", generator=g)
    B = torch.randn(batch, N, K, device="cuda", generator=g)
    if abf:
        A = A.to(torch.bfloat16)
    if bbf:
        B = B.to(torch.bfloat16)
    bias = torch.randn(M, device="cuda", generator=g)
    C = torch.empty(batch, M, N, device="cuda")
    pre = torch.empty_like(C)
    args = (M, N, K, A, (K, 1, M * K), B, (1, K, N * K), C, (N, 1, M * N))
    gemm(*args, batch=batch, alpha=0.75, bias=bias, bias_mode=2, act=1, pre=pre)
    ref = 0.75 * torch.bmm(_bf(A.float()), _bf(B.float()).transpose(1, 2).contiguous()) \
        + bias[:, None]
    torch.cuda.synchronize()
    assert rel_l2(pre.cpu().numpy(), ref.cpu().numpy()) < 2e-6
    C2 = torch.empty_like(C)
    lib_opts.set("gemm_generic", 1)
    gemm(M, N, K, A, (K, 1, M * K), B, (1, K, N * K), C2, (N, 1, M * N), batch=batch,
         alpha=0.75, bias=bias, bias_mode=2, act=1)
    torch.cuda.synchronize()
    assert rel_l2(C.cpu().numpy(), C2.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("abf,bbf", [(False, False), (True, True), (False, True)])
@pytest.mark.parametrize("M,N,K", [(32, 1000, 96), (130, 70, 100), (6, 257, 128)])
def test_gemm_tile_layouts_match_generic(M, N, K, abf, bbf, ta, tb, lib_opts):
    """Layout-specialised tile kernel (K- or row-contiguous operands, K tails)
    against the element-wise kernel, bitwise, and against FP32 products."""
    from diffnano_amd.fno import gemm
    g = torch.Generator(device="cuda").manual_seed(M * 3 + N + K)
    A = torch.randn(K, M, device="cuda", generator=g) if ta else \
        torch.randn(M, K, device="cuda", generator=g)
    B = torch.randn(N, K, device="cuda", generator=g) if tb else \
        torch.randn(K, N, device="cuda", generator=g)
    if abf:
        A = A.to(torch.bfloat16)
    if bbf:
        B = B.to(torch.bfloat16)
    sA = (1, M, 0) if ta else (K, 1, 0)
    sB = (1, K, 0) if tb else (N, 1, 0)
    Cin = torch.randn(M, N, device="cuda", generator=g)
    C1 = torch.empty(M, N, device="cuda")
    C2 = torch.empty(M, N, device="cuda")
    gemm(M, N, K, A, sA, B, sB, C1, (N, 1, 0), beta=0.5, Cin=Cin)
    lib_opts.set("gemm_generic", 1)
    gemm(M, N, K, A, sA, B, sB, C2, (N, 1, 0), beta=0.5, Cin=Cin)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)
    Am = _bf(A.float()).t() if ta else _bf(A.float())
    Bm = _bf(B.float()).t() if tb else _bf(B.float())
    ref = Am.contiguous() @ Bm.contiguous() + 0.5 * Cin
    assert rel_l2(C1.cpu().numpy(), ref.cpu().numpy()) < 2e-6


@pytest.mark.parametrize("M,N,K", [(256, 1024, 16384), (96, 64, 4096)])
def test_gemm_nt_split_k(M, N, K):
    from diffnano_amd.fno import gemm_nt
    g = torch.Generator(device="cuda").manual_seed(K)
    A = torch.randn(M, K, device="cuda", generator=g)
    B = torch.randn(N, K, device="cuda", generator=g).to(torch.bfloat16)
    Cin = torch.randn(M, N, device="cuda", generator=g)
    out = torch.empty(M, N, device="cuda")
    gemm_nt(M, N, K, A, B, out, beta=0.5, Cin=Cin)
    ref = _bf(A) @ B.float().t() + 0.5 * Cin
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), ref.cpu().numpy()) < 2e-6


def _small_model(layers=2):
    from diffnano_amd.fno import FNO2d
    return FNO2d(32, 32, in_ch=3, out_ch=6, width=8, modes=(4, 4), layers=layers, hidden=16,
                 device="cuda", seed=1)


def test_fno_forward_matches_fft_oracle():
    m = _small_model(layers=3)
    torch.manual_seed(2)
    x = torch.randn(4, 3, 32, 32)
    out = m(x.cuda()).detach().double().cpu()
    p64 = {k: v.detach().double().cpu() for k, v in m.params.items()}
    ref = fno_oracle.fno_forward(p64, x.double(), 3, 4, 4)
    err = rel_l2(out.numpy(), ref.numpy())
    print("fno fwd rel L2", err)
    assert err < 2e-2


def test_fno_backward_matches_fft_oracle():
    m = _small_model(layers=2)
    torch.manual_seed(3)
    x = torch.randn(4, 3, 32, 32)
    w = torch.randn(4, 6, 32, 32)
    out = m(x.cuda())
    (out * w.cuda()).sum().backward()
    p64 = {k: v.detach().double().cpu().requires_grad_(True) for k, v in m.params.items()}
    ref = fno_oracle.fno_forward(p64, x.double(), 2, 4, 4)
    (ref * w.double()).sum().backward()
    worst = 0.0
    for k in m.names:
        e = rel_l2(m.params[k].grad.double().cpu().numpy(), p64[k].grad.numpy())
        worst = max(worst, e)
        print(f"grad {k}: {e:.3e}")
        assert e < 5e-2, k
    assert worst < 5e-2


@pytest.mark.parametrize("abf,bbf", [(False, True), (True, True), (False, False)])
@pytest.mark.parametrize("M,N,K,batch", [(256, 256, 2048, 1), (300, 522, 2112, 2),
                                         (512, 1024, 4096, 1), (520, 768, 1024, 1)])
def test_gemm_large_tile_kernel_matches_small_and_fp32(M, N, K, batch, abf, bbf, lib_opts):
    """The 256 x 256 MFMA 32x32x16 kernel (M, N >= 256, K >= 2048, K-contiguous operands):
    ragged M and N (vector and element epilogue paths), both bias modes,
    GELU with the pre-activation, beta * Cin, against FP32 products of the
    BF16-rounded operands and against the 128 x 128 kernel."""
    from diffnano_amd.fno import gemm
    g = torch.Generator(device="cuda").manual_seed(M + 3 * N + K)
    A = torch.randn(batch, M, K, device="cuda", generator=g)
    B = torch.randn(batch, N, K, device="cuda", generator=g)
    if abf:
        A = A.to(torch.bfloat16)
    if bbf:
        B = B.to(torch.bfloat16)
    Cin = torch.randn(batch, M, N, device="cuda", generator=g)
    ref0 = torch.bmm(_bf(A.float()), _bf(B.float()).transpose(1, 2).contiguous())
    for mode, bias in ((1, torch.randn(N, device="cuda", generator=g)),
                       (2, torch.randn(M, device="cuda", generator=g))):
        C = torch.empty(batch, M, N, device="cuda")
        pre = torch.empty_like(C)
        args = (M, N, K, A, (K, 1, M * K), B, (1, K, N * K), C, (N, 1, M * N))
        gemm(*args, batch=batch, alpha=0.75, beta=0.5, Cin=Cin, bias=bias, bias_mode=mode,
             act=1, pre=pre)
        bb = bias[None, None, :] if mode == 1 else bias[None, :, None]
        ref = 0.75 * ref0 + 0.5 * Cin + bb
        torch.cuda.synchronize()
        assert rel_l2(pre.cpu().numpy(), ref.cpu().numpy()) < 2e-6, mode
        assert rel_l2(C.cpu().numpy(), torch.nn.functional.gelu(ref).cpu().numpy()) < 2e-6
        C2 = torch.empty_like(C)
        lib_opts.set("gemm_small", 1)
        gemm(M, N, K, A, (K, 1, M * K), B, (1, K, N * K), C2, (N, 1, M * N), batch=batch,
             alpha=0.75, beta=0.5, Cin=Cin, bias=bias, bias_mode=mode, act=1)
        lib_opts.reset("gemm_small")
        torch.cuda.synchronize()
        assert rel_l2(C.cpu().numpy(), C2.cpu().numpy()) < 1e-6
        if abf and bbf and K >= 2048:
            # BF16 operands run on the direct-to-LDS kernel (K >= 1024); the
            # register-staged kernel (K >= 2048) does the same MFMA sequence
            # per accumulator: bitwise equal
            C3, pre3 = torch.empty_like(C), torch.empty_like(C)
            lib_opts.set("gemm_big_reg", 1)
            gemm(M, N, K, A, (K, 1, M * K), B, (1, K, N * K), C3, (N, 1, M * N), batch=batch,
                 alpha=0.75, beta=0.5, Cin=Cin, bias=bias, bias_mode=mode, act=1, pre=pre3)
            lib_opts.reset("gemm_big_reg")
            torch.cuda.synchronize()
            assert torch.equal(C, C3) and torch.equal(pre, pre3)


@pytest.mark.parametrize("M,N,K,batch", [(256, 256, 2048, 1), (300, 522, 2112, 2),
                                         (8192, 1024, 4096, 2)])
def test_lds_kernel_bf16_operand_equals_fp32_operand(M, N, K, batch, lib_opts):
    """The long-K product with A given as BF16 (the rounded copy the FNO's
    producers write) on the direct-to-LDS kernel equals, bit for bit, the
    product with the FP32 A that the register-staged kernel rounds on load:
    ragged tiles, a split-K batch and the spectral layer's own shape."""
    from diffnano_amd.fno import gemm
    g = torch.Generator(device="cuda").manual_seed(M + N + K + batch)
    A = torch.randn(batch, M, K, device="cuda", generator=g)
    B = torch.randn(batch, N, K, device="cuda", generator=g).to(torch.bfloat16)
    outs = []
    for a in (A, A.to(torch.bfloat16)):
        C = torch.full((batch, M, N), float("nan"), device="cuda")
        gemm(M, N, K, a, (K, 1, M * K), B, (1, K, N * K), C, (N, 1, M * N), batch=batch)
        outs.append(C)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1])
    ref = torch.bmm(_bf(A)[:1], B[:1].float().transpose(1, 2).contiguous())
    assert rel_l2(outs[1][:1].cpu().numpy(), ref.cpu().numpy()) < 2e-6


@pytest.mark.parametrize("M,K", [(32, 9), (32, 32), (128, 32), (6, 128), (32, 128)])
@pytest.mark.parametrize("path", ["default", "valu", "generic"])
def test_bf16_copy_of_the_mixing_output(M, K, path, lib_opts):
    """c16 receives C rounded to nearest even (torch's float -> bfloat16),
    on every kernel that takes the mixing products; other products reject it."""
    from diffnano_amd.fno import gemm
    N = 3000
    g = torch.Generator(device="cuda").manual_seed(M + 17 * K)
    W = torch.randn(M, K, device="cuda", generator=g)
    X = torch.randn(K, N, device="cuda", generator=g)
    Cin = torch.randn(M, N, device="cuda", generator=g)
    bias = torch.randn(M, device="cuda", generator=g)
    if path == "valu":
        lib_opts.set("gemm_pw_valu", 1)
    if path == "generic":
        lib_opts.set("gemm_generic", 1)
    C = torch.empty(M, N, device="cuda")
    c16 = torch.zeros(M, N, device="cuda", dtype=torch.bfloat16)
    gemm(M, N, K, W, (K, 1, 0), X, (N, 1, 0), C, (N, 1, 0), beta=1.0, Cin=Cin, bias=bias,
         bias_mode=2, act=1, c16=c16)
    torch.cuda.synchronize()
    assert torch.equal(c16, C.to(torch.bfloat16))
    big = torch.randn(256, 2048, device="cuda")
    with pytest.raises(RuntimeError, match="BF16 copy"):
        gemm(256, 256, 2048, big, (2048, 1, 0), big, (1, 2048, 0),
             torch.empty(256, 256, device="cuda"), (256, 1, 0),
             c16=torch.empty(256, 256, device="cuda", dtype=torch.bfloat16))


@pytest.mark.parametrize("n", [4096, 1001])
def test_gelu_backward_bf16_copy(n):
    from diffnano_amd.fno import gelu_bwd
    g = torch.Generator(device="cuda").manual_seed(n)
    dy = torch.randn(n, device="cuda", generator=g)
    pre = torch.randn(n, device="cuda", generator=g)
    dx, d16 = torch.empty_like(dy), torch.empty(n, device="cuda", dtype=torch.bfloat16)
    gelu_bwd(dy, pre, dx, d16)
    x = pre.double()
    ref = dy.double() * (0.5 * (1 + torch.erf(x / 2 ** 0.5)) + x * torch.exp(-x * x / 2) /
                         (2 * torch.pi) ** 0.5)
    torch.cuda.synchronize()
    assert rel_l2(dx.cpu().numpy(), ref.cpu().numpy()) < 1e-6
    assert torch.equal(d16, dx.to(torch.bfloat16))


def test_fno_bf16_copies_change_nothing():
    """The forward-DFT products read BF16 copies of h and d(pre) written by
    their producers (the direct-to-LDS kernel at this size: 128 x 128 grid,
    K = 16384): output and every parameter gradient bitwise equal to the
    products rounding the FP32 operands themselves."""
    from diffnano_amd.fno import FNO2d
    m = FNO2d(128, 128, in_ch=3, out_ch=6, width=32, modes=(16, 16), layers=2, hidden=32,
              device="cuda", seed=4)
    m.fused_inverse = False  # the layer kernel needs the copies; compare the two-kernel path
    torch.manual_seed(5)
    x = torch.randn(8, 3, 128, 128, device="cuda")
    w = torch.randn(8, 6, 128, 128, device="cuda")
    res = []
    for copies in (True, False):
        m.bf16_copies = copies
        for v in m.params.values():
            v.grad = None
        out = m(x)
        (out * w).sum().backward()
        torch.cuda.synchronize()
        res.append((out.detach().clone(), {k: v.grad.clone() for k, v in m.params.items()}))
    m.bf16_copies = True
    assert torch.equal(res[0][0], res[1][0])
    for k in m.names:
        assert torch.equal(res[0][1][k], res[1][1][k]), k


@pytest.mark.parametrize("out_ch,B,H", [(6, 8, 128), (3, 2, 40), (32, 1, 24)])
def test_fused_projection_head_matches_two_kernels(out_ch, B, H):
    """Inference with the projection head in one kernel (dnx_fno_proj) against
    the two channel-mixing kernels: the hidden layer is bitwise the same
    product, so the outputs differ only by the order of the last FP32 sum
    (and cells past a multiple of 32 are handled)."""
    from diffnano_amd.fno import FNO2d
    m = FNO2d(H, H, in_ch=3, out_ch=out_ch, width=32, modes=(4, 4), layers=2, hidden=128,
              device="cuda", seed=6)
    torch.manual_seed(7)
    x = torch.randn(B, 3, H, H, device="cuda")
    with torch.no_grad():
        m.fused_head = True
        a = m(x).clone()
        m.fused_head = False
        b = m(x).clone()
    m.fused_head = True
    err = rel_l2(a.double().cpu().numpy(), b.double().cpu().numpy())
    print("fused head rel L2", err)
    assert err < 1e-5
    assert float(b.abs().max()) > 0


def test_fused_projection_head_rejects_bad_shapes():
    import diffnano_amd._lib as L
    t = torch.zeros(1024, device="cuda")
    p = t.data_ptr()
    assert L.lib().dnx_fno_proj(p, 0, p, p, p, p, p, 32, 64, 6, 32, None) != 0
    assert b"fno_proj" in L.lib().dnx_last_error()
    assert L.lib().dnx_fno_proj(p, 0, p, p, p, p, p, 32, 128, 7, 32, None) != 0


@pytest.mark.parametrize("B,H", [(8, 128), (3, 40)])
def test_bf16_only_inference_is_bitwise_the_fp32_activation_path(B, H):
    """Inference writing the layers' activations only as BF16 copies (the
    pointwise kernel's FP32 output skipped, the projection head reading BF16)
    gives the bits of the path that writes and reads FP32 activations."""
    from diffnano_amd.fno import FNO2d
    m = FNO2d(H, H, in_ch=9, out_ch=6, width=32, modes=(4, 4), layers=3, hidden=128,
              device="cuda", seed=8)
    torch.manual_seed(9)
    x = torch.randn(B, 9, H, H, device="cuda")
    with torch.no_grad():
        m.bf16_inference = True
        a = m(x).clone()
        m.bf16_inference = False
        b = m(x).clone()
    m.bf16_inference = True
    assert torch.equal(a, b)
    assert float(a.abs().max()) > 0


@pytest.mark.parametrize("B,H,train", [(8, 128, False), (16, 64, True), (8, 32, True)])
def test_fused_inverse_and_pointwise_match_two_kernels(B, H, train):
    """Each layer's inverse DFT and pointwise part in one kernel against the
    inverse-DFT product followed by the pointwise kernel: the same sums,
    added in the accumulators instead of after an FP32 round, so outputs
    (and in training the parameter gradients) agree to FP32 rounding."""
    from diffnano_amd.fno import FNO2d
    m = FNO2d(H, H, in_ch=9, out_ch=6, width=32, modes=(8, 8), layers=3, hidden=128,
              device="cuda", seed=12)
    torch.manual_seed(13)
    x = torch.randn(B, 9, H, H, device="cuda")
    w = torch.randn(B, 6, H, H, device="cuda")
    res = []
    for fused in (True, False):
        m.fused_inverse = fused
        for v in m.params.values():
            v.grad = None
        if train:
            out = m(x)
            (out * w).sum().backward()
            res.append((out.detach().clone(), {k: v.grad.clone() for k, v in m.params.items()}))
        else:
            with torch.no_grad():
                res.append((m(x).clone(), {}))
    m.fused_inverse = True
    e = rel_l2(res[0][0].double().cpu().numpy(), res[1][0].double().cpu().numpy())
    print("fused inverse: output rel L2", e)
    # an FP32-rounding difference in one layer can flip the BF16 rounding of
    # a few of the next layer's inputs (2^-9 relative each)
    assert e < 1e-4 and float(res[1][0].abs().max()) > 0
    # the gradients (BF16 products throughout, against the FP64 oracle within
    # 2e-2) carry those flips back through every layer: 1.4e-4 seen on the
    # lifting weights
    for k in res[0][1]:
        eg = rel_l2(res[0][1][k].double().cpu().numpy(), res[1][1][k].double().cpu().numpy())
        assert eg < 1e-3, (k, eg)


@pytest.mark.parametrize("B,P,K,act,outs", [(8, 256, 64, 1, "all"), (16, 512, 1024, 0, "out"),
                                           (8, 1024, 256, 1, "bf16"), (8, 512, 128, 0, "bwd")])
def test_fused_inverse_kernel_matches_fp64(B, P, K, act, outs):
    """dnx_fno_inverse_pw on random operands against the formula in FP64 on
    the same BF16-rounded values (out, pre and the BF16 copy; "bwd": W
    transposed, no bias, times GELU' of a pre-activation)."""
    import diffnano_amd._lib as L
    torch.manual_seed(14)
    C = 32
    Y = torch.randn(C * B, K, device="cuda").to(torch.bfloat16)
    GT = (torch.randn(P, K, device="cuda") / K ** 0.5).to(torch.bfloat16)
    h16 = torch.randn(C, B * P, device="cuda").to(torch.bfloat16)
    W = torch.randn(C, C, device="cuda") / C ** 0.5
    b = torch.randn(C, device="cuda")
    bwd = outs == "bwd"
    out = (torch.full((C, B * P), float("nan"), device="cuda")
           if outs in ("all", "out", "bwd") else None)
    pre = torch.full((C, B * P), float("nan"), device="cuda") if outs == "all" else None
    o16 = (torch.zeros(C, B * P, device="cuda", dtype=torch.bfloat16)
           if outs in ("all", "bf16", "bwd") else None)
    gp = torch.randn(C, B * P, device="cuda") if bwd else None
    ptr = lambda t: t.data_ptr() if t is not None else None  # noqa: E731
    L.check(L.lib().dnx_fno_inverse_pw(Y.data_ptr(), GT.data_ptr(), h16.data_ptr(), W.data_ptr(),
                                       int(bwd), None if bwd else b.data_ptr(), ptr(out),
                                       ptr(pre), ptr(o16), act, ptr(gp), C, B, P, K, None))
    torch.cuda.synchronize()
    S = (Y.double() @ GT.double().T).reshape(C, B * P)            # rows (c, b) -> [c][b P]
    Wb = W.to(torch.bfloat16).double()
    if bwd:
        z = S + Wb.T @ h16.double()
        g = gp.double()
        gd = 0.5 * (1 + torch.erf(g / 2 ** 0.5)) + g * torch.exp(-0.5 * g * g) / (2 * math.pi) ** 0.5
        ref = z * gd
    else:
        z = S + Wb @ h16.double() + b.double()[:, None]
        ref = torch.nn.functional.gelu(z) if act else z
    if pre is not None:
        assert rel_l2(pre.double().cpu().numpy(), z.cpu().numpy()) < 1e-5
    if out is not None:
        assert rel_l2(out.double().cpu().numpy(), ref.cpu().numpy()) < 1e-5
    if o16 is not None:
        assert rel_l2(o16.double().cpu().numpy(), ref.cpu().numpy()) < 5e-3


def test_fused_inverse_rejects_bad_shapes():
    import diffnano_amd._lib as L
    t = torch.zeros(4096, device="cuda")
    p = t.data_ptr()
    assert L.lib().dnx_fno_inverse_pw(p, p, p, p, 0, p, p, None, None, 1, None, 32, 7, 256, 64,
                                      None) != 0
    assert b"fno_inverse_pw" in L.lib().dnx_last_error()
    assert L.lib().dnx_fno_inverse_pw(p, p, p, p, 0, p, p, None, None, 1, None, 32, 8, 100, 64,
                                      None) != 0
    assert L.lib().dnx_fno_inverse_pw(p, p, p, p, 1, None, p, None, None, 1, p, 32, 8, 256, 64,
                                      None) != 0   # GELU' and act together


def test_gemm_without_fp32_output_only_on_pointwise_shapes():
    """C = NULL (only the BF16 copy written) is taken by the pointwise MFMA
    kernel and rejected elsewhere."""
    from diffnano_amd.fno import gemm
    torch.manual_seed(10)
    W = torch.randn(32, 32, device="cuda")
    X = torch.randn(32, 4096, device="cuda")
    ref = torch.empty(32, 4096, device="cuda")
    r16 = torch.empty(32, 4096, device="cuda", dtype=torch.bfloat16)
    gemm(32, 4096, 32, W, (32, 1, 0), X, (4096, 1, 0), ref, (4096, 1, 0), act=1, c16=r16)
    o16 = torch.empty(32, 4096, device="cuda", dtype=torch.bfloat16)
    gemm(32, 4096, 32, W, (32, 1, 0), X, (4096, 1, 0), None, (4096, 1, 0), act=1, c16=o16)
    assert torch.equal(o16, r16)
    W2 = torch.randn(48, 32, device="cuda")
    with pytest.raises(RuntimeError, match="C may be NULL"):
        gemm(48, 4096, 32, W2, (32, 1, 0), X, (4096, 1, 0), None, (4096, 1, 0),
             c16=torch.empty(48, 4096, device="cuda", dtype=torch.bfloat16))


@pytest.mark.parametrize("s,n,beta", [(37, 100, 0.0), (1024, 1024, 0.5), (3, 70000, 1.0),
                                      (1, 5, 2.0)])
def test_sum_partials_matches_fp64(s, n, beta):
    """The split-K reduction (16 threads per output for few outputs, one per
    four outputs for many) against an FP64 sum."""
    from diffnano_amd.fno import sum_partials
    torch.manual_seed(11)
    part = torch.randn(s, n, device="cuda")
    out = torch.randn(n, device="cuda")
    ref = part.double().sum(0) + beta * out.double()
    sum_partials(part.reshape(-1), s, n, out, beta=beta)
    assert rel_l2(out.double().cpu().numpy(), ref.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("conj", [False, True])
def test_mode_mix_kernels_match_fp64(conj):
    """Per-mode complex channel mixing and its weight gradient (FP32 kernels)
    against FP64 products, with spectra rows longer than 2M (padding)."""
    from diffnano_amd.fno import mode_mix, mode_wgrad
    c, B, M, ld = 32, 24, 48, 100
    g = torch.Generator(device="cuda").manual_seed(7)
    X = torch.randn(c * B, ld, device="cuda", generator=g)
    dY = torch.randn(c * B, ld, device="cuda", generator=g)
    Wr = torch.randn(c, c, M, device="cuda", generator=g)
    Wi = torch.randn(c, c, M, device="cuda", generator=g)
    Y = torch.full((c * B, ld), 7.0, device="cuda")
    mode_mix(X, Wr, Wi, Y, B, M, ld, c, conj=conj)
    x = X.double().cpu().view(c, B, ld)
    xc = torch.complex(x[..., :M], x[..., M:2 * M])                   # [p][b][mu]
    w = torch.complex(Wr.double().cpu(), (-1 if conj else 1) * Wi.double().cpu())  # [p][q][mu]
    ref = torch.einsum("pbm,pqm->qbm", xc, w)
    y = Y.double().cpu().view(c, B, ld)
    assert rel_l2(y[..., :M].numpy(), ref.real.numpy()) < 1e-5
    assert rel_l2(y[..., M:2 * M].numpy(), ref.imag.numpy()) < 1e-5
    assert torch.all(y[..., 2 * M:] == 7.0)  # padding untouched
    Yb = torch.zeros((c * B, ld), device="cuda", dtype=torch.bfloat16)
    mode_mix(X, Wr, Wi, Yb, B, M, ld, c, conj=conj)  # BF16 output: Y rounded to nearest even
    assert torch.equal(Yb[:, :2 * M], Y[:, :2 * M].to(torch.bfloat16))
    ns = 3
    pr = torch.empty(ns, c, c, M, device="cuda")
    pi = torch.empty_like(pr)
    mode_wgrad(X, dY, pr, pi, B, M, ld, c, ns)
    d = dY.double().cpu().view(c, B, ld)
    dc = torch.complex(d[..., :M], d[..., M:2 * M])
    gref = torch.einsum("pbm,qbm->pqm", xc.conj(), dc)
    assert rel_l2(pr.sum(0).double().cpu().numpy(), gref.real.numpy()) < 1e-5
    assert rel_l2(pi.sum(0).double().cpu().numpy(), gref.imag.numpy()) < 1e-5


def test_fno_fast_mix_matches_generic_and_oracle():
    """Width-32 FNO: the dedicated mixing kernels against the GEMM-based
    mixing path and the FP64 FFT oracle, forward and gradients."""
    from diffnano_amd.fno import FNO2d
    torch.manual_seed(4)
    x = torch.randn(8, 3, 32, 32)
    w = torch.randn(8, 6, 32, 32)
    res = []
    for fast in (True, False):
        m = FNO2d(32, 32, in_ch=3, out_ch=6, width=32, modes=(4, 4), layers=2, hidden=16,
                  device="cuda", seed=5)
        m.fast_mix = fast
        out = m(x.cuda())
        (out * w.cuda()).sum().backward()
        res.append((out.detach().double().cpu(),
                    {k: m.params[k].grad.double().cpu() for k in m.names}, m))
    (o1, g1, m1), (o2, g2, _) = res
    assert rel_l2(o1.numpy(), o2.numpy()) < 2e-2
    p64 = {k: v.detach().double().cpu().requires_grad_(True) for k, v in m1.params.items()}
    ref = fno_oracle.fno_forward(p64, x.double(), 2, 4, 4)
    (ref * w.double()).sum().backward()
    assert rel_l2(o1.numpy(), ref.detach().numpy()) < 2e-2
    for k in m1.names:
        assert rel_l2(g1[k].numpy(), p64[k].grad.numpy()) < 5e-2, k
        assert rel_l2(g1[k].numpy(), g2[k].numpy()) < 5e-2, k


@pytest.mark.parametrize("O,I,n,ns", [(40, 9, 64 * 37, 37), (32, 32, 64 * 64, 8), (6, 128, 4096, 1)])
def test_pointwise_weight_and_bias_gradient_kernel(O, I, n, ns):
    """dnx_fno_wgrad: ragged 32 x 32 tiles, several column splits, against
    FP64 sums."""
    import diffnano_amd._lib as L
    g = torch.Generator(device="cuda").manual_seed(O + I + n)
    dY = torch.randn(O, n, device="cuda", generator=g)
    X = torch.randn(I, n, device="cuda", generator=g)
    part = torch.empty(ns, O * I, device="cuda")
    pb = torch.empty(ns, O, device="cuda")
    L.check(L.lib().dnx_fno_wgrad(dY.data_ptr(), X.data_ptr(), part.data_ptr(), pb.data_ptr(),
                                  O, I, n, ns, L.stream_handle(dY.device)))
    ref = dY.double().cpu() @ X.double().cpu().t()
    # FP32 running sums over up to 4096 columns per thread: ~sqrt(n) eps
    assert rel_l2(part.sum(0).view(O, I).double().cpu().numpy(), ref.numpy()) < 1e-5
    assert rel_l2(pb.sum(0).double().cpu().numpy(), dY.double().cpu().sum(1).numpy()) < 1e-5
