"""GPU parity of the BF16 MFMA convolution kernels against the float64 conv
oracle.  Inputs and weights are drawn already rounded to BF16, so both sides
see identical operands; the kernels accumulate in FP32 and round the output
to BF16, so outputs are compared at a BF16 tolerance (2^-8 relative to the
largest entry plus the FP32 accumulation error)."""
import numpy as np
import pytest

from oracle import conv_oracle as C

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _bf16_dev(a):
    torch = _torch()
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()
    return t.to(torch.bfloat16).view(torch.int16)


def _from_bf16(t):
    torch = _torch()
    return t.view(torch.bfloat16).float().cpu().numpy().astype(np.float64)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


@pytest.mark.parametrize("shape,cin,cout", [((5, 7, 9), 16, 16), ((9, 5, 17), 16, 32), ((7, 7, 7), 32, 64),
                                            ((5, 9, 5), 64, 32), ((3, 3, 3), 64, 64)])
def test_conv_fwd(shape, cin, cout):
    from pyfantom_amd import _lib
    torch = _torch()
    print(f"conv fwd {shape} {cin}->{cout}", flush=True)
    rng = np.random.default_rng(0)
    x = C.bf16_round(rng.standard_normal(shape + (cin,)))
    w = C.bf16_round(rng.standard_normal((27, cout, cin)) / np.sqrt(27 * cin))
    b = rng.standard_normal(cout).astype(np.float32).astype(np.float64)
    skip = C.bf16_round(rng.standard_normal(shape + (cout,)))
    xd, wd, sd = _bf16_dev(x), _bf16_dev(w), _bf16_dev(skip)
    bd = torch.from_numpy(b.astype(np.float32)).cuda()
    out = torch.empty(int(np.prod(shape)) * cout, dtype=torch.int16, device="cuda")
    of = torch.empty(int(np.prod(shape)) * cout, dtype=torch.float32, device="cuda")
    for relu, use_skip in ((0, False), (1, True)):
        _lib.call("fantom_conv3d_fwd_bf16", *shape, cin, cout, xd.data_ptr(), wd.data_ptr(), bd.data_ptr(),
                  sd.data_ptr() if use_skip else None, out.data_ptr(), of.data_ptr(), relu)
        ref = C.conv3d(x, w, b) + (skip if use_skip else 0)
        if relu:
            ref = np.maximum(ref, 0)
        ref = ref.reshape(-1)
        assert _rel(of.cpu().numpy().astype(np.float64), ref) < 1e-5
        assert _rel(_from_bf16(out).reshape(-1), ref) < 2 ** -8


@pytest.mark.parametrize("shape,cin,cout", [((41, 45, 37), 16, 16), ((24, 200, 130), 16, 16), ((22, 70, 49), 16, 32),
                                            ((41, 45, 37), 32, 16), ((20, 130, 70), 32, 64), ((20, 130, 70), 64, 32),
                                            ((17, 66, 61), 64, 64)])
def test_conv_fwd_planes(shape, cin, cout):
    """Volumes of at least 2^16 voxels take the kernel that walks planes
    along D (LDS ring of three planes, chunks of planes per workgroup): edge
    tiles in W and H, one-plane and multi-plane chunks, bias, skip, ReLU and
    the FP32 copy on and off."""
    from pyfantom_amd import _lib
    torch = _torch()
    print(f"conv fwd (planes) {shape} {cin}->{cout}", flush=True)
    rng = np.random.default_rng(1)
    x = C.bf16_round(rng.standard_normal(shape + (cin,)))
    w = C.bf16_round(rng.standard_normal((27, cout, cin)) / np.sqrt(27 * cin))
    b = rng.standard_normal(cout).astype(np.float32).astype(np.float64)
    skip = C.bf16_round(rng.standard_normal(shape + (cout,)))
    xd, wd, sd = _bf16_dev(x), _bf16_dev(w), _bf16_dev(skip)
    bd = torch.from_numpy(b.astype(np.float32)).cuda()
    n = int(np.prod(shape)) * cout
    out = torch.empty(n, dtype=torch.int16, device="cuda")
    of = torch.full((n,), np.nan, dtype=torch.float32, device="cuda")
    base = C.conv3d(x, w, None)
    for relu, use_skip, use_bias, f32 in ((0, False, True, True), (1, True, True, False), (1, False, False, True)):
        _lib.call("fantom_conv3d_fwd_bf16", *shape, cin, cout, xd.data_ptr(), wd.data_ptr(),
                  bd.data_ptr() if use_bias else None, sd.data_ptr() if use_skip else None, out.data_ptr(),
                  of.data_ptr() if f32 else None, relu)
        ref = base + (b if use_bias else 0) + (skip if use_skip else 0)
        if relu:
            ref = np.maximum(ref, 0)
        ref = ref.reshape(-1)
        if f32:
            assert _rel(of.cpu().numpy().astype(np.float64), ref) < 1e-5
        assert _rel(_from_bf16(out).reshape(-1), ref) < 2 ** -8


@pytest.mark.parametrize("shape,cin,cout", [((5, 7, 9), 16, 32), ((7, 5, 5), 64, 32), ((5, 5, 5), 32, 64)])
def test_conv_backward(shape, cin, cout):
    from pyfantom_amd import _lib
    torch = _torch()
    print(f"conv bwd {shape} {cin}->{cout}", flush=True)
    rng = np.random.default_rng(1)
    x = C.bf16_round(rng.standard_normal(shape + (cin,)))
    w = C.bf16_round(rng.standard_normal((27, cout, cin)) / np.sqrt(27 * cin))
    dy = C.bf16_round(rng.standard_normal(shape + (cout,)))
    xd, wd, dyd = _bf16_dev(x), _bf16_dev(w), _bf16_dev(dy)
    n = int(np.prod(shape))
    # input gradient: forward with the flipped, transposed weights
    wt = torch.empty(27 * cout * cin, dtype=torch.int16, device="cuda")
    _lib.call("fantom_flip_transpose_bf16", cout, cin, wd.data_ptr(), wt.data_ptr())
    dx = torch.empty(n * cin, dtype=torch.int16, device="cuda")
    dxf = torch.empty(n * cin, dtype=torch.float32, device="cuda")
    _lib.call("fantom_conv3d_fwd_bf16", *shape, cout, cin, dyd.data_ptr(), wt.data_ptr(), None, None, dx.data_ptr(),
              dxf.data_ptr(), 0)
    ref = C.conv3d_bwd_data(dy, w).reshape(-1)
    assert _rel(dxf.cpu().numpy().astype(np.float64), ref) < 1e-5
    # weight gradient
    nparts = _lib.load().fantom_conv3d_wgrad_parts(*shape, cin, cout, None)
    part = torch.empty(nparts * 27 * cout * cin, dtype=torch.float32, device="cuda")
    dw = torch.empty(27 * cout * cin, dtype=torch.float32, device="cuda")
    _lib.call("fantom_conv3d_wgrad_bf16", *shape, cin, cout, xd.data_ptr(), dyd.data_ptr(), part.data_ptr(),
              dw.data_ptr())
    refw = C.conv3d_bwd_weight(x, dy).reshape(-1)
    assert _rel(dw.cpu().numpy().astype(np.float64), refw) < 1e-5
    # bias gradient
    bpart = torch.empty(256 * cout, dtype=torch.float32, device="cuda")
    db = torch.empty(cout, dtype=torch.float32, device="cuda")
    _lib.call("fantom_bias_grad_bf16", n, cout, dyd.data_ptr(), bpart.data_ptr(), db.data_ptr())
    assert _rel(db.cpu().numpy().astype(np.float64), dy.reshape(-1, cout).sum(0)) < 1e-5


@pytest.mark.parametrize("shape,ch", [((9, 5, 7), 16), ((7, 11, 5), 32), ((5, 5, 9), 64)])
def test_vol_transfers(shape, ch):
    from pyfantom_amd import _lib
    torch = _torch()
    rng = np.random.default_rng(2)
    x = C.bf16_round(rng.standard_normal(shape + (ch,)))
    cs = tuple((s - 1) // 2 + 1 for s in shape)
    xc = C.bf16_round(rng.standard_normal(cs + (ch,)))
    xd, xcd = _bf16_dev(x), _bf16_dev(xc)
    r = torch.empty(int(np.prod(cs)) * ch, dtype=torch.int16, device="cuda")
    _lib.call("fantom_vol_restrict_bf16", *shape, ch, xd.data_ptr(), r.data_ptr(), 0.125)
    assert _rel(_from_bf16(r).reshape(-1), C.restrict_vol(x, 0.125).reshape(-1)) < 2 ** -8
    p = torch.empty(int(np.prod(shape)) * ch, dtype=torch.int16, device="cuda")
    _lib.call("fantom_vol_prolong_bf16", *shape, ch, xcd.data_ptr(), p.data_ptr(), 1.0, 0)
    assert _rel(_from_bf16(p).reshape(-1), C.prolong_vol(xc, shape).reshape(-1)) < 2 ** -8
    # accumulating form: out += 0.5 * P xc
    base = C.bf16_round(rng.standard_normal(shape + (ch,)))
    p2 = _bf16_dev(base)
    _lib.call("fantom_vol_prolong_bf16", *shape, ch, xcd.data_ptr(), p2.data_ptr(), 0.5, 1)
    ref = base + 0.5 * C.prolong_vol(xc, shape)
    assert _rel(_from_bf16(p2).reshape(-1), ref.reshape(-1)) < 2 ** -7
