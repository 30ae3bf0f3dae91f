"""Volume::downsample and filter_volume: the oracle pinned by closed forms
(CPU), the GPU kernels against it (gpu)."""
import numpy as np
import pytest
import torch

from oracle import volume_ops as OVO


def test_oracle_downsample_closed_forms():
    res = (4, 4, 2)
    vol = np.arange(32, dtype=np.uint8)
    mx = OVO.downsample(vol, res, OVO.MAX)
    mn = OVO.downsample(vol, res, OVO.MIN)
    mean = OVO.downsample(vol, res, OVO.MEAN)
    # output (0,0,0) reduces texels (x,y,z) in {0,1}^3: indices 0,1,4,5,16,17,20,21
    nb = np.array([0, 1, 4, 5, 16, 17, 20, 21])
    assert mx[0] == nb.max() and mn[0] == nb.min() and mean[0] == nb.sum() // 8
    assert len(mx) == 2 * 2 * 1


def test_oracle_filter_closed_forms():
    res = (9, 9, 9)
    # width 1: identity
    v = np.random.default_rng(0).random(729)
    assert np.allclose(OVO.filter_volume(v, res, 1), v)
    # a constant volume stays constant where the whole kernel is inside
    c = OVO.filter_volume(np.full(729, 0.5), res, 5).reshape(9, 9, 9)
    assert np.allclose(c[2:7, 2:7, 2:7], 0.5)
    assert c[0, 0, 0] < 0.5  # the zero border pulls the corner down
    # the shader's weight: s^2 multiplies the squared distance
    w = OVO.gaussian_weights(3)
    s2 = (1.5 / 2.4) ** 2
    assert w[1, 1, 2] / w[1, 1, 1] == pytest.approx(np.exp(-0.5 * s2))


@pytest.mark.gpu
def test_downsample_gpu(device):
    from vkhr_amd import capi
    from vkhr_amd.capi import call, ptr
    res = (64, 32, 16)
    vol = np.random.default_rng(1).integers(0, 256, np.prod(res)).astype(np.uint8)
    src = torch.from_numpy(vol).to(device)
    for filt in (OVO.MAX, OVO.MIN, OVO.MEAN):
        out = torch.empty(np.prod(res) // 8, dtype=torch.uint8, device=device)
        call("vkhr_volume_downsample", ptr(src), *res, filt, ptr(out), capi.stream_of())
        assert np.array_equal(out.cpu().numpy(), OVO.downsample(vol, res, filt)), filt
    with pytest.raises(capi.VkhrError):
        call("vkhr_volume_downsample", ptr(src), 63, 32, 16, 0, ptr(out), capi.stream_of())


@pytest.mark.gpu
@pytest.mark.parametrize("width", [1, 3, 5])
def test_filter_gpu(device, width):
    from vkhr_amd import capi
    from vkhr_amd.capi import call, ptr
    res = (24, 20, 16)
    vol = np.random.default_rng(2).random(np.prod(res)).astype(np.float32)
    src = torch.from_numpy(vol).to(device)
    out = torch.empty_like(src)
    call("vkhr_volume_filter", ptr(src), *res, width, ptr(out), capi.stream_of())
    want = OVO.filter_volume(vol, res, width)
    got = out.cpu().numpy().astype(np.float64)
    assert np.abs(got - want).max() <= 1e-5 * max(1.0, np.abs(want).max())
