"""Image minima on the GPU (im - im.min() of src/fluidimage/works/piv/
singlepass.py:206): the vectorised pass over contiguous aligned images and
the row-by-row pass over strided ones give numpy's minimum exactly."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
@pytest.mark.parametrize("shape", [(37, 53), (64, 96), (1, 7), (300, 411)])
def test_image_minima(dtype, shape):
    import torch

    from fluidimage_amd.device import ImageSet, to_device_image

    print("case: minima", dtype.__name__, shape, flush=True)
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    dev = torch.device("cuda")
    hi = 255 if dtype == np.uint8 else 4095
    arrs = []
    for k in range(5):
        a = (rng.random(shape) * hi + 3).astype(dtype)
        # the minimum somewhere inside, at the very end, or in the tail
        pos = [(shape[0] // 2, shape[1] // 3), (shape[0] - 1, shape[1] - 1), (0, 0)][k % 3]
        a[pos] = dtype(1) if dtype != np.float32 else np.float32(-2.5)
        arrs.append(a)
    flat = ImageSet([to_device_image(a, dev) for a in arrs])
    assert flat._flat_dtype >= 0
    big = [to_device_image(np.pad(a, ((0, 0), (0, 3)), constant_values=hi), dev) for a in arrs]
    views = ImageSet([t[:, :shape[1]] for t in big])  # pitch != width: row kernel
    assert views._flat_dtype == -1
    torch.cuda.synchronize()
    expect = np.array([a.min() for a in arrs], np.float32)
    assert np.array_equal(flat.mins.cpu().numpy(), expect)
    assert np.array_equal(views.mins.cpu().numpy(), expect)
