"""Error of a rendered image against a high-sample reference, the measure the
guided-vs-plain comparison is quoted in: relMSE = mean over pixels and RGB
channels of (I - R)^2 / (R^2 + 0.01)."""
import torch

from . import TracerConfig, render_accumulate


def relmse(img, ref):
    """img, ref: (h, w, 4) tensors of per-pixel MEANS (same device)."""
    d = (img[..., :3].double() - ref[..., :3].double())
    return float((d * d / (ref[..., :3].double() ** 2 + 0.01)).mean())


def render_plain(tree, scene, conf, sample_offset=0):
    """Per-pixel radiance SUMS of conf.num_pixel_samples plain samples."""
    acc = torch.zeros((conf.height(), conf.width, 4), dtype=torch.float32, device="cuda")
    render_accumulate(tree, scene["camera"], scene["lights"], conf, acc, sample_offset)
    return acc


def reference_image(tree, scene, width, depth, flags, spp, seed, batch_spp=256, height=None):
    """Mean image of spp plain samples (in batches of batch_spp pixel samples,
    accumulated in FP64 on the device)."""
    h = height or width
    total = torch.zeros((h, width, 4), dtype=torch.float64, device="cuda")
    done = 0
    while done < spp:
        k = min(batch_spp, spp - done)
        conf = TracerConfig(width=width, aspect=width / h, max_recursion_depth=depth, num_pixel_samples=k,
                            num_monte_carlo_samples=1, flags=flags, seed=seed)
        total += render_plain(tree, scene, conf, sample_offset=done).double()
        done += k
    return (total / spp).float()
