"""Error of a rendered image against a high-sample reference, the measure the
guided-vs-plain comparison is quoted in: relMSE = mean over pixels and RGB
channels of (I - R)^2 / (R^2 + 0.01).

A finite-sample reference R carries its own noise: for independent unbiased
I and R, E[(I - R)^2] = MSE(I) + Var(R).  The reference is therefore rendered
in K independent batches, Var(R) per pixel is estimated from the spread of
the batch means (divided by K), and relmse(..., ref_var=) subtracts it, which
estimates the relMSE of I against the exact image without the reference's
noise floor.
"""
import torch

from . import TracerConfig, render_accumulate


def relmse(img, ref, ref_var=None):
    """img, ref: (h, w, 4) tensors of per-pixel MEANS (same device); ref_var:
    per-pixel variance of ref's mean (optional)."""
    d = img[..., :3].double() - ref[..., :3].double()
    num = d * d
    if ref_var is not None:
        num = num - ref_var[..., :3].double()
    return float((num / (ref[..., :3].double() ** 2 + 0.01)).mean())


def render_plain(tree, scene, conf, sample_offset=0):
    """Per-pixel radiance SUMS of conf.num_pixel_samples plain samples."""
    acc = torch.zeros((conf.height(), conf.width, 4), dtype=torch.float32, device="cuda")
    render_accumulate(tree, scene["camera"], scene["lights"], conf, acc, sample_offset)
    return acc


def reference_image(tree, scene, width, depth, flags, spp, seed, batch_spp=256, height=None,
                    with_var=False):
    """Mean image of spp plain samples (rounded up to a multiple of batch_spp)
    rendered in batches of batch_spp pixel samples (accumulated in FP64 on the
    device).  with_var=True also returns
    the per-pixel variance of that mean, estimated from the batch means."""
    h = height or width
    spp = -(-spp // batch_spp) * batch_spp  # equal batches
    s1 = torch.zeros((h, width, 4), dtype=torch.float64, device="cuda")
    s2 = torch.zeros_like(s1)
    done, k = 0, 0
    while done < spp:
        n = min(batch_spp, spp - done)
        conf = TracerConfig(width=width, aspect=width / h, max_recursion_depth=depth, num_pixel_samples=n,
                            num_monte_carlo_samples=1, flags=flags, seed=seed)
        b = render_plain(tree, scene, conf, sample_offset=done).double() / n
        s1 += b
        s2 += b * b
        done += n
        k += 1
    mean = s1 / k
    if not with_var:
        return mean.float()
    var_of_mean = (s2 / k - mean * mean).clamp_min(0.0) * (k / max(k - 1, 1)) / k
    return mean.float(), var_of_mean.float()
