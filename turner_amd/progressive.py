"""Progressive rendering until a relMSE target: the plain cosine tracer and
the learned-guided tracer, timed the same way.

Both add passes of `spp_per_pass` samples per pixel (fresh random streams per
pass through sample_offset) into one FP32 accumulator and stop when the
running mean reaches the target relMSE against a reference image.  Only the
render (and, for the guided tracer, the online training that follows each
training pass) is timed; evaluating the error is not.  The guided tracer
trains on the radiance samples of its first `train_passes` passes
(train_samples per pass) and keeps every pass in the image: one-sample MIS
with the exact mixture pdf keeps each pass unbiased whatever the state of
the network.
"""
import time

import torch

from . import TracerConfig, render_accumulate
from .relmse import relmse


def _conf(width, height, depth, flags, spp, seed):
    return TracerConfig(width=width, aspect=width / height, max_recursion_depth=depth, num_pixel_samples=spp,
                        num_monte_carlo_samples=1, flags=flags, seed=seed)


def time_to_relmse(tree, scene, ref, depth, flags, target=0.01, guide=None, spp_per_pass=64,
                   train_passes=16, train_samples=1 << 18, train_spp=2, max_seconds=60.0, seed=1,
                   max_spp=1 << 16, ref_var=None, keep_training_passes=False, train_every=0,
                   passes_per_render=1, min_pass=0, image_mis=None, defer_training=False):
    """Returns a dict: seconds (timed render + training), spp, relmse,
    reached (bool), rays, history [(seconds, spp, relmse)].  ref_var: the
    reference's per-pixel variance (relmse subtracts it).
    keep_training_passes=False leaves the passes rendered while the guide is
    still training out of the image once a later pass exists (a choice fixed
    in advance, so the image stays unbiased); their time still counts.
    train_every=k > 0 keeps training after that phase on every k-th image
    pass (those passes are rendered before the update and stay in the
    image).  passes_per_render=p renders the data of p training passes in
    one render of p * train_spp samples per pixel and trains p passes on it
    (each on its own uniform subset of train_samples records); a small render
    costs about as much as a larger one, its slowest rays bound it.
    min_pass=q > 0 sizes each image pass after the first from the error so
    far: relMSE falls as 1/spp, so about spp * err / target samples reach the
    target; the next pass takes the missing samples rounded up to a multiple
    of q, at most spp_per_pass (the same rule for both tracers).
    image_mis (guided): the guide's share of the one-sample MIS mixture from
    the first image pass on (the training passes keep the guide's own
    setting); every pass stays unbiased whatever the share.
    defer_training (guided): each training pass's records train during the
    next training pass's render (TN_RENDER_DEFER_TRAINING); the last set
    trains before the first image pass (tn_guide_flush), so every image pass
    samples from the final weights.  Every call returns with its training
    done, so the timed windows hold all of it."""
    from .guide import render_guided
    h, w = ref.shape[0], ref.shape[1]
    acc = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
    tacc = acc if keep_training_passes else torch.zeros_like(acc)
    spent, spp, rays, hist = 0.0, 0, 0, []
    spp_img = 0  # samples in acc
    err = float("inf")
    k = 0
    ppr = max(1, int(passes_per_render))
    train_renders = -(-train_passes // ppr) if train_passes > 0 else 0
    while spent < max_seconds and spp < max_spp:
        training = guide is not None and k < train_renders
        late = (guide is not None and not training and train_every > 0
                and (k - train_renders) % train_every == 0)
        extra = min(ppr, train_passes - k * ppr) - 1 if training else 0  # more passes on this render's data
        n = train_spp * (extra + 1) if training else spp_per_pass
        if not training and min_pass > 0 and spp_img > 0 and err > target and err != float("inf"):
            missing = spp_img * err / target - spp_img
            n = int(min(spp_per_pass, max(min_pass, -(-missing // min_pass) * min_pass)))
        if guide is not None and image_mis is not None and k == train_renders:
            guide.set_options(guide.lr, guide.batch, image_mis, guide.epochs)
        conf = _conf(w, h, depth, flags, n, seed)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if guide is not None and defer_training and k == train_renders:
            guide.flush()  # the last training pass's records, before the image passes
        if guide is None:
            st = render_accumulate(tree, scene["camera"], scene["lights"], conf, acc, sample_offset=spp)
        else:
            st = render_guided(tree, scene["camera"], scene["lights"], conf, guide, tacc if training else acc,
                               train_samples=train_samples if (training or late) else 0, sample_offset=spp,
                               defer=defer_training and training)
            for _ in range(extra):
                guide.train_recorded(train_samples)
        torch.cuda.synchronize()
        spent += time.perf_counter() - t0
        spp += n
        if acc is tacc or not training:
            spp_img += n
        rays += st["rays"] + st["shadow_rays"]
        k += 1
        img = acc / spp_img if spp_img else tacc / spp
        err = relmse(img, ref, ref_var)
        hist.append((spent, spp, err))
        if err <= target:
            break
    return {"seconds": spent, "spp": spp, "relmse": err, "reached": err <= target, "rays": rays,
            "history": hist}
