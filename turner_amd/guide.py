"""Learned path guiding on the MI355X (tn_guide_* of include/turner_amd.h).

No reference counterpart: turner's pathtracer draws indirect directions
uniformly (lib/sampling.h:20-32).  A Guide is an MLP (78 -> 64 -> 64 -> 64 ->
40, ReLU) from a hit point and its normal to an 8-lobe von Mises-Fisher
mixture, trained online with Adam on radiance samples the tracer records
(render_guided(train_samples=...)), and used as the proposal of one-sample
MIS with the cosine lobe.

Batch arguments are float32 device tensors (torch) of shape (n, 3) / (n,).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import TN_GUIDE_OUTPUTS, TN_GUIDE_PARAMS, check, lib, ptr

__all__ = ["Guide", "render_guided", "GUIDE_PARAMS", "GUIDE_OUTPUTS"]

GUIDE_PARAMS = TN_GUIDE_PARAMS
GUIDE_OUTPUTS = TN_GUIDE_OUTPUTS


def _stream(t):
    import torch
    return torch.cuda.current_stream(t.device).cuda_stream


def _dev(t, cols=None):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32):
        raise TypeError("expected a float32 device tensor")
    t = t.contiguous()
    if cols is not None and (t.dim() != 2 or t.shape[1] != cols):
        raise ValueError(f"expected shape (n, {cols})")
    return t


class Guide:
    """tn_guide: parameters, Adam state and training set on the device."""

    def __init__(self, bmin, bmax, seed=1, lr=1e-3, batch=16384, mis_guided=0.5, epochs=1):
        h = C.c_void_p()
        bmin = np.ascontiguousarray(bmin, np.float32)
        bmax = np.ascontiguousarray(bmax, np.float32)
        check(lib().tn_guide_create(ptr(bmin), ptr(bmax), seed, C.byref(h)))
        self.handle = h
        self.set_options(lr, batch, mis_guided, epochs)

    @classmethod
    def for_tree(cls, tree, **kw):
        box = np.asarray(tree.box(), np.float32)
        return cls(box[:3], box[3:], **kw)

    def close(self):
        if getattr(self, "handle", None):
            lib().tn_guide_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_options(self, lr=1e-3, batch=16384, mis_guided=0.5, epochs=1):
        check(lib().tn_guide_set_options(self.handle, lr, batch, mis_guided, epochs))
        self.lr, self.batch, self.mis_guided, self.epochs = lr, batch, mis_guided, epochs

    def params(self):
        theta = np.zeros(GUIDE_PARAMS, np.float32)
        check(lib().tn_guide_get_params(self.handle, ptr(theta)))
        return theta

    def set_params(self, theta):
        theta = np.ascontiguousarray(theta, np.float32)
        if theta.shape != (GUIDE_PARAMS,):
            raise ValueError(f"expected {GUIDE_PARAMS} parameters")
        check(lib().tn_guide_set_params(self.handle, ptr(theta)))

    def forward(self, p, n):
        import torch
        p, n = _dev(p, 3), _dev(n, 3)
        out = torch.empty((p.shape[0], GUIDE_OUTPUTS), dtype=torch.float32, device=p.device)
        check(lib().tn_guide_forward(self.handle, ptr(p), ptr(n), p.shape[0], ptr(out), _stream(p)))
        return out

    def pdf(self, p, n, w):
        import torch
        p, n, w = _dev(p, 3), _dev(n, 3), _dev(w, 3)
        out = torch.empty(p.shape[0], dtype=torch.float32, device=p.device)
        check(lib().tn_guide_pdf(self.handle, ptr(p), ptr(n), ptr(w), p.shape[0], ptr(out),
                                 _stream(p)))
        return out

    def sample(self, p, n, u):
        import torch
        p, n, u = _dev(p, 3), _dev(n, 3), _dev(u, 3)
        w = torch.empty_like(p)
        pdf = torch.empty(p.shape[0], dtype=torch.float32, device=p.device)
        check(lib().tn_guide_sample(self.handle, ptr(p), ptr(n), ptr(u), p.shape[0], ptr(w),
                                    ptr(pdf), _stream(p)))
        return w, pdf

    def loss_grad(self, p, n, w, c):
        import torch
        p, n, w, c = _dev(p, 3), _dev(n, 3), _dev(w, 3), _dev(c)
        grad = torch.empty(GUIDE_PARAMS, dtype=torch.float32, device=p.device)
        loss = C.c_float()
        check(lib().tn_guide_loss_grad(self.handle, ptr(p), ptr(n), ptr(w), ptr(c), p.shape[0],
                                       ptr(grad), C.byref(loss), _stream(p)))
        return loss.value, grad

    def train_step(self, p, n, w, c, want_loss=True):
        p, n, w, c = _dev(p, 3), _dev(n, 3), _dev(w, 3), _dev(c)
        loss = C.c_float()
        check(lib().tn_guide_train_step(self.handle, ptr(p), ptr(n), ptr(w), ptr(c), p.shape[0],
                                        C.byref(loss) if want_loss else None, _stream(p)))
        return loss.value if want_loss else None

    def train_recorded(self, train_samples=1 << 18, stream=None):
        """One more training pass on the records of the last training render
        (a new uniform subset of train_samples of them)."""
        import torch
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        check(lib().tn_guide_train_recorded(self.handle, train_samples, s))

    def flush(self, stream=None):
        """Trains the records a deferred training render left (tn_guide_flush)."""
        import torch
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        check(lib().tn_guide_flush(self.handle, s))

    def info(self):
        ns, st, loss = C.c_uint32(), C.c_uint32(), C.c_float()
        check(lib().tn_guide_info(self.handle, C.byref(ns), C.byref(st), C.byref(loss)))
        return {"train_samples": ns.value, "steps": st.value, "loss": loss.value}

    def train_set(self, cap=1 << 22):
        """The training set of the last training render as a dict of host
        arrays: p, c, n, cos_pdf, w, path (the sample index within the render
        call, modulo 2^26), level."""
        n = self.info()["train_samples"]
        n = min(n, cap)
        buf = np.zeros((max(n, 1), 12), np.float32)
        got = check(lib().tn_guide_get_train_set(self.handle, ptr(buf), n))
        buf = buf[:got]
        tag = buf[:, 11].copy().view(np.uint32)
        return {"p": buf[:, 0:3], "c": buf[:, 3], "n": buf[:, 4:7], "cos_pdf": buf[:, 7],
                "w": buf[:, 8:11], "path": tag >> 6, "level": tag & 63}


def render_guided(tree, cam, lights, conf, guide, accum, train_samples=0, sample_offset=0, defer=False):
    """One guided render call adding per-pixel radiance sums into accum (a
    (h, w, 4) float32 device tensor).  train_samples > 0 (needs
    num_monte_carlo_samples == 1) also records radiance samples and trains
    the guide on train_samples of them afterwards -- or, with defer=True
    (TN_RENDER_DEFER_TRAINING), during the next deferred call, beside its
    render (guide.flush() trains them at once).  Returns the stats dict."""
    from . import _camera, _lights, _params
    w, h = conf.width, conf.height()
    accum = _dev(accum)
    if accum.numel() != w * h * 4:
        raise ValueError("accum must hold width * height * 4 floats")
    p = _params(conf, w, h)
    p.sample_offset = sample_offset
    if defer:
        p.flags |= _lib.TN_RENDER_DEFER_TRAINING
    st = _lib.RenderStats()
    check(lib().tn_render_guided(tree.handle, C.byref(_camera(cam)), _lights(lights), len(lights),
                                 C.byref(p), guide.handle, train_samples, ptr(accum), C.byref(st),
                                 _stream(accum)))
    return {"rays": st.rays, "primary_rays": st.primary_rays, "shadow_rays": st.shadow_rays,
            "ms": st.ms}
