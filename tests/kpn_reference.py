"""Plain PyTorch fp32 restatement of the kernel-predicting denoiser, the
reference the HIP kernels are checked against. The GPU stores activations,
weights and input features in bf16; `bf16=True` rounds the same tensors at the
same places (round to nearest even), so the remaining difference is the FP32
summation order. `dtype=torch.float64` runs the same network with FP64
sums (bf16 rounding at the same places): the yardstick for how far any FP32
summation order, torch's or the GPU's, lands from the exactly summed
network."""
import numpy as np
import torch
import torch.nn.functional as F

from raymotion_amd.denoise import LAYERS, TAPS, layout


def _r(x, bf16, dtype=torch.float32):
    return x.to(torch.bfloat16).to(dtype) if bf16 else x.to(dtype)


def torch_weights(params, bf16=True, dtype=torch.float32):
    """[(weight (cout, cin, 3, 3), bias (cout,))] from the flat parameter vector."""
    p = params if isinstance(params, torch.Tensor) else torch.as_tensor(np.asarray(params, np.float32))
    out = []
    for L in layout():
        w = p[L.w_off:L.w_off + L.cout * L.kp].reshape(L.cout, L.kp)[:, :9 * L.cin]
        w = w.reshape(L.cout, 3, 3, L.cin).permute(0, 3, 1, 2)  # K = (3 ky + kx) * cin + ci
        out.append((_r(w, bf16, dtype), p[L.b_off:L.b_off + L.cout].to(dtype)))
    return out


def forward(params, feat, noisy, bf16=True, dtype=torch.float32):
    """feat (N, H, W, 16), noisy (N, H, W, 3) -> denoised (N, H, W, 3)."""
    x = _r(torch.as_tensor(feat, dtype=torch.float32).permute(0, 3, 1, 2), bf16, dtype)
    ws = torch_weights(params, bf16, dtype)
    for i, (w, b) in enumerate(ws):
        x = F.conv2d(x, w, b, padding=1)
        if i < len(ws) - 1:
            x = _r(torch.relu(x), bf16, dtype)
    logits = x[:, :TAPS]
    wts = torch.softmax(logits, dim=1)  # (N, 25, H, W)
    rad = torch.as_tensor(noisy, dtype=dtype).permute(0, 3, 1, 2)
    pad = F.pad(rad, (2, 2, 2, 2), mode="replicate")
    n, _, h, w = rad.shape
    taps = torch.stack([pad[:, :, ky:ky + h, kx:kx + w] for ky in range(5) for kx in range(5)], dim=1)
    out = (wts.unsqueeze(2) * taps).sum(dim=1)
    return out.permute(0, 2, 3, 1)


def loss(out, reference):
    ref = torch.as_tensor(reference, dtype=torch.float32)
    return torch.mean((torch.log1p(torch.clamp(out, min=0)) - torch.log1p(torch.clamp(ref, min=0))) ** 2)


def gradient(params, feat, noisy, reference):
    """(loss, flat gradient in the parameter layout) with bf16-rounded forward."""
    p = torch.tensor(np.asarray(params, np.float32), requires_grad=True)
    out = forward(p, feat, noisy, bf16=True)
    val = loss(out, reference)
    val.backward()
    return float(val.detach()), p.grad.numpy()


__all__ = ["forward", "loss", "gradient", "torch_weights", "LAYERS"]
