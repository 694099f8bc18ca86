"""FP32 torch-CPU reference of the pressure-guess U-Net (TEST INFRASTRUCTURE).

The network has no counterpart in the reference solver; this module fixes its definition so the
HIP kernels (mobydiff_amd/csrc/mdx_unet.hip) can be checked against a plain PyTorch fp32 model
of the same ops.  Only the test suite imports it.

Architecture (3 levels, 16/32/64 channels, 3x3x3 convolutions, zero padding, ReLU):
  enc1: conv(16->16) conv(16->16)                [input channels 0,1 used, 2..15 zero]
  pool: 2x2x2 average
  enc2: conv(16->32) conv(32->32)
  pool
  bott: conv(32->64) conv(64->64)
  up:   nearest x2, concat [up(64), enc2(32)] -> conv(96->32) conv(32->32)
  up:   nearest x2, concat [up(32), enc1(16)] -> conv(48->16) conv(16->16)
  head: conv1x1(16->1), linear
Tensors are (N, C, D, H, W) = (batch, channel, z, y, x) here; the HIP code stores activations
channels-last, (z, y, x, c).
"""
import torch
import torch.nn.functional as F

LAYERS = [  # (name, cin, cout, ksize)
    ("e1a", 16, 16, 3), ("e1b", 16, 16, 3),
    ("e2a", 16, 32, 3), ("e2b", 32, 32, 3),
    ("b3a", 32, 64, 3), ("b3b", 64, 64, 3),
    ("d2a", 96, 32, 3), ("d2b", 32, 32, 3),
    ("d1a", 48, 16, 3), ("d1b", 16, 16, 3),
    ("head", 16, 1, 1),
]


def init_params(seed=0, dtype=torch.float32):
    """He-normal weights (fan-in), zero bias; weight layout (tap, cin, cout), tap = (dz*3+dy)*3+dx."""
    g = torch.Generator().manual_seed(seed)
    params = {}
    for name, cin, cout, k in LAYERS:
        taps = k ** 3
        std = (2.0 / (taps * cin)) ** 0.5
        params[name + ".w"] = (torch.randn(taps, cin, cout, generator=g) * std).to(dtype)
        params[name + ".b"] = torch.zeros(cout, dtype=dtype)
    return params


def _conv(x, w, b, k):
    """x (N,C,D,H,W); w (taps, cin, cout) -> torch weight (cout, cin, kz, ky, kx)."""
    cin, cout = w.shape[1], w.shape[2]
    wt = w.permute(2, 1, 0).reshape(cout, cin, k, k, k)
    return F.conv3d(x, wt, b, padding=k // 2)


def forward(params, x, keep=False):
    """x: (N, 16, D, H, W) fp32.  Returns (N, 1, D, H, W); with keep, also the activations."""
    acts = {}

    def cv(name, t, relu=True):
        k = 3 if name != "head" else 1
        y = _conv(t, params[name + ".w"], params[name + ".b"], k)
        y = F.relu(y) if relu else y
        acts[name] = y
        return y

    e1 = cv("e1b", cv("e1a", x))
    e2 = cv("e2b", cv("e2a", F.avg_pool3d(e1, 2)))
    b3 = cv("b3b", cv("b3a", F.avg_pool3d(e2, 2)))
    u2 = F.interpolate(b3, scale_factor=2, mode="nearest")
    d2 = cv("d2b", cv("d2a", torch.cat([u2, e2], dim=1)))
    u1 = F.interpolate(d2, scale_factor=2, mode="nearest")
    d1 = cv("d1b", cv("d1a", torch.cat([u1, e1], dim=1)))
    out = cv("head", d1, relu=False)
    return (out, acts) if keep else out


def _bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def forward_bf16(params, x):
    """Same network with the HIP kernels' storage precision: 3x3x3 weights, the input and every
    stored activation rounded to bfloat16 (round to nearest even), accumulation in fp32, the 1x1
    head in fp32 on bf16 features."""
    def cv(name, t):
        w = _bf(params[name + ".w"])
        return _bf(F.relu(_conv(t, w, params[name + ".b"], 3)))

    x = _bf(x)
    e1 = cv("e1b", cv("e1a", x))
    e2 = cv("e2b", cv("e2a", _bf(F.avg_pool3d(e1, 2))))
    b3 = cv("b3b", cv("b3a", _bf(F.avg_pool3d(e2, 2))))
    d2 = cv("d2b", cv("d2a", torch.cat([F.interpolate(b3, scale_factor=2, mode="nearest"), e2], dim=1)))
    d1 = cv("d1b", cv("d1a", torch.cat([F.interpolate(d2, scale_factor=2, mode="nearest"), e1], dim=1)))
    return _conv(d1, params["head.w"], params["head.b"], 1)


def flat_to_params(flat):
    """The HIP library's flat FP32 layout (per layer W(tap,cin,cout) then b) -> dict of tensors."""
    params, off = {}, 0
    for name, cin, cout, k in LAYERS:
        n = k ** 3 * cin * cout
        params[name + ".w"] = torch.as_tensor(flat[off:off + n]).reshape(k ** 3, cin, cout).clone()
        off += n
        params[name + ".b"] = torch.as_tensor(flat[off:off + cout]).clone()
        off += cout
    return params


def params_to_flat(params):
    import numpy as np
    return np.concatenate([np.concatenate([params[n + ".w"].reshape(-1).numpy(), params[n + ".b"].numpy()])
                           for n, _, _, _ in LAYERS]).astype(np.float32)


def loss_and_grads(params, x, target):
    """MSE loss and parameter gradients (autograd, fp32)."""
    p = {k: v.detach().clone().requires_grad_(True) for k, v in params.items()}
    out = forward(p, x)
    loss = torch.mean((out - target) ** 2)
    loss.backward()
    return loss.detach(), {k: v.grad.detach() for k, v in p.items()}


def adam_step(params, grads, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """Adam (Kingma & Ba 2015, Algorithm 1) with bias correction; returns new (params, m, v)."""
    out_p, out_m, out_v = {}, {}, {}
    for k in params:
        mk = b1 * m[k] + (1 - b1) * grads[k]
        vk = b2 * v[k] + (1 - b2) * grads[k] ** 2
        mh = mk / (1 - b1 ** t)
        vh = vk / (1 - b2 ** t)
        out_p[k] = params[k] - lr * mh / (torch.sqrt(vh) + eps)
        out_m[k], out_v[k] = mk, vk
    return out_p, out_m, out_v


# ----------------------------------------------------------------------------- strided coupling
# A network with stride s runs on the solver grid coarsened by s in every direction.  Its input is
# the restriction of the solver's fields (mean over each voxel's s^3 cells), its output reaches the
# solver grid by trilinear interpolation between voxel centres; a cell beyond the outermost centre
# takes that voxel's value.  Here both operators are written as separable matrices (numpy, FP64).
def restrict_mean(fine, s):
    """(FZ, FY, FX) -> (FZ/s, FY/s, FX/s): the mean of each s x s x s block"""
    import numpy as np
    fz, fy, fx = fine.shape
    return np.asarray(fine, dtype=np.float64).reshape(fz // s, s, fy // s, s, fx // s, s).mean(axis=(1, 3, 5))


def prolong_matrix(nf, s):
    """(nf x nf/s) interpolation matrix of one axis: fine cell i sits at coarse coordinate
    (i + 1/2)/s - 1/2; linear weights between the two nearest voxel centres, clamped at the ends"""
    import numpy as np
    nc = nf // s
    P = np.zeros((nf, nc))
    for i in range(nf):
        xc = (i + 0.5) / s - 0.5
        i0 = int(np.floor(xc))
        t = xc - i0
        if i0 < 0:
            P[i, 0] = 1.0
        elif i0 + 1 > nc - 1:
            P[i, nc - 1] = 1.0
        else:
            P[i, i0] += 1.0 - t
            P[i, i0 + 1] += t
    return P


def prolong_trilinear(coarse, s):
    """(CZ, CY, CX) -> (s CZ, s CY, s CX) by the tensor product of the three axes' matrices"""
    import numpy as np
    cz, cy, cx = coarse.shape
    Pz, Py, Px = prolong_matrix(cz * s, s), prolong_matrix(cy * s, s), prolong_matrix(cx * s, s)
    return np.einsum("ia,jb,kc,abc->ijk", Pz, Py, Px, np.asarray(coarse, dtype=np.float64), optimize=True)
