"""Plain PyTorch FP32 reference of the spline-coupling flow (test infrastructure).

Architecture (identical to the device kernels): 3-D, 6 rational-quadratic spline couplings on
[0,1] (Durkan, Bekasov, Murray, Papamakarios, "Neural Spline Flows", NeurIPS 2019), 16 bins,
layer l transforms coordinate l % 3 conditioned on the other two through an MLP
2 -> 64 -> 64 -> 49 with ReLU.  Weights are rounded to bfloat16 and so are the hidden
activations that feed the 64-wide layers (the MFMA operands); accumulation is FP32.  The
normalizing direction y -> z gives log q(y) = log|det dz/dy|; sampling inverts it.
"""
import math

import torch

L, K, H, NOUT = 6, 16, 64, 49
NPL = 2 * H + H + H * H + H + NOUT * H + NOUT     # 7537
MINW = MINH = MIND = 1e-3
COND = [(1, 2), (0, 2), (0, 1)]
D_INIT = math.log(math.expm1(1.0 - MIND))         # softplus^-1(1 - MIND): identity knots


BF16 = True     # round weights and 64-wide activations to bfloat16 as the device does


def bf(x):
    return x.to(torch.bfloat16).to(x.dtype) if BF16 else x


def unpack(params, l):
    p = params[l * NPL:(l + 1) * NPL]
    o = 0
    W1 = p[o:o + 2 * H].view(H, 2); o += 2 * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].view(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + NOUT * H].view(NOUT, H); o += NOUT * H
    b3 = p[o:o + NOUT]
    return W1, b1, W2, b2, W3, b3


def mlp(c, params, l):
    W1, b1, W2, b2, W3, b3 = unpack(params, l)
    h1 = torch.relu(c @ bf(W1).T + b1)
    h2 = torch.relu(bf(h1) @ bf(W2).T + b2)
    return bf(h2) @ bf(W3).T + b3


def knots(o):
    w = MINW + (1 - K * MINW) * torch.softmax(o[:, :K], dim=1)
    h = MINH + (1 - K * MINH) * torch.softmax(o[:, K:2 * K], dim=1)
    d = MIND + torch.nn.functional.softplus(o[:, 2 * K:])
    cx = torch.cat([torch.zeros_like(w[:, :1]), torch.cumsum(w, 1)[:, :-1], torch.ones_like(w[:, :1])], 1)
    cy = torch.cat([torch.zeros_like(h[:, :1]), torch.cumsum(h, 1)[:, :-1], torch.ones_like(h[:, :1])], 1)
    return w, h, d, cx, cy


def spline_fwd(x, o):
    w, h, d, cx, cy = knots(o)
    k = torch.searchsorted(cx[:, 1:K].detach().contiguous(), x.detach()[:, None].contiguous(), right=True).squeeze(1)
    k = k.clamp(0, K - 1)[:, None]
    xk, wk, yk, hk = cx.gather(1, k), w.gather(1, k), cy.gather(1, k), h.gather(1, k)
    dk, dk1 = d.gather(1, k), d.gather(1, k + 1)
    xi = ((x[:, None] - xk) / wk)
    s = hk / wk
    t = xi * (1 - xi)
    den = s + (dk1 + dk - 2 * s) * t
    y = yk + hk * (s * xi * xi + dk * t) / den
    dydx = s * s * (dk1 * xi * xi + 2 * s * t + dk * (1 - xi) ** 2) / (den * den)
    return y.squeeze(1), torch.log(dydx).squeeze(1)


def spline_inv(y, o):
    w, h, d, cx, cy = knots(o)
    k = torch.searchsorted(cy[:, 1:K].contiguous(), y[:, None].contiguous(), right=True).squeeze(1).clamp(0, K - 1)[:, None]
    xk, wk, yk, hk = cx.gather(1, k), w.gather(1, k), cy.gather(1, k), h.gather(1, k)
    dk, dk1 = d.gather(1, k), d.gather(1, k + 1)
    s = hk / wk
    dy = y[:, None] - yk
    a = hk * (s - dk) + dy * (dk1 + dk - 2 * s)
    b = hk * dk - dy * (dk1 + dk - 2 * s)
    c = -s * dy
    disc = torch.clamp(b * b - 4 * a * c, min=0)
    xi = (2 * c) / (-b - torch.sqrt(disc))
    x = xk + xi * wk
    t = xi * (1 - xi)
    den = s + (dk1 + dk - 2 * s) * t
    dydx = s * s * (dk1 * xi * xi + 2 * s * t + dk * (1 - xi) ** 2) / (den * den)
    return x.squeeze(1), torch.log(dydx).squeeze(1)


def forward(y, params):
    """Normalizing direction: y (N,3) -> z, log|det dz/dy|."""
    x = [y[:, 0], y[:, 1], y[:, 2]]
    ld = torch.zeros_like(y[:, 0])
    for l in range(L):
        d = l % 3
        c = torch.stack([x[COND[d][0]], x[COND[d][1]]], 1)
        o = mlp(c, params, l)
        x[d], lj = spline_fwd(x[d], o)
        ld = ld + lj
    return torch.stack(x, 1), ld


def inverse(z, params):
    """Sampling direction: z -> y, log|det dy/dz|."""
    x = [z[:, 0], z[:, 1], z[:, 2]]
    ld = torch.zeros_like(z[:, 0])
    for l in reversed(range(L)):
        d = l % 3
        c = torch.stack([x[COND[d][0]], x[COND[d][1]]], 1)
        o = mlp(c, params, l)
        x[d], lj = spline_inv(x[d], o)
        ld = ld - lj
    return torch.stack(x, 1), ld


def identity_params(seed=0):
    """Hidden layers random (uniform +-1/sqrt(fan_in)), last layers zero, derivative biases at
    softplus^-1(1 - MIND): every spline is the identity."""
    g = torch.Generator().manual_seed(seed)
    ps = []
    for _ in range(L):
        W1 = (torch.rand(H, 2, generator=g) * 2 - 1) / math.sqrt(2)
        W2 = (torch.rand(H, H, generator=g) * 2 - 1) / math.sqrt(H)
        b3 = torch.cat([torch.zeros(2 * K), torch.full((K + 1,), D_INIT)])
        ps += [W1.flatten(), torch.zeros(H), W2.flatten(), torch.zeros(H), torch.zeros(NOUT * H), b3]
    return torch.cat(ps)
