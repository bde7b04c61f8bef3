"""Learned pruning of the state space (the pruned execution of a plan).

A small convolutional net reads each instance's log unaries, max-pooled 4x
in x and y ([P * R] channels at H/4 x W/4), and predicts how the exact
marginal of every part spreads over its states: a probability per (part,
rotation, 4 x 4 cell), normalised over each part's cells.  Per part, the
64 x 64 tiles of its (rotation, y, x) planes are kept in decreasing order of
predicted mass until the predicted mass of the tiles left out is at most
1e-4 (1e-3 over the ten parts of an instance: the error budget of the BF16
method); an instance whose predicted dropped mass exceeds 1e-3 runs the exact
sweep instead (the fallback).  A fixed keep threshold of 1e-6 per state does
not fit these marginals: at 300 x 300 x 24 states the average state holds
4.6e-7, and exact marginals lose 3-12 % of their mass to that rule.  The keep
flags and fallback flags feed ps_plan_execute_pruned.

Net: three 3 x 3 convolutions (64, 32, 32 channels, ReLU) and a 1 x 1 output
layer (P * R channels), BF16 weights and activations with FP32 accumulation
(torch's convolutions on the device's matrix cores), trained on exact
marginals (prune_train.py).  No counterpart in the reference: its sweep is
exact over the whole state space (objectdetect_findrot.cpp:582-667).
"""
import os

import numpy as np

POOL = 4           # unary downsampling in x and y
TILE = 64          # keep-flag tile (the column pass's 64 x 64 output tile)
CELLS = TILE // POOL
PART_BUDGET = 1e-4  # predicted mass a part may lose to its dropped tiles
FALLBACK_MASS = 1e-3
INPUT_SCALE = 0.1   # pooled log unaries (0 .. -30) scaled for the first layer
FLOOR = -30.0      # pooled log unaries below max - 30 are clamped
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "prune_net.pt")


def _torch():
    import torch
    return torch


def make_net(P=10, R=24):
    torch = _torch()
    nn = torch.nn

    class PruneNet(nn.Module):
        def __init__(self):
            super().__init__()
            c = P * R
            self.P, self.R = P, R
            self.c1 = nn.Conv2d(c, 64, 3, padding=1)
            self.c2 = nn.Conv2d(64, 32, 3, padding=1)
            self.c3 = nn.Conv2d(32, 32, 3, padding=1)
            self.head = nn.Conv2d(32, c, 1)

        def forward(self, x):
            """x [B, P*R, h, w] -> log predicted cell mass [B, P, R, h, w]
            (log-softmax over each part's R * h * w cells)."""
            f = torch.relu(self.c1(x))
            f = torch.relu(self.c2(f))
            f = torch.relu(self.c3(f))
            z = self.head(f).float()
            B, _, h, w = z.shape
            z = z.view(B, self.P, self.R * h * w)
            return torch.log_softmax(z, dim=2).view(B, self.P, self.R, h, w)

    return PruneNet()


def load_net(device, path=WEIGHTS, P=10, R=24):
    """The trained net in BF16 on `device` (weights loaded without executing
    anything from the file)."""
    torch = _torch()
    net = make_net(P, R)
    state = torch.load(path, map_location="cpu", weights_only=True)
    net.load_state_dict({k: v.float() for k, v in state.items()})
    return net.to(device=device, dtype=torch.bfloat16).eval()


def features(unary_block, s):
    """Pooled, per-part max-shifted log unaries of scale s of a run of
    equal-size instances ([n][P][S][R][H][W] CUDA tensor) ->
    [n, P*R, ceil(H/4), ceil(W/4)] float."""
    torch = _torch()
    u = unary_block[:, :, s]  # [n, P, R, H, W]
    n, P, R, H, W = u.shape
    m = u.reshape(n, P, -1).amax(dim=2).view(n, P, 1, 1, 1)
    x = torch.clamp(u - m, min=FLOOR)
    ph, pw = (-H) % POOL, (-W) % POOL
    if ph or pw:
        x = torch.nn.functional.pad(x, (0, pw, 0, ph), value=FLOOR)
    return torch.nn.functional.max_pool2d(x.reshape(n, P * R, H + ph, W + pw), POOL) * INPUT_SCALE


def cell_mass(marginal_inst, s):
    """Exact cell masses (training labels): the normalised marginal of every
    part of scale s summed over 4 x 4 cells -> [P, R, ceil(H/4), ceil(W/4)]."""
    torch = _torch()
    b = marginal_inst[:, s].double()  # [P, R, H, W]
    P, R, H, W = b.shape
    m = b.reshape(P, -1).amax(dim=1).view(P, 1, 1, 1)
    e = torch.exp(b - m)
    e = e / e.reshape(P, -1).sum(dim=1).view(P, 1, 1, 1)
    ph, pw = (-H) % POOL, (-W) % POOL
    if ph or pw:
        e = torch.nn.functional.pad(e, (0, pw, 0, ph))
    e = e.reshape(P, R, (H + ph) // POOL, POOL, (W + pw) // POOL, POOL).sum(dim=(3, 5))
    return e.float()


def keep_from_mass(mass, H, W, part_budget=PART_BUDGET):
    """Keep flags [..., P, R, ceil(H/64), ceil(W/64)] (uint8) from cell
    masses [..., P, R, h, w], and the mass of the dropped tiles per part
    [..., P]: per part, tiles in decreasing order of mass until the mass left
    out is at most part_budget (at least one tile is kept)."""
    torch = _torch()
    lead = mass.shape[:-4]
    P, R, h, w = mass.shape[-4:]
    ty, tx = (H + TILE - 1) // TILE, (W + TILE - 1) // TILE
    mp = torch.nn.functional.pad(mass, (0, tx * CELLS - w, 0, ty * CELLS - h))
    tile_sum = mp.reshape(*lead, P, R, ty, CELLS, tx, CELLS).sum(dim=(-3, -1))
    flat = tile_sum.reshape(-1, R * ty * tx)
    srt, order = flat.sort(dim=1, descending=True)
    before = torch.cumsum(srt, dim=1) - srt  # mass of the tiles ahead of each in the order
    total = flat.sum(dim=1, keepdim=True)
    keep_sorted = before < total - part_budget
    keep_sorted[:, 0] = True
    keep = torch.zeros_like(keep_sorted).scatter_(1, order, keep_sorted)
    dropped = (flat * (~keep)).sum(dim=1).reshape(*lead, P)
    return keep.reshape(*lead, P, R, ty, tx).to(torch.uint8), dropped


class Pruner:
    """Keep flags and fallback flags of a plan's batch from its unaries.
    Instances are processed in runs of equal size, one net call per run and
    scale."""

    def __init__(self, plan, device, net=None, part_budget=PART_BUDGET, fallback_mass=FALLBACK_MASS):
        torch = _torch()
        self.plan = plan
        # let the convolution library time its algorithms for these shapes once
        torch.backends.cudnn.benchmark = True
        self.net = net if net is not None else load_net(device, P=plan.P, R=plan.R)
        self.part_budget = part_budget
        self.fallback_mass = fallback_mass
        self.offsets, total = plan.keep_layout()
        self.keep = torch.ones(max(1, total), dtype=torch.uint8, device=device)
        self.fallback = torch.zeros(plan.n_instances, dtype=torch.int32, device=device)
        self.predicted_dropped = torch.zeros(plan.n_instances, dtype=torch.float32, device=device)
        self.runs = []  # (first instance, count, H, W): consecutive instances of one size
        for i, hw in enumerate(plan.sizes):
            if self.runs and self.runs[-1][2:] == hw and self.runs[-1][0] + self.runs[-1][1] == i:
                self.runs[-1] = (self.runs[-1][0], self.runs[-1][1] + 1) + hw
            else:
                self.runs.append((i, 1) + hw)

    def update(self, unary, masses=None):
        """Predict (or, with `masses` [inst][S] of [P, R, h, w], take) the cell
        masses of every instance and scale and fill the keep and fallback
        flags on the device."""
        torch = _torch()
        plan = self.plan
        P, S, R = plan.P, plan.S, plan.R
        with torch.no_grad():
            for i0, n, H, W in self.runs:
                ty, tx = (H + TILE - 1) // TILE, (W + TILE - 1) // TILE
                per = P * S * R * ty * tx
                # equal-size consecutive instances: their flags are contiguous
                block = self.keep[self.offsets[i0]:self.offsets[i0] + n * per].view(n, P, S, R, ty, tx)
                drop = torch.zeros(n, dtype=torch.float32, device=unary.device)
                for s in range(S):
                    if masses is None:
                        ub = unary[plan.unary_offsets[i0]:plan.unary_offsets[i0] + n * P * S * R * H * W]
                        x = features(ub.view(n, P, S, R, H, W), s)
                        m = torch.exp(self.net(x.to(torch.bfloat16)).float())
                    else:
                        m = torch.stack([masses[i][s] for i in range(i0, i0 + n)])
                    k, d = keep_from_mass(m, H, W, self.part_budget)
                    block[:, :, s].copy_(k)
                    drop += d.sum(dim=1)
                self.predicted_dropped[i0:i0 + n] = drop
            self.fallback.copy_((self.predicted_dropped > self.fallback_mass).to(torch.int32))

    def kept_fraction(self):
        return float(self.keep.float().mean())
