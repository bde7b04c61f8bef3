"""Train the pruning net (prune.py) on exact marginals, on the device.

    python -m partapp_amd.prune_train --instances 1024 --epochs 4 --out partapp_amd/prune_net.pt

Training data: uneven instances like the batching workload (H, W uniform in
{160, 168, ..., 480}, 24 rotations x 3 scales, 10-part person tree, a
per-image spatial model), with seeds disjoint from the benchmark's; labels:
the cell masses of the exact marginals of the BF16 sweep (its error against
FP64, ~1.5e-4 relative L1, is far below the masses the keep rule resolves).
Loss: the cross entropy of the predicted cell distribution of every part
against the exact one.  The weights are stored in BF16.
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def say(*a):
    print(*a, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=1024)
    ap.add_argument("--group", type=int, default=8, help="instances of one size per plan")
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=100000)
    ap.add_argument("--out", default=os.path.join(HERE, "prune_net.pt"))
    args = ap.parse_args()

    import numpy as np
    import torch

    from partapp_amd.api import Plan, config_exp_param, person_model, synth_joints, synth_unary_dev
    from partapp_amd.prune import cell_mass, features, keep_from_mass, make_net

    dev = torch.device("cuda:0")
    torch.manual_seed(args.seed)
    ep = config_exp_param(2)
    S, R, P = ep.num_scale_steps, ep.num_rotation_steps, 10
    rng = np.random.default_rng(args.seed)
    data = []  # (features [n, P*R, h, w] bf16, labels [n, P, R, h, w] float32, H, W)
    t0 = time.time()
    n_groups = (args.instances + args.group - 1) // args.group
    for g in range(n_groups):
        H, W = 8 * int(rng.integers(20, 61)), 8 * int(rng.integers(20, 61))
        seeds = [args.seed + args.group * g + k for k in range(args.group)]
        joints = [synth_joints(s) for s in seeds]
        model = person_model(ep, joints[0])
        plan = Plan(model, [(H, W)] * len(seeds), joints)
        plan.set_method("bf16mix")
        un = torch.empty(plan.unary_numel, dtype=torch.float32, device=dev)
        numel = P * S * R * H * W
        for i, s in enumerate(seeds):
            synth_unary_dev(s, P, S, R, H, W, un[i * numel:(i + 1) * numel])
        marg = torch.empty_like(un)
        plan.execute(un, marg)
        torch.cuda.synchronize()
        ub = un.view(len(seeds), P, S, R, H, W)
        for s in range(S):
            x = features(ub, s).to(torch.bfloat16)
            y = torch.stack([cell_mass(plan.instance_view(marg, i), s) for i in range(len(seeds))])
            data.append((x, y, H, W))
        plan.close()
        del un, marg
        if g % 16 == 0:
            say("data: group %d/%d (%dx%d), %.1f s" % (g + 1, n_groups, H, W, time.time() - t0))
    # the last eighth of the samples is held out
    n_hold = max(1, len(data) // 8)
    train, hold = data[:-n_hold], data[-n_hold:]
    say("samples: %d train, %d held out (groups of %d instances)" % (len(train), len(hold), args.group))

    net = make_net(P, R).to(dev)
    opt = torch.optim.Adam(net.parameters(), lr=args.lr)

    def loss_of(x, y):
        logp = net(x.float())
        return -(y * logp).sum(dim=(2, 3, 4)).mean()

    step = 0
    for epoch in range(args.epochs):
        order = np.random.default_rng(args.seed + epoch).permutation(len(train))
        tot = 0.0
        for j in order:
            x, y, _, _ = train[j]
            loss = loss_of(x, y)
            opt.zero_grad()
            loss.backward()
            opt.step()
            tot += float(loss.detach())
            step += 1
            if step % 100 == 0:
                say("epoch %d step %d: mean loss %.4f, %.1f s" % (epoch, step, tot / 100, time.time() - t0))
                tot = 0.0
        if epoch >= args.epochs // 2:
            for group in opt.param_groups:
                group["lr"] *= 0.7

    # held-out statistics of the keep rule with the BF16 net, and with the
    # exact cell masses (the best any predictor could do at this granularity)
    netb = net.to(torch.bfloat16).eval()
    stats = {"net": ([], [], []), "exact": ([], [], [])}
    with torch.no_grad():
        for x, y, H, W in hold:
            for name, m in (("net", torch.exp(netb(x).float())), ("exact", y)):
                k, d = keep_from_mass(m, H, W)
                ty, tx = k.shape[-2:]
                yt = torch.nn.functional.pad(y, (0, tx * 16 - y.shape[-1], 0, ty * 16 - y.shape[-2]))
                ts = yt.reshape(*yt.shape[:3], ty, 16, tx, 16).sum(dim=(4, 6))
                stats[name][0].append(float(k.float().mean()))
                stats[name][1].extend(d.sum(dim=1).tolist())
                stats[name][2].extend((ts * (1 - k.float())).sum(dim=(2, 3, 4)).sum(dim=1).tolist())
    for name, (kept, pred, true) in stats.items():
        say("held out, %s masses: kept tiles %.3f, predicted dropped mass median %.3g (max %.3g), exact dropped "
            "mass median %.3g (max %.3g), samples above 1e-3: %.3f" % (
                name, float(np.mean(kept)), float(np.median(pred)), max(pred), float(np.median(true)), max(true),
                float(np.mean(np.array(true) > 1e-3))))
    torch.save({k: v.detach().to("cpu", torch.bfloat16) for k, v in netb.state_dict().items()}, args.out)
    say("saved %s (%.1f KB)" % (args.out, os.path.getsize(args.out) / 1024))


if __name__ == "__main__":
    main()
