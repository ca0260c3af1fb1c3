"""Train the learned preconditioner and write its weight blob.

    python -m ariel_amd.learned.train --out DIR [--minutes 8]

1. Harvest (residual, solution) pairs. Run the dam-break, dam-break + sphere
   and pool + falling cube scenes through the library with MIC(0)-PCG; after
   each harvested step take its cell classes and level set (the matrix), and
   right-hand sides b for it: the step's own divergence and FLUID-masked noise
   smoothed by 0..32 box-filter passes. Solve each with the library's
   MIC(0)-PCG to 1e-6 (``FlipSim.solve_rhs``): x*. Every b gives the pair
   (b, x*) and, from a Jacobi-PCG run on the same system, the pairs
   (r_k, x* - x_k) at several iterations k: A (x* - x_k) = r_k, and r_k has the
   roughness and surface concentration of the residuals a preconditioner
   actually meets inside PCG. The step's own solve adds its warm-started
   system: r0 = b - A p_prev with the error P - p_prev (the library's solve
   from the previous pressure p_prev), and Jacobi-PCG pairs on it.
2. Fit the network with the library's own training kernels (``Trainer``:
   BF16 MFMA forward, FP32 backward, Adam on FP32 master weights) by
   minimising the A-angle loss between its output and the solution, A built
   from the cell classes and level set (``net.laplacian``); cosine learning
   rate. Between rounds, add the residuals of flexible PCG driven by the
   current network (the distribution the network sees when it is used).
3. Check: held-out loss, flexible PCG iterations in torch, then through the
   library with the new weights against MIC(0)-PCG; write ``weights.bin`` and
   ``report.json`` into DIR.

Needs a GPU (the library and torch on cuda:0).
"""
import argparse
import json
import math
import os
import sys
import time

import torch
import torch.nn.functional as F

from .. import FlipSim, _abi, scenes
from . import layers_to_flat
from . import net as lnet
from .trainer import Trainer

SCENES = {
    "dam": scenes.dam_break,
    "sphere": scenes.dam_break_sphere,
    "pool": scenes.pool_and_cube,
}
KSET = (1, 2, 3, 5, 8, 12, 18, 27, 40)
T0 = time.time()


def log(msg):
    print("[%7.1fs] %s" % (time.time() - T0, msg), flush=True)


def dot(a, b):
    return (a.double() * b.double()).sum()


class Pairs:
    """Pairs grouped by grid size; geometry (diag, fluid, w) shared per snapshot."""

    def __init__(self):
        self.geo = {}     # n -> list of (diag, fluid, w)
        self.items = {}   # n -> list of (geo index, r, e) with A e = r
        self.rhs = {}     # n -> list of (geo index, b, x*) for later rounds

    def add_geo(self, n, diag, fluid):
        g = self.geo.setdefault(n, [])
        g.append((diag, fluid, lnet.diag_weight(diag, fluid)))
        return len(g) - 1

    def add(self, n, gi, r, e):
        fl = self.geo[n][gi][1]
        c = torch.sqrt(dot(r, r) / fl.sum().clamp(min=1)).float()
        if not torch.isfinite(c) or float(c) <= 0.0 or not bool(torch.isfinite(e).all()):
            return
        self.items.setdefault(n, []).append((gi, r / c, e / c))

    def add_residual(self, n, gi, rk, e, b):
        """(r_k, x* - x_k) while |r_k| > 1e-4 |b|: x* solves to 1e-6, so the
        pair's own residual |A e - r_k| stays within 1% of |r_k|."""
        if float(dot(rk, rk)) > 1e-8 * float(dot(b, b)):
            self.add(n, gi, rk, e)

    def count(self):
        return sum(len(v) for v in self.items.values())


def pcg(b, diag, fluid, precond, iters, record=(), tol=0.0):
    """(Flexible) PCG from x = 0 in torch; Polak-Ribiere beta. Yields
    (k, r_k, x_k) for k in record; returns (through StopIteration) the
    iteration at which |r| <= tol |b|, iters if none, or -k on a breakdown."""
    x = torch.zeros_like(b)
    r = b.clone()
    z = precond(r)
    p = z.clone()
    rho = dot(r, z)
    b2 = dot(b, b)
    for k in range(1, iters + 1):
        if not (rho > 0):
            return -k
        q = lnet.apply_laplacian(p, diag, fluid)
        sig = dot(p, q)
        if not (sig > 0):
            return -k
        a = (rho / sig).float()
        x = x + a * p
        rn = r - a * q
        if k in record:
            yield k, rn, x
        r2 = dot(rn, rn)
        if tol > 0 and r2 <= tol * tol * b2:
            return k
        zn = precond(rn)
        zr = dot(zn, rn)
        beta = ((zr - dot(zn, r)) / rho).float()
        rho = zr
        p = zn + beta * p
        r = rn
    return iters


def run_pcg(*args, **kw):
    g = pcg(*args, **kw)
    try:
        while True:
            next(g)
    except StopIteration as e:
        return e.value


def jacobi(diag, fluid):
    inv = torch.where(fluid > 0, 1.0 / torch.where(fluid > 0, diag, torch.ones_like(diag)),
                      torch.zeros_like(diag))
    return lambda r: r * inv


def net_precond(trainer, diag, fluid):
    """z = network(r) through the library's kernels (the trainer's weights)."""
    def apply(r):
        z = torch.empty_like(r)
        return trainer.forward(r.contiguous(), diag, fluid, z)
    return apply


def rhs_set(b, fluid, count, gen):
    """count right-hand sides: b itself, then masked noise with 0..32 smoothing passes."""
    n = b.shape[-1]
    out = [b]
    passes = [0, 1, 2, 4, 8, 16, 32]
    for i in range(count - 1):
        x = torch.randn((1, 1, n, n, n), device=b.device, generator=gen) * fluid
        for _ in range(passes[i % len(passes)]):
            x = F.avg_pool3d(x, 3, stride=1, padding=1) * fluid
        out.append(x)
    return out


def harvest(pairs, spec, rhs_per_step, gen, dev):
    """spec: list of (scene name, n, steps, every)."""
    for name, n, steps, every in spec:
        sim = FlipSim(SCENES[name](n))
        sim.Init()
        t = time.time()
        for st in range(steps):
            harvest_step = (st + 1) % every == 0
            # the warm start of this step's solve: the previous step's pressure
            P_prev = sim.get_field(_abi.FIELD_P) if harvest_step else None
            sim.Step(1)
            if not harvest_step:
                continue
            flags = sim.get_flags()
            if not (flags == lnet.FLUID).any():
                continue
            L = sim.get_field(_abi.FIELD_L)
            D = sim.get_field(_abi.FIELD_D)
            P = sim.get_field(_abi.FIELD_P)
            ft = torch.from_numpy(flags.reshape(n, n, n)).to(dev)
            Lt = torch.from_numpy(L.reshape(n, n, n)).to(dev)
            diag, fluid = lnet.laplacian(ft, Lt)
            diag = diag.view(1, 1, n, n, n)
            fluid = fluid.view(1, 1, n, n, n)
            gi = pairs.add_geo(n, diag, fluid)
            b0 = torch.from_numpy(D).to(dev).view(1, 1, n, n, n) * fluid
            # the step's own warm-started system: r0 = b - A p_prev, error P - p_prev
            # (P solves it to 1e-6 of |r0|)
            pp = torch.from_numpy(P_prev).to(dev).view(1, 1, n, n, n) * fluid
            r0 = b0 - lnet.apply_laplacian(pp, diag, fluid)
            e0 = torch.from_numpy(P).to(dev).view(1, 1, n, n, n) * fluid - pp
            if float(dot(r0, r0)) > 0.0:
                pairs.add(n, gi, r0, e0)
                pairs.rhs.setdefault(n, []).append((gi, r0, e0))
                for k, rk, xk in pcg(r0, diag, fluid, jacobi(diag, fluid), max(KSET), KSET):
                    pairs.add_residual(n, gi, rk, e0 - xk, r0)
            for j, b in enumerate(rhs_set(b0, fluid, rhs_per_step, gen)):
                x, sst = sim.solve_rhs(b.reshape(-1).cpu().numpy())
                if sst.solve_status != 0:
                    continue
                xs = torch.from_numpy(x).to(dev).view(1, 1, n, n, n) * fluid
                pairs.add(n, gi, b, xs)
                if j < 2:
                    pairs.rhs.setdefault(n, []).append((gi, b, xs))
                for k, rk, xk in pcg(b, diag, fluid, jacobi(diag, fluid), max(KSET), KSET):
                    pairs.add_residual(n, gi, rk, xs - xk, b)
            sim.set_field(_abi.FIELD_P, P)
            sim.set_field(_abi.FIELD_D, D)
            if (st + 1) % (every * 5) == 0:
                log("harvest %s n=%d step %d/%d pairs %d (%.1fs)" % (name, n, st + 1, steps,
                                                                   pairs.count(), time.time() - t))
        sim.close()
        log("harvest %s n=%d done: %d pairs total" % (name, n, pairs.count()))


def collect_fpcg(pairs, rhs, model, limit, gen):
    """Residuals of flexible PCG driven by the current network, with their exact
    errors x* - x_k, for up to `limit` of the given right-hand sides."""
    added = 0
    for n, lst in rhs.items():
        if not lst:
            continue
        perm = torch.randperm(len(lst), generator=gen).tolist()[:max(1, limit // len(rhs))]
        for i in perm:
            gi, b, xs = lst[i]
            diag, fluid, _w = pairs.geo[n][gi]
            before = pairs.count()
            for k, rk, xk in pcg(b, diag, fluid, net_precond(model, diag, fluid), max(KSET), KSET):
                pairs.add_residual(n, gi, rk, xs - xk, b)
            added += pairs.count() - before
    return added


def check_operator(pairs, n):
    """The torch operator reproduces the library's A x* = b on a harvested pair."""
    gi, b, xs = pairs.rhs[n][0]
    diag, fluid, _w = pairs.geo[n][gi]
    ax = lnet.apply_laplacian(xs, diag, fluid)
    err = float((ax - b).norm() / b.norm())
    log("operator check n=%d: |A x* - b|/|b| = %.2e" % (n, err))
    return err


def batches(pairs, held, bs, gen):
    """Shuffled batches of one grid size each, training pairs only."""
    order = []
    for n, items in pairs.items.items():
        idx = [i for i in range(len(items)) if (n, i) not in held]
        perm = torch.randperm(len(idx), generator=gen).tolist()
        b = bs.get(n, 1)
        for k in range(0, len(idx) - b + 1, b):
            order.append((n, [idx[j] for j in perm[k:k + b]]))
    perm = torch.randperm(len(order), generator=gen).tolist()
    return [order[i] for i in perm]


def stack(pairs, n, ids):
    items = pairs.items[n]
    geo = [pairs.geo[n][items[i][0]] for i in ids]
    r = torch.cat([items[i][1] for i in ids])
    e = torch.cat([items[i][2] for i in ids])
    diag = torch.cat([g[0] for g in geo])
    fluid = torch.cat([g[1] for g in geo])
    w = torch.cat([g[2] for g in geo])
    return r, e, diag, fluid, w


def evaluate(model, pairs, held, bs):
    """Mean loss over the held-out pairs (no update)."""
    tot, cnt = 0.0, 0
    by = {}
    for n, i in sorted(held):
        by.setdefault(n, []).append(i)
    for n, ids in by.items():
        b = bs.get(n, 1)
        for k in range(0, len(ids) - b + 1, b):
            r, e, diag, fluid, _w = stack(pairs, n, ids[k:k + b])
            tot += model.step(r, e, diag, fluid, update=False) * b
            cnt += b
    return tot / max(cnt, 1)


def torch_fpcg_check(model, pairs, held_geo):
    """Iterations to 1e-6 in torch: network FPCG vs Jacobi PCG, held-out snapshots."""
    out = []
    for n, gi in held_geo:
        lst = [t for t in pairs.rhs.get(n, []) if t[0] == gi]
        if not lst:
            continue
        _gi, b, _xs = lst[0]
        diag, fluid, _w = pairs.geo[n][gi]
        it_net = run_pcg(b, diag, fluid, net_precond(model, diag, fluid), 600, tol=1e-6)
        it_jac = run_pcg(b, diag, fluid, jacobi(diag, fluid), 3000, tol=1e-6)
        out.append({"n": n, "geo": gi, "net_fpcg": it_net, "jacobi_pcg": it_jac})
        log("torch check n=%d geo %d: net FPCG %d it, Jacobi PCG %d it" % (n, gi, it_net, it_jac))
    return out


def library_check(blob, spec):
    """Flexible PCG with the new weights vs MIC(0)-PCG, on the divergence of a step."""
    out = []
    for name, n, steps in spec:
        sim = FlipSim(SCENES[name](n, precond=_abi.PRECOND_LEARNED))
        sim.set_learned_weights(blob)
        sim.Init()
        sim.Step(steps)
        mic = FlipSim(SCENES[name](n))
        mic.Init()
        mic.set_flags(sim.get_flags())
        mic.set_field(_abi.FIELD_L, sim.get_field(_abi.FIELD_L))
        D = sim.get_field(_abi.FIELD_D)
        xl, sl = sim.solve_rhs(D)
        xm, sm = mic.solve_rhs(D)
        res = {"scene": name, "n": n, "step": steps,
               "learned_iters": sl.iterations, "learned_ms": sl.solve_ms,
               "fallback": sl.precond_fallback, "fallback_iter": sl.fallback_iter,
               "mic_iters": sm.iterations, "mic_ms": sm.solve_ms,
               "learned_true_rel": sl.true_residual_norm / max(sl.rhs_norm, 1e-30),
               "mic_true_rel": sm.true_residual_norm / max(sm.rhs_norm, 1e-30)}
        sim.close()
        mic.close()
        # inside a simulation: warm-started solves of the next steps
        its = {}
        for tag, kw in (("learned", {"precond": _abi.PRECOND_LEARNED}), ("mic", {})):
            s2 = FlipSim(SCENES[name](n, **kw))
            if kw:
                s2.set_learned_weights(blob)
            s2.Init()
            s2.Step(steps)
            acc = []
            for _ in range(3):
                s2.Step(1)
                acc.append(s2.stats().iterations)
            its[tag] = float(sum(acc) / len(acc))
            s2.close()
        res["steps_learned_iters"] = its["learned"]
        res["steps_mic_iters"] = its["mic"]
        log("library check %s" % json.dumps(res))
        out.append(res)
    return out


def parse_spec(s):
    out = []
    for part in s.split(","):
        name, n, steps, every = part.split(":")
        out.append((name, int(n), int(steps), int(every)))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--spec", default="dam:32:120:4,dam:64:120:6,sphere:64:120:6,pool:64:120:6,"
                                      "sphere:128:40:8,pool:128:40:8")
    ap.add_argument("--rhs", type=int, default=8, help="right-hand sides per harvested step")
    ap.add_argument("--minutes", type=float, default=8.0, help="training time budget")
    ap.add_argument("--rounds", type=int, default=3, help="training rounds (FPCG residuals between)")
    ap.add_argument("--collect", type=int, default=300, help="right-hand sides per FPCG collection")
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None, help="start from this weight blob")
    ap.add_argument("--check", default="sphere:128:20,pool:128:20")
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    dev = torch.device("cuda:0")
    torch.manual_seed(a.seed)
    gen = torch.Generator(device=dev).manual_seed(a.seed)
    cgen = torch.Generator().manual_seed(a.seed)

    pairs = Pairs()
    harvest(pairs, parse_spec(a.spec), a.rhs, gen, dev)
    op_err = {n: check_operator(pairs, n) for n in pairs.rhs}

    # hold out the last two snapshots of each size (all their pairs)
    held_geo = [(n, len(g) - 1) for n, g in pairs.geo.items()] + \
               [(n, len(g) - 2) for n, g in pairs.geo.items() if len(g) > 1]
    hg = set(held_geo)

    def held_set():
        return {(n, i) for n, items in pairs.items.items() for i, t in enumerate(items)
                if (n, t[0]) in hg}

    train_rhs = {n: [t for t in lst if (n, t[0]) not in hg] for n, lst in pairs.rhs.items()}
    if a.init:
        model = Trainer(open(a.init, "rb").read())
    else:
        init = lnet.PressureNet(seed=a.seed)
        model = Trainer(layers_to_flat([p.detach().numpy() for p in init.w]))
    bs = {32: 8, 48: 4, 64: 4, 96: 1, 128: 1, 256: 1}
    budget = a.minutes * 60.0
    held = held_set()
    log("pairs %d (held out %d); held-out loss before training: %.4f"
        % (pairs.count(), len(held), evaluate(model, pairs, held, bs)))
    trained = 0.0
    skipped = 0
    step, epoch, last = 0, 0, time.time()
    hist = []
    for rnd in range(a.rounds):
        r_end = budget * (rnd + 1) / a.rounds
        while trained < r_end:
            epoch += 1
            run, cnt = 0.0, 0
            t_ep = time.time()
            for n, ids in batches(pairs, held, bs, cgen):
                t_b = time.time()
                frac = min(trained / budget, 1.0)
                lr = a.lr * 0.5 * (1.0 + math.cos(math.pi * frac))
                r, e, diag, fluid, _w = stack(pairs, n, ids)
                loss = model.step(r, e, diag, fluid, lr=lr, clip=1.0)
                if not math.isfinite(loss):
                    skipped += 1
                    continue
                run += loss
                cnt += 1
                step += 1
                trained += time.time() - t_b
                if time.time() - last > 20.0:
                    last = time.time()
                    log("round %d epoch %d step %d loss %.4f lr %.2e skipped %d" % (
                        rnd, epoch, step, run / max(cnt, 1), lr, skipped))
                if trained >= r_end:
                    break
            ho = evaluate(model, pairs, held, bs)
            hist.append({"round": rnd, "epoch": epoch, "steps": step,
                         "train_loss": run / max(cnt, 1), "held_out": ho,
                         "seconds": time.time() - t_ep})
            log("round %d epoch %d done: train %.4f held-out %.4f" % (rnd, epoch, run / max(cnt, 1), ho))
        if rnd + 1 < a.rounds:
            added = collect_fpcg(pairs, train_rhs, model, a.collect, cgen)
            log("round %d: added %d FPCG residual pairs (total %d)" % (rnd, added, pairs.count()))
            torch_fpcg_check(model, pairs, held_geo[:2])

    fp = torch_fpcg_check(model, pairs, held_geo)
    blob = model.blob()
    with open(os.path.join(a.out, "weights.bin"), "wb") as f:
        f.write(blob)
    check = []
    if a.check:
        spec_c = [(p.split(":")[0], int(p.split(":")[1]), int(p.split(":")[2]))
                  for p in a.check.split(",")]
        check = library_check(blob, spec_c)
    report = {"pairs": pairs.count(), "by_size": {str(n): len(v) for n, v in pairs.items.items()},
              "operator_check": {str(k): v for k, v in op_err.items()}, "history": hist,
              "train_seconds": trained, "skipped_batches": skipped, "torch_fpcg": fp, "library_check": check, "args": vars(a)}
    with open(os.path.join(a.out, "report.json"), "w") as f:
        json.dump(report, f, indent=1)
    log("wrote %s" % a.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
