"""End-to-end surrogate experiment: generate FD runs with the batched solver,
train the network on the GPU with the library's BF16 kernels, and evaluate
it on held-out sources against FD.

    python -m cgfd3d_amd.surrogate_experiment --train 768 --test 128

The default form is the per-receiver network (one row per source-receiver
pair); --per-source selects one row per source with all receivers' traces.

Prints one JSON line: FD and surrogate seismogram sets per second (a set is
the 3 x nrecv x nt traces of one source), held-out relative L2 misfits of
the surrogate and of the baselines without it (nearest training source in
parameter space, and the untrained network), and the training record.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def train(net, x, y, tn2, epochs, batch, lr_max, seed, time_budget=None, log_every=None, m=None):
    """Minibatch Adam with a cosine learning-rate decay; y and tn2 are device
    tensors, x (and m, the moment tensors of a moment-tensor network) host
    arrays.  Returns the per-epoch mean training loss."""
    import torch
    n = x.shape[0]
    rng = np.random.default_rng(seed)
    xd = torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
    md = None if m is None else torch.from_numpy(np.ascontiguousarray(m, np.float32)).to("cuda:0")
    steps_per_epoch = n // batch
    total = epochs * steps_per_epoch
    hist = []
    t0 = time.perf_counter()
    step = 0
    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(n)).to("cuda:0")
        acc = torch.zeros((), device="cuda:0")
        for s in range(steps_per_epoch):
            idx = perm[s * batch:(s + 1) * batch]
            lr = lr_max * 0.5 * (1.0 + math.cos(math.pi * step / max(1, total)))
            loss = net.train_step(xd[idx], y[idx], tn2[idx], lr,
                                  m=None if md is None else md[idx])
            acc += loss.sum()
            step += 1
        hist.append(float(acc) / (steps_per_epoch * batch))
        if log_every and (ep % log_every == 0 or ep == epochs - 1):
            log(f"epoch {ep}: train loss {hist[-1]:.4f} ({time.perf_counter() - t0:.1f} s)")
        if time_budget and time.perf_counter() - t0 > time_budget:
            log(f"training time budget reached after {ep + 1} epochs")
            break
    return hist


def predict(net, x, m=None, nsamp_per_call=64):
    """Host predictions [n, nt * channels] of the network for feature rows x
    (and moment tensors m of a moment-tensor network)."""
    outs = []
    n = x.shape[0]
    for c0 in range(0, n, nsamp_per_call):
        xb = x[c0:c0 + nsamp_per_call]
        mb = None if m is None else m[c0:c0 + nsamp_per_call]
        cnt = xb.shape[0]
        pad = (-cnt) % 8
        if pad:
            xb = np.concatenate([xb, np.zeros((pad, xb.shape[1]), xb.dtype)])
            if mb is not None:
                mb = np.concatenate([mb, np.zeros((pad, mb.shape[1]), mb.dtype)])
        y = net.forward(xb, m=mb).float().cpu().numpy().reshape(cnt + pad, -1)[:cnt]
        outs.append(y)
    return np.concatenate(outs)


def inference_rate(net, x, m=None, nsamp=256, reps=10):
    """Surrogate seismogram sets per second: forward passes of nsamp sources
    timed with HIP events on the launch stream."""
    import torch
    xb = np.resize(x, (nsamp, x.shape[1])).astype(np.float32)
    xd = torch.from_numpy(xb).to("cuda:0")
    md = None if m is None else torch.from_numpy(
        np.resize(m, (nsamp, m.shape[1])).astype(np.float32)).to("cuda:0")
    net.forward(xd, m=md)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        net.forward(xd, m=md)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return nsamp / (ms / 1e3), ms


def nearest_neighbour(xtr, ytr, xte):
    """Baseline without the network: the seismograms of the training source
    nearest in the scaled parameter space (the 7 scaled parameters)."""
    d = ((xte[:, None, :7] - xtr[None, :, :7]) ** 2).sum(axis=2)
    return ytr[np.argmin(d, axis=1)]


def per_receiver(args, grid, ptr, pte, seis_tr, seis_te, fd_rate, n_fd_here):
    """The per-receiver form: rows are (source, receiver) pairs, the decoder
    emits the 3-component trace of one receiver (6 moment-tensor basis sets),
    the loss of a row is normalised by its source's whole trace set, so the
    sum over a source's rows is that source's relative squared misfit."""
    import torch
    from . import surrogate as sg
    from . import surrogate_data as sd
    nr, nt = grid.nrecv, grid.nt
    xtr, xte = sd.receiver_features(ptr, grid), sd.receiver_features(pte, grid)
    mtr = np.repeat(sd.moment_tensors(ptr), nr, axis=0)
    mte = np.repeat(sd.moment_tensors(pte), nr, axis=0)
    ytr, yte = sd.to_receiver_layout(seis_tr), sd.to_receiver_layout(seis_te)
    # arrival-aligned targets: each trace moved earlier by a lower bound of
    # its first arrival, predictions moved back (an exact reparametrisation:
    # the trace is zero before the bound)
    sh_tr = sd.arrival_shifts(ptr, grid) if args.align else np.zeros(len(ytr), np.int64)
    sh_te = sd.arrival_shifts(pte, grid) if args.align else np.zeros(len(yte), np.int64)
    if args.align:
        ytr = sd.shift_rows(ytr, sh_tr)
    scale = 1.0 / float(np.sqrt((ytr.astype(np.float64) ** 2).mean()))
    ref = sd.to_target_layout(seis_te).reshape(len(pte), -1)
    nn = sd.relative_misfit(nearest_neighbour(sd.features(ptr, grid),
                                              sd.to_target_layout(seis_tr).reshape(len(ptr), -1),
                                              sd.features(pte, grid)), ref)
    tn2_src = ((seis_tr.astype(np.float64) * scale) ** 2).reshape(len(ptr), -1).sum(axis=1)
    tn2 = torch.from_numpy(np.repeat(tn2_src, nr).astype(np.float32)).to("cuda:0")
    yd = torch.from_numpy((ytr * scale).astype(np.float32)).to("cuda:0")

    def assemble(rows):
        return sd.to_target_layout(rows.reshape(len(pte), nr, nt, 3).transpose(0, 1, 3, 2)) \
            .reshape(len(pte), -1)

    def fit(width, epochs=args.epochs, lr=args.lr):
        layers = sg.default_layers(1, nt, 6, width)
        net = sg.SurrogateNet(layers, seed=args.seed, mt_terms=6, out_shape=(nt, 3))
        t = time.perf_counter()
        hist = train(net, xtr, yd, tn2, epochs, args.batch * nr, lr, args.seed,
                     time_budget=args.train_seconds, log_every=50, m=mtr)
        torch.cuda.synchronize()
        pred = predict(net, xte, mte, nsamp_per_call=64 * nr) / scale
        if args.align:
            pred = sd.shift_rows(pred.reshape(len(pte) * nr, nt, 3), sh_te, back=True)
        mis = sd.relative_misfit(assemble(pred), ref)
        ntr = min(128, len(ptr))
        ptr_pred = predict(net, xtr[:ntr * nr], mtr[:ntr * nr], nsamp_per_call=64 * nr) / scale
        if args.align:
            ptr_pred = sd.shift_rows(ptr_pred.reshape(ntr * nr, nt, 3), sh_tr[:ntr * nr],
                                     back=True)
        mis_tr = sd.relative_misfit(
            sd.to_target_layout(ptr_pred.reshape(ntr, nr, nt, 3).transpose(0, 1, 3, 2))
            .reshape(ntr, -1), sd.to_target_layout(seis_tr[:ntr]).reshape(ntr, -1))
        rec = dict(width=width, epochs=epochs, lr=lr, train_runs=len(ptr), aligned=args.align,
                   heldout_misfit_median=float(np.median(mis)),
                   heldout_misfit_mean=float(mis.mean()),
                   heldout_frac_le_5pct=float((mis <= 0.05).mean()),
                   heldout_frac_le_50pct=float((mis <= 0.5).mean()),
                   train_misfit_median=float(np.median(mis_tr)), epochs_run=len(hist),
                   train_seconds=time.perf_counter() - t,
                   final_train_loss=hist[-1] if hist else None)
        log(f"per-receiver width {width}, {epochs} epochs, lr {lr}: held-out median "
            f"{rec['heldout_misfit_median']:.3f}, train {rec['train_misfit_median']:.3f}")
        return net, rec, mis

    sweep = []
    for item in [v for v in args.sweep.split(";") if v]:
        kw = dict(kv.split("=") for kv in item.split(","))
        sweep.append(fit(int(kw.get("width", args.width)), int(kw.get("epochs", args.epochs)),
                         float(kw.get("lr", args.lr)))[1])
    net, rec, mis = fit(args.width)
    rate, ms = inference_rate(net, xte[:256 * nr], mte[:256 * nr], nsamp=256 * nr, reps=5)
    out = {"metric": "surrogate vs FD seismogram sets/s", "form": "per-receiver decoder",
           "surrogate_sets_per_s": rate / nr, "surrogate_ms_per_256": ms,
           "fd_sets_per_s": fd_rate, "speedup": rate / nr / fd_rate if fd_rate else None,
           "grid": [grid.ni, grid.nj, grid.nk], "nt": nt, "nrecv": nr,
           "test_runs": len(pte), "fd_runs_this_job": n_fd_here, **rec,
           "nearest_neighbour_misfit_median": float(np.median(nn)), "sweep": sweep,
           "params": net.n_params(), "moment_tensor_output": True,
           "dtype": "bf16 weights/activations, fp32 accumulation"}
    if args.save:
        np.savez(args.save, misfit=mis, nn_misfit=nn, params_test=pte)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=768)
    ap.add_argument("--test", type=int, default=128)
    ap.add_argument("--fd-batch", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--train-seconds", type=float, default=240.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--save", type=str, default="")
    ap.add_argument("--subsets", type=str, default="",
                    help="comma-separated training-set sizes (nested prefixes of the "
                         "training runs): one network each, reported under 'subsets'")
    ap.add_argument("--width", type=int, default=1, help="decoder channel multiplier")
    ap.add_argument("--sweep", type=str, default="",
                    help="extra trainings on the same data, ';'-separated key=value lists "
                         "over lr, epochs, batch, width (e.g. 'lr=3e-3;width=2,lr=1e-3')")
    ap.add_argument("--generate", type=int, default=0, metavar="N",
                    help="only run N training sources (seed --seed) and store them, band-"
                         "limited, in --store; prints the FD rate and the storage error")
    ap.add_argument("--store", type=str, default="")
    ap.add_argument("--test-store", type=str, default="",
                    help="held-out runs from a store of a --generate job instead of new FD runs")
    ap.add_argument("--extra", type=str, default="",
                    help="comma-separated stores of earlier --generate jobs, added to the "
                         "training runs")
    ap.add_argument("--per-source", action="store_true",
                    help="one decoder output per source with every receiver's traces (the "
                         "default form runs one shared decoder per receiver trace, the MLP "
                         "also taking the receiver position and the source-receiver offset)")
    ap.add_argument("--no-mt", action="store_true",
                    help="plain output layer (no moment-tensor contraction)")
    ap.add_argument("--align", action="store_true",
                    help="per-receiver form: train on arrival-aligned traces (each moved "
                         "earlier by a lower bound of its first arrival)")
    ap.add_argument("--small", action="store_true",
                    help="a 64x64x48 grid with 256 steps (pipeline check)")
    args = ap.parse_args()
    import torch
    from . import surrogate as sg
    from . import surrogate_data as sd

    t = time.perf_counter()
    if args.small:
        grid = sd.SurrogateGrid(ni=64, nj=64, nk=48, nt=256, n_side=4)
    else:
        grid = sd.SurrogateGrid()
    log(f"grid {grid.wl.name} built in {time.perf_counter() - t:.1f} s")
    if args.generate:
        ptr, ijk_tr = grid.sample_sources(args.generate, args.seed)
        gen = sd.FdGenerator(grid, args.fd_batch, progress=True)
        seis = gen.run(ptr, ijk_tr)
        gen.close()
        z, nt = sd.compress_runs(seis, grid.dt)
        err = sd.relative_misfit(sd.decompress_runs(z, nt).reshape(len(seis), -1),
                                 seis.reshape(len(seis), -1))
        np.savez(args.store, params=ptr, ijk=ijk_tr, z=z, nt=nt, dt=grid.dt, seed=args.seed)
        print(json.dumps({"generated": args.generate, "seed": args.seed,
                          "fd_sets_per_s": gen.runs / gen.fd_seconds,
                          "store_bytes": int(z.nbytes), "storage_misfit_median": float(np.median(err)),
                          "storage_misfit_max": float(err.max())}), flush=True)
        return
    ptr, ijk_tr = grid.sample_sources(args.train, args.seed)
    gen = sd.FdGenerator(grid, args.fd_batch, progress=True)
    t = time.perf_counter()
    seis_tr = gen.run(ptr, ijk_tr)
    log(f"{args.train} training runs in {time.perf_counter() - t:.1f} s "
        f"(device {gen.fd_seconds:.1f} s)")
    if args.test_store:
        d = np.load(args.test_store)
        pte = d["params"]
        seis_te = sd.decompress_runs(d["z"], int(d["nt"]))
        log(f"{len(pte)} held-out runs from {args.test_store} (seed {int(d['seed'])})")
    else:
        pte, ijk_te = grid.sample_sources(args.test, args.seed + 1000)
        t = time.perf_counter()
        seis_te = gen.run(pte, ijk_te)
        log(f"{args.test} held-out runs in {time.perf_counter() - t:.1f} s")
    gen.close()
    # FD rate of this job's runs (None when every run came from stores)
    fd_rate = gen.runs / gen.fd_seconds if gen.runs else None
    n_fd_here = len(ptr)
    for path in [v for v in args.extra.split(",") if v]:
        d = np.load(path)
        if int(d["nt"]) != grid.nt:
            raise ValueError(f"{path}: {int(d['nt'])} steps, the grid runs {grid.nt}")
        ptr = np.concatenate([d["params"], ptr])
        seis_tr = np.concatenate([sd.decompress_runs(d["z"], int(d["nt"])), seis_tr])
        log(f"added {len(d['params'])} stored runs from {path} (seed {int(d['seed'])})")

    if not args.per_source:
        return per_receiver(args, grid, ptr, pte, seis_tr, seis_te, fd_rate, n_fd_here)
    mt = 0 if args.no_mt else 6
    xtr, xte = sd.features(ptr, grid, not mt), sd.features(pte, grid, not mt)
    mtr = sd.moment_tensors(ptr) if mt else None
    mte = sd.moment_tensors(pte) if mt else None
    ytr, yte = sd.to_target_layout(seis_tr), sd.to_target_layout(seis_te)
    scale = 1.0 / float(np.sqrt((ytr.astype(np.float64) ** 2).mean()))
    nn = sd.relative_misfit(nearest_neighbour(sd.features(ptr, grid), ytr.reshape(len(ytr), -1),
                                              sd.features(pte, grid)),
                            yte.reshape(len(yte), -1))

    def fit(ntrain, lr=args.lr, epochs=args.epochs, batch=args.batch, width=args.width):
        layers = sg.default_layers(grid.nrecv, grid.nt, mt, width)
        net = sg.SurrogateNet(layers, seed=args.seed, mt_terms=mt,
                              out_shape=(grid.nt, 3 * grid.nrecv))
        yd = torch.from_numpy((ytr[:ntrain] * scale).astype(np.float32)).to("cuda:0")
        tn2 = torch.from_numpy(((ytr[:ntrain] * scale).reshape(ntrain, -1).astype(np.float64) ** 2)
                               .sum(axis=1).astype(np.float32)).to("cuda:0")
        untrained = sd.relative_misfit(predict(net, xte, mte) / scale, yte.reshape(len(yte), -1))
        t = time.perf_counter()
        hist = train(net, xtr[:ntrain], yd, tn2, epochs, batch, lr, args.seed,
                     time_budget=args.train_seconds, log_every=50,
                     m=None if mtr is None else mtr[:ntrain])
        torch.cuda.synchronize()
        train_s = time.perf_counter() - t
        pred = predict(net, xte, mte) / scale
        mis = sd.relative_misfit(pred, yte.reshape(len(yte), -1))
        ntr = min(128, ntrain)
        mis_tr = sd.relative_misfit(
            predict(net, xtr[:ntr], None if mtr is None else mtr[:ntr]) / scale,
            ytr[:ntr].reshape(ntr, -1))
        del yd, tn2
        return net, dict(train_runs=ntrain, lr=lr, epochs=epochs, batch=batch, width=width,
                         heldout_misfit_median=float(np.median(mis)),
                         heldout_misfit_mean=float(mis.mean()),
                         heldout_frac_le_5pct=float((mis <= 0.05).mean()),
                         train_misfit_median=float(np.median(mis_tr)),
                         untrained_misfit_median=float(np.median(untrained)),
                         epochs_run=len(hist), train_seconds=train_s,
                         final_train_loss=hist[-1] if hist else None), mis, pred

    subsets = []
    for n_sub in [int(v) for v in args.subsets.split(",") if v]:
        if n_sub < len(xtr):
            _, rec, _, _ = fit(n_sub)
            log(f"subset {n_sub}: held-out median misfit {rec['heldout_misfit_median']:.3f}")
            subsets.append(rec)
    sweep = []
    for item in [v for v in args.sweep.split(";") if v]:
        kw = {}
        for kv in item.split(","):
            key, val = kv.split("=")
            kw[key.strip()] = float(val) if key.strip() == "lr" else int(val)
        _, rec, _, _ = fit(len(xtr), **kw)
        log(f"sweep {kw}: held-out median misfit {rec['heldout_misfit_median']:.3f}, "
            f"train {rec['train_misfit_median']:.3f}")
        sweep.append(rec)
    net, rec, mis, pred = fit(len(xtr))
    log(f"network: {net.n_params() / 1e6:.2f} M parameters")
    rate, ms = inference_rate(net, xte, mte)
    out = {
        "metric": "surrogate vs FD seismogram sets/s",
        "surrogate_sets_per_s": rate, "surrogate_ms_per_256": ms,
        "fd_sets_per_s": fd_rate, "speedup": rate / fd_rate if fd_rate else None,
        "grid": [grid.ni, grid.nj, grid.nk], "nt": grid.nt, "nrecv": grid.nrecv,
        "test_runs": len(pte), "fd_runs_this_job": n_fd_here, **rec,
        "nearest_neighbour_misfit_median": float(np.median(nn)),
        "subsets": subsets, "sweep": sweep,
        "params": net.n_params(), "moment_tensor_output": bool(mt),
        "dtype": "bf16 weights/activations, fp32 accumulation",
    }
    if args.save:
        os.makedirs(os.path.dirname(os.path.abspath(args.save)), exist_ok=True)
        np.savez(args.save, misfit=mis, nn_misfit=nn, params_test=pte,
                 pred0=pred[:2], ref0=yte[:2])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
