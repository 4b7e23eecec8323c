"""GPU: what the trained surrogate gains over the method without it.

A tiny fixed grid (48 x 48 x 40, 128 steps, 4 x 4 surface receivers), FD
runs of seeded sources from the batched solver as training and held-out
data, the moment-tensor network trained by the library's kernels.  Without
the network the best one can do with the same FD runs is to reuse the
seismograms of the nearest training source (or predict nothing, misfit 1);
the trained network must beat both on held-out sources, and it must be far
faster than the FD solver it replaces."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_trained_surrogate_beats_nearest_neighbour_and_fd_speed():
    import torch
    from cgfd3d_amd import surrogate as sg
    from cgfd3d_amd import surrogate_data as sd
    from cgfd3d_amd.surrogate_experiment import (inference_rate, nearest_neighbour, predict,
                                                 train)
    grid = sd.SurrogateGrid(ni=48, nj=48, nk=40, nt=128, n_side=4, pml=8)
    ptr, ijk_tr = grid.sample_sources(1024, 5)
    pte, ijk_te = grid.sample_sources(32, 6)
    gen = sd.FdGenerator(grid, 64)
    seis_tr = gen.run(ptr, ijk_tr)
    seis_te = gen.run(pte, ijk_te)
    fd_rate = gen.runs / gen.fd_seconds
    gen.close()
    ytr, yte = sd.to_target_layout(seis_tr), sd.to_target_layout(seis_te)
    assert np.isfinite(ytr).all() and np.abs(ytr).max() > 0
    scale = 1.0 / float(np.sqrt((ytr.astype(np.float64) ** 2).mean()))
    xtr, xte = sd.features(ptr, grid, False), sd.features(pte, grid, False)
    mtr, mte = sd.moment_tensors(ptr), sd.moment_tensors(pte)
    layers = sg.default_layers(grid.nrecv, grid.nt, 6)
    net = sg.SurrogateNet(layers, seed=3, mt_terms=6, out_shape=(grid.nt, 3 * grid.nrecv))
    yd = torch.from_numpy((ytr * scale).astype(np.float32)).to("cuda:0")
    tn2 = torch.from_numpy(((ytr * scale).reshape(len(ytr), -1).astype(np.float64) ** 2)
                           .sum(axis=1).astype(np.float32)).to("cuda:0")
    train(net, xtr, yd, tn2, 200, 32, 1e-3, 3, m=mtr)
    ref = yte.reshape(len(yte), -1)
    mis = np.median(sd.relative_misfit(predict(net, xte, mte) / scale, ref))
    nn = np.median(sd.relative_misfit(
        nearest_neighbour(sd.features(ptr, grid), ytr.reshape(len(ytr), -1),
                          sd.features(pte, grid)), ref))
    rate, _ = inference_rate(net, xte, mte)
    print(f"held-out median misfit: surrogate {mis:.3f}, nearest training source {nn:.3f}; "
          f"seismogram sets/s: surrogate {rate:.3e}, FD {fd_rate:.3e}")
    assert mis < 0.9 and mis < nn
    assert rate > 100 * fd_rate
