"""CPU checks that pin the output-sub-volume oracle (snapshots, slices):
the packing against a point-by-point loop, and the stress-to-strain
conversion against Hooke's law."""
import numpy as np
import pytest

from oracle import output_oracle as oo


def test_pack_matches_point_loop():
    rng = np.random.default_rng(0)
    w = rng.standard_normal((9, 7, 8, 11)).astype(np.float32)
    box = (1, 4, 3, 1, 3, 2, 0, 3, 3)          # i 1,4,7,10  j 1,3,5  k 0,3,6
    comps = [8, 0, 4]
    got = oo.pack_subvolume(w, comps, box)
    flat = []
    for c in comps:                             # io_funcs.c:1536-1551 loop order
        for n3 in range(3):
            for n2 in range(3):
                for n1 in range(4):
                    flat.append(w[c, 0 + 3 * n3, 1 + 2 * n2, 1 + 3 * n1])
    assert np.array_equal(got.reshape(-1), np.array(flat, np.float32))


def test_slices_are_one_point_boxes():
    rng = np.random.default_rng(1)
    w = rng.standard_normal((3, 6, 7, 8)).astype(np.float32)
    # slice x at i=5 over j, k (io_funcs.c:1022-1045): k outer, j inner
    sx = oo.pack_subvolume(w, [0, 1, 2], (5, 1, 1, 0, 7, 1, 0, 6, 1))
    assert np.array_equal(sx[:, :, :, 0], w[:, :, :, 5])
    sz = oo.pack_subvolume(w, [2], (0, 8, 1, 0, 7, 1, 4, 1, 1))
    assert np.array_equal(sz[0, 0], w[2, 4])


def _stresses(lam, mu, E):
    """Hooke's law in float64: T = lam tr(E) I + 2 mu E; E as (xx, yy, zz,
    yz, xz, xy) tensor components."""
    exx, eyy, ezz, eyz, exz, exy = E
    tr = exx + eyy + ezz
    return [lam * tr + 2 * mu * exx, lam * tr + 2 * mu * eyy, lam * tr + 2 * mu * ezz,
            2 * mu * eyz, 2 * mu * exz, 2 * mu * exy]


def test_strain_inverts_hooke():
    rng = np.random.default_rng(2)
    shape = (4, 5, 6)
    rho = rng.uniform(1500, 3000, shape)
    vs = rng.uniform(500, 3500, shape)
    vp = vs * rng.uniform(1.6, 2.2, shape)
    mu = (rho * vs ** 2).astype(np.float32)
    lam = (rho * vp ** 2 - 2 * rho * vs ** 2).astype(np.float32)
    # distinct strain components so a swapped or dropped term shows
    E = [rng.uniform(-1e-4, 1e-4, shape) * s for s in (1.0, 2.0, 3.0, 4.0, 5.0, 6.0)]
    T = [t.astype(np.float32) for t in _stresses(lam.astype(np.float64), mu.astype(np.float64), E)]
    box = (0, 6, 1, 0, 5, 1, 0, 4, 1)
    got = oo.stress_to_strain_eliso(lam, mu, T, box).astype(np.float64)
    want = [E[0], E[1], E[2], E[4], E[3], E[5]]          # Exx, Eyy, Ezz, Exz, Eyz, Exy
    for g, e in zip(got, want):
        assert np.allclose(g, e, rtol=2e-4, atol=1e-9 * np.abs(e).max())
    # the same check fails for a shear strain without the factor one half and
    # for Exz and Eyz in each other's place
    assert not np.allclose(2 * got[3], want[3], rtol=2e-4)
    assert not np.allclose(got[3], E[3], rtol=1e-2)


def test_strain_of_pure_pressure():
    """Isotropic stress -p I gives the volumetric strain -p / (3 lam + 2 mu)
    on each diagonal component and no shear."""
    lam = np.full((2, 2, 2), 3.0e10, np.float32)
    mu = np.full((2, 2, 2), 2.5e10, np.float32)
    p = np.float32(1.0e6)
    z = np.zeros_like(lam)
    T = [-p + z, -p + z, -p + z, z, z, z]
    got = oo.stress_to_strain_eliso(lam, mu, T, (0, 2, 1, 0, 2, 1, 0, 2, 1))
    assert np.allclose(got[:3], -1.0e6 / (3 * 3.0e10 + 2 * 2.5e10), rtol=1e-6)
    assert not got[3:].any()
