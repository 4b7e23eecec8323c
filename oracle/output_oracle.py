"""Test-only checker for the output sub-volumes (snapshots and slices): a
numpy restatement of the reference's snapshot packing and stress-to-strain
conversion.  Only tests import it.

The reference sources for these two functions (forward/io_funcs.c) include
netcdf.h, which this image lacks, so they are not compiled into oracle/_ref.
Each function is pinned by mathematics instead (tests/test_output_oracle.py):
  pack_subvolume        numpy strided slicing of the same box, and a
                        point-by-point loop on a small grid;
  stress_to_strain_eliso
                        Hooke's law applied to the result gives the stresses
                        back (float64 within float32 rounding), on stresses
                        made from known strains, with distinct shear values
                        so that a swapped or dropped component fails.
"""
import numpy as np


def pack_subvolume(w, comps, box):
    """io_snap_pack_buff (forward/io_funcs.c:1525-1556) for each component:
    w is (ncomp_total, nz, ny, nx); box = (i1, ni, di, j1, nj, dj, k1, nk, dk);
    returns (len(comps), nk, nj, ni), i fastest."""
    i1, ni, di, j1, nj, dj, k1, nk, dk = (int(v) for v in box)
    out = np.empty((len(comps), nk, nj, ni), np.float32)
    for c, comp in enumerate(comps):
        out[c] = w[comp, k1:k1 + (nk - 1) * dk + 1:dk, j1:j1 + (nj - 1) * dj + 1:dj,
                   i1:i1 + (ni - 1) * di + 1:di]
    return out


def stress_to_strain_eliso(lam, mu, T, box):
    """io_snap_stress_to_strain_eliso (forward/io_funcs.c:1456-1522) with its
    rounding: E1, E2, E3 evaluated in double (the literals 3.0, 2.0, 1.0 are
    double) and stored as float, the normal strains in float arithmetic, the
    shear strains as 0.5 * E3 * T in double.  lam, mu: (nz, ny, nx); T: the six
    stresses Txx, Tyy, Tzz, Tyz, Txz, Txy, each (nz, ny, nx).  Returns
    (6, nk, nj, ni) in the order io_snap_nc_put writes the strain variables
    (forward/io_funcs.c:1226-1252): Exx, Eyy, Ezz, Exz, Eyz, Exy."""
    sub = lambda a: pack_subvolume(a[None], [0], box)[0]
    l, m = sub(lam), sub(mu)
    txx, tyy, tzz, tyz, txz, txy = (sub(t) for t in T)
    f64 = np.float64
    den = 3.0 * l.astype(f64) + 2.0 * m.astype(f64)
    E1 = ((l + m).astype(f64) / (m.astype(f64) * den)).astype(np.float32)
    E2 = ((-l).astype(f64) / (2.0 * m.astype(f64) * den)).astype(np.float32)
    E3 = (1.0 / m.astype(f64)).astype(np.float32)
    E0 = E2 * ((txx + tyy) + tzz)
    d = E2 - E1
    shear = lambda t: (0.5 * E3.astype(f64) * t.astype(f64)).astype(np.float32)
    return np.stack([E0 - d * txx, E0 - d * tyy, E0 - d * tzz, shear(txz), shear(tyz),
                     shear(txy)]).astype(np.float32)
