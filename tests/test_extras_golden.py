"""The oracle's ablexp sponge, Graves-Qs attenuation and force-source RHS
against the reference's outputs stored in tests/golden/extras.npz (written
by `python tests/golden/make_golden.py extras` from the reference's own
bdry_t.c and sv_curv_col_el.c), so these stay pinned where the reference
tree is absent."""
import os

import numpy as np

from oracle import pyoracle as po
from tests.golden import extras_cases as ec
from tests.golden.make_golden import checksum

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "extras.npz"),
            allow_pickle=False)


def test_sponge_matches_golden():
    assert checksum(ec.sponge(po.load_oracle().orc_ablexp_apply)) == str(G["ablexp"])


def test_graves_matches_golden():
    assert checksum(ec.graves(po.load_oracle().orc_graves_qs)) == str(G["graves_qs"])


def test_force_rhs_matches_golden():
    assert checksum(ec.force_oracle()) == str(G["force_rhs"])
