"""GPU parity for the Hénon–Heiles configuration's pieces: the energy-band box predicate, the
subdivision loop restricted to the band with UniformGrid(64, seed 0) and 50 Störmer–Verlet
steps, its transition matrix, and the subspace iteration for the 4 leading eigenvalues."""
import numpy as np
import pytest

import b12_amd as gx
from helpers import random_tree, same_tree
from oracle_api import OracleTree

pytestmark = pytest.mark.gpu

HH_C, HH_R = [0.0, 0.0, 0.0, 0.0], [0.6, 0.6, 0.5, 0.5]
EMIN, EMAX = 0.11, 0.12
BAND = 0x80  # build-defined flag for the band predicate


def band_steps(t, grid, mp, depth):
    """subdivide(SD); keep the boxes meeting the band; keep the boxes hit by an image."""
    counts = []
    t.setFlags(gx.SD, 0)
    for _ in range(depth):
        a = t.subdivide(gx.SD)
        b = t.setFlagsByEnergyBand(EMIN, EMAX, BAND)
        c = t.remove(BAND)
        t.unsetFlags(BAND, -1)
        d = t.setFlagsByIterators(t.makePointIteratorPairOverLeaves(grid, mp), gx.HIT, -1)
        e = t.remove(gx.HIT)
        t.unsetFlags(gx.HIT, -1)
        counts.append((a, b, c, d, e, t.size()))
    return counts


def sorted_eigs(H):
    w = np.linalg.eigvals(H)
    return w[np.lexsort((w.imag, w.real))]


def test_energy_band_predicate_matches_oracle():
    print("[hh] band predicate", flush=True)
    rng = np.random.default_rng(31)
    for depth in (12, 20, 28):
        bits, dep, flags = random_tree(rng, 4, 4000, max_depth=depth)
        g = gx.ImplicitBoxTree(HH_C, HH_R)
        o = OracleTree(HH_C, HH_R)
        g.upload(bits, dep, flags)
        o.upload(bits, dep, flags)
        for emin, emax in [(EMIN, EMAX), (0.0, 0.05), (-1.0, -0.5)]:
            assert g.setFlagsByEnergyBand(emin, emax, BAND) == o.setFlagsByEnergyBand(emin, emax, BAND)
            assert same_tree(g.download(), o.download())
            assert g.unsetFlags(BAND, -1) == o.unsetFlags(BAND, -1)


def test_band_subdivision_matrix_and_subspace_match_oracle():
    """The configuration's loop to depth 14: per-step counts and covers bit-exact, the matrix
    bit-exact, the 4 leading Ritz values within 1e-9."""
    print("[hh] band subdivision d14", flush=True)
    grid, mp = gx.UniformGrid(64, 0), gx.HenonHeiles(2, 0.02, 50)
    g = gx.ImplicitBoxTree(HH_C, HH_R)
    o = OracleTree(HH_C, HH_R)
    cg = band_steps(g, grid, mp, 14)
    co = band_steps(o, grid, mp, 14)
    assert cg == co
    assert same_tree(g.download(), o.download())
    mg = g.transitionMatrixForLeaves(grid, mp)
    mo = o.transitionMatrixForLeaves(grid, mp)
    for a, b in zip(mg.download(), mo.download()):
        assert np.array_equal(a, b)
    print(f"  boxes {g.size()} nnz {mg.nnz}", flush=True)
    Hg, Xg = mg.subspace_iteration(4, 60, want_basis=True)
    Ho, Xo = mo.subspace_iteration(4, 60, want_basis=True)
    eg, eo = sorted_eigs(Hg), sorted_eigs(Ho)
    print("  ritz values", np.round(eg, 12), flush=True)
    assert np.allclose(eg, eo, rtol=0, atol=1e-9)
    assert np.allclose(Xg @ Xg.T, np.eye(4), atol=1e-10)
    assert np.all(np.abs(eg) <= 1.0 + 1e-12)


@pytest.mark.parametrize("nvec", [1, 3, 8])
def test_subspace_iteration_vector_counts_match_oracle(nvec):
    """Other subspace sizes on the Hénon depth-14 matrix: Ritz values within 1e-9 of the oracle;
    one vector gives the leading eigenvalue the power iteration converges to."""
    print(f"[subspace] nvec={nvec}", flush=True)
    grid, mp = gx.InnerGrid(4), gx.Henon(1.4, 0.3, 1)
    g = gx.ImplicitBoxTree([0.0, 0.0], [1.5, 0.5])
    o = OracleTree([0.0, 0.0], [1.5, 0.5])
    for t in (g, o):
        t.setFlags(gx.SD, 0)
        t.subdivision_steps(grid, mp, 14)
    mg = g.transitionMatrixForLeaves(grid, mp)
    mo = o.transitionMatrixForLeaves(grid, mp)
    Hg, Xg = mg.subspace_iteration(nvec, 80, want_basis=True)
    Ho, _ = mo.subspace_iteration(nvec, 80)
    assert np.allclose(sorted_eigs(Hg), sorted_eigs(Ho), rtol=0, atol=1e-9)
    if nvec == 1:
        # one vector is the power method normalised in the 2-norm from the same uniform start:
        # its basis is the 80th power iterate scaled to unit 2-norm
        mu, it, _ = mg.power_iteration(80)
        assert it == 80
        assert np.allclose(Xg[0], mu / np.linalg.norm(mu), rtol=1e-10, atol=1e-15)
