"""Pins of the screener oracle (CPU only).

* the MLP's analytic gradient equals central finite differences (float64);
* the norm-product bound majorises every chi0s block norm of the oracle on mixed-species k-point
  systems at small and large tau, and is not vacuous (within a factor 1e4 of the largest block);
* a bound from the G_IJ(R) blocks alone (a plausible mistake) fails for some block.
"""
import numpy as np
import pytest

from librpa_amd import synth
from oracle import mlp, rpa, tfgrids


def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    p = mlp.init_params(3) + 0.05 * rng.standard_normal(mlp.NPARAM)
    X = rng.standard_normal((37, mlp.NFEAT))
    t = rng.standard_normal(37)
    loss, g = mlp.loss_and_grad(p, X, t)
    assert abs(loss - np.mean((mlp.forward(p, X) - t) ** 2)) < 1e-14
    idx = np.concatenate([rng.choice(mlp.NPARAM, 60, replace=False), [mlp.NPARAM - 1, 0, 512, 4672, 4736]])
    h = 1e-6
    for i in idx:
        e = np.zeros(mlp.NPARAM)
        e[i] = h
        fd = (mlp.loss_and_grad(p + e, X, t)[0] - mlp.loss_and_grad(p - e, X, t)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 * max(1.0, abs(g[i])), i


def _blocks(s, tau):
    cs = rpa.fold_cs(s)
    Rs = rpa.r_grid(s.period)
    pairs = rpa.upper_pairs(s.natom)
    Gp = rpa.gf_R(s, tau, Rs)
    Gn = rpa.gf_R(s, -tau, Rs)
    chi = rpa.chi0s_factorized(s, cs, Gp, Gn, pairs, Rs)
    return cs, chi, mlp.g_block_norms(s, Gp, pairs, Rs), mlp.g_block_norms(s, Gn, pairs, Rs)


@pytest.mark.parametrize("kw", [
    dict(natom=2, period=(2, 1, 1), nao=(3, 2), nabf=(4, 3), nocc=(1, 1), a=7.0),
    dict(natom=3, period=(1, 2, 2), nao=(2, 3), nabf=(3, 4), nocc=(1, 1), a=6.5, seed=11),
])
@pytest.mark.parametrize("tau", [0.05, 3.0, 40.0])
def test_norm_product_bound_majorises_blocks(kw, tau):
    s = synth.tiny(**kw)
    cs, chi, gp, gm = _blocks(s, tau)
    c = mlp.cs_norm_sums(s, cs)
    ratios = []
    for k, blk in chi.items():
        b = mlp.plain_bound(c, gp, gm, *k)
        nrm = np.linalg.norm(blk)
        assert nrm <= b * (1 + 1e-12), (k, nrm, b)
        ratios.append(b / max(nrm, 1e-300))
    big = max(chi, key=lambda k: np.linalg.norm(chi[k]))
    assert mlp.plain_bound(c, gp, gm, *big) < 1e4 * np.linalg.norm(chi[big])


def test_bound_with_ij_blocks_only_fails():
    # a plausible mistake: bound by the G_IJ(R) blocks alone, c_I c_J g+_IJ g-_IJ (forgetting
    # that the labels also pair G_KL / G_IL / G_KJ blocks of other atoms) undercuts a block
    s = synth.tiny(natom=2, period=(2, 1, 1), nao=(3, 2), nabf=(4, 3), nocc=(1, 1), a=7.0)
    cs, chi, gp, gm = _blocks(s, 0.05)
    c = mlp.cs_norm_sums(s, cs)
    worst = max(np.linalg.norm(b) / (c[k[0]] * c[k[1]] * gp[k] * gm[k]) for k, b in chi.items())
    assert worst > 1.0
