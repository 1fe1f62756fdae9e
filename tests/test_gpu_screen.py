"""Block-norm screener on the device: MLP kernels against the oracle, the features the chi0
build emits, exactness of screened builds, and what the learned screener gains over the
norm-product bound (dEc is measured against the unscreened FP64 build)."""
import numpy as np
import pytest

import librpa_amd as L
from librpa_amd import synth
from oracle import mlp, rpa


def _say(msg):
    print(msg, flush=True)


def _run(s, keep=False, mode=L.SCREEN_OFF, thr=1e-8, model=None, prec=L.FP64):
    h = L.Handler()
    L.load_system(h, s)
    h.set_precision(prec, L.FP64)
    h.set_keep_chi0(keep)
    h.set_screening(mode, thr, model)
    o = L.Options(nfreq=s.nfreq, tfgrids_type=L.MINIMAX, parallel_routing=L.LIBRI)
    ec, _ = h.get_rpa_correlation_energy(o, s.nk)
    return h, o, ec


@pytest.mark.gpu
def test_mlp_kernels_match_oracle(lib):
    _say("mlp kernels vs oracle")
    rng = np.random.default_rng(5)
    p = (mlp.init_params(1) + 0.05 * rng.standard_normal(mlp.NPARAM)).astype(np.float32)
    X = rng.standard_normal((777, mlp.NFEAT)).astype(np.float32)
    t = rng.standard_normal(777).astype(np.float32)
    p64, X64, t64 = p.astype(np.float64), X.astype(np.float64), t.astype(np.float64)
    y = L.mlp_predict(p, X)
    yr = mlp.forward(p64, X64)
    assert np.abs(y - yr).max() < 1e-5 * max(1.0, np.abs(yr).max())
    loss, g = L.mlp_loss_grad(p, X, t)
    lr_, gr = mlp.loss_and_grad(p64, X64, t64)
    _say(f"  loss {loss:.8f} oracle {lr_:.8f}; max |dg| {np.abs(g - gr).max():.2e} of {np.abs(gr).max():.2e}")
    assert abs(loss - lr_) < 1e-5 * lr_
    assert np.abs(g - gr).max() < 1e-4 * np.abs(gr).max()


@pytest.mark.gpu
def test_training_fits_smooth_target(lib):
    _say("mlp training")
    rng = np.random.default_rng(6)
    X = rng.uniform(-2, 2, (4000, mlp.NFEAT)).astype(np.float32)
    t = (3.0 * X[:, 3] - 2.0 * X[:, 6] + np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2]).astype(np.float32)
    model, loss = L.screen_train(X, t, iters=1500, lr=3e-3, seed=2)
    _say(f"  final standardised mse {loss:.2e}")
    assert loss < 2e-3
    # the exported model predicts the raw target through its own scaling
    k = L.SCREEN_NPARAM
    mean, scale = model[k:k + 8], model[k + 8:k + 16]
    tm, ts = model[k + 16], model[k + 17]
    pred = L.mlp_predict(model[:k], (X - mean) / scale) * ts + tm
    assert np.sqrt(np.mean((pred - t) ** 2)) < 0.1 * np.std(t)


@pytest.mark.gpu
def test_screen_samples_match_oracle(lib):
    _say("screen samples vs oracle")
    s = synth.tiny(natom=3, period=(1, 2, 2), nao=(2, 3), nabf=(3, 4), nocc=(1, 1), a=6.5, seed=11, nfreq=8)
    h, o, _ = _run(s, keep=True)
    X, t = h.get_screen_samples()
    grid = h.get_tfgrids(o)
    cs = rpa.fold_cs(s)
    c = mlp.cs_norm_sums(s, cs)
    Rs = rpa.r_grid(s.period)
    pairs = rpa.upper_pairs(s.natom)
    rows = []
    for tau in grid["tau"]:
        Gp, Gn = rpa.gf_R(s, tau, Rs), rpa.gf_R(s, -tau, Rs)
        chi = rpa.chi0s_factorized(s, cs, Gp, Gn, pairs, Rs)
        gp, gm = mlp.g_block_norms(s, Gp, pairs, Rs), mlp.g_block_norms(s, Gn, pairs, Rs)
        for (I, J, R), b in chi.items():
            nrm = np.linalg.norm(b)
            if nrm > 0:
                rows.append((np.log(nrm), mlp.block_features(s, tau, gp, gm, I, J, R, c)))
    assert len(rows) == len(t)
    rows.sort(key=lambda r: r[0])
    order = np.argsort(t)
    tr = np.array([r[0] for r in rows])
    Xr = np.array([r[1] for r in rows])
    assert np.abs(t[order] - tr).max() < 1e-5 * max(1.0, np.abs(tr).max())
    assert np.abs(X[order] - Xr).max() < 1e-4 * max(1.0, np.abs(Xr).max())


@pytest.mark.gpu
def test_bound_screening_is_exact_on_kept_blocks(lib):
    _say("bound screening cell2")
    s = synth.cell2()
    h0, o, ec0 = _run(s, keep=True)
    # threshold at the median block norm: the large-tau blocks, orders of magnitude smaller,
    # fall under it even with the bound's slack
    _, t = h0.get_screen_samples()
    thr = float(np.exp(np.median(t)))
    h1, _, ec1 = _run(s, keep=True, mode=L.SCREEN_BOUND, thr=thr)
    st = h1.get_screen_stats()
    _say(f"  threshold {thr:.2e}: skipped {st['skipped']:.0f} of {st['blocks']:.0f} blocks, dEc {ec1 - ec0:.3e}")
    assert st["skipped"] > 0
    nz = 0
    for it in range(s.nfreq):
        for (I, J) in rpa.upper_pairs(s.natom):
            for R in rpa.r_grid(s.period):
                a = h0.get_chi0s_block(it, I, J, R, 60, 60)
                b = h1.get_chi0s_block(it, I, J, R, 60, 60)
                if not b.any():
                    nz += 1
                    assert np.linalg.norm(a) < thr  # a valid bound never skips a block above it
                else:
                    assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max()
    assert nz == st["skipped"]


@pytest.mark.gpu
def test_bound_screening_with_mixed_precision(lib):
    # the mixed mode's FP32 tau points take their G-block norms from the FP32 G panel; screened
    # and unscreened mixed builds must agree on every kept block (to FP32 rounding at the FP32
    # tau points: the contraction's K parts accumulate atomically, in no fixed order), and no
    # skipped block may exceed the threshold (up to FP32 rounding of the norms)
    _say("bound screening cell2, mixed precision")
    s = synth.cell2()
    h0, o, ec0 = _run(s, keep=True, prec=L.MIXED)
    h64, _, _ = _run(s, keep=True)
    _, t = h64.get_screen_samples()
    thr = float(np.exp(np.median(t)))
    h1, _, ec1 = _run(s, keep=True, mode=L.SCREEN_BOUND, thr=thr, prec=L.MIXED)
    grid = h0.get_tfgrids(o)
    st0 = h0.get_stats()
    low = set(t for t in range(s.nfreq) if np.abs(grid["cos_t2f"][:, t]).max() <= st0["mixed_amax"])
    assert len(low) == st0["n_fp32_tau"]
    st = h1.get_screen_stats()
    _say(f"  threshold {thr:.2e}: skipped {st['skipped']:.0f} of {st['blocks']:.0f} blocks, dEc {ec1 - ec0:.3e}")
    assert st["skipped"] > 0
    nz = 0
    for it in range(s.nfreq):
        for (I, J) in rpa.upper_pairs(s.natom):
            for R in rpa.r_grid(s.period):
                a = h0.get_chi0s_block(it, I, J, R, 60, 60)
                b = h1.get_chi0s_block(it, I, J, R, 60, 60)
                if not b.any():
                    nz += 1
                    assert np.linalg.norm(a) < thr * (1 + 1e-5)
                else:
                    assert np.abs(a - b).max() <= (1e-6 if it in low else 1e-13) * np.abs(a).max()
    assert nz == st["skipped"]
    assert abs(ec1 - ec0) / s.natom < 1e-5


@pytest.mark.gpu
def test_learned_screening_gain_cell16(lib):
    # trained on the block norms of one 16-atom cell (seed 1), applied to a held-out one (seed 6:
    # other jitter, eigenvalues, eigenvectors, Cs and v); dEc against the held-out cell's own
    # unscreened build
    _say("learned screening cell16, held-out system")
    train = synth.cell16()
    ht, _, _ = _run(train, keep=True)
    X, t = ht.get_screen_samples()
    model, loss = L.screen_train(X, t, iters=1500, lr=3e-3, seed=0)
    del ht
    s = synth.cell16(seed=6)
    h0, _, ec0 = _run(s)
    thr = 1e-8
    hb, _, ec_b = _run(s, mode=L.SCREEN_BOUND, thr=thr)
    h2, _, ec_l = _run(s, mode=L.SCREEN_LEARNED, thr=thr, model=model)
    sb, sl = hb.get_screen_stats(), h2.get_screen_stats()
    fb, fl = sb["skipped"] / sb["blocks"], sl["skipped"] / sl["blocks"]
    _say(f"  train mse {loss:.2e} on {len(t)} samples; skipped blocks: bound {fb:.3f}, learned {fl:.3f}; "
         f"flops skipped: bound {sb['flops_skipped'] / sb['flops']:.3f}, learned {sl['flops_skipped'] / sl['flops']:.3f}; "
         f"dEc/atom bound {(ec_b - ec0) / s.natom:.2e}, learned {(ec_l - ec0) / s.natom:.2e}")
    assert fl > 0 and fl >= fb
    assert abs(ec_b - ec0) / s.natom < 1e-5
    assert abs(ec_l - ec0) / s.natom < 1e-5
