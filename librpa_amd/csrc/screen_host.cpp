// Block-norm screening of the chi0 build: plan-time block tables, per-tau features and skip
// flags, training-sample export and the MLP training loop.
//
// A chi0 output block (I, J, R) at tau is skipped when the norm-product bound
//   ||chi0s(I,J,R)||_F <= c_I c_J (g+_IJ gm- + g-_IJ gm+ + 2 gm+ gm-)
// (c_A: sum of the Frobenius norms of atom A's folded Cs blocks; g: Frobenius norms of the
// G_IJ(R, +-tau) blocks, gm: their maxima) is below the threshold, or, in learned mode, when
// the MLP's prediction of ln ||chi0s(I,J,R)||_F is.  Every term of the chi0 label sums
// (thirdparty/LibRI/include/RI/physics/RPA.hpp:104-138) is a contraction <C^I_mu, G C^J_nu G'^T>,
// which the bound majorises term by term.  Skipped blocks are exactly zero in chi0(R, tau).
#include <algorithm>
#include <cmath>
#include <random>

#include "rpa.h"
#include "screen.h"

namespace lra {

void RpaEngine::screen_plan(const Dataset &ds) {
    const int nout = (int)outs.size();
    const long long gpm_r = (long long)NAO * 2 * NAO;
    gblocks.assign(2 * (size_t)nout, BlockRef{});
    feat_static.assign(5 * (size_t)nout, 0.0f);
    std::vector<long long> eo(nout);
    std::vector<int> es(nout);
    const bool geo = ds.lat_set && ds.pos.size() >= 3 * (size_t)natom;
    for (int o = 0; o < nout; ++o) {
        const auto &O = outs[o];
        for (int s = 0; s < 2; ++s)
            gblocks[2 * (size_t)o + s] = {O.r * gpm_r + ao[O.I] * 2LL * NAO + 2 * ao[O.J] + (long long)s * nb[O.J],
                                          nb[O.I], nb[O.J]};
        double d = 0;
        if (geo) {
            double v[3];
            for (int c = 0; c < 3; ++c) {
                v[c] = ds.pos[3 * (size_t)O.J + c] - ds.pos[3 * (size_t)O.I + c];
                for (int a = 0; a < 3; ++a) v[c] += cells[O.r][a] * ds.latvec[3 * a + c];
            }
            d = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        }
        float *f = &feat_static[5 * (size_t)o];
        f[0] = (float)(d / 10.0);
        f[1] = (float)(nb[O.I] / 27.0);
        f[2] = (float)(nb[O.J] / 27.0);
        f[3] = (float)std::log(std::max(cnorm[O.I], 1e-300));
        f[4] = (float)std::log(std::max(cnorm[O.J], 1e-300));
        eo[o] = O.eoff;
        es[o] = na[O.I] * na[O.J];
    }
    d_gblocks = upload(gblocks, stream);
    d_eoff = upload(eo, stream);
    d_esize = upload(es, stream);
    d_gnorm.alloc(sizeof(double) * 2 * std::max(1, nout));
    d_feat.alloc(sizeof(float) * MLP_NFEAT * std::max(1, nout));
    d_pred.alloc(sizeof(float) * std::max(1, nout));
    for (auto &c : chunks) {
        if (c.x_owner.size() != c.xb.probs.size() || (!c.ta.probs.empty() && c.ta_ptr.size() != c.ta.probs.size() + 1))
            LRA_THROW("screening maps out of step with the problem lists");
        for (int k = 0; k < 4; ++k)
            if (c.fin_owner[k].size() != c.fin[k].probs.size()) LRA_THROW("screening maps out of step");
        auto prep = [&](GemmBatch &b, bool zero) {
            b.skipv.assign(b.probs.size(), 0);
            b.d_skip.alloc(std::max<size_t>(1, b.probs.size()));
            b.skip_zero = zero;
            b.use_skip = false;
        };
        prep(c.xb, true);
        prep(c.ta, false);
        for (auto &f : c.fin) prep(f, false);
    }
}

void RpaEngine::screen_tau(int it, double tau) {
    const int nout = (int)outs.size();
    if (nout == 0) return;
    auto &prof = profiler();
    prof.begin("chi0_screen", stream);
    if (cur == Prec::F64)
        launch_block_norms<double>(d_Gpm.as<double>(), 2LL * NAO, d_gblocks.as<BlockRef>(), 2 * nout,
                                   d_gnorm.as<double>(), stream);
    else
        launch_block_norms<float>(d_Gpm.as<float>(), 2LL * NAO, d_gblocks.as<BlockRef>(), 2 * nout,
                                  d_gnorm.as<double>(), stream);
    gnorm_h.resize(2 * (size_t)nout);
    HIP_CHECK(hipMemcpyAsync(gnorm_h.data(), d_gnorm.get(), sizeof(double) * 2 * nout, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    double gpx = 0, gmx = 0;
    for (int o = 0; o < nout; ++o) {
        gpx = std::max(gpx, gnorm_h[2 * (size_t)o]);
        gmx = std::max(gmx, gnorm_h[2 * (size_t)o + 1]);
    }
    feat_h.resize((size_t)MLP_NFEAT * nout);
    const float ltau = (float)std::log(tau), lgm = (float)std::log(gpx * gmx + 1e-300);
    for (int o = 0; o < nout; ++o) {
        const float *fs = &feat_static[5 * (size_t)o];
        float *f = &feat_h[(size_t)MLP_NFEAT * o];
        f[0] = fs[0]; f[1] = fs[1]; f[2] = fs[2]; f[3] = ltau; f[4] = fs[3]; f[5] = fs[4];
        f[6] = (float)std::log(gnorm_h[2 * (size_t)o] * gnorm_h[2 * (size_t)o + 1] + 1e-300);
        f[7] = lgm;
    }
    if (collect) std::copy(feat_h.begin(), feat_h.end(), features.begin() + (size_t)it * MLP_NFEAT * nout);
    if (screen_mode == 0) {
        prof.end("chi0_screen", stream);
        return;
    }
    skip_out.assign(nout, 0);
    for (int o = 0; o < nout; ++o) {
        const auto &O = outs[o];
        const double bound = cnorm[O.I] * cnorm[O.J] *
                             (gnorm_h[2 * (size_t)o] * gmx + gnorm_h[2 * (size_t)o + 1] * gpx + 2.0 * gpx * gmx);
        if (bound < screen_thr) skip_out[o] = 1;
    }
    if (screen_mode == 2) {
        // standardise with the model's scaling, predict ln ||chi0s|| on the device
        const float *mean = screen_model.data() + MLP_NPARAM, *scale = mean + MLP_NFEAT;
        const float tmean = screen_model[MLP_NPARAM + 2 * MLP_NFEAT], tscale = screen_model[MLP_NPARAM + 2 * MLP_NFEAT + 1];
        std::vector<float> xs(feat_h.size());
        for (int o = 0; o < nout; ++o)
            for (int k = 0; k < MLP_NFEAT; ++k)
                xs[(size_t)MLP_NFEAT * o + k] = (feat_h[(size_t)MLP_NFEAT * o + k] - mean[k]) / scale[k];
        HIP_CHECK(hipMemcpyAsync(d_feat.get(), xs.data(), sizeof(float) * xs.size(), hipMemcpyHostToDevice, stream));
        launch_mlp_forward(d_model.as<float>(), d_feat.as<float>(), nout, d_pred.as<float>(), stream);
        pred_h.resize(nout);
        HIP_CHECK(hipMemcpyAsync(pred_h.data(), d_pred.get(), sizeof(float) * nout, hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        const double lthr = std::log(screen_thr);
        for (int o = 0; o < nout; ++o)
            if ((double)pred_h[o] * tscale + tmean < lthr) skip_out[o] = 1;
    }
    // per-problem flags of every batch that writes or reads an output block
    double fl = 0, fls = 0, nsk = 0;
    auto flops_of = [](const GemmProb &p) { return 2.0 * p.M * p.N * (double)p.K; };
    for (auto &c : chunks) {
        for (size_t p = 0; p < c.xb.probs.size(); ++p) c.xb.skipv[p] = skip_out[c.x_owner[p]];
        for (size_t p = 0; p < c.ta.probs.size(); ++p) {
            unsigned char all = 1;
            for (int q = c.ta_ptr[p]; q < c.ta_ptr[p + 1] && all; ++q) all = skip_out[c.ta_outs[q]];
            c.ta.skipv[p] = all;
        }
        for (int k = 0; k < 4; ++k)
            for (size_t p = 0; p < c.fin[k].probs.size(); ++p) c.fin[k].skipv[p] = skip_out[c.fin_owner[k][p]];
        auto push = [&](GemmBatch &b) {
            for (size_t p = 0; p < b.probs.size(); ++p) {
                fl += flops_of(b.probs[p]);
                if (b.skipv[p]) fls += flops_of(b.probs[p]);
            }
            if (!b.probs.empty())
                HIP_CHECK(hipMemcpyAsync(b.d_skip.get(), b.skipv.data(), b.skipv.size(), hipMemcpyHostToDevice, stream));
            b.use_skip = true;
        };
        push(c.xb);
        push(c.ta);
        for (auto &f : c.fin) push(f);
    }
    for (int o = 0; o < nout; ++o) nsk += skip_out[o];
    scr_blocks += nout;
    scr_skipped += nsk;
    scr_flops += fl;
    scr_flops_skipped += fls;
    prof.end("chi0_screen", stream);
}

void RpaEngine::chi0_norms(std::vector<double> &out) {
    if (host_staged) LRA_THROW("chi0 block norms need the device-resident chi0 store (too large here)");
    const int nout = (int)outs.size();
    out.assign((size_t)ntau_done * nout, 0.0);
    if (out.empty()) return;
    DevBuf d((size_t)ntau_done * nout * sizeof(double));
    if (prec == Prec::F64)
        launch_chi0_block_norms<double>(d_chi0s.as<double>(), E_tot, ntau_done, d_eoff.as<long long>(),
                                        d_esize.as<int>(), nout, d.as<double>(), stream);
    else
        launch_chi0_block_norms<float>(d_chi0s.as<float>(), E_tot, ntau_done, d_eoff.as<long long>(),
                                       d_esize.as<int>(), nout, d.as<double>(), stream);
    HIP_CHECK(hipMemcpyAsync(out.data(), d.get(), sizeof(double) * out.size(), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
}

double mlp_train(const float *X, const float *t, int n, int iters, double lr, int seed, float *model) {
    if (n < 2) LRA_THROW("mlp_train: need at least two samples");
    if (iters < 1) LRA_THROW("mlp_train: iters must be positive");
    // feature / target standardisation (a constant feature keeps scale 1)
    std::vector<double> mean(MLP_NFEAT, 0.0), var(MLP_NFEAT, 0.0);
    double tm = 0, tv = 0;
    for (int s = 0; s < n; ++s) {
        for (int k = 0; k < MLP_NFEAT; ++k) mean[k] += X[(size_t)s * MLP_NFEAT + k];
        tm += t[s];
    }
    for (int k = 0; k < MLP_NFEAT; ++k) mean[k] /= n;
    tm /= n;
    for (int s = 0; s < n; ++s) {
        for (int k = 0; k < MLP_NFEAT; ++k) {
            const double d = X[(size_t)s * MLP_NFEAT + k] - mean[k];
            var[k] += d * d;
        }
        tv += (t[s] - tm) * (t[s] - tm);
    }
    std::vector<float> sc(MLP_NFEAT);
    for (int k = 0; k < MLP_NFEAT; ++k) {
        const double sd = std::sqrt(var[k] / n);
        sc[k] = sd > 1e-6 ? (float)sd : 1.0f;
    }
    const double tsd = std::sqrt(tv / n) > 1e-12 ? std::sqrt(tv / n) : 1.0;
    std::vector<float> xs((size_t)n * MLP_NFEAT), ts(n);
    for (int s = 0; s < n; ++s) {
        for (int k = 0; k < MLP_NFEAT; ++k)
            xs[(size_t)s * MLP_NFEAT + k] = (float)((X[(size_t)s * MLP_NFEAT + k] - mean[k]) / sc[k]);
        ts[s] = (float)((t[s] - tm) / tsd);
    }
    // fan-in scaled uniform weights, zero biases
    std::vector<float> p(MLP_NPARAM, 0.0f);
    std::mt19937_64 rng((unsigned long long)seed);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (int i = 0; i < MLP_NFEAT * MLP_H; ++i) p[i] = (float)(u(rng) / std::sqrt((double)MLP_NFEAT));
    const int ow2 = MLP_NFEAT * MLP_H + MLP_H, ow3 = ow2 + MLP_H * MLP_H + MLP_H;
    for (int i = 0; i < MLP_H * MLP_H; ++i) p[ow2 + i] = (float)(u(rng) / std::sqrt((double)MLP_H));
    for (int i = 0; i < MLP_H; ++i) p[ow3 + i] = (float)(u(rng) / std::sqrt((double)MLP_H));
    hipStream_t s = 0;
    DevBuf dX = upload(xs, s), dt = upload(ts, s), dp = upload(p, s);
    DevBuf dg(sizeof(float) * MLP_NPARAM), dm(sizeof(float) * MLP_NPARAM), dv(sizeof(float) * MLP_NPARAM),
        dl(sizeof(double));
    HIP_CHECK(hipMemsetAsync(dm.get(), 0, dm.bytes(), s));
    HIP_CHECK(hipMemsetAsync(dv.get(), 0, dv.bytes(), s));
    MlpWork w;
    for (int step = 1; step <= iters; ++step) {
        // learning rate decays geometrically to lr / 20 over the run
        const float lr_s = (float)(lr * std::pow(0.05, (double)(step - 1) / iters));
        mlp_loss_grad(dp.as<float>(), dX.as<float>(), dt.as<float>(), n, dg.as<float>(), dl.as<double>(), w, s);
        launch_adam(dp.as<float>(), dg.as<float>(), dm.as<float>(), dv.as<float>(), MLP_NPARAM, lr_s, step, s);
    }
    mlp_loss_grad(dp.as<float>(), dX.as<float>(), dt.as<float>(), n, dg.as<float>(), dl.as<double>(), w, s);
    double loss = 0;
    HIP_CHECK(hipMemcpy(&loss, dl.get(), sizeof(double), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(model, dp.get(), sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
    for (int k = 0; k < MLP_NFEAT; ++k) {
        model[MLP_NPARAM + k] = (float)mean[k];
        model[MLP_NPARAM + MLP_NFEAT + k] = sc[k];
    }
    model[MLP_NPARAM + 2 * MLP_NFEAT] = (float)tm;
    model[MLP_NPARAM + 2 * MLP_NFEAT + 1] = (float)tsd;
    return loss;
}

}  // namespace lra
