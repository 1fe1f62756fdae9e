// C-ABI implementation of LibRPA's public API (include/librpa_*.h) on top of
// the MI355X engines.  Semantics follow LibRPA src/api/*.cpp; each entry point
// names the reference function it replaces.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/librpa.h"
#include "dataset.h"
#include "dielec.h"
#include "gw.h"
#include "rpa.h"

using namespace lra;

namespace {

std::vector<std::shared_ptr<Dataset>> g_instances(1);  // slot 0 is a sentinel (instance_manager.cpp:15)
bool g_record_errors = false;
std::string g_last_error;
std::mutex g_mu;

Dataset &ds_of(LibrpaHandler *h) {
    if (!h) LRA_THROW("null handler");
    const int id = h->instance_id_;
    if (id <= 0 || id >= (int)g_instances.size() || !g_instances[id]) LRA_THROW("invalid handler");
    return *g_instances[id];
}

template <class F, class R>
R guarded(F &&f, R fallback) {
    if (!g_record_errors) return f();
    try {
        return f();
    } catch (const std::exception &e) {
        std::lock_guard<std::mutex> lk(g_mu);
        g_last_error = e.what();
        return fallback;
    }
}

template <class F>
void guarded_void(F &&f) {
    guarded([&]() { f(); return 0; }, 0);
}

void check_mf_dims(const Dataset &d) {
    if (!d.mf.dims_set) LRA_THROW("mean-field dimensions not set: call librpa_set_scf_dimension first");
}

void set_wfc_common(LibrpaHandler *h, int ispin, int ik, int nstates, int nbasis, const double *re, const double *im,
                    const double *ri) {
    auto &d = ds_of(h);
    check_mf_dims(d);
    auto &mf = d.mf;
    if (ispin < 0 || ispin >= mf.nspin || ik < 0 || ik >= mf.nk) LRA_THROW("spin or k index out of range");
    if (nstates != mf.nbands || nbasis != mf.nbasis)
        LRA_THROW("distributed eigenvectors are not supported: pass all states and basis functions");
    auto &w = mf.wfc[(size_t)ispin * mf.nk + ik];
    w.resize((size_t)nstates * nbasis);
    for (size_t i = 0; i < w.size(); ++i) w[i] = ri ? cplx(ri[2 * i], ri[2 * i + 1]) : cplx(re[i], im[i]);
    d.rpa.reset();
    d.gw.reset();
}

void set_vq_common(Dataset &d, std::map<std::tuple<int, int, int>, std::vector<cplx>> &store, int ik, int I, int J,
                   int nmu, int nnu, const double *re, const double *im, double thr) {
    std::vector<cplx> b((size_t)nmu * nnu);
    double mx = 0;
    for (size_t i = 0; i < b.size(); ++i) {
        b[i] = cplx(re[i], im[i]);
        mx = std::max(mx, std::fabs(re[i]));
    }
    if (mx >= thr) store[{ik, I, J}] = std::move(b);  // input.cpp:580-583
    d.rpa.reset();
    d.gw.reset();
}

void set_vq_2d(Dataset &d, std::map<std::tuple<int, int, int>, std::vector<cplx>> &store, int ik, int mu0, int mu1,
               int nu0, int nu1, const double *re, const double *im) {
    // scatter a dense [mu0, mu1) x [nu0, nu1) block into upper atom-pair blocks
    if (d.nb_aux.empty()) LRA_THROW("auxiliary basis not set");
    const int nat = (int)d.nb_aux.size();
    std::vector<int> off(nat + 1, 0);
    for (int a = 0; a < nat; ++a) off[a + 1] = off[a] + d.nb_aux[a];
    auto atom_of = [&](int g) { return (int)(std::upper_bound(off.begin(), off.end(), g) - off.begin()) - 1; };
    const int ncol = nu1 - nu0;
    for (int g = mu0; g < mu1; ++g)
        for (int h = nu0; h < nu1; ++h) {
            const int I = atom_of(g), J = atom_of(h);
            if (I > J) continue;
            auto &blk = store[{ik, I, J}];
            if (blk.empty()) blk.assign((size_t)d.nb_aux[I] * d.nb_aux[J], cplx(0, 0));
            const size_t idx = (size_t)(g - mu0) * ncol + (h - nu0);
            blk[(size_t)(g - off[I]) * d.nb_aux[J] + (h - off[J])] = cplx(re[idx], im[idx]);
        }
    d.rpa.reset();
    d.gw.reset();
}

void prepare_grid(Dataset &d, const LibrpaOptions &o) {
    // initialize_ds_tfgrids (src/api/dataset_helper.cpp:11-30)
    check_mf_dims(d);
    if (!d.mf.eig_set) LRA_THROW("eigenvalues / occupations not set");
    double emin_mf, emax_mf;
    e_min_max(d.mf, emin_mf, emax_mf);
    double emin = o.tfgrids_freq_min, emax = o.tfgrids_freq_max;
    if (o.tfgrids_type == Minimax || o.tfgrids_type == TFGRID_UNSET || o.tfgrids_type == GaussLegendre ||
        o.tfgrids_type == GaussChebyshevI || o.tfgrids_type == GaussChebyshevII) {
        emax = o.minimax_emax > 0 ? o.minimax_emax : emax_mf;
        emin = o.minimax_emin > 0 ? o.minimax_emin : emin_mf;
    }
    const std::vector<double> key = {(double)o.tfgrids_type, (double)o.nfreq, emin, emax, o.minimax_regulation,
                                     o.tfgrids_freq_min, o.tfgrids_freq_interval, o.tfgrids_time_min,
                                     o.tfgrids_time_interval};
    if (key == d.grid_key && d.grid.n == o.nfreq) return;
    d.grid = make_grid((int)o.tfgrids_type, o.nfreq, emin, emax, o.minimax_regulation, o.tfgrids_freq_min,
                       o.tfgrids_freq_interval, o.tfgrids_time_min, o.tfgrids_time_interval);
    d.grid_key = key;
}

std::string json_escape(const std::string &s) {
    std::string r;
    for (char c : s) { if (c == '"' || c == '\\') r += '\\'; r += c; }
    return r;
}

}  // namespace

namespace lra {
void lra_prepare_grid(Dataset &d, const LibrpaOptions &o) { prepare_grid(d, o); }
}  // namespace lra

extern "C" {

// ---------------------------------------------------------------- global (librpa_global.h)
const char *librpa_get_build_info(void) {
    return "librpa-mi355x: HIP/CDNA4 (gfx950) space-time chi0 -> RPA Ec and G0W0 Sigma; single GPU, single rank";
}
int librpa_get_major_version(void) { return 0; }
int librpa_get_minor_version(void) { return 9; }
int librpa_get_patch_version(void) { return 0; }

void librpa_init_global(LibrpaSwitch, const char *, LibrpaSwitch) {
    guarded_void([&]() {
        int n = 0;
        HIP_CHECK(hipGetDeviceCount(&n));
        if (n < 1) LRA_THROW("no GPU visible");
        HIP_CHECK(hipSetDevice(0));
        (void)mlp_noalloc_setup();  // address of the allocation-free entry's work array on device 0
    });
}

void librpa_init_global_fortran(int *, LibrpaSwitch a, const char *p, LibrpaSwitch b) { librpa_init_global(a, p, b); }

void librpa_finalize_global(void) {
    guarded_void([&]() {
        g_instances.assign(1, nullptr);  // instance_manager.cpp:65-70
        profiler().clear();
    });
}

void librpa_test(void) {
    guarded_void([&]() {
        GemmBatch b;
        b.add(make_prob(16, 16, 16, 0, 16, 0, 16, 0, 16));
        b.finalize(T64x64);
        std::vector<double> a(256, 1.0);
        DevBuf A = upload(a), C(256 * 8);
        gemm_run(Prec::F64, b, A.get(), A.get(), C.get(), 1.0, 0.0, 0);
        HIP_CHECK(hipMemcpy(a.data(), C.get(), 256 * 8, hipMemcpyDeviceToHost));
        for (double v : a)
            if (v != 16.0) LRA_THROW("librpa_test: device GEMM self-check failed");
    });
}

void librpa_print_profile(void) {
    guarded_void([&]() { printf("%s", profiler().report().c_str()); fflush(stdout); });
}

// ---------------------------------------------------------------- handler (librpa_handler.h)
LibrpaHandler *librpa_create_handler(MPI_Comm) {
    return guarded(
        [&]() -> LibrpaHandler * {
            g_instances.push_back(std::make_shared<Dataset>());
            const int id = (int)g_instances.size() - 1;
            auto *h = static_cast<LibrpaHandler *>(::operator new(sizeof(LibrpaHandler)));
            *const_cast<int *>(&h->instance_id_) = id;
            return h;
        },
        (LibrpaHandler *)nullptr);
}

LibrpaHandler *librpa_create_handler_fortran(int *f_comm) {
    return librpa_create_handler((MPI_Comm)(f_comm ? *f_comm : 0));
}

void librpa_destroy_handler(LibrpaHandler *h) {
    if (!h) return;
    guarded_void([&]() {
        const int id = h->instance_id_;
        if (id > 0 && id < (int)g_instances.size()) g_instances[id].reset();
        ::operator delete(h);
    });
}

// ---------------------------------------------------------------- options (librpa_options.h)
void librpa_set_output_dir(LibrpaOptions *opts, const char *output_dir) {
    std::string s(output_dir ? output_dir : ".");
    if (s.empty() || s.back() != '/') s += '/';
    std::strncpy(opts->output_dir, s.c_str(), LIBRPA_MAX_STRLEN - 1);
    opts->output_dir[LIBRPA_MAX_STRLEN - 1] = '\0';
}

void librpa_init_options(LibrpaOptions *o) {
    // defaults of src/api/options.cpp:11-67
    std::memset(o, 0, sizeof(*o));
    librpa_set_output_dir(o, ".");
    o->parallel_routing = AUTO;
    o->output_level = LIBRPA_VERBOSE_INFO;
    o->vq_threshold = 0.0;
    o->use_kpara_scf_eigvec = LIBRPA_SWITCH_OFF;
    o->tfgrids_type = TFGRID_UNSET;
    o->nfreq = 6;
    o->tfgrids_freq_min = 0.005;
    o->tfgrids_freq_interval = 0.0;
    o->tfgrids_freq_max = 1000.0;
    o->tfgrids_time_min = 0.005;
    o->tfgrids_time_interval = 0.0;
    o->minimax_emin = -1.0;
    o->minimax_emax = -1.0;
    o->minimax_regulation = 0.0;
    o->use_fullcoul_eps = LIBRPA_SWITCH_ON;
    o->use_fullcoul_exx = LIBRPA_SWITCH_OFF;
    o->use_fullcoul_wc = LIBRPA_SWITCH_OFF;
    o->n_bands_chi0 = -1;
    o->n_bands_sigc = -1;
    o->gf_threshold = 0.0;
    o->use_scalapack_ecrpa = LIBRPA_SWITCH_ON;
    o->use_shrink_abfs = LIBRPA_SWITCH_OFF;
    o->use_shrink_chi = LIBRPA_SWITCH_OFF;
    o->n_params_anacon = -1;
    o->use_scalapack_gw_wc = LIBRPA_SWITCH_ON;
    o->use_cholesky_gw_wc = LIBRPA_SWITCH_OFF;
    o->replace_w_head = LIBRPA_SWITCH_OFF;
    o->option_dielect_func = 0;
    o->use_2d_dielectric = LIBRPA_SWITCH_OFF;
    o->sqrt_coulomb_threshold = 0.0;
    o->load_sigc_from_file = LIBRPA_SWITCH_OFF;
    o->option_output_Wc_Rf_mat = 0;
}

// ---------------------------------------------------------------- input (librpa_input.h)
void librpa_set_scf_dimension(LibrpaHandler *h, int nspins, int nkpts, int nstates, int nbasis, int nspinor) {
    guarded_void([&]() {
        auto &d = ds_of(h);
        auto &mf = d.mf;
        mf.nspin = nspins; mf.nk = nkpts; mf.nbands = nstates; mf.nbasis = nbasis; mf.nspinor = nspinor;
        mf.eps.assign((size_t)nspins * nkpts * nstates, 0.0);
        mf.wg.assign(mf.eps.size(), 0.0);
        mf.wfc.assign((size_t)nspins * nkpts, {});
        mf.dims_set = true;
        d.rpa.reset();
        d.gw.reset();
    });
}

void librpa_set_wg_ekb_efermi(LibrpaHandler *h, int nspins, int nkpts, int nstates, const double *wg,
                              const double *ekb, double efermi) {
    guarded_void([&]() {
        auto &d = ds_of(h);
        check_mf_dims(d);
        auto &mf = d.mf;
        if (nspins != mf.nspin || nkpts != mf.nk || nstates != mf.nbands) LRA_THROW("dimension mismatch");
        const size_t n = (size_t)nspins * nkpts * nstates;
        for (size_t i = 0; i < n; ++i) { mf.eps[i] = ekb[i]; mf.wg[i] = wg[i] / nkpts; }  // input.cpp:78-84
        mf.efermi = efermi;
        mf.eig_set = true;
        d.rpa.reset();
        d.gw.reset();
    });
}

void librpa_set_wfc(LibrpaHandler *h, int ispin, int ik, int nstates_local, int nbasis_local, const double *wfc_real,
                    const double *wfc_imag) {
    guarded_void([&]() { set_wfc_common(h, ispin, ik, nstates_local, nbasis_local, wfc_real, wfc_imag, nullptr); });
}

void librpa_set_wfc_packed(LibrpaHandler *h, int ispin, int ik, int nstates_local, int nbasis_local,
                           const double *wfc_ri) {
    guarded_void([&]() { set_wfc_common(h, ispin, ik, nstates_local, nbasis_local, nullptr, nullptr, wfc_ri); });
}

void librpa_set_wfc_spinor(LibrpaHandler *h, int, int, int, const double *, const double *, const double *,
                           const double *) {
    guarded_void([&]() { (void)ds_of(h); LRA_THROW("spinor eigenvectors (SOC) are outside this build's path"); });
}

void librpa_set_wfc_spinor_packed(LibrpaHandler *h, int, int, int, const double *, const double *) {
    guarded_void([&]() { (void)ds_of(h); LRA_THROW("spinor eigenvectors (SOC) are outside this build's path"); });
}

void librpa_set_ao_basis_wfc(LibrpaHandler *h, int natoms, const size_t *nbs) {
    guarded_void([&]() {
        auto &d = ds_of(h);
        d.nb_wfc.assign(nbs, nbs + natoms);
        d.rpa.reset();
        d.gw.reset();
    });
}

void librpa_set_ao_basis_aux(LibrpaHandler *h, int natoms, const size_t *nbs) {
    guarded_void([&]() {
        auto &d = ds_of(h);
        d.nb_aux.assign(nbs, nbs + natoms);
        d.rpa.reset();
        d.gw.reset();
    });
}

void librpa_set_latvec_and_G(LibrpaHandler *h, const double lat_mat[9], const double G_mat[9]) {
    guarded_void([&]() {
        auto &d = ds_of(h);
        std::memcpy(d.latvec, lat_mat, sizeof(d.latvec));
        std::memcpy(d.Gvec, G_mat, sizeof(d.Gvec));
        d.lat_set = true;
    });
}

void librpa_set_atoms(LibrpaHandler *h, int natoms, const int *types, const double *posi_cart) {
    guarded_void([&]() {
        auto &d = ds_of(h);
        d.types.assign(types, types + natoms);
        d.pos.assign(posi_cart, posi_cart + 3 * (size_t)natoms);
        // fractional coordinates only when the lattice is already known (input.cpp:344-347,
        // geometry.cpp:108-126): frac = cart * latvec^{-1}
        d.frac.clear();
        if (d.lat_set) {
            const double *a = d.latvec;
            const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                               a[2] * (a[3] * a[7] - a[4] * a[6]);
            if (std::fabs(det) < 1e-300) LRA_THROW("singular lattice vectors");
            const double inv[9] = {(a[4] * a[8] - a[5] * a[7]) / det, (a[2] * a[7] - a[1] * a[8]) / det,
                                   (a[1] * a[5] - a[2] * a[4]) / det, (a[5] * a[6] - a[3] * a[8]) / det,
                                   (a[0] * a[8] - a[2] * a[6]) / det, (a[2] * a[3] - a[0] * a[5]) / det,
                                   (a[3] * a[7] - a[4] * a[6]) / det, (a[1] * a[6] - a[0] * a[7]) / det,
                                   (a[0] * a[4] - a[1] * a[3]) / det};
            d.frac.assign(3 * (size_t)natoms, 0.0);
            for (int i = 0; i < natoms; ++i)
                for (int c = 0; c < 3; ++c)
                    for (int r = 0; r < 3; ++r) d.frac[3 * (size_t)i + c] += posi_cart[3 * i + r] * inv[3 * r + c];
        }
        d.gw.reset();
    });
}

void librpa_set_kgrids_kvec(LibrpaHandler *h, int nk1, int nk2, int nk3, const double *kvecs) {
    guarded_void([&]() {
        auto &d = ds_of(h);
        if (!d.lat_set) LRA_THROW("lattice vectors must be set before the k grid");
        d.period[0] = nk1; d.period[1] = nk2; d.period[2] = nk3;
        const int n = nk1 * nk2 * nk3;
        d.kcart.assign(kvecs, kvecs + 3 * (size_t)n);
        d.kfrac.assign(3 * (size_t)n, 0.0);
        // kfrac = latvec * k / 2pi (pbc.cpp:79-102), -0 cleaned to 0
        for (int k = 0; k < n; ++k)
            for (int a = 0; a < 3; ++a) {
                double v = 0;
                for (int b = 0; b < 3; ++b) v += d.latvec[3 * a + b] * kvecs[3 * k + b] / (2 * M_PI);
                if (std::fabs(v) < 1e-8) v = 0.0;
                d.kfrac[3 * k + a] = v;
            }
        d.ibz_map.resize(n);
        for (int k = 0; k < n; ++k) d.ibz_map[k] = k;
        d.rpa.reset();
        d.gw.reset();
    });
}

void librpa_set_ibz_mapping(LibrpaHandler *h, int nkpts, const int *map_ibzk) {
    guarded_void([&]() {
        auto &d = ds_of(h);
        if (nkpts != (int)d.kfrac.size() / 3) LRA_THROW("irk_point_id_mapping size not equal to n_k_points");
        d.ibz_map.assign(map_ibzk, map_ibzk + nkpts);
    });
}

void librpa_set_lri_coeff(LibrpaHandler *h, LibrpaParallelRouting, int I, int J, int nbasis_i, int nbasis_j,
                          int naux_mu, const int R[3], const double *Cs_in) {
    guarded_void([&]() {
        auto &d = ds_of(h);
        if (d.nb_wfc.empty()) LRA_THROW("wave function basis not set, call (librpa_)set_ao_basis_wfc first");
        if (d.nb_aux.empty()) LRA_THROW("auxiliary basis not set, call (librpa_)set_ao_basis_aux first");
        if (d.nb_aux[I] != naux_mu || d.nb_wfc[I] != nbasis_i || d.nb_wfc[J] != nbasis_j)
            LRA_THROW("Cs block dimensions do not match the basis sizes");
        const size_t n = (size_t)nbasis_i * nbasis_j * naux_mu;
        d.cs[{I, J, R[0], R[1], R[2]}] = std::vector<double>(Cs_in, Cs_in + n);
        d.rpa.reset();
        d.gw.reset();
    });
}

void librpa_set_aux_bare_coulomb_k_atom_pair(LibrpaHandler *h, int ik, int I, int J, int naux_mu, int naux_nu,
                                             const double *re, const double *im, double thr) {
    guarded_void([&]() { auto &d = ds_of(h); set_vq_common(d, d.vq, ik, I, J, naux_mu, naux_nu, re, im, thr); });
}

void librpa_set_aux_cut_coulomb_k_atom_pair(LibrpaHandler *h, int ik, int I, int J, int naux_mu, int naux_nu,
                                            const double *re, const double *im, double thr) {
    guarded_void([&]() { auto &d = ds_of(h); set_vq_common(d, d.vq_cut, ik, I, J, naux_mu, naux_nu, re, im, thr); });
}

void librpa_set_aux_bare_coulomb_k_2d_block(LibrpaHandler *h, int ik, int mu_begin, int mu_end, int nu_begin,
                                            int nu_end, const double *re, const double *im) {
    guarded_void([&]() { auto &d = ds_of(h); set_vq_2d(d, d.vq, ik, mu_begin, mu_end, nu_begin, nu_end, re, im); });
}

void librpa_set_aux_cut_coulomb_k_2d_block(LibrpaHandler *h, int ik, int mu_begin, int mu_end, int nu_begin,
                                           int nu_end, const double *re, const double *im) {
    guarded_void([&]() { auto &d = ds_of(h); set_vq_2d(d, d.vq_cut, ik, mu_begin, mu_end, nu_begin, nu_end, re, im); });
}

void librpa_set_dielect_func_imagfreq(LibrpaHandler *h, int nfreq, const double *om, const double *df) {
    guarded_void([&]() {
        auto &d = ds_of(h);
        d.dielect_omega.assign(om, om + nfreq);
        d.dielect_val.assign(df, df + nfreq);
    });
}

void librpa_set_band_kvec(LibrpaHandler *h, int n, const double *kfrac) {
    guarded_void([&]() {
        auto &d = ds_of(h);
        d.band.nk = n;
        d.band.kfrac.assign(kfrac, kfrac + 3 * (size_t)n);
    });
}

void librpa_set_band_occ_eigval(LibrpaHandler *h, int ns, int nk, int nst, const double *occ, const double *eig) {
    guarded_void([&]() {
        auto &d = ds_of(h);
        d.band.nspin = ns; d.band.nk = nk; d.band.nstates = nst;
        d.band.occ.assign(occ, occ + (size_t)ns * nk * nst);
        d.band.eig.assign(eig, eig + (size_t)ns * nk * nst);
        d.band.wfc.assign((size_t)ns * nk, {});
    });
}

void librpa_set_wfc_band(LibrpaHandler *h, int ispin, int ik, int nst, int nbas, const double *re, const double *im) {
    guarded_void([&]() {
        auto &d = ds_of(h);
        if (d.band.wfc.empty()) LRA_THROW("set band occupations / eigenvalues first");
        auto &w = d.band.wfc[(size_t)ispin * d.band.nk + ik];
        w.resize((size_t)nst * nbas);
        for (size_t i = 0; i < w.size(); ++i) w[i] = cplx(re[i], im[i]);
    });
}

void librpa_set_wfc_band_packed(LibrpaHandler *h, int ispin, int ik, int nst, int nbas, const double *ri) {
    guarded_void([&]() {
        auto &d = ds_of(h);
        if (d.band.wfc.empty()) LRA_THROW("set band occupations / eigenvalues first");
        auto &w = d.band.wfc[(size_t)ispin * d.band.nk + ik];
        w.resize((size_t)nst * nbas);
        for (size_t i = 0; i < w.size(); ++i) w[i] = cplx(ri[2 * i], ri[2 * i + 1]);
    });
}

void librpa_set_wfc_band_spinor(LibrpaHandler *h, int, int, int, const double *, const double *, const double *,
                                const double *) {
    guarded_void([&]() { (void)ds_of(h); LRA_THROW("spinor eigenvectors (SOC) are outside this build's path"); });
}

void librpa_set_wfc_band_spinor_packed(LibrpaHandler *h, int, int, int, const double *, const double *) {
    guarded_void([&]() { (void)ds_of(h); LRA_THROW("spinor eigenvectors (SOC) are outside this build's path"); });
}

void librpa_reset_band_data(LibrpaHandler *h) {
    guarded_void([&]() { ds_of(h).band = BandData(); });
}

// ---------------------------------------------------------------- compute (librpa_compute.h)
void librpa_get_imaginary_frequency_grids(LibrpaHandler *h, const LibrpaOptions *p_opts, double *omegas,
                                          double *weights) {
    guarded_void([&]() {
        auto &d = ds_of(h);
        prepare_grid(d, *p_opts);
        std::memcpy(omegas, d.grid.omega.data(), sizeof(double) * p_opts->nfreq);
        std::memcpy(weights, d.grid.omega_w.data(), sizeof(double) * p_opts->nfreq);
    });
}

double librpa_get_rpa_correlation_energy(LibrpaHandler *h, const LibrpaOptions *p_opts, int n_ibz_kpoints,
                                         double *re, double *im) {
    return guarded(
        [&]() -> double {
            auto &d = ds_of(h);
            const auto &o = *p_opts;
            if (o.use_shrink_abfs == LIBRPA_SWITCH_ON) LRA_THROW("ABF shrinking is outside this build's path");
            prepare_grid(d, o);
            // the tau-invariant plan (packed Cs, problem lists, G-build operands) is reused
            // until a setter changes the inputs
            auto eng = d.rpa;
            if (!eng || eng->prec != d.chi0_prec || eng->mixed != d.chi0_mixed || eng->planned_ntau != d.grid.n) {
                d.rpa.reset();
                d.gw.reset();
                eng = std::make_shared<RpaEngine>();
                eng->mixed = d.chi0_mixed;
                eng->verbose = o.output_level >= LIBRPA_VERBOSE_DEBUG;
                eng->plan(d, d.chi0_prec, d.ec_prec, d.grid.n);
            }
            eng->ec_prec = d.ec_prec;
            eng->mixed_low = d.chi0_mixed_low;
            eng->verbose = o.output_level >= LIBRPA_VERBOSE_DEBUG;
            eng->screen_mode = d.screen_mode;
            eng->screen_thr = d.screen_thr;
            eng->screen_model = d.screen_model;
            // training samples need the chi0 store on the device (not kept when it is host-staged)
            eng->collect = d.keep_chi0 && d.screen_mode == 0 && !eng->host_staged;
            eng->build_chi0(d, d.grid, o.n_bands_chi0);
            auto r = eng->ec(d, d.grid);
            if ((int)r.qcontrib.size() != n_ibz_kpoints && o.output_level >= LIBRPA_VERBOSE_WARN)
                fprintf(stderr, "WARNING: parsed n_ibz_kpoints is not consistent with generated CorrEne object: %d != %zu\n",
                        n_ibz_kpoints, r.qcontrib.size());
            for (size_t q = 0; q < r.qcontrib.size() && (int)q < n_ibz_kpoints; ++q) {
                re[q] = r.qcontrib[q].real();
                im[q] = r.qcontrib[q].imag();
            }
            d.stats[0] = eng->flops_chi0_alg * d.mf.nspin;
            d.stats[1] = eng->flops_chi0_exec * d.mf.nspin;
            d.stats[2] = eng->t_chi0;
            d.stats[3] = eng->t_ec;
            d.stats[4] = (double)eng->outs.size();
            d.stats[5] = (double)eng->n_cs_blocks;
            d.stats[6] = eng->NABF;
            d.stats[7] = eng->NAO;
            d.stats[8] = eng->mixed_amax;
            d.stats[9] = eng->n_lowp_tau;
            d.rpa = eng;
            return r.ec_re;
        },
        std::nan(""));
}

void librpa_build_exx(LibrpaHandler *h, const LibrpaOptions *) {
    guarded_void([&]() { (void)ds_of(h); LRA_THROW("exact exchange is outside this build's path (pass vexx to the G0W0 calls)"); });
}

void librpa_get_exx_pot_kgrid(LibrpaHandler *h, const LibrpaOptions *, const int, const int, const int *, int, int,
                              double *) {
    guarded_void([&]() { (void)ds_of(h); LRA_THROW("exact exchange is outside this build's path"); });
}

void librpa_get_exx_pot_band_k(LibrpaHandler *h, const LibrpaOptions *, const int, const int, const int *, int, int,
                               double *) {
    guarded_void([&]() { (void)ds_of(h); LRA_THROW("exact exchange is outside this build's path"); });
}

void librpa_build_g0w0_sigma(LibrpaHandler *h, const LibrpaOptions *p_opts) {
    guarded_void([&]() { gw_build_sigma(ds_of(h), *p_opts); });
}

void librpa_get_g0w0_sigc_kgrid(LibrpaHandler *h, const LibrpaOptions *p_opts, const int n_spins, const int n_kpts_this,
                                const int *iks_this, int i_state_low, int i_state_high, const double *vxc,
                                const double *vexx, double *sigc_re, double *sigc_im) {
    guarded_void([&]() {
        gw_sigc_kgrid(ds_of(h), *p_opts, n_spins, n_kpts_this, iks_this, i_state_low, i_state_high, vxc, vexx,
                      sigc_re, sigc_im);
    });
}

void librpa_get_g0w0_sigc_band_k(LibrpaHandler *h, const LibrpaOptions *, const int, const int, const int *, int, int,
                                 const double *, const double *, double *, double *) {
    guarded_void([&]() { (void)ds_of(h); LRA_THROW("band-path self-energies are outside this build's path"); });
}

// ---------------------------------------------------------------- extensions (librpa_amd.h)
void librpa_amd_set_precision(LibrpaHandler *h, int chi0_prec, int ec_prec) {
    guarded_void([&]() {
        auto &d = ds_of(h);
        if (chi0_prec < 0 || chi0_prec > 4 || ec_prec < 0 || ec_prec > 2) LRA_THROW("unknown precision");
        d.chi0_mixed = chi0_prec >= 3;
        d.chi0_mixed_low = chi0_prec == 4 ? Prec::BF16X3 : Prec::F32;
        d.chi0_prec = d.chi0_mixed ? Prec::F64 : static_cast<Prec>(chi0_prec);
        d.ec_prec = static_cast<Prec>(ec_prec);
    });
}

void librpa_amd_set_error_mode(int record) { g_record_errors = record != 0; }

int librpa_amd_get_last_error(char *buf, int len) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_last_error.empty()) return 0;
    if (buf && len > 0) {
        std::strncpy(buf, g_last_error.c_str(), len - 1);
        buf[len - 1] = '\0';
    }
    g_last_error.clear();
    return 1;
}

void librpa_amd_set_screening(LibrpaHandler *h, int mode, double threshold, const float *model) {
    guarded_void([&]() {
        auto &d = ds_of(h);
        if (mode < 0 || mode > 2) LRA_THROW("screening mode must be 0, 1 or 2");
        if (mode > 0 && !(threshold > 0)) LRA_THROW("screening threshold must be positive");
        if (mode == 2 && !model) LRA_THROW("learned screening needs a model");
        d.screen_mode = mode;
        d.screen_thr = threshold;
        if (mode == 2) d.screen_model.assign(model, model + MLP_NMODEL);
        else d.screen_model.clear();
    });
}

int librpa_amd_get_screen_stats(LibrpaHandler *h, double *out) {
    return guarded(
        [&]() -> int {
            auto &d = ds_of(h);
            if (!d.rpa) LRA_THROW("no chi0 build yet");
            out[0] = d.rpa->scr_blocks;
            out[1] = d.rpa->scr_skipped;
            out[2] = d.rpa->scr_flops;
            out[3] = d.rpa->scr_flops_skipped;
            return 4;
        },
        -1);
}

int librpa_amd_get_screen_samples(LibrpaHandler *h, float *X, float *t) {
    return guarded(
        [&]() -> int {
            auto &d = ds_of(h);
            if (d.rpa && d.rpa->host_staged)
                LRA_THROW("no samples: this system's chi0 store is host-staged (too large for the device)");
            if (!d.rpa || d.rpa->features.empty())
                LRA_THROW("no samples: build chi0 with librpa_amd_set_keep_chi0(h, 1) and screening off");
            auto &e = *d.rpa;
            std::vector<double> nrm;
            e.chi0_norms(nrm);
            const size_t nout = e.outs.size();
            int n = 0;
            for (size_t i = 0; i < nrm.size(); ++i) {
                if (!(nrm[i] > 0)) continue;
                if (X) std::memcpy(X + (size_t)n * MLP_NFEAT, &e.features[i * MLP_NFEAT], sizeof(float) * MLP_NFEAT);
                if (t) t[n] = (float)std::log(nrm[i]);
                ++n;
            }
            (void)nout;
            return n;
        },
        -1);
}

double librpa_amd_screen_train(const float *X, const float *t, int n, int iters, double lr, int seed, float *model) {
    return guarded([&]() -> double { return mlp_train(X, t, n, iters, lr, seed, model); }, std::nan(""));
}

int librpa_amd_mlp_predict(const float *params, const float *X, int n, float *y) {
    return guarded(
        [&]() -> int {
            std::vector<float> p(params, params + MLP_NPARAM), x(X, X + (size_t)n * MLP_NFEAT);
            DevBuf dp = upload(p), dx = upload(x), dy(sizeof(float) * std::max(1, n));
            launch_mlp_forward(dp.as<float>(), dx.as<float>(), n, dy.as<float>(), 0);
            HIP_CHECK(hipMemcpy(y, dy.get(), sizeof(float) * n, hipMemcpyDeviceToHost));
            return 0;
        },
        1);
}

int librpa_amd_mlp_loss_grad(const float *params, const float *X, const float *t, int n, float *grad, double *loss) {
    return guarded(
        [&]() -> int {
            std::vector<float> p(params, params + MLP_NPARAM), x(X, X + (size_t)n * MLP_NFEAT), tt(t, t + n);
            DevBuf dp = upload(p), dx = upload(x), dt = upload(tt), dg(sizeof(float) * MLP_NPARAM), dl(sizeof(double));
            MlpWork w;
            mlp_loss_grad(dp.as<float>(), dx.as<float>(), dt.as<float>(), n, dg.as<float>(), dl.as<double>(), w, 0);
            HIP_CHECK(hipMemcpy(grad, dg.get(), sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
            HIP_CHECK(hipMemcpy(loss, dl.get(), sizeof(double), hipMemcpyDeviceToHost));
            return 0;
        },
        1);
}

int librpa_amd_mlp_loss_grad_noalloc(const float *d_params, const float *d_X, const float *d_t, int n, float *d_grad,
                                     double *d_loss, void *stream) {
    return guarded(
        [&]() -> int {
            mlp_loss_grad_noalloc(d_params, d_X, d_t, n, d_grad, d_loss, static_cast<hipStream_t>(stream));
            return 0;
        },
        1);
}

void librpa_amd_set_keep_chi0(LibrpaHandler *h, int keep) {
    guarded_void([&]() { ds_of(h).keep_chi0 = keep != 0; });
}

int librpa_amd_get_chi0s_block(LibrpaHandler *h, int itau, int I, int J, const int R[3], double *out) {
    return guarded(
        [&]() -> int {
            auto &d = ds_of(h);
            if (!d.rpa) LRA_THROW("no retained chi0: call librpa_amd_set_keep_chi0(h, 1) before the Ec call");
            d.rpa->chi0s_block(itau, I, J, d.rpa->fold_index(R[0], R[1], R[2]), out);
            return 0;
        },
        1);
}

int librpa_amd_get_wc_block(LibrpaHandler *h, int itau, int I, int J, const int R[3], double *out) {
    return guarded(
        [&]() -> int {
            auto &d = ds_of(h);
            if (!d.gw) LRA_THROW("no G0W0 data: call librpa_build_g0w0_sigma first");
            auto &g = *d.gw;
            if (itau < 0 || itau >= g.ntau || I < 0 || J < 0 || I >= g.natom || J >= g.natom)
                LRA_THROW("index out of range");
            g.w_block(itau, I, J, g.rpa->fold_index(R[0], R[1], R[2]), out);
            return 0;
        },
        1);
}

int librpa_amd_get_sigma_block(LibrpaHandler *h, int ispin, int ifreq, int I, int J, const int R[3], double *re,
                               double *im) {
    return guarded(
        [&]() -> int {
            auto &d = ds_of(h);
            if (!d.gw) LRA_THROW("no G0W0 data: call librpa_build_g0w0_sigma first");
            auto &g = *d.gw;
            if (ispin < 0 || ispin >= g.nspin || ifreq < 0 || ifreq >= g.ntau || I < 0 || J < 0 || I >= g.natom ||
                J >= g.natom)
                LRA_THROW("index out of range");
            g.sigma_block(ispin, ifreq, I, J, g.rpa->fold_index(R[0], R[1], R[2]), re, im);
            return 0;
        },
        1);
}

int librpa_amd_get_sigc_ks(LibrpaHandler *h, int ispin, int ik, int i_state, double *re, double *im) {
    return guarded(
        [&]() -> int {
            auto &d = ds_of(h);
            if (!d.gw) LRA_THROW("no G0W0 data: call librpa_build_g0w0_sigma first");
            auto &g = *d.gw;
            if (ispin < 0 || ispin >= g.nspin || ik < 0 || ik >= g.nk || i_state < 0 || i_state >= d.mf.nbands)
                LRA_THROW("index out of range");
            auto v = g.ks_diag(d, ispin, ik, {i_state});
            for (size_t w = 0; w < v[0].size(); ++w) {
                re[w] = v[0][w].real();
                im[w] = v[0][w].imag();
            }
            return (int)v[0].size();
        },
        -1);
}

int librpa_amd_get_chi0q_block(LibrpaHandler *h, int ifreq, int iq, int I, int J, double *re, double *im) {
    return guarded(
        [&]() -> int {
            auto &d = ds_of(h);
            if (!d.rpa) LRA_THROW("no retained chi0: call librpa_amd_set_keep_chi0(h, 1) before the Ec call");
            d.rpa->chi0q_block(d, d.grid, ifreq, iq, I, J, re, im);
            return 0;
        },
        1);
}

int librpa_amd_get_tfgrids(LibrpaHandler *h, const LibrpaOptions *p_opts, double *tau, double *tau_w, double *omega,
                           double *omega_w, double *cos_t2f, double *cos_f2t, double *sin_t2f) {
    return guarded(
        [&]() -> int {
            auto &d = ds_of(h);
            prepare_grid(d, *p_opts);
            const auto &g = d.grid;
            const size_t n = g.n;
            if (tau) std::memcpy(tau, g.tau.data(), n * 8);
            if (tau_w) std::memcpy(tau_w, g.tau_w.data(), n * 8);
            if (omega) std::memcpy(omega, g.omega.data(), n * 8);
            if (omega_w) std::memcpy(omega_w, g.omega_w.data(), n * 8);
            if (cos_t2f) std::memcpy(cos_t2f, g.cos_t2f.data(), n * n * 8);
            if (cos_f2t) std::memcpy(cos_f2t, g.cos_f2t.data(), n * n * 8);
            if (sin_t2f) std::memcpy(sin_t2f, g.sin_t2f.data(), n * n * 8);
            return g.n;
        },
        -1);
}

int librpa_amd_interpolate_dielec_func(int option, int n_in, const double *omega_in, const double *df_in,
                                       int n_out, const double *omega_out, double *df_out) {
    return guarded(
        [&]() -> int {
            if (n_in < 0 || n_out < 0) LRA_THROW("dielectric function: negative size");
            const std::vector<double> om(omega_in, omega_in + n_in), df(df_in, df_in + n_in);
            const std::vector<double> tgt(omega_out, omega_out + n_out);
            const auto v = lra::interpolate_dielec(option, om, df, tgt);
            std::memcpy(df_out, v.data(), v.size() * sizeof(double));
            return n_out;
        },
        -1);
}

int librpa_amd_profile_json(char *buf, int len) {
    return guarded(
        [&]() -> int {
            std::ostringstream os;
            os << "{";
            bool first = true;
            for (const auto &kv : profiler().records()) {
                if (!first) os << ",";
                first = false;
                os << "\"" << json_escape(kv.first) << "\":[" << kv.second.ms << "," << kv.second.calls << ","
                   << kv.second.flops << "," << kv.second.bytes << "]";
            }
            os << "}";
            const std::string s = os.str();
            if (buf && len > 0) {
                std::strncpy(buf, s.c_str(), len - 1);
                buf[len - 1] = '\0';
            }
            return (int)s.size();
        },
        -1);
}

void librpa_amd_profile_reset(void) {
    guarded_void([&]() { profiler().clear(); });
}

int librpa_amd_gemm(int prec, int trans_a, int trans_b, int M, int N, int K, double alpha, const double *A, int lda,
                    const double *B, int ldb, double beta, double *Cm, int ldc, int tile_cfg) {
    return guarded(
        [&]() -> int {
            if (tile_cfg < 0 || tile_cfg >= NUM_TILE_CFG) LRA_THROW("bad tile configuration");
            GemmBatch b;
            b.trans_a = trans_a != 0;
            b.trans_b = trans_b != 0;
            b.add(make_prob(M, N, K, 0, lda, 0, ldb, 0, ldc));
            b.finalize(tile_cfg);
            const size_t na = (size_t)(trans_a ? K : M) * lda, nb = (size_t)(trans_b ? N : K) * ldb,
                         nc = (size_t)M * ldc;
            if (prec == 0) {
                DevBuf dA(na * 8), dB(nb * 8), dC(nc * 8);
                HIP_CHECK(hipMemcpy(dA.get(), A, na * 8, hipMemcpyHostToDevice));
                HIP_CHECK(hipMemcpy(dB.get(), B, nb * 8, hipMemcpyHostToDevice));
                HIP_CHECK(hipMemcpy(dC.get(), Cm, nc * 8, hipMemcpyHostToDevice));
                gemm_run(Prec::F64, b, dA.get(), dB.get(), dC.get(), alpha, beta, 0);
                HIP_CHECK(hipMemcpy(Cm, dC.get(), nc * 8, hipMemcpyDeviceToHost));
            } else {
                std::vector<float> a(A, A + na), bb(B, B + nb), c(Cm, Cm + nc);
                DevBuf dA = upload(a), dB = upload(bb), dC = upload(c);
                gemm_run(static_cast<Prec>(prec), b, dA.get(), dB.get(), dC.get(), alpha, beta, 0);
                HIP_CHECK(hipMemcpy(c.data(), dC.get(), nc * 4, hipMemcpyDeviceToHost));
                for (size_t i = 0; i < nc; ++i) Cm[i] = c[i];
            }
            return 0;
        },
        1);
}

int librpa_amd_get_stats(LibrpaHandler *h, double *out, int n) {
    return guarded(
        [&]() -> int {
            auto &d = ds_of(h);
            const int m = std::min(n, 10);
            for (int i = 0; i < m; ++i) out[i] = d.stats[i];
            return m;
        },
        -1);
}

}  // extern "C"
