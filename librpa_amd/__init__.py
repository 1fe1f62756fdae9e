"""Python host mirror of LibRPA's C API on the MI355X library.

``librpa_mi355x.so`` (built by ``librpa_amd.build``) exports LibRPA's C
functions (include/librpa_*.h) and this build's extensions
(include/librpa_amd.h).  This module binds them with ctypes and mirrors the
reference's C++ wrapper ``librpa::Handler`` / ``librpa::Options``
(LibRPA include/librpa.hpp:57-381): same method names, argument meaning and
error behaviour (errors surface as exceptions).

There is no CPU fallback: if the library is missing, ``load()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os
import sys as _sys
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librpa_mi355x.so")
_lib = None

FP64, FP32, BF16X3, MIXED, MIXED_BF16 = 0, 1, 2, 3, 4
PREC = {"fp64": FP64, "fp32": FP32, "bf16x3": BF16X3, "mixed": MIXED, "mixed-bf16": MIXED_BF16}

# LibrpaTimeFreqGrid / LibrpaParallelRouting values (librpa_enums.h)
MINIMAX = 3
RTAU = 0
ATOMPAIR = 1
LIBRI = 2
AUTO = -51
UNSET = -101


class LibrpaError(RuntimeError):
    pass


class Options(C.Structure):
    """LibrpaOptions, field order of LibRPA include/librpa_options.h:33-212."""
    _fields_ = [
        ("output_dir", C.c_char * 200),
        ("parallel_routing", C.c_int),
        ("output_level", C.c_int),
        ("vq_threshold", C.c_double),
        ("use_kpara_scf_eigvec", C.c_int),
        ("tfgrids_type", C.c_int),
        ("nfreq", C.c_int),
        ("tfgrids_freq_min", C.c_double),
        ("tfgrids_freq_interval", C.c_double),
        ("tfgrids_freq_max", C.c_double),
        ("tfgrids_time_min", C.c_double),
        ("tfgrids_time_interval", C.c_double),
        ("minimax_emin", C.c_double),
        ("minimax_emax", C.c_double),
        ("minimax_regulation", C.c_double),
        ("use_fullcoul_exx", C.c_int),
        ("use_fullcoul_eps", C.c_int),
        ("use_fullcoul_wc", C.c_int),
        ("n_bands_chi0", C.c_int),
        ("n_bands_sigc", C.c_int),
        ("gf_threshold", C.c_double),
        ("use_scalapack_ecrpa", C.c_int),
        ("use_shrink_abfs", C.c_int),
        ("use_shrink_chi", C.c_int),
        ("n_params_anacon", C.c_int),
        ("use_scalapack_gw_wc", C.c_int),
        ("use_cholesky_gw_wc", C.c_int),
        ("replace_w_head", C.c_int),
        ("option_dielect_func", C.c_int),
        ("use_2d_dielectric", C.c_int),
        ("load_sigc_from_file", C.c_int),
        ("sqrt_coulomb_threshold", C.c_double),
        ("libri_chi0_threshold_C", C.c_double),
        ("libri_chi0_threshold_G", C.c_double),
        ("libri_exx_threshold_C", C.c_double),
        ("libri_exx_threshold_D", C.c_double),
        ("libri_exx_threshold_V", C.c_double),
        ("libri_g0w0_threshold_C", C.c_double),
        ("libri_g0w0_threshold_G", C.c_double),
        ("libri_g0w0_threshold_Wc", C.c_double),
        ("output_gw_sigc_mat", C.c_int),
        ("output_gw_sigc_mat_rt", C.c_int),
        ("output_gw_sigc_mat_rf", C.c_int),
        ("option_output_Wc_Rf_mat", C.c_int),
    ]

    def __init__(self, **kw):
        super().__init__()
        load().librpa_init_options(C.byref(self))
        for k, v in kw.items():
            setattr(self, k, v)

    def set_output_dir(self, d: str):
        load().librpa_set_output_dir(C.byref(self), d.encode())


class _Handler(C.Structure):
    _fields_ = [("instance_id_", C.c_int)]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_hp = C.POINTER(_Handler)
_op = C.POINTER(Options)
_fp = C.POINTER(C.c_float)

_SIGS = {
    "librpa_get_build_info": (C.c_char_p, []),
    "librpa_get_major_version": (C.c_int, []),
    "librpa_get_minor_version": (C.c_int, []),
    "librpa_get_patch_version": (C.c_int, []),
    "librpa_init_global": (None, [C.c_int, C.c_char_p, C.c_int]),
    "librpa_finalize_global": (None, []),
    "librpa_test": (None, []),
    "librpa_print_profile": (None, []),
    "librpa_create_handler": (_hp, [C.c_int]),
    "librpa_destroy_handler": (None, [_hp]),
    "librpa_init_options": (None, [_op]),
    "librpa_set_output_dir": (None, [_op, C.c_char_p]),
    "librpa_set_scf_dimension": (None, [_hp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "librpa_set_wg_ekb_efermi": (None, [_hp, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_double]),
    "librpa_set_wfc": (None, [_hp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]),
    "librpa_set_wfc_packed": (None, [_hp, C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
    "librpa_set_ao_basis_wfc": (None, [_hp, C.c_int, C.POINTER(C.c_size_t)]),
    "librpa_set_ao_basis_aux": (None, [_hp, C.c_int, C.POINTER(C.c_size_t)]),
    "librpa_set_latvec_and_G": (None, [_hp, _dp, _dp]),
    "librpa_set_atoms": (None, [_hp, C.c_int, _ip, _dp]),
    "librpa_set_kgrids_kvec": (None, [_hp, C.c_int, C.c_int, C.c_int, _dp]),
    "librpa_set_ibz_mapping": (None, [_hp, C.c_int, _ip]),
    "librpa_set_lri_coeff": (None, [_hp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _dp]),
    "librpa_set_aux_bare_coulomb_k_atom_pair": (None, [_hp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp,
                                                       C.c_double]),
    "librpa_set_aux_cut_coulomb_k_atom_pair": (None, [_hp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp,
                                                      C.c_double]),
    "librpa_set_aux_bare_coulomb_k_2d_block": (None, [_hp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]),
    "librpa_set_band_kvec": (None, [_hp, C.c_int, _dp]),
    "librpa_set_band_occ_eigval": (None, [_hp, C.c_int, C.c_int, C.c_int, _dp, _dp]),
    "librpa_set_wfc_band": (None, [_hp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]),
    "librpa_reset_band_data": (None, [_hp]),
    "librpa_get_imaginary_frequency_grids": (None, [_hp, _op, _dp, _dp]),
    "librpa_get_rpa_correlation_energy": (C.c_double, [_hp, _op, C.c_int, _dp, _dp]),
    "librpa_build_exx": (None, [_hp, _op]),
    "librpa_build_g0w0_sigma": (None, [_hp, _op]),
    "librpa_get_g0w0_sigc_kgrid": (None, [_hp, _op, C.c_int, C.c_int, _ip, C.c_int, C.c_int, _dp, _dp, _dp, _dp]),
    "librpa_amd_set_precision": (None, [_hp, C.c_int, C.c_int]),
    "librpa_amd_set_error_mode": (None, [C.c_int]),
    "librpa_amd_get_last_error": (C.c_int, [C.c_char_p, C.c_int]),
    "librpa_amd_set_keep_chi0": (None, [_hp, C.c_int]),
    "librpa_amd_get_chi0s_block": (C.c_int, [_hp, C.c_int, C.c_int, C.c_int, _ip, _dp]),
    "librpa_amd_get_chi0q_block": (C.c_int, [_hp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]),
    "librpa_amd_get_tfgrids": (C.c_int, [_hp, _op, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "librpa_amd_interpolate_dielec_func": (C.c_int, [C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, _dp]),
    "librpa_amd_get_wc_block": (C.c_int, [_hp, C.c_int, C.c_int, C.c_int, _ip, _dp]),
    "librpa_amd_get_sigma_block": (C.c_int, [_hp, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _dp, _dp]),
    "librpa_amd_get_sigc_ks": (C.c_int, [_hp, C.c_int, C.c_int, C.c_int, _dp, _dp]),
    "librpa_amd_profile_json": (C.c_int, [C.c_char_p, C.c_int]),
    "librpa_amd_profile_reset": (None, []),
    "librpa_amd_get_stats": (C.c_int, [_hp, _dp, C.c_int]),
    "librpa_amd_gemm": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int,
                                  _dp, C.c_int, C.c_double, _dp, C.c_int, C.c_int]),
    "librpa_amd_set_screening": (None, [_hp, C.c_int, C.c_double, _fp]),
    "librpa_amd_get_screen_stats": (C.c_int, [_hp, _dp]),
    "librpa_amd_get_screen_samples": (C.c_int, [_hp, _fp, _fp]),
    "librpa_amd_screen_train": (C.c_double, [_fp, _fp, C.c_int, C.c_int, C.c_double, C.c_int, _fp]),
    "librpa_amd_mlp_predict": (C.c_int, [_fp, _fp, C.c_int, _fp]),
    "librpa_amd_mlp_loss_grad": (C.c_int, [_fp, _fp, _fp, C.c_int, _fp, _dp]),
    "librpa_amd_mlp_loss_grad_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                                   C.c_void_p, C.c_void_p]),
}

# block-norm screener (librpa_amd.h): MLP sizes
SCREEN_NFEAT = 8
SCREEN_NPARAM = 4801
SCREEN_MODEL = SCREEN_NPARAM + 2 * SCREEN_NFEAT + 2
SCREEN_OFF, SCREEN_BOUND, SCREEN_LEARNED = 0, 1, 2
MLP_NOALLOC_MAX = 32768  # LIBRPA_AMD_MLP_NOALLOC_MAX


def load(path: Optional[str] = None):
    """Load the HIP library (no fallback: raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise LibrpaError(f"{p} is missing: run `python -m librpa_amd.build` (or __graft_entry__.build())")
    lib = C.CDLL(p)
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib.librpa_amd_set_error_mode(1)
    _lib = lib
    return lib


def _check():
    buf = C.create_string_buffer(4096)
    if load().librpa_amd_get_last_error(buf, 4096):
        raise LibrpaError(buf.value.decode(errors="replace"))


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


def interpolate_dielec_func(option, omega_in, df_in, omega_out):
    """eps_M(i omega) given at omega_in, on omega_out (option_dielect_func 0, 1 or 2): the values
    the G0W0 head replacement uses, from the library's host code."""
    om, po = _d(omega_in)
    df, pd = _d(df_in)
    tg, pt = _d(omega_out)
    out = np.zeros(len(tg))
    n = load().librpa_amd_interpolate_dielec_func(int(option), len(om), po, pd, len(tg), pt, out.ctypes.data_as(_dp))
    _check()
    if n != len(tg):
        raise LibrpaError("librpa_amd_interpolate_dielec_func failed")
    return out


def screen_train(X, t, iters=2000, lr=3e-3, seed=0):
    """Train the screening MLP on the device; returns (model, final standardised MSE)."""
    X, xp = _f(X)
    t, tp = _f(t)
    model = np.zeros(SCREEN_MODEL, dtype=np.float32)
    loss = load().librpa_amd_screen_train(xp, tp, len(t), iters, lr, seed, model.ctypes.data_as(_fp))
    _check()
    return model, loss


def mlp_predict(params, X):
    p, pp = _f(params)
    X, xp = _f(X)
    y = np.zeros(len(X), dtype=np.float32)
    load().librpa_amd_mlp_predict(pp, xp, len(X), y.ctypes.data_as(_fp))
    _check()
    return y


def mlp_loss_grad(params, X, t):
    p, pp = _f(params)
    X, xp = _f(X)
    t, tp = _f(t)
    g = np.zeros(SCREEN_NPARAM, dtype=np.float32)
    loss = np.zeros(1)
    load().librpa_amd_mlp_loss_grad(pp, xp, tp, len(t), g.ctypes.data_as(_fp), loss.ctypes.data_as(_dp))
    _check()
    return float(loss[0]), g


def mlp_loss_grad_noalloc(d_params, d_X, d_t, n, d_grad, d_loss, stream=None):
    """librpa_amd_mlp_loss_grad_noalloc on device addresses (ints); enqueued on ``stream`` (a
    hipStream_t address, None for the null stream), not synchronised.  Raises on error."""
    rc = load().librpa_amd_mlp_loss_grad_noalloc(d_params, d_X, d_t, int(n), d_grad, d_loss, stream)
    _check()
    if rc != 0:
        raise LibrpaError(f"librpa_amd_mlp_loss_grad_noalloc returned {rc}")


def init_global():
    load().librpa_init_global(0, b"stdout", 1)
    _check()


def finalize_global():
    load().librpa_finalize_global()
    _check()


def gemm(A, B, C=None, alpha=1.0, beta=0.0, prec=FP64, trans_a=False, trans_b=False, tile=0):
    """C = alpha op(A) op(B) + beta C through the device GEMM engine (host arrays)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    M = A.shape[1] if trans_a else A.shape[0]
    K = A.shape[0] if trans_a else A.shape[1]
    N = B.shape[0] if trans_b else B.shape[1]
    C = np.zeros((M, N)) if C is None else np.ascontiguousarray(C, dtype=np.float64).copy()
    load().librpa_amd_gemm(prec, int(trans_a), int(trans_b), M, N, K, alpha, A.ctypes.data_as(_dp), A.shape[1],
                           B.ctypes.data_as(_dp), B.shape[1], beta, C.ctypes.data_as(_dp), N, tile)
    _check()
    return C


def profile() -> dict:
    import json
    lib = load()
    n = lib.librpa_amd_profile_json(None, 0)
    buf = C.create_string_buffer(n + 16)
    lib.librpa_amd_profile_json(buf, n + 16)
    _check()
    return json.loads(buf.value.decode())


def profile_reset():
    load().librpa_amd_profile_reset()


class Handler:
    """Mirror of librpa::Handler (LibRPA include/librpa.hpp:57-156)."""

    def __init__(self, comm: int = 0x44000000):
        lib = load()
        self._lib = lib
        self._h = lib.librpa_create_handler(comm)
        _check()

    def free(self):
        if self._h:
            self._lib.librpa_destroy_handler(self._h)
            self._h = None
            _check()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def _call(self, name, *args):
        r = getattr(self._lib, name)(self._h, *args)
        _check()
        return r

    # ---- input setters (librpa_input.h)
    def set_scf_dimension(self, nspins, nkpts, nstates, nbasis, nspinor=1):
        self._call("librpa_set_scf_dimension", nspins, nkpts, nstates, nbasis, nspinor)

    def set_wg_ekb_efermi(self, nspins, nkpts, nstates, wg, ekb, efermi):
        wg, pw = _d(wg)
        ekb, pe = _d(ekb)
        self._call("librpa_set_wg_ekb_efermi", nspins, nkpts, nstates, pw, pe, efermi)

    def set_wfc(self, ispin, ik, nstates, nbasis, wfc_re, wfc_im):
        a, pa = _d(wfc_re)
        b, pb = _d(wfc_im)
        self._call("librpa_set_wfc", ispin, ik, nstates, nbasis, pa, pb)

    def set_ao_basis_wfc(self, nbs):
        arr = (C.c_size_t * len(nbs))(*[int(x) for x in nbs])
        self._call("librpa_set_ao_basis_wfc", len(nbs), arr)

    def set_ao_basis_aux(self, nbs):
        arr = (C.c_size_t * len(nbs))(*[int(x) for x in nbs])
        self._call("librpa_set_ao_basis_aux", len(nbs), arr)

    def set_latvec_and_G(self, lat, G):
        a, pa = _d(np.asarray(lat).reshape(9))
        b, pb = _d(np.asarray(G).reshape(9))
        self._call("librpa_set_latvec_and_G", pa, pb)

    def set_atoms(self, types, pos):
        t = np.ascontiguousarray(types, dtype=np.int32)
        p, pp = _d(np.asarray(pos).reshape(-1))
        self._call("librpa_set_atoms", len(t), t.ctypes.data_as(_ip), pp)

    def set_kgrids_kvec(self, nk1, nk2, nk3, kvecs):
        k, pk = _d(np.asarray(kvecs).reshape(-1))
        self._call("librpa_set_kgrids_kvec", nk1, nk2, nk3, pk)

    def set_ibz_mapping(self, mapping):
        m = np.ascontiguousarray(mapping, dtype=np.int32)
        self._call("librpa_set_ibz_mapping", len(m), m.ctypes.data_as(_ip))

    def set_dielect_func_imagfreq(self, omegas, df):
        om, po = _d(omegas)
        v, pv = _d(df)
        self._call("librpa_set_dielect_func_imagfreq", len(om), po, pv)

    def set_lri_coeff(self, routing, I, J, nbasis_i, nbasis_j, naux_mu, R, cs):
        r = (C.c_int * 3)(*[int(x) for x in R])
        c, pc = _d(cs)
        self._call("librpa_set_lri_coeff", routing, I, J, nbasis_i, nbasis_j, naux_mu, r, pc)

    def set_aux_bare_coulomb_k_atom_pair(self, ik, I, J, naux_mu, naux_nu, vre, vim, threshold=0.0):
        a, pa = _d(vre)
        b, pb = _d(vim)
        self._call("librpa_set_aux_bare_coulomb_k_atom_pair", ik, I, J, naux_mu, naux_nu, pa, pb, threshold)

    def set_aux_cut_coulomb_k_atom_pair(self, ik, I, J, naux_mu, naux_nu, vre, vim, threshold=0.0):
        a, pa = _d(vre)
        b, pb = _d(vim)
        self._call("librpa_set_aux_cut_coulomb_k_atom_pair", ik, I, J, naux_mu, naux_nu, pa, pb, threshold)

    # ---- compute (librpa_compute.h)
    def get_imaginary_frequency_grids(self, opts: Options):
        om = np.zeros(opts.nfreq)
        w = np.zeros(opts.nfreq)
        self._call("librpa_get_imaginary_frequency_grids", C.byref(opts), om.ctypes.data_as(_dp),
                   w.ctypes.data_as(_dp))
        return om, w

    def get_rpa_correlation_energy(self, opts: Options, n_ibz_kpoints: int):
        re = np.zeros(n_ibz_kpoints)
        im = np.zeros(n_ibz_kpoints)
        ec = self._call("librpa_get_rpa_correlation_energy", C.byref(opts), n_ibz_kpoints,
                        re.ctypes.data_as(_dp), im.ctypes.data_as(_dp))
        return ec, re + 1j * im

    def build_g0w0_sigma(self, opts: Options):
        self._call("librpa_build_g0w0_sigma", C.byref(opts))

    def get_g0w0_sigc_kgrid(self, opts, n_spins, iks, i_state_low, i_state_high, vxc, vexx):
        iks = np.ascontiguousarray(iks, dtype=np.int32)
        n = n_spins * len(iks) * (i_state_high - i_state_low)
        vx, pvx = _d(vxc)
        ve, pve = _d(vexx)
        re = np.zeros(n)
        im = np.zeros(n)
        self._call("librpa_get_g0w0_sigc_kgrid", C.byref(opts), n_spins, len(iks), iks.ctypes.data_as(_ip),
                   i_state_low, i_state_high, pvx, pve, re.ctypes.data_as(_dp), im.ctypes.data_as(_dp))
        return re + 1j * im

    # ---- extensions (librpa_amd.h)
    def set_precision(self, chi0: int = FP64, ec: int = FP64):
        self._call("librpa_amd_set_precision", chi0, ec)

    def set_keep_chi0(self, keep: bool = True):
        self._call("librpa_amd_set_keep_chi0", 1 if keep else 0)

    def get_chi0s_block(self, itau, I, J, R, naI, naJ):
        out = np.zeros((naI, naJ))
        r = (C.c_int * 3)(*[int(x) for x in R])
        self._call("librpa_amd_get_chi0s_block", itau, I, J, r, out.ctypes.data_as(_dp))
        return out

    def get_chi0q_block(self, ifreq, iq, I, J, naI, naJ):
        re = np.zeros((naI, naJ))
        im = np.zeros((naI, naJ))
        self._call("librpa_amd_get_chi0q_block", ifreq, iq, I, J, re.ctypes.data_as(_dp), im.ctypes.data_as(_dp))
        return re + 1j * im

    def get_wc_block(self, itau, I, J, R, naI, naJ):
        out = np.zeros((naI, naJ))
        r = (C.c_int * 3)(*[int(x) for x in R])
        self._call("librpa_amd_get_wc_block", itau, I, J, r, out.ctypes.data_as(_dp))
        return out

    def get_sigma_block(self, ispin, ifreq, I, J, R, nbI, nbJ):
        re = np.zeros((nbI, nbJ))
        im = np.zeros((nbI, nbJ))
        r = (C.c_int * 3)(*[int(x) for x in R])
        self._call("librpa_amd_get_sigma_block", ispin, ifreq, I, J, r, re.ctypes.data_as(_dp),
                   im.ctypes.data_as(_dp))
        return re + 1j * im

    def get_sigc_ks(self, ispin, ik, state, nfreq):
        re = np.zeros(nfreq)
        im = np.zeros(nfreq)
        self._call("librpa_amd_get_sigc_ks", ispin, ik, state, re.ctypes.data_as(_dp), im.ctypes.data_as(_dp))
        return re + 1j * im

    def get_tfgrids(self, opts: Options):
        n = opts.nfreq
        arrs = [np.zeros(n) for _ in range(4)] + [np.zeros((n, n)) for _ in range(3)]
        self._call("librpa_amd_get_tfgrids", C.byref(opts), *[a.ctypes.data_as(_dp) for a in arrs])
        keys = ["tau", "tau_w", "omega", "omega_w", "cos_t2f", "cos_f2t", "sin_t2f"]
        return dict(zip(keys, arrs))

    def set_screening(self, mode=SCREEN_OFF, threshold=1e-8, model=None):
        m = None if model is None else np.ascontiguousarray(model, dtype=np.float32)
        self._call("librpa_amd_set_screening", mode, threshold, None if m is None else m.ctypes.data_as(_fp))

    def get_screen_stats(self):
        out = np.zeros(4)
        self._call("librpa_amd_get_screen_stats", out.ctypes.data_as(_dp))
        return dict(zip(["blocks", "skipped", "flops", "flops_skipped"], out.tolist()))

    def get_screen_samples(self):
        n = self._call("librpa_amd_get_screen_samples", None, None)
        X = np.zeros((n, SCREEN_NFEAT), dtype=np.float32)
        t = np.zeros(n, dtype=np.float32)
        self._call("librpa_amd_get_screen_samples", X.ctypes.data_as(_fp), t.ctypes.data_as(_fp))
        return X, t

    def get_stats(self):
        out = np.zeros(10)
        self._call("librpa_amd_get_stats", out.ctypes.data_as(_dp), 10)
        keys = ["flops_chi0_alg", "flops_chi0_exec", "t_chi0", "t_ec", "n_out_blocks", "n_cs_blocks", "n_abf",
                "n_basis", "mixed_amax", "n_fp32_tau"]
        return dict(zip(keys, out.tolist()))


def load_system(h: Handler, sys, routing: int = LIBRI, vq: bool = True, vq_cut=False, verbose: bool = False):
    """Push a librpa_amd.synth.System through the C API (mirrors the driver's read_* calls).

    ``vq_cut``: True passes the Coulomb matrix again as the truncated one used for W (G0W0);
    an array (nk, n_abf, n_abf) passes that matrix instead."""
    ns, nk, nb, nbas = sys.nspin, sys.nk, sys.nbands, sys.nbasis
    h.set_scf_dimension(ns, nk, nb, nbas)
    h.set_wg_ekb_efermi(ns, nk, nb, sys.occ.reshape(-1), sys.eps.reshape(-1), sys.efermi)
    for s in range(ns):
        for k in range(nk):
            w = sys.wfc[s, k]
            h.set_wfc(s, k, nb, nbas, w.real, w.imag)
    h.set_ao_basis_wfc(sys.nao)
    h.set_ao_basis_aux(sys.nabf)
    h.set_latvec_and_G(sys.latvec.reshape(9), sys.recvec.reshape(9))
    h.set_atoms(sys.types, sys.positions)
    h.set_kgrids_kvec(*sys.period, sys.kcart)
    for (I, J, R, blk) in sys.cs:
        h.set_lri_coeff(routing, I, J, int(sys.nao[I]), int(sys.nao[J]), int(sys.nabf[I]), R, blk)
    if vq and sys.vq is not None:
        off = np.concatenate([[0], np.cumsum(sys.nabf)])
        for ik in range(nk):
            if verbose:
                print(f"[load_system] Coulomb blocks of k {ik + 1} / {nk}", file=_sys.stderr, flush=True)
            V = sys.vq[ik]
            for I in range(sys.natom):
                for J in range(I, sys.natom):
                    b = V[off[I]:off[I + 1], off[J]:off[J + 1]]
                    h.set_aux_bare_coulomb_k_atom_pair(ik, I, J, b.shape[0], b.shape[1], b.real, b.imag, 0.0)
                    if isinstance(vq_cut, np.ndarray):
                        c = vq_cut[ik][off[I]:off[I + 1], off[J]:off[J + 1]]
                        h.set_aux_cut_coulomb_k_atom_pair(ik, I, J, c.shape[0], c.shape[1], c.real, c.imag, 0.0)
                    elif vq_cut:
                        h.set_aux_cut_coulomb_k_atom_pair(ik, I, J, b.shape[0], b.shape[1], b.real, b.imag, 0.0)
