"""ctypes mirror of the C-ABI in include/gfm_hip.h.

The structures here must match the header field for field; the test
``tests/test_abi.py`` checks their sizes against ``sizeof`` reported by the
compiled libraries.
"""
import ctypes as ct

import numpy as np

D = ct.c_double
I32 = ct.c_int32


class GfmCtrl(ct.Structure):
    _fields_ = [
        ("mp", I32), ("np", I32), ("lp", I32), ("nr0", I32),
        ("ncells", I32), ("i_ref", I32), ("j_ref", I32), ("k_ref", I32),
        ("sweep_order", I32), ("precision", I32), ("want_pre_residuals", I32), ("want_irc", I32),
        ("i1st_rad_done", I32), ("isothermal_walls", I32), ("gray_radiation", I32), ("pcorr_block_corr", I32),
        ("itr_gas", I32), ("itr_gend", I32), ("maxgi", I32), ("want_plane_flows", I32),
        ("bgcon", D), ("factor", D), ("fadj", D),
        ("t0", D), ("trdg", D), ("h_0", D), ("r0", D), ("area0", D), ("vol0", D),
        ("reyg", D), ("reyg2", D), ("eun", D), ("epsr", D), ("taur", D), ("ug0", D), ("dnst0", D),
        ("qr0", D), ("q0", D), ("gmfr", D), ("tk0", D), ("epsi0", D),
        ("gra", D * 3),
        ("p_ref", D),
        ("t_mx", D), ("t_mn", D), ("u_mx", D), ("u_mn", D), ("p_mx", D), ("p_mn", D), ("f_mx", D), ("f_mn", D),
        ("rf", D * 21), ("rfl", D * 21), ("rf_hsource", D),
        ("ntimesg", I32 * 20),
        ("gfin0", D),
        ("t_wall", D), ("h_g", D), ("kappa_gray", D), ("sigma_sb", D),
        ("res_mass", D), ("res_mass_max", D),
        ("imx", I32), ("jmx", I32), ("kmx", I32), ("pad2", I32),
        ("p_resid_pre", D), ("p_resid", D), ("pcorr_resid", D), ("h_resid_pre", D), ("h_resid", D),
        ("u_resid_pre", D * 3), ("u_resid", D * 3),
        ("gf_resid_pre", D * 10), ("gf_resid", D * 10),
        ("resg", D * 40),
        ("gfex", D), ("gfin", D), ("fuel_out", D), ("tmean_ex", D), ("q_react", D), ("q_exhaust", D), ("q_wall", D),
        ("tave", D), ("tave_vol", D), ("tmax_c", D),
        ("t_av_exit", D), ("t_x_exit", D),
        ("pcorr_mode", I32), ("pcorr_guess", I32), ("pcorr_max_sweeps", I32), ("pcorr_dump", I32),
        ("pcorr_sweeps", I32), ("pcorr_fallback", I32),
        ("pcorr_tol", D), ("pcorr_resid_start", D), ("pcorr_alpha", D), ("pcorr_gain", D),
        ("pcorr_guess_accept", D),
        ("q_wall_bottom", D), ("q_wall_top", D),
        ("bv_f", D * 2), ("ba_f", D * 6),
    ]

    def copy(self):
        c = GfmCtrl()
        ct.memmove(ct.byref(c), ct.byref(self), ct.sizeof(GfmCtrl))
        return c


DP = ct.POINTER(ct.c_double)


class GfmHostArrays(ct.Structure):
    _fields_ = [
        ("x", DP), ("dx", DP), ("r", DP), ("dr", DP), ("rs", DP), ("z", DP), ("dz", DP),
        ("ibcell", ct.POINTER(ct.c_int16)),
        ("p", DP), ("t", DP), ("dnst", DP), ("theta", DP), ("tmu", DP), ("gcp", DP), ("glam", DP),
        ("gama", DP), ("gte", DP), ("crate", DP), ("hrate", DP), ("bs_hprev", DP),
        ("ug", DP), ("gf", DP),
        ("rtim", DP), ("delrt", DP), ("hrt", DP), ("delr", DP), ("timed", DP), ("hrd", DP),
    ]


FIELD3 = ("p", "t", "dnst", "theta", "tmu", "gcp", "glam", "gama", "gte", "crate", "hrate", "bs_hprev")
GRID1 = ("x", "dx", "r", "dr", "rs", "z", "dz")
KIN = ("rtim", "delrt", "hrt", "delr", "timed", "hrd")


class GfmMsCtrl(ct.Structure):
    """gfm_ms_ctrl: the minor-species inputs and outputs (MSFLOW/GSLV3M)."""
    _fields_ = [
        ("maxms", I32), ("react", I32), ("ntimesm", I32), ("itr_ms", I32), ("igs0", I32), ("pad0", I32),
        ("bmcon", D), ("ssigma", D), ("rfm", D * 8),
        ("aform", D), ("aoxid", D), ("esf_r", D), ("eso_r", D), ("t0_sf", D), ("t0_so", D),
        ("rf_soot_form_source", D), ("rf_soot_oxid_source", D), ("soot_cap", D), ("dnst_soot", D), ("vol_tot", D),
        ("iterations", I32), ("converged", I32), ("smf_clipped_cnt", I32), ("pad1", I32),
        ("resid_ms_pre", D * 8), ("resid_ms", D * 8),
        ("bal_ms", D), ("soot_tot", D), ("smf_max", D), ("smf_max_noclip", D), ("rr_f", D), ("rr_c", D),
        ("gfm_0", D * 7), ("gfmex", D), ("ave_chg", D), ("svf_r", D),
    ]

    def copy(self):
        c = GfmMsCtrl()
        ct.memmove(ct.byref(c), ct.byref(self), ct.sizeof(GfmMsCtrl))
        return c


class GfmMsArrays(ct.Structure):
    _fields_ = [("gfm", DP), ("smf", DP), ("bs_soot_prev", DP), ("sfp_soot_prev", DP), ("rrms", DP)]


MS_FIELDS = ("gfm", "smf", "bs_soot_prev", "sfp_soot_prev", "rrms")


def ms_arrays(arrs):
    """GfmMsArrays pointing at the numpy arrays of a minor-species state
    (keep ``arrs`` alive while the structure is used)."""
    a = GfmMsArrays()
    for k in MS_FIELDS:
        v = arrs[k]
        assert v.dtype == np.float64 and v.flags["F_CONTIGUOUS"], k
        setattr(a, k, v.ctypes.data_as(DP))
    return a


NWL_MAX = 16  # GFM_NWL_MAX
RADSRC_NONE, RADSRC_GRAY, RADSRC_FIELD = 0, 1, 2  # GFM_RADSRC_*


class GfmRadCtrl(ct.Structure):
    """gfm_rad_ctrl: inputs and outputs of a radiation pass (RAD_EMIS/RAD0/QRNF)."""
    _fields_ = [
        ("nwl", I32), ("have_crown", I32), ("wl", D * NWL_MAX), ("ri", D),
        ("eps_w", D), ("eps_m", D), ("eps_c", D), ("height_to_ceiling", D), ("dnst_soot", D),
        ("l_ab", I32), ("n_emit", I32), ("carry_unset", I32), ("angle_oob", I32),
        ("tot_qe_h2o_co2", D), ("tot_qe_soot", D), ("qe_sum_from_open", D),
        ("qa_sum_to_melt", D), ("qa_sum_to_wall", D), ("qa_sum_to_open", D), ("qa_sum_to_exit", D),
        ("qa_sum_to_inlet", D), ("qeg0", D), ("qeg_g", D), ("qeg_w", D), ("qeg_s", D), ("qeg_io", D),
    ]

    def copy(self):
        c = GfmRadCtrl()
        ct.memmove(ct.byref(c), ct.byref(self), ct.sizeof(GfmRadCtrl))
        return c


class GfmRadArrays(ct.Structure):
    _fields_ = [("qe", DP), ("qa", DP), ("qal", DP), ("akl", DP)]


RAD_FIELDS = ("qe", "qa", "qal", "akl")
RAD_SUMS = ("tot_qe_h2o_co2", "tot_qe_soot", "qe_sum_from_open", "qa_sum_to_melt", "qa_sum_to_wall",
            "qa_sum_to_open", "qa_sum_to_exit", "qa_sum_to_inlet", "qeg0", "qeg_g", "qeg_w", "qeg_s", "qeg_io")


def rad_arrays(arrs):
    """GfmRadArrays pointing at the numpy arrays of a radiation state (a
    missing or None entry is a null pointer)."""
    a = GfmRadArrays()
    for k in RAD_FIELDS:
        v = arrs.get(k)
        if v is None:
            continue
        assert v.dtype == np.float64 and v.flags["F_CONTIGUOUS"], k
        setattr(a, k, v.ctypes.data_as(DP))
    return a


NEX_MAX = 32      # GFM_NEX_MAX
NWALG_MAX = 64    # GFM_NWALG_MAX
I16P = ct.POINTER(ct.c_int16)


class GfmWallCtrl(ct.Structure):
    """gfm_wall_ctrl: inputs, carried state and outputs of the wall exchange (C/Qwall.f90)."""
    _fields_ = [
        ("maxri1", I32), ("maxri2", I32), ("minri2", I32), ("preset_vf", I32),
        ("have_crown", I32), ("nex", I32), ("nwal0", I32), ("isoot_cal", I32),
        ("eps_w", D), ("eps_m", D), ("eps_c", D), ("height_to_ceiling", D),
        ("t_a", D), ("h_a", D), ("w_d", D), ("w_k", D), ("h_g", D),
        ("wall_radiosity_tolerance", D), ("radiation_boundary_tolerance", D),
        ("exh_box", (I32 * 6) * NEX_MAX),
        ("exh_orient", I32 * NEX_MAX), ("exh_direct", I32 * NEX_MAX), ("exh_type", I32 * NEX_MAX),
        ("exh_frac", D * NEX_MAX), ("exh_fixed", D * NEX_MAX),
        ("w1_d", D * NWALG_MAX), ("w1_k", D * NWALG_MAX), ("w1_ha", D * NWALG_MAX), ("w1_ta", D * NWALG_MAX),
        ("q_melt_req", D), ("rf_aform", D), ("rf_afover", D),
        ("i1st_rad_done", I32), ("iterw", I32),
        ("avg_wall_T", D), ("aform", D), ("qa_s0", D), ("qa_s1", D),
        ("b_cnt", I32), ("iwall_count", I32), ("imelt_count", I32), ("inlet_count", I32), ("iexit_count", I32),
        ("pad0", I32),
        ("wall_area_total", D), ("melt_area_total", D), ("inlet_area_total", D), ("exit_area_total", D),
        ("boundary_area_total", D),
        ("sweeps", I32), ("outer", I32),
        ("q_incident_total", D), ("q_incident_wall", D), ("q_incident_inlet", D), ("q_incident_exit", D),
        ("q_incident_melt", D), ("qc_avg", D),
        ("qconv_total", D), ("qconv_avg", D), ("q_wall_loss", D), ("qamb_avg", D), ("q_melt", D), ("qrs_avg", D),
        ("q_rad_inlet", D), ("q_rad_outlet", D), ("qm_flame", D), ("qm_bnds", D), ("qm_conv", D),
        ("rel_rad_chg_last", D),
    ]

    def copy(self):
        c = GfmWallCtrl()
        ct.memmove(ct.byref(c), ct.byref(self), ct.sizeof(GfmWallCtrl))
        return c


WALL_I16 = ("b_typ", "b_i", "b_j", "b_k", "b_orient", "b_direct", "bc_i", "bc_j", "bc_k", "b_exh")
WALL_F64 = ("patch_area", "b_face_area", "emis", "radiosity", "q_amb", "q_conv", "qc", "qeb", "b_t")
WALL_OTHER = ("vf", "vf_sum", "qrs", "wall_rad_t", "conv_log")


class GfmWallArrays(ct.Structure):
    _fields_ = ([(k, I16P) for k in WALL_I16] + [(k, DP) for k in WALL_F64] + [(k, DP) for k in WALL_OTHER] +
                [("log_cap", I32), ("pad", I32)])


def wall_arrays(arrs):
    """GfmWallArrays pointing at the numpy arrays of a wall state (a missing
    or None entry is a null pointer)."""
    a = GfmWallArrays()
    for k in WALL_I16:
        v = arrs.get(k)
        if v is not None:
            assert v.dtype == np.int16 and v.flags["C_CONTIGUOUS"], k
            setattr(a, k, v.ctypes.data_as(I16P))
    for k in WALL_F64 + WALL_OTHER:
        v = arrs.get(k)
        if v is not None:
            assert v.dtype == np.float64 and v.flags["F_CONTIGUOUS"], k
            setattr(a, k, v.ctypes.data_as(DP))
    log = arrs.get("conv_log")
    a.log_cap = 0 if log is None else int(log.size)
    return a


def _ptr(a, ctype=ct.c_double):
    assert a.flags["F_CONTIGUOUS"] or a.ndim == 1, "arrays must be Fortran-ordered"
    return a.ctypes.data_as(ct.POINTER(ctype))


def host_arrays(arrs):
    """Build a GfmHostArrays pointing at the numpy arrays in ``arrs``.

    The caller must keep ``arrs`` alive while the structure is used."""
    h = GfmHostArrays()
    for k in GRID1 + FIELD3 + KIN + ("ug", "gf"):
        a = arrs[k]
        assert a.dtype == np.float64, k
        setattr(h, k, _ptr(a))
    assert arrs["ibcell"].dtype == np.int16
    h.ibcell = _ptr(arrs["ibcell"], ct.c_int16)
    return h


def copy_arrays(arrs):
    return {k: (np.array(v, order="F", copy=True) if isinstance(v, np.ndarray) else v) for k, v in arrs.items()}
