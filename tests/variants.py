"""Model variants exercising each feature of the flux path (overrides of the
reference's test model)."""

VARIANTS = {
    "poly4_ld": {"ldc1_2": 0.2, "ldc1_3": -0.1, "ldc1_4": 0.05, "ldc2_2": 0.1, "ldc2_3": 0.05,
                 "ldc2_4": -0.02},
    "claret_ld": {"limb1": "Claret", "limb2": "Claret", "ldc1_1": 0.55, "ldc1_2": -0.35,
                  "ldc1_3": 0.65, "ldc1_4": -0.25, "ldc2_1": 0.7, "ldc2_2": -0.45, "ldc2_3": 0.85,
                  "ldc2_4": -0.35},
    "beaming": {"velocity_scale": 300.0, "beam_factor1": 1.5, "beam_factor2": 1.2},
    "beaming_spin": {"velocity_scale": 300.0, "beam_factor1": 1.5, "beam_factor2": 1.2,
                     "spin1": 3.0, "spin2": 0.8},
    "lensing": {"glens1": 1, "velocity_scale": 300.0},
    "mucrit": {"mucrit1": 0.1, "mucrit2": 0.05},
    "ephemeris": {"t0": 0.013, "period": 0.97, "pdot": 2e-4, "deltat": 0.003},
    "poly_factor": {"slope": 0.1, "quad": -0.05, "cube": 0.02, "third": 2e-11},
    "mirror": {"mirror": 1, "absorb": 0.8},
    "star_spots": {"stsp11_long": 30.0, "stsp11_lat": 20.0, "stsp11_fwhm": 40.0, "stsp11_tcen": 9000.0,
                   "stsp21_long": 10.0, "stsp21_lat": -10.0, "stsp21_fwhm": 30.0, "stsp21_tcen": 5000.0},
    "uniform_eq_spot": {"uesp_long1": 20.0, "uesp_long2": 80.0, "uesp_lathw": 15.0, "uesp_taper": 10.0,
                        "uesp_temp": 10000.0},
    "disc_edge_spot": {"add_disc": 1, "opaque": 0, "rdisc1": 0.1, "rdisc2": 0.3, "nrad": 30,
                       "temp_disc": 6000, "texp_disc": -0.75, "temp_edge": 5000, "absorb_edge": 0.5,
                       "add_spot": 1, "nspot": 50, "radius_spot": 0.3, "expon_spot": 1.0},
    "opaque_disc_calibration": {"add_disc": 1, "opaque": 1, "rdisc1": 0.1, "rdisc2": 0.3, "nrad": 30,
                                "temp_disc": 6000, "texp_disc": -0.75, "add_spot": 1, "nspot": 50,
                                "radius_spot": 0.3, "expon_spot": 1.0},
    "npole_infill": {"npole": 1, "nlatfill": 2, "nlngfill": 2, "nlat2c": 40},
    "roche_primary": {"roche1": 1, "r1": 0.2, "nlat1f": 16, "nlat1c": 10, "nlat2f": 24, "nlat2c": 16},
    "no_eclipses": {"eclipse1": 0, "eclipse2": 0},
}


# Disc, edge and bright spot for the per-component (iscale) fits.
ISCALE_DISC = {"add_disc": 1, "opaque": 0, "rdisc1": 0.1, "rdisc2": 0.3, "nrad": 30, "temp_disc": 6000,
               "texp_disc": -0.75, "temp_edge": 5000, "absorb_edge": 0.5, "add_spot": 1, "nspot": 50,
               "radius_spot": 0.3, "expon_spot": 1.0}
