"""Inputs of the golden fixtures (tests/golden/make_fixtures.py writes the
outputs from the compiled reference; the tests re-run the same inputs)."""
import numpy as np

from mytrim_amd import configs
from mytrim_amd.problem import (CUT, GEO_LAYERS, PBC, Bins, Element, Geometry, Material, Primaries,
                                Problem)

# (Z, mass) of ions; mass 0 takes the table mass (material.C:256-259)
IONS = [(1, 1.008), (2, 4.0026), (5, 10.81), (6, 12.011), (14, 28.086), (26, 55.845),
        (29, 63.546), (47, 107.87), (54, 131.29), (74, 183.84), (92, 238.03), (29, 0.0)]
# (name, rho, [(Z, m, t)])
TARGETS = [("O", 1.429, [(8, 15.999, 1.0)]), ("C", 2.253, [(6, 12.011, 1.0)]),
           ("Si", 2.33, [(14, 28.086, 1.0)]), ("Fe", 7.87, [(26, 55.845, 1.0)]),
           ("Cu", 8.96, [(29, 63.546, 1.0)]), ("Ag", 10.5, [(47, 107.87, 1.0)]),
           ("Xe", 3.5, [(54, 131.29, 1.0)]), ("W", 19.3, [(74, 183.84, 1.0)]),
           ("U", 19.05, [(92, 238.03, 1.0)]), ("Ge", 5.32, [(32, 72.63, 1.0)]),
           ("UO2", 10.96, [(92, 238.03, 1.0), (8, 15.999, 2.0)])]
ENERGIES = np.logspace(0, 8, 97)  # 1 eV .. 100 MeV


def c4_uniform(w, n_hist):
    """UO2 + Xe bubbles with a single 300 keV Xe primary species."""
    c4 = configs.c4_mixed_uo2(n_hist, w=w)
    return c4.with_(primaries=Primaries.single(54, 131.29, 3.0e5, (w / 2, w / 2, w / 2),
                                               (0.6, 0.8, 0.0)))


def trace_cases():
    return {"c1": configs.c1_he_fe(4), "c2": configs.c2_cu_cu(4), "c3": configs.c3_ag_layers(4),
            "c4": c4_uniform(2000.0, 4), "c5": configs.c5_c_w(4),
            "moliere": configs.c2_cu_cu(4).with_(potential=1),
            "ckr": configs.c5_c_w(4).with_(potential=2),
            "wire": configs.wire_b_si(4), "bwire": configs.burried_wire_p_si(4)}


def log_cases():
    """name -> (problem, TrimVacMap z1 z2 z3) for the event-stream fixtures"""
    return {"vac_cu_10keV": (configs.vacancy_case("cu", 1e4, 6), (29, -1, -1)),
            "wire": (configs.wire_b_si(12), (14, 5, -1)),
            "c4": (c4_uniform(2000.0, 3).with_(primaries=Primaries.single(
                54, 131.29, 3.0e4, (1000.0, 1000.0, 1000.0), (0.6, 0.8, 0.0))), (92, 8, 54))}


def tally_cases(quick=False):
    """name -> (problem, processes used to write it)"""
    s = 0.1 if quick else 1.0
    return {
        "c1": (configs.c1_he_fe(2000), 1),
        "c2": (configs.c2_cu_cu(int(4000 * s)), 8),
        "c3": (configs.c3_ag_layers(int(800 * s)), 8),
        "c4": (c4_uniform(4000.0, int(400 * s)), 8),
        "c5": (configs.c5_c_w(int(2000 * s)), 8),
        "c2_primaries_kp": (configs.c2_cu_cu(int(4000 * s)).with_(follow_max_gen=1, kp=True), 4),
        "vac_cu_1keV": (configs.vacancy_case("cu", 1e3, 4000), 1),
        "vac_cu_10keV": (configs.vacancy_case("cu", 1e4, 2000), 2),
        "vac_cu_100keV": (configs.vacancy_case("cu", 1e5, int(800 * s)), 8),
        "vac_si_10keV": (configs.vacancy_case("si", 1e4, 2000), 2),
        "vac_xe_10keV": (configs.vacancy_case("xe", 1e4, 2000), 2),
        "wire": (configs.wire_b_si(int(2000 * s)), 8),
        "bwire": (configs.burried_wire_p_si(int(400 * s)), 8),
    }


def stat_cases():
    """name -> problem for the statistical fixtures (tests/golden/stat_*.npz):
    larger samples of the workloads, with per-history tallies and batched
    depth histograms, for the two-sample tests of tests/tally_stats.py.  Each
    holds about 1e8 collisions or more (at most 32 000 histories).  The
    mixed C4 workload runs on its benchmark box with its per-history H, He, C
    and Xe primaries."""
    return {
        "c1": configs.c1_he_fe(32000),
        "c2": configs.c2_cu_cu(4800),
        "c3": configs.c3_ag_layers(3200),
        "c4mix": configs.c4_mixed_uo2(12800),
        "c5": configs.c5_c_w(25600),
        "c2_primaries_kp": configs.c2_cu_cu(32000).with_(follow_max_gen=1, kp=True),
        "vac_cu_10keV": configs.vacancy_case("cu", 1e4, 32000),
        "wire": configs.wire_b_si(32000),
        "bwire": configs.burried_wire_p_si(9600),
    }


STAT_BATCHES = 32   # reference batches per statistical fixture (equal sizes)


def xe_si_sio2(n_hist, seed=2344):
    """30 keV Xe into 100 A Si over 400 A SiO2, 500 x 100 x 100 A: the layer
    stack of a SampleDynamic run; many ions leave through either face."""
    t = (100.0, 400.0)
    w = (sum(t), 100.0, 100.0)
    return Problem(
        name="xe_si_sio2_30keV",
        materials=[Material(2.33, [Element(14, 28.086, 1.0, 25.0, 3.0)]),
                   Material(2.2, [Element(14, 28.086, 1.0, 25.0, 3.0), Element(8, 15.999, 2.0, 28.0, 3.0)])],
        geometry=Geometry(GEO_LAYERS, w, (CUT, PBC, PBC), layers=t),
        primaries=Primaries.single(54, 131.0, 3.0e4, (0.0, w[1] / 2, w[2] / 2)),
        n_hist=n_hist, seed=seed,
        range_bins=Bins(0.0, 10.0, 100), vac_bins=Bins(0.0, 10.0, 100),
        lat_bw=10.0, lat_nb=40, lat_y0=w[1] / 2, lat_z0=w[2] / 2)


def transport_cases():
    """name -> problem for the layer-transport fixtures (SampleDynamic's atom
    moves with the composition held fixed)"""
    return {"c3": configs.c3_ag_layers(24), "xe_si_sio2": xe_si_sio2(200)}
