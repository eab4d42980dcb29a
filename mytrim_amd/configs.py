"""The benchmark and validation workloads, as :class:`Problem` objects.

C1..C5 are the five cascade workloads this engine is measured on; the wire
problems follow apps/mytrim_wire2.C; the ``vacancy_*`` problems restate the infinite-medium set-up of the reference's
vacancy validation inputs (validation/vacancy_count/*_vac.json run through
apps/runmytrim.C: SampleLayers(thickness, 100, 100) with the default all-PBC
boundaries and primaries started at (0, 50, 50) along +x).
"""
from __future__ import annotations

import numpy as np

from .problem import (CUT, GEO_BURRIED_WIRE, GEO_CLUSTERS, GEO_LAYERS, GEO_SOLID, GEO_WIRE, INF, PBC,
                      Bins, Element, Geometry, Material, Primaries, Problem, log_uniform)


def c1_he_fe(n_hist: int = 2000, seed: int = 2344) -> Problem:
    """100 keV He into a 1 um Fe cube (7.87 g/cm3), Ed 40 eV, Eb 3 eV."""
    w = 10000.0
    return Problem(
        name="c1_he_fe_100keV",
        materials=[Material(7.87, [Element(26, 55.845, 1.0, 40.0, 3.0)])],
        geometry=Geometry(GEO_SOLID, (w, w, w), (CUT, PBC, PBC)),
        primaries=Primaries.single(2, 4.0026, 1.0e5, (0.0, w / 2, w / 2)),
        n_hist=n_hist, seed=seed,
        range_bins=Bins(0.0, 10.0, 1000), vac_bins=Bins(0.0, 10.0, 1000),
        lat_bw=10.0, lat_nb=400, lat_y0=w / 2, lat_z0=w / 2)


def c2_cu_cu(n_hist: int = 1000000, seed: int = 2344) -> Problem:
    """500 keV Cu into a 2 um Cu cube (8.96 g/cm3), Ed 30 eV, Eb 3 eV."""
    w = 20000.0
    return Problem(
        name="c2_cu_cu_500keV",
        materials=[Material(8.96, [Element(29, 63.546, 1.0, 30.0, 3.0)])],
        geometry=Geometry(GEO_SOLID, (w, w, w), (CUT, PBC, PBC)),
        primaries=Primaries.single(29, 63.546, 5.0e5, (0.0, w / 2, w / 2)),
        n_hist=n_hist, seed=seed,
        range_bins=Bins(0.0, 10.0, 2000), vac_bins=Bins(0.0, 10.0, 2000),
        lat_bw=10.0, lat_nb=400, lat_y0=w / 2, lat_z0=w / 2)


def c3_ag_layers(n_hist: int = 1000000, seed: int = 2344) -> Problem:
    """300 keV Ag into 50 nm Ag / 200 nm Cu / 5 um Si (10.5 / 8.96 / 2.33 g/cm3)."""
    t = (500.0, 2000.0, 50000.0)
    w = (sum(t), 10000.0, 10000.0)
    return Problem(
        name="c3_ag_layers_300keV",
        materials=[Material(10.5, [Element(47, 107.87, 1.0, 25.0, 3.0)]),
                   Material(8.96, [Element(29, 63.546, 1.0, 25.0, 3.0)]),
                   Material(2.33, [Element(14, 28.086, 1.0, 25.0, 3.0)])],
        geometry=Geometry(GEO_LAYERS, w, (CUT, PBC, PBC), layers=t),
        primaries=Primaries.single(47, 107.87, 3.0e5, (0.0, w[1] / 2, w[2] / 2)),
        n_hist=n_hist, seed=seed,
        range_bins=Bins(0.0, 10.0, 2000), vac_bins=Bins(0.0, 10.0, 2000),
        lat_bw=10.0, lat_nb=400, lat_y0=w[1] / 2, lat_z0=w[2] / 2)


def random_bubbles(w: float, n: int, rmin: float, rmax: float, sep: float,
                   seed: int) -> np.ndarray:
    """Non-overlapping spheres with surface separation >= sep in a periodic cube,
    placed by rejection like sampleClusters::addRandomClusters
    (sample_clusters.C:227-252), with radii uniform in [rmin, rmax].

    Candidates are drawn in chunks and taken in order: a candidate is kept
    when it overlaps neither a sphere kept before the chunk nor an earlier
    kept candidate of the chunk (minimum-image distance under
    r_a + r_b + sep), until n are kept."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    out = np.zeros((0, 4))

    def overlaps(a, b):
        d = a[:, :3] - b[:, :3]
        d -= np.round(d / w) * w
        return np.einsum("ij,ij->i", d, d) < (a[:, 3] + b[:, 3] + sep) ** 2

    while len(out) < n:
        k = min(4 * (n - len(out)) + 64, 200000)
        cand = np.column_stack([rng.uniform(0, w, (k, 3)), rng.uniform(rmin, rmax, k)])
        bad = np.zeros(k, bool)
        if len(out):
            tree = cKDTree(out[:, :3], boxsize=w)
            near = tree.query_ball_point(cand[:, :3], cand[:, 3] + rmax + sep)
            ci = np.repeat(np.arange(k), [len(x) for x in near])
            if len(ci):
                oj = np.concatenate([np.asarray(x, np.int64) for x in near])
                bad[ci[overlaps(out[oj], cand[ci])]] = True
        # conflicts among the chunk's own candidates (pairs i < j)
        pairs = cKDTree(cand[:, :3], boxsize=w).query_pairs(2 * rmax + sep, output_type="ndarray")
        later = [[] for _ in range(k)]
        if len(pairs):
            i, j = pairs[:, 0], pairs[:, 1]
            lo, hi = np.minimum(i, j), np.maximum(i, j)
            hit = overlaps(cand[lo], cand[hi])
            for a, b in zip(lo[hit], hi[hit]):
                later[b].append(a)
        kept = np.zeros(k, bool)
        need = n - len(out)
        taken = []
        for c in range(k):
            if bad[c] or any(kept[a] for a in later[c]):
                continue
            kept[c] = True
            taken.append(c)
            if len(taken) >= need:
                break
        out = np.vstack([out, cand[taken]])
    return out[:n]


def c4_mixed_uo2(n_hist: int = 1000000, seed: int = 2344, w: float = 10000.0,
                 bubble_density: float = 1.0e-7, geometry_seed: int = -1) -> Problem:
    """H, He, C, Xe ions, energies log-uniform in 10 keV..2 MeV, started at
    uniform random points with isotropic directions in periodic UO2
    (10.96 g/cm3) holding Xe bubbles (3.5 g/cm3, r in 10..50 A, 1e17 cm^-3).
    The primaries and the random streams come from `seed`, the bubbles from
    `geometry_seed` (default seed + 1)."""
    rng = np.random.default_rng(seed)
    n_b = int(round(bubble_density * w ** 3))
    bubbles = random_bubbles(w, n_b, 10.0, 50.0, 5.0, seed + 1 if geometry_seed < 0 else geometry_seed)
    species = np.array([(1, 1.008), (2, 4.0026), (6, 12.011), (54, 131.29)])
    pick = rng.integers(0, 4, n_hist)
    d = rng.normal(size=(n_hist, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    prim = Primaries(species[pick, 0].astype(np.int32), species[pick, 1].copy(),
                     log_uniform(rng, 1.0e4, 2.0e6, n_hist), rng.uniform(0, w, (n_hist, 3)), d)
    kn = max(1, int(w / 100.0))
    return Problem(
        name="c4_mixed_uo2_bubbles",
        materials=[Material(10.96, [Element(92, 238.03, 1.0, 40.0, 3.0),
                                    Element(8, 15.999, 2.0, 20.0, 3.0)]),
                   Material(3.5, [Element(54, 131.29, 1.0, 25.0, 3.0)])],
        geometry=Geometry(GEO_CLUSTERS, (w, w, w), (PBC, PBC, PBC), kn=(kn, kn, kn),
                          clusters=bubbles),
        primaries=prim, n_hist=n_hist, seed=seed,
        range_bins=Bins(-w, 20.0, 1000), vac_bins=Bins(0.0, 10.0, 1000),
        lat_bw=20.0, lat_nb=1000, lat_y0=0.0, lat_z0=0.0)


def c5_c_w(n_hist: int = 1000000, seed: int = 2344) -> Problem:
    """2 MeV C into a 2 um W cube (19.3 g/cm3), Ed 40 eV, Eb 3 eV."""
    w = 20000.0
    return Problem(
        name="c5_c_w_2MeV",
        materials=[Material(19.3, [Element(74, 183.84, 1.0, 40.0, 3.0)])],
        geometry=Geometry(GEO_SOLID, (w, w, w), (CUT, PBC, PBC)),
        primaries=Primaries.single(6, 12.011, 2.0e6, (0.0, w / 2, w / 2)),
        n_hist=n_hist, seed=seed,
        range_bins=Bins(0.0, 10.0, 2000), vac_bins=Bins(0.0, 10.0, 2000),
        lat_bw=10.0, lat_nb=400, lat_y0=w / 2, lat_z0=w / 2)


def vacancy_case(ion: str, energy_eV: float, n_hist: int, seed: int = 2344) -> Problem:
    """Infinite medium of validation/vacancy_count/{cu_on_cu,si_on_c,xe_on_u}_vac.json."""
    ions = {"cu": ((29, 63.546), (8.92, Element(29, 63.546))),
            "si": ((14, 27.977), (2.253, Element(6, 12.01))),
            "xe": ((54, 131.904), (19.043, Element(92, 238.0)))}
    (Z, m), (rho, el) = ions[ion]
    t = 100000.0
    return Problem(
        name="vacancy_%s_%gkeV" % (ion, energy_eV / 1000),
        materials=[Material(rho, [el])],
        geometry=Geometry(GEO_LAYERS, (t, 100.0, 100.0), (PBC, PBC, PBC), layers=(t,)),
        primaries=Primaries.single(Z, m, energy_eV, (0.0, 50.0, 50.0)),
        n_hist=n_hist, seed=seed,
        range_bins=Bins(0.0, 10.0, 1000), vac_bins=Bins(0.0, 10.0, 1000),
        lat_bw=10.0, lat_nb=400, lat_y0=50.0, lat_z0=50.0)


ALL = {"c1": c1_he_fe, "c2": c2_cu_cu, "c3": c3_ag_layers, "c4": c4_mixed_uo2, "c5": c5_c_w}


def wire_b_si(n_hist: int = 100000, seed: int = 2344, d: float = 500.0, length: float = 2000.0) -> Problem:
    """320 keV B onto the top face of a Si nanowire (SampleWire: a cylinder of
    diameter d along z, length `length`), as apps/mytrim_wire2.C:70, 92-93, 120-128
    at theta = 0: CUT on every axis, the ion enters at the centre of the top."""
    return Problem(
        name="wire_b_si_320keV",
        materials=[Material(2.329, [Element(14, 28.0, 1.0)])],
        geometry=Geometry(GEO_WIRE, (d, d, length), (CUT, CUT, CUT)),
        primaries=Primaries.single(5, 11.0, 320.0e3, (d / 2, d / 2, 0.0), (0.0, 0.0, 1.0)),
        n_hist=n_hist, seed=seed,
        range_bins=Bins(0.0, 10.0, 400), vac_bins=Bins(0.0, 10.0, 400),
        lat_bw=10.0, lat_nb=100, lat_y0=d / 2, lat_z0=0.0)


def burried_wire_p_si(n_hist: int = 100000, seed: int = 2344, d: float = 500.0,
                      length: float = 2000.0) -> Problem:
    """250 keV P through the 250 A SiO2 cover onto a buried Si nanowire
    (SampleBurriedWire), as apps/mytrim_wire2.C:74, 89, 120-142, 182-189 with the
    ion started above the wire axis; every axis INF."""
    return Problem(
        name="burried_wire_p_si_250keV",
        materials=[Material(2.329, [Element(14, 28.0, 1.0)]),
                   Material(2.634, [Element(14, 28.0, 1.0), Element(8, 16.0, 2.0)])],
        geometry=Geometry(GEO_BURRIED_WIRE, (d, d, length), (INF, INF, INF)),
        primaries=Primaries.single(15, 31.0, 250.0e3, (d / 2, d / 2, -250.0), (0.0, 0.0, 1.0)),
        n_hist=n_hist, seed=seed,
        range_bins=Bins(-250.0, 10.0, 400), vac_bins=Bins(-250.0, 10.0, 400),
        lat_bw=10.0, lat_nb=100, lat_y0=d / 2, lat_z0=-250.0)
