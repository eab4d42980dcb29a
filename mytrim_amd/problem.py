"""Cascade problem description shared by the Python bindings, tests and bench.

A :class:`Problem` holds what a MyTRIM driver sets up before its drain loop:
the materials (``MaterialBase`` + ``Element``, material.h:34-75,
element.h:29-41), the sample geometry (``SampleSolid`` / ``SampleLayers`` /
``sampleClusters``, sample*.h), the ``SimconfType`` knobs (simconf.C:44-49),
the primary ions and the tally bins.  Nothing here computes physics.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np

PBC, INF, CUT = 0, 1, 2
GEO_SOLID, GEO_LAYERS, GEO_CLUSTERS, GEO_WIRE, GEO_BURRIED_WIRE = 0, 1, 2, 3, 4
UNIVERSAL, MOLIERE, CKR = 0, 1, 2


@dataclasses.dataclass
class Element:
    Z: int
    m: float
    t: float = 1.0
    Edisp: float = 25.0   # element.C:25 default
    Elbind: float = 3.0


@dataclasses.dataclass
class Material:
    rho: float            # g/cm^3
    elements: List[Element]


@dataclasses.dataclass
class Geometry:
    kind: int = GEO_SOLID
    w: Sequence[float] = (10000.0, 10000.0, 10000.0)
    bc: Sequence[int] = (PBC, PBC, PBC)
    layers: Sequence[float] = ()                  # SampleLayers::layerThickness
    kn: Sequence[int] = (1, 1, 1)                 # sampleClusters spatial hash
    clusters: Optional[np.ndarray] = None         # (n, 4): x y z r


@dataclasses.dataclass
class Primaries:
    """Either one ion replicated ``n_hist`` times or one entry per history."""
    Z: np.ndarray
    m: np.ndarray
    E: np.ndarray
    pos: np.ndarray   # (k, 3)
    dir: np.ndarray   # (k, 3)

    @property
    def uniform(self) -> bool:
        return len(self.Z) == 1

    @staticmethod
    def single(Z, m, E, pos, dir=(1.0, 0.0, 0.0)) -> "Primaries":
        return Primaries(np.array([Z], np.int32), np.array([m], np.float64),
                         np.array([E], np.float64),
                         np.array([pos], np.float64), np.array([dir], np.float64))


@dataclasses.dataclass
class Bins:
    x0: float = 0.0
    bw: float = 10.0
    nb: int = 100


@dataclasses.dataclass
class Problem:
    materials: List[Material]
    geometry: Geometry
    primaries: Primaries
    n_hist: int
    seed: int = 1
    hist0: int = 0
    tmin: float = 0.2
    tau: float = 0.0
    cw: float = 0.001
    scale: float = 1.0
    Ef: float = 3.0
    potential: int = UNIVERSAL
    follow_max_gen: int = -1      # < 0: TrimBase (follow all); else TrimPrimaries/TrimRecoils
    kp: bool = False              # modified Kinchin-Pease at gen == follow_max_gen
    range_bins: Bins = dataclasses.field(default_factory=lambda: Bins(0.0, 10.0, 1000))
    vac_bins: Bins = dataclasses.field(default_factory=lambda: Bins(0.0, 10.0, 1000))
    lat_bw: float = 10.0
    lat_nb: int = 400
    lat_y0: float = 0.0
    lat_z0: float = 0.0
    name: str = ""

    def with_(self, **kw) -> "Problem":
        return dataclasses.replace(self, **kw)

    def batch(self, first: int, count: int) -> "Problem":
        """Histories first .. first+count-1 of this problem (hist0 + first on):
        the same random streams, and for per-history primaries the rows of
        those histories."""
        if first < 0 or count < 0 or first + count > self.n_hist:
            raise ValueError("batch %d+%d outside %d histories" % (first, count, self.n_hist))
        p = self.primaries
        if not p.uniform:
            sl = slice(first, first + count)
            p = Primaries(p.Z[sl], p.m[sl], p.E[sl], p.pos[sl], p.dir[sl])
        return dataclasses.replace(self, primaries=p, hist0=self.hist0 + first, n_hist=count)

    def primary_energies(self) -> np.ndarray:
        """E0 of every history of the problem."""
        p = self.primaries
        return np.full(self.n_hist, float(p.E[0])) if p.uniform else np.asarray(p.E[:self.n_hist], float)

    # flat views ----------------------------------------------------------
    def flat_materials(self):
        rho = np.array([m.rho for m in self.materials], np.float64)
        off = [0]
        Z, mm, t, ed, eb = [], [], [], [], []
        for mat in self.materials:
            for e in mat.elements:
                Z.append(e.Z)
                mm.append(e.m)
                t.append(e.t)
                ed.append(e.Edisp)
                eb.append(e.Elbind)
            off.append(len(Z))
        return (rho, np.array(off, np.int32), np.array(Z, np.int32),
                np.array(mm, np.float64), np.array(t, np.float64),
                np.array(ed, np.float64), np.array(eb, np.float64))

    def spec_text(self, pkafile: Optional[str] = None) -> str:
        """Key/value text read by the reference fixture harness.  Per-history
        primaries go to `pkafile` (rows Z m E x y z dx dy dz, row i for global
        history i), which the caller writes with :meth:`primary_rows`."""
        g = self.geometry
        kind = {GEO_SOLID: "solid", GEO_LAYERS: "layers", GEO_CLUSTERS: "clusters", GEO_WIRE: "wire",
                GEO_BURRIED_WIRE: "burriedwire"}[g.kind]
        out = ["geo " + kind,
               "w %r %r %r" % tuple(float(x) for x in g.w),
               "bc %d %d %d" % tuple(g.bc)]
        for mat in self.materials:
            s = "mat %r %d" % (float(mat.rho), len(mat.elements))
            for e in mat.elements:
                s += " %d %r %r %r %r" % (e.Z, float(e.m), float(e.t), float(e.Edisp), float(e.Elbind))
            out.append(s)
        if g.layers:
            out.append("layers " + " ".join(repr(float(t)) for t in g.layers))
        if g.kind == GEO_CLUSTERS:
            out.append("kn %d %d %d" % tuple(g.kn))
            for c in g.clusters:
                out.append("cluster %r %r %r %r" % tuple(float(v) for v in c))
        p = self.primaries
        if not p.uniform:
            if pkafile is None:
                raise ValueError("per-history primaries need a pkafile")
            out.append("pkafile " + pkafile)
        out += ["tmin %r" % self.tmin, "tau %r" % self.tau, "cw %r" % self.cw,
                "scale %r" % self.scale, "Ef %r" % self.Ef, "potential %d" % self.potential,
                "pka %d %r %r %r %r %r %r %r %r" % (
                    int(p.Z[0]), float(p.m[0]), float(p.E[0]), *[float(v) for v in p.pos[0]],
                    *[float(v) for v in p.dir[0]]),
                "nhist %d" % self.n_hist, "hist0 %d" % self.hist0, "seed %d" % self.seed,
                "follow %d" % self.follow_max_gen, "kp %d" % int(self.kp),
                "range %r %r %d" % (self.range_bins.x0, self.range_bins.bw, self.range_bins.nb),
                "vac %r %r %d" % (self.vac_bins.x0, self.vac_bins.bw, self.vac_bins.nb),
                "lat %r %r %r %d" % (self.lat_y0, self.lat_z0, self.lat_bw, self.lat_nb)]
        return "\n".join(out) + "\n"

    def primary_rows(self) -> str:
        """Rows of the harness's pkafile for histories 0 .. hist0+n_hist-1 of a
        problem whose primaries start at history 0."""
        p = self.primaries
        if self.hist0 != 0:
            raise ValueError("primary_rows: problem must start at history 0")
        return "".join("%d %r %r %r %r %r %r %r %r\n" % (
            int(p.Z[i]), float(p.m[i]), float(p.E[i]), *[float(v) for v in p.pos[i]],
            *[float(v) for v in p.dir[i]]) for i in range(self.n_hist))


def log_uniform(rng: np.random.Generator, lo: float, hi: float, n: int) -> np.ndarray:
    return np.exp(rng.uniform(math.log(lo), math.log(hi), n))
