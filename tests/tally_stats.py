"""Two-sample comparison of the cascade tallies of two independent runs.

One sample is the compiled reference (std::mt19937 streams, one seed per
history; per-history tallies and batched depth histograms in
tests/golden/stat_*.npz, written by tests/golden/make_fixtures.py --stat).
The other is the device draw scheme (Philox streams, direct azimuth), run on
the GPU or in the oracle, in K equal batches of histories.  Every quantity is
a ratio of sums over independent units (histories, or batches of them), so
its variance comes from the units themselves.

The tallies compared are those of the reference's tally hooks: vacancies
(trim.C:439-443), collisions (one checkPKAState per collision, trim.C:418),
the electronic-loss sum EelTotal (trim.C:286-289), the phonon ledger of
TrimPhononOut (trim.C:503-527), the vacancy depth bins of TrimVacCount
(apps/src/TrimVacCount.C:31-53) and the end point of the primary ion
(apps/mytrim_layers.C:171-186).

Acceptance criteria (`check`):

1. Scalars: vacancies per ion, collisions per history, Eel/E0, Enuc/E0 and the
   mean projected range; z = (b - a) / sqrt(var a + var b).  The first four
   are strongly correlated (a larger cascade has more of each), so a chance
   fluctuation moves them together and "each within 2 sigma" cannot be
   asserted one by one.  The criteria:
   * the four jointly: chi2 = d' C^-1 d with C the covariance of the
     difference (linearised residuals of the ratio estimators over histories
     and batches), p >= 1e-3 (chi-square, 4 degrees of freedom);
   * every |z| <= 3.3: the two-sided Bonferroni bound for the 45 scalar
     comparisons of one suite (9 workloads x 5 scalars) at a 5 % family-wise
     rate;
   * the number of scalars within 2 sigma is reported.
2. End-point histograms (projected range, lateral y, lateral z; at most one
   entry per history): adjacent bins merged until each expects >= 10 entries
   in the smaller sample (from the pooled counts), plus one category for the
   histories without an entry.  Pearson's chi-square test of homogeneity of
   the 2 x C table: p >= 1e-3.  Per merged bin the standardised difference
   of the two proportions (pooled variance): at most q(B) of the B bins
   beyond 2 sigma, q(B) the 99.9 % quantile of Binomial(B, 0.0455), and none
   beyond z_max(B) = Phi^-1(1 - 0.0005 / B), the Bonferroni bound over the B
   bins at 1e-3 per histogram (3.9 sigma at B = 10, 4.7 at B = 365).
3. Depth histograms (vacancies and replacements by depth: many entries per
   history, correlated within a history): at most 10 merged bins of about
   equal reference mass; the per-history means of each batch; Hotelling's
   two-sample T^2 with the pooled covariance of the batch means: p >= 1e-3
   (F distribution with B and nu - B + 1 degrees of freedom, nu = K_ref + K
   - 2).  Per merged bin z with the pooled variance: none beyond z_max(B)
   (the Bonferroni bound holds for correlated bins; the count of bins beyond
   2 sigma of 2 does not, and is only reported).
"""
from __future__ import annotations

import math
import os

import numpy as np
from scipy import stats

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

Z_REPORT = 2.0       # scalars within this many sigma are counted
Z_SCALAR_MAX = 3.3   # Bonferroni bound, 45 comparisons at 5 % family-wise
P_MIN = 1e-3         # chi-square / Hotelling p-value floor
MIN_EXPECTED = 10    # expected entries of the smaller sample per merged end-point bin
DEPTH_BINS = 10      # merged depth bins


def load_ref(name: str) -> dict:
    with np.load(os.path.join(GOLD, "stat_%s.npz" % name), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def ratio(Y, X):
    """R = sum Y / sum X over independent units and its variance (the
    linearised ratio estimator, n/(n-1) sum (Y - R X)^2 / (sum X)^2)."""
    R, e = ratio_resid(Y, X)
    return R, float(e @ e)


def ratio_resid(Y, X):
    """R = sum Y / sum X and the per-unit linearised residuals e, scaled so
    that e'e is the variance of R and e_a'e_b the covariance of two such
    estimates over the same units."""
    Y = np.asarray(Y, float)
    X = np.broadcast_to(np.asarray(X, float), Y.shape)
    n = len(Y)
    sx = X.sum()
    R = Y.sum() / sx
    if n < 2:
        # one unit (a single large run): its variance is not estimated; the
        # callers use this only where it is far below the other sample's
        return R, np.zeros(n)
    return R, math.sqrt(n / (n - 1)) * (Y - R * X) / sx


def q999(B: int) -> int:
    return int(stats.binom.ppf(0.999, B, 2 * stats.norm.sf(2.0)))


def z_max(B: int) -> float:
    return float(stats.norm.isf(0.0005 / B))


class Batches:
    """The other sample: K equal batches of histories, each a result dict of
    the engine (or the oracle) plus the batch's E0."""

    def __init__(self, results, n_per_batch, E0s):
        self.r = results
        self.m = n_per_batch
        self.K = len(results)
        self.n = self.K * self.m
        self.E0 = np.asarray(E0s, float)

    def col(self, k):
        return np.array([r[k] for r in self.r], float)

    def hist(self, k):
        return np.sum([np.asarray(r[k], np.int64) for r in self.r], axis=0)

    def hist_batches(self, k):
        return np.array([np.asarray(r[k], np.int64) for r in self.r])


def run_batches(run, prob, K):
    """Run `prob` in K equal batches through `run(problem) -> result dict`."""
    if prob.n_hist % K:
        raise ValueError("%d histories do not split into %d batches" % (prob.n_hist, K))
    m = prob.n_hist // K
    res, E0 = [], []
    for i in range(K):
        b = prob.batch(i * m, m)
        res.append(run(b))
        E0.append(float(b.primary_energies().sum()))
    return Batches(res, m, E0)


JOINT = ("vac_per_ion", "coll_per_hist", "Eel_over_E0", "Enuc_over_E0")


def scalars(ref: dict, oth: Batches):
    """({name: (reference value, other value, z)}, joint chi2, its p)"""
    one = np.ones(oth.K) * oth.m
    pairs = {
        "vac_per_ion": (ratio_resid(ref["h_vac"], 1.0), ratio_resid(oth.col("vac"), one)),
        "coll_per_hist": (ratio_resid(ref["h_coll"], 1.0), ratio_resid(oth.col("n_coll"), one)),
        "Eel_over_E0": (ratio_resid(ref["h_Eel"], ref["h_E0"]), ratio_resid(oth.col("Eel"), oth.E0)),
        "Enuc_over_E0": (ratio_resid(ref["h_Enuc"], ref["h_E0"]),
                         ratio_resid(oth.col("Enuc"), oth.E0)),
    }
    out = {}
    for k, ((a, ea), (b, eb)) in pairs.items():
        s = math.sqrt(float(ea @ ea + eb @ eb))
        out[k] = (a, b, (b - a) / s if s > 0 else (0.0 if a == b else math.inf))
    Ea = np.column_stack([pairs[k][0][1] for k in JOINT])
    Eb = np.column_stack([pairs[k][1][1] for k in JOINT])
    d = np.array([pairs[k][1][0] - pairs[k][0][0] for k in JOINT])
    C = Ea.T @ Ea + Eb.T @ Eb
    # scale to unit diagonal before solving (the scalars differ by orders of magnitude)
    sd = np.sqrt(np.diag(C))
    live = sd > 0
    dn, Cn = (d / np.where(live, sd, 1))[live], (C / np.outer(np.where(live, sd, 1), np.where(live, sd, 1)))[np.ix_(live, live)]
    chi2 = float(dn @ np.linalg.lstsq(Cn, dn, rcond=1e-12)[0])
    p = float(stats.chi2.sf(chi2, int(live.sum())))
    rn = int(ref["range_n"])
    if rn >= 30 and oth.col("range_n").sum() >= 30:
        m = float(ref["range_sum"]) / rn
        v = (float(ref["range_sum2"]) / rn - m * m) / (rn - 1)
        b, vb = ratio(oth.col("range_sum"), oth.col("range_n"))
        if oth.K < 2:
            vb = 0.0
        out["mean_range"] = (m, b, (b - m) / math.sqrt(v + vb))
    return out, chi2, p


def merge_min_count(counts, min_count):
    """Groups of adjacent bins, each holding >= min_count of `counts` (the
    last partial group joins the one before it); bins before the first
    entry join the first group."""
    groups, cur, acc = [], [], 0
    for i, c in enumerate(counts):
        cur.append(i)
        acc += c
        if acc >= min_count:
            groups.append(cur)
            cur, acc = [], 0
    if cur:
        if groups:
            groups[-1].extend(cur)
        else:
            groups.append(cur)
    return groups


def endpoint_hist(ref_counts, n_ref, oth_counts, n_oth):
    """Two-sample homogeneity test of an end-point histogram (criterion 2)."""
    a = np.asarray(ref_counts, np.int64)
    b = np.asarray(oth_counts, np.int64)
    # expected entries of the smaller sample, from the pooled counts
    frac = min(n_ref, n_oth) / float(n_ref + n_oth)
    if (a.sum() + b.sum()) * frac < 2 * MIN_EXPECTED:
        return None
    groups = merge_min_count((a + b) * frac, MIN_EXPECTED)
    ga = np.array([a[g].sum() for g in groups] + [n_ref - a.sum()])
    gb = np.array([b[g].sum() for g in groups] + [n_oth - b.sum()])
    keep = (ga + gb) > 0
    table = np.vstack([ga[keep], gb[keep]])
    chi2, p, dof, _ = stats.chi2_contingency(table, correction=False)
    pa, pb = ga[:-1] / n_ref, gb[:-1] / n_oth
    pp = (ga[:-1] + gb[:-1]) / (n_ref + n_oth)
    z = (pb - pa) / np.sqrt(np.maximum(pp * (1 - pp), 1e-300) * (1 / n_ref + 1 / n_oth))
    return {"chi2": float(chi2), "dof": int(dof), "p": float(p), "z": z, "B": len(groups)}


def equal_mass_groups(total, B):
    """Assign each bin to one of B groups of about equal mass of `total`."""
    t = np.asarray(total, float)
    s = t.sum()
    mid = (np.cumsum(t) - 0.5 * t) / s
    g = np.minimum((mid * B).astype(int), B - 1)
    # bins before the first and after the last entry go to the edge groups
    nz = np.nonzero(t)[0]
    g[:nz[0]] = g[nz[0]]
    g[nz[-1] + 1:] = g[nz[-1]]
    uniq = np.unique(g)
    return [np.nonzero(g == u)[0] for u in uniq]


def depth_hist(ref_batches, m_ref, oth_batches, m_oth):
    """Hotelling two-sample T^2 on merged depth bins (criterion 3)."""
    R = np.asarray(ref_batches, float)
    O = np.asarray(oth_batches, float)
    if R.sum() < 100:
        return None
    groups = equal_mass_groups(R.sum(0), DEPTH_BINS)
    Rg = np.column_stack([R[:, g].sum(1) for g in groups]) / m_ref
    Og = np.column_stack([O[:, g].sum(1) for g in groups]) / m_oth
    Kr, Ko = len(Rg), len(Og)
    n_r, n_o = Kr * m_ref, Ko * m_oth
    Sr = m_ref * np.atleast_2d(np.cov(Rg, rowvar=False, ddof=1))
    So = m_oth * np.atleast_2d(np.cov(Og, rowvar=False, ddof=1))
    nu = Kr + Ko - 2
    S = ((Kr - 1) * Sr + (Ko - 1) * So) / nu
    live = np.diag(S) > 0
    d = (Og.mean(0) - Rg.mean(0))[live]
    C = S[np.ix_(live, live)] * (1.0 / n_r + 1.0 / n_o)
    B = int(live.sum())
    T2 = float(d @ np.linalg.solve(C, d))
    F = (nu - B + 1) / (nu * B) * T2
    p = float(stats.f.sf(F, B, nu - B + 1))
    z = d / np.sqrt(np.diag(C))
    return {"T2": T2, "F": F, "p": p, "z": z, "B": B}


def check(name: str, ref: dict, oth: Batches, depth=("vac_hist", "repl_hist")):
    """Apply criteria 1-3; returns (failures, report lines)."""
    fails, lines = [], []
    sc, chi2, p = scalars(ref, oth)
    for k, (a, b, z) in sc.items():
        lines.append("%s %-14s ref %.6g  other %.6g  z %+.2f" % (name, k, a, b, z))
    within = sum(abs(z) <= Z_REPORT for (_, _, z) in sc.values())
    lines.append("%s scalars        within 2 sigma %d/%d; joint chi2 %.2f / %d  p %.3g" % (
        name, within, len(sc), chi2, len(JOINT), p))
    if p < P_MIN:
        fails.append("joint scalars p %.3g" % p)
    big = [k for k, (_, _, z) in sc.items() if not abs(z) <= Z_SCALAR_MAX]
    if big:
        fails.append("scalars beyond %g sigma: %s" % (Z_SCALAR_MAX, big))
    n_r = int(ref["n_hist"])
    for k in ("range_hist", "lat_y_hist", "lat_z_hist"):
        t = endpoint_hist(ref[k], n_r, oth.hist(k), oth.n)
        if t is None:
            lines.append("%s %-14s too few entries" % (name, k))
            continue
        nz = int(np.sum(np.abs(t["z"]) > 2.0))
        lines.append("%s %-14s chi2 %.1f / %d  p %.3g  bins>2sig %d/%d  max|z| %.2f" % (
            name, k, t["chi2"], t["dof"], t["p"], nz, t["B"], float(np.max(np.abs(t["z"])))))
        if t["p"] < P_MIN or nz > q999(t["B"]) or np.max(np.abs(t["z"])) > z_max(t["B"]):
            fails.append(k)
    for k in depth:
        t = depth_hist(ref[k + "_batches"], n_r // int(ref["batches"]), oth.hist_batches(k), oth.m)
        if t is None:
            lines.append("%s %-14s too few entries" % (name, k))
            continue
        nz = int(np.sum(np.abs(t["z"]) > 2.0))
        lines.append("%s %-14s T2 %.1f  B %d  p %.3g  bins>2sig %d/%d  max|z| %.2f" % (
            name, k, t["T2"], t["B"], t["p"], nz, t["B"], float(np.max(np.abs(t["z"])))))
        if t["p"] < P_MIN or np.max(np.abs(t["z"])) > z_max(t["B"]):
            fails.append(k)
    return fails, lines
