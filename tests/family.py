"""The 3-sigma comparison of a family of entries against the reference's runs.

With standard errors estimated from a finite number of runs on each side, every entry's
z = |mean - mean_ref| / sqrt(s^2/n + s_ref^2/n_ref) follows Student's t, so a fixed 3.0 cut
on each of many entries fails by chance now and then.  The family passes when the number of
entries beyond 3 stays within the 99.9th percentile of Binomial(n, P(|t| > 3)) and the largest
z within the bound a family of n entries passes with probability 99.9 %.
"""
import numpy as np
from scipy import stats


def family_ok(z, dof):
    """(passes, detail) for a family of 3-sigma comparisons whose z values are Student-t(dof)."""
    z = np.asarray(z).ravel()
    n = z.size
    allowed = stats.binom.ppf(0.999, n, 2 * stats.t.sf(3.0, dof))
    z_max = stats.t.isf(0.0005 / n, dof)
    over = int(np.sum(z > 3.0))
    return over <= allowed and z.max() <= z_max, (n, over, allowed, round(float(z.max()), 2), round(float(z_max), 2))
