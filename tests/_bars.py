"""Accuracy bars shared by the GPU tests.

Mixed precision (FP32 transforms and products, FP64 compensated sums) is held
to 1e-5 relative error per polygon against the FP64 path.  The rounding
errors of the filtered fields enter a polygon's sum S = sum_x prod D(x) on the
scale of sum_x |prod D(x)| = kappa |S|, so where B crosses zero (kappa large)
the relative error of S grows with kappa whatever the algorithm: FP32 data
carry sum |prod D| to about u32 = 2^-24 only.  A polygon therefore meets the
bar when its relative error is at most 1e-5, or at most two FP32 units of
sum |prod D| (rel <= 2 u32 kappa); the second form is the looser one only for
kappa > 1e-5 / (2 u32) = 84.
"""
import numpy as np

MIXED_REL = 1e-5
U32 = 2.0 ** -24


def mixed_within_bar(rel, kappa):
    """Per polygon: rel <= 1e-5 or rel <= 2 u32 kappa."""
    rel = np.asarray(rel, dtype=np.float64)
    kappa = np.asarray(kappa, dtype=np.float64)
    return rel <= np.maximum(MIXED_REL, 2.0 * U32 * kappa)


def report(name, rel, kappa):
    rel = np.asarray(rel)
    kappa = np.asarray(kappa)
    i = int(np.argmax(rel))
    j = int(np.argmax(rel / kappa))
    print(f"{name}: mixed rel err max {rel.max():.3e} (kappa {kappa[i]:.3e}), median {np.median(rel):.3e}, "
          f"kappa max {kappa.max():.3e}, {int(np.sum(rel > MIXED_REL))} above 1e-5, "
          f"largest rel/(u32 kappa) {rel[j] / (U32 * kappa[j]):.3f}")
