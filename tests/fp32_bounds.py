"""The error bound an FP32 state meets against FP64, in the form the tests
assert it.

Rounding the stored state to FP32 injects noise of about eps32 = 6e-8 of
each field per half-step.  That noise is white in space, so the transport
spreads it over every moment order at a similar absolute level, while the
solution's own high-order moments are orders of magnitude smaller than its
low ones (the angular distribution is smooth).  The error of a moment field
is therefore bounded in absolute terms, by a small multiple of eps32 times
the largest field's norm, and its error relative to its own norm grows as
the field gets smaller.  The bound checked here:

  * relative L2 over all moments (the whole state) <= 1e-5;
  * relative L2 per moment field <= 1e-5 for every field whose norm is at
    least 1e-2 of the largest field's;
  * L2 error of every field <= 1e-6 of the largest field's norm.
"""
import numpy as np

ALL_MOMENTS = 1e-5
PER_FIELD = 1e-5
PER_FIELD_ABOVE = 1e-2
OF_LARGEST = 1e-6


def fp32_errors(got, ref):
    """got, ref: sequences of arrays per moment field (FP32 result, FP64
    reference).  Returns the error measures of the bound."""
    norms = np.array([np.linalg.norm(r) for r in ref])
    errs = np.array([np.linalg.norm(np.asarray(g, dtype=float) - r) for g, r in zip(got, ref)])
    big = norms.max()
    rel = np.where(norms > 0, errs / np.where(norms > 0, norms, 1.0), np.where(errs > 0, np.inf, 0.0))
    large = norms >= PER_FIELD_ABOVE * big
    return {
        "all_moments": float(np.sqrt((errs ** 2).sum() / (norms ** 2).sum())),
        "per_field_max": float(rel.max()),
        "per_field_max_norm_over_largest": float(norms[int(rel.argmax())] / big),
        "per_field_max_above_1e-2": float(rel[large].max()),
        "of_largest_max": float(errs.max() / big),
    }


def check_fp32_bound(got, ref):
    e = fp32_errors(got, ref)
    assert e["all_moments"] <= ALL_MOMENTS, e
    assert e["per_field_max_above_1e-2"] <= PER_FIELD, e
    assert e["of_largest_max"] <= OF_LARGEST, e
    return e
