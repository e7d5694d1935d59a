"""Per-entry tolerances for FP32 results against FP64 references.

An FP32 evaluation whose longest chain of dependent roundings has n steps
differs from the exact value by at most about n * 2^-24 * M, where M is the
same computation carried out on the absolute values of every term (the
oracle's magnitude adjoints give M entry by entry).  The check is

    |gpu - ref| <= rtol * |ref| + n * 2^-24 * M      for every entry,

so an entry is allowed an absolute error only in proportion to its own
terms, not to the largest entry of the result.
"""
import numpy as np

U32 = 2.0 ** -24


def ea_chain_ops(obj_min, obj_max, step):
    """Dependent roundings of the EA adjoint: T and I are running products
    and sums over the ray's samples (at most diag / step + 1 of them), each
    sample's gradient term and exp2 attenuation a few more, and a voxel sums
    at most 64 sample terms.  2 * samples + 64."""
    diag = float(np.linalg.norm(np.asarray(obj_max, np.float64) - np.asarray(obj_min, np.float64)))
    return 2 * (int(np.ceil(diag / step)) + 1) + 64


def splat_ops(radius):
    """A point's gradient sums at most (2 ceil(r) + 1)^3 voxel terms of a few roundings each."""
    w = 2 * int(np.ceil(radius)) + 1
    return w ** 3 + 16


APRON_OPS = 16   # an interior voxel adds itself and at most 7 apron copies


def check_entries(got, ref, mag, n_ops, what, rtol=1e-5):
    """Assert |got - ref| <= rtol*|ref| + n_ops*2^-24*mag entry by entry;
    returns the largest error as a fraction of its bound (for the log)."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    mag = np.asarray(mag, np.float64)
    assert got.shape == ref.shape == mag.shape, (got.shape, ref.shape, mag.shape)
    assert (mag >= np.abs(ref) * (1 - 1e-12) - 1e-300).all(), f"{what}: magnitude below |reference|"
    bound = rtol * np.abs(ref) + n_ops * U32 * mag
    err = np.abs(got - ref)
    bad = err > bound
    ratio = float((err / np.where(bound > 0, bound, 1.0)).max()) if err.size else 0.0
    assert not bad.any(), (f"{what}: {int(bad.sum())} of {bad.size} entries off, worst at {np.argmax(err - bound)}: "
                           f"got {got.flat[np.argmax(err - bound)]:.9g} ref {ref.flat[np.argmax(err - bound)]:.9g}, "
                           f"worst error/bound {ratio:.3f}")
    return ratio
