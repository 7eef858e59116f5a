import numpy as np


def rel_err(a, b):
    """‖a − b‖₂ / ‖b‖₂ (the flux-moment error metric)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def group_rel_errs(phi, ref):
    return [rel_err(phi[g], ref[g]) for g in range(ref.shape[0])]
