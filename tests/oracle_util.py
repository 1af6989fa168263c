"""Helpers shared by the oracle tests (test infrastructure)."""
import numpy as np


def as_build_arrays(res):
    """Present an oracle result (e.g. the FP32 one) in the layout of parity.gpu_arrays."""
    out = {
        "K": res.K,
        "s2i": res["sorted_to_initial"],
        "nb_count": res["nb_count"],
        "nb_ids": res["nb_ids"],
        "status": res["status"],
        "concave_ids": res["concave_ids"],
        "concave_pos": res["concave_pos"],
        "concave_nb": res["concave_nb"],
        "max_nb": 32,
        "counts": {"concave_with_neighbors_nb": res.iw},
    }
    if res.mode == 0:
        out.update(full_circles=res["full_circles"], sectors=res["sectors"], convex_patches=res["convex_patches"],
                   segments=res["segments"], concave_pos_w=res["concave_pos"][:, 3])
    return out


def tetrahedron(edge=3.0, r=1.6):
    v = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], np.float64) * edge / (2 * np.sqrt(2))
    return np.concatenate([v, np.full((4, 1), r)], 1).astype(np.float32)
