"""``distCUDA2(points) -> mean squared distance to the 3 nearest other points``,
as used at R/scene/gaussian_model.py:175-178 and :730-733 (the caller clamps at
1e-7 and takes log(sqrt(.)) as the initial log-scale).  Exact, on the GPU."""
from gvc.sampling import knn3_mean_sq


def distCUDA2(points):
    return knn3_mean_sq(points)
