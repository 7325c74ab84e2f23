"""Drop-in for ``simple_knn`` (not in the reference tree; its one call,
``from simple_knn._C import distCUDA2``, is at R/scene/gaussian_model.py:16)."""
