"""Group data (SU(2) Clebsch-Gordan tensors)."""
