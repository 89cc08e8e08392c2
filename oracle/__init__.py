"""Test infrastructure only: the FP64 CPU oracle of the smooth-particle path."""
