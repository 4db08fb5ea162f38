"""Numerical algorithms on the device, named as spectre.NumericalAlgorithms."""
