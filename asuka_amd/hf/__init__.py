"""Hartree-Fock drivers over the direct J/K engine."""
