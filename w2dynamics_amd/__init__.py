"""MI355X-native CT-HYB sampling core (drop-in for w2dynamics' src/ctqmc_fortran
Z-space sampling path)."""

__version__ = "0.1.0"
