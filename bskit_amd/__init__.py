"""bskit_amd: MI355X-native FFT bispectrum estimator (drop-in for bskit's FFTBispectrum)."""
from .bins import *  # noqa: F401,F403
