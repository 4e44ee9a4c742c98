"""MI355X-native communication-avoiding QR (TSQR + Householder reconstruction
+ compact-WY trailing update) behind CANDMC's alg/QR interface."""
from ._lib import CQRError, load  # noqa: F401

__version__ = "0.1.0"
