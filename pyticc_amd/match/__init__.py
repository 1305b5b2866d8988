"""Asymptotic matching on the MI355X backend."""

from pyticc_amd.match.asymptotic import get_Bmat_BF_to_SF, transform_logD_BF_to_SF
from pyticc_amd.match.bessel import modified_bessel_K_logD, riccati_bessel_jy
from pyticc_amd.match.smatrix import get_Smat

__all__ = ["get_Bmat_BF_to_SF", "get_Smat", "modified_bessel_K_logD", "riccati_bessel_jy", "transform_logD_BF_to_SF"]
