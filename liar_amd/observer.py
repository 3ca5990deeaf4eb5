"""Standard observer: CIE 1931 2-degree colour matching functions, 360-830 nm
at 5 nm (CVRL, http://www-cvrl.ucsd.edu/cmfs.htm), loaded like LiAR's
``liar.Observer.setStandard`` (src/liar/__init__.py:51-69)."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from . import _native as N

DATA = Path(__file__).resolve().parent / "data" / "cie1931_2deg_observer.tsv"
_cache = None
_installed = False


def standard_observer():
    """(wavelengths nm (n,), xyz (n, 3)) float32."""
    global _cache
    if _cache is None:
        rows = [l.split("\t") for l in DATA.read_text().splitlines() if l and not l.startswith("#")]
        head, body = rows[0], rows[1:]
        cols = {name.strip(): np.array([float(r[i]) for r in body]) for i, name in enumerate(head)}
        w = cols["w"].astype(np.float32)
        xyz = np.stack([cols["xbar"], cols["ybar"], cols["zbar"]], 1).astype(np.float32)
        _cache = (w, np.ascontiguousarray(xyz))
    return _cache


def ensure_standard_observer():
    global _installed
    if not _installed:
        w, xyz = standard_observer()
        N.check(N.lib().lr_set_standard_observer(N.ptr(w), N.ptr(xyz), len(w)), "lr_set_standard_observer")
        _installed = True
