"""Standard observer: CIE 1931 2-degree colour matching functions, 360-830 nm
at 5 nm (CVRL, http://www-cvrl.ucsd.edu/cmfs.htm), loaded at import as
LiAR's package does (src/liar/__init__.py:51-69), and ``Observer`` with
``setStandard`` to replace it (src/liar/kernel/observer.cpp:57-129)."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from . import _native as N

DATA = Path(__file__).resolve().parent / "data" / "cie1931_2deg_observer.tsv"
_bundled = None
_standard = None   # (w nm float32 (n,), xyz float32 (n, 3)) installed by Observer.setStandard
_installed = None  # the table last handed to the native library


class Observer:
    """``liar.Observer(wavelengths, sensitivities)``: wavelengths in metres
    or nanometres (any value > 1: nanometres, the loader's rule), one
    (xbar, ybar, zbar) triple per wavelength.  Y is normalised by the
    library so that a unit spectrum has Y = 1 (observer.cpp:105-123)."""

    def __init__(self, wavelengths, sensitivities):
        w = np.asarray(wavelengths, np.float64).reshape(-1)
        if not (w > 1).any():
            w = w * 1e9
        xyz = np.asarray(sensitivities, np.float64).reshape(-1, 3)
        if len(w) < 2 or len(xyz) != len(w):
            raise ValueError("Observer needs >= 2 wavelengths and one (x, y, z) triple per wavelength")
        if not (np.diff(w) > 0).all():
            raise ValueError("Observer wavelengths must be strictly increasing")
        self.wavelengths = w.astype(np.float32)  # nm
        self.sensitivities = np.ascontiguousarray(xyz, np.float32)

    @staticmethod
    def setStandard(observer: "Observer") -> None:
        """Makes ``observer`` the standard observer of every later render
        (``liar.Observer.setStandard``)."""
        global _standard
        if not isinstance(observer, Observer):
            raise TypeError("setStandard takes an Observer")
        _standard = (observer.wavelengths, observer.sensitivities)

    @staticmethod
    def standard() -> "Observer":
        w, xyz = standard_observer()
        return Observer(w, xyz)


def _load_bundled():
    global _bundled
    if _bundled is None:
        rows = [l.split("\t") for l in DATA.read_text().splitlines() if l and not l.startswith("#")]
        head, body = rows[0], rows[1:]
        cols = {name.strip(): np.array([float(r[i]) for r in body]) for i, name in enumerate(head)}
        w = cols["w"].astype(np.float32)
        xyz = np.stack([cols["xbar"], cols["ybar"], cols["zbar"]], 1).astype(np.float32)
        _bundled = (w, np.ascontiguousarray(xyz))
    return _bundled


def standard_observer():
    """The current standard observer: (wavelengths nm (n,), xyz (n, 3))
    float32 -- the bundled CIE 1931 table unless Observer.setStandard
    replaced it."""
    return _standard if _standard is not None else _load_bundled()


def ensure_standard_observer():
    """Hands the current standard observer to the native library (once per
    change)."""
    global _installed
    cur = standard_observer()
    if _installed is not cur:
        w, xyz = cur
        N.check(N.lib().lr_set_standard_observer(N.ptr(w), N.ptr(xyz), len(w)), "lr_set_standard_observer")
        _installed = cur
