"""Spectral power distribution files (mirrors ``liar.tools.spd``,
src/liar/tools/spd.py:27-75): whitespace-separated columns, the first the
wavelength (nanometres if any value exceeds 1, else metres), every other one
a spectrum; lines starting with '#' or ';' are comments."""
from __future__ import annotations

import io
from pathlib import Path

import numpy as np

from ..spectra import Sampled

DATA = Path(__file__).resolve().parent.parent / "data"


def load(fp):
    """A list of Sampled spectra, one per value column of the file object."""
    rows = []
    for raw in fp:
        line = raw.strip()
        if not line or line[0] in "#;":
            continue
        rows.append([float(x) for x in line.split()])
    if not rows:
        raise ValueError("no spectral data")
    width = len(rows[0])
    if width < 2 or any(len(r) != width for r in rows):
        raise ValueError("every row needs the same number (>= 2) of columns")
    table = np.asarray(rows, np.float64)
    return [Sampled(table[:, 0], table[:, k]) for k in range(1, width)]


def loads(s):
    """``load`` of a string."""
    return load(io.StringIO(s))


def load_builtin_refractive_index(material):
    """(n, k) Sampled spectra of a bundled complex refractive index: 'Ag',
    'Al', 'Au' or 'Cu' (Hagemann, Gudat & Kunz 1974, DESY SR-74/7)."""
    path = DATA / f"{material}.spd"
    if not path.is_file():
        raise ValueError(f"Material '{material}' not found in built-in data.")
    with open(path) as fp:
        cols = load(fp)
    if len(cols) != 2:
        raise ValueError(f"{path.name}: expected the columns w, n, k")
    return cols[0], cols[1]


def dump(obj, fp):
    """Writes one Sampled or a list of Sampled on equal wavelengths as
    columns (wavelengths in nanometres)."""
    spectra = [obj] if isinstance(obj, Sampled) else list(obj)
    if not spectra:
        raise ValueError("Must have a list of at least one spectrum")
    w = np.asarray(spectra[0].wavelengths, np.float64)
    for s in spectra[1:]:
        if not np.array_equal(np.asarray(s.wavelengths, np.float64), w):
            raise ValueError("All spectra must be sampled on the same wavelengths. Resample before dumping.")
    for i, wl in enumerate(w):
        fp.write(" ".join([repr(float(wl))] + [repr(float(s.values[i])) for s in spectra]) + "\n")
