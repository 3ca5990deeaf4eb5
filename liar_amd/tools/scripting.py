"""Script helpers (mirrors ``liar.tools.scripting``,
src/liar/tools/scripting.py:26-120): command-line render options and a render
target.  There is no display output here, so the target is an Image."""
from __future__ import annotations


def makeRenderTarget(width, height, filename=None, title=None, display=True, fStops=None, rgbSpace=None,
                     toneMapping=None):
    """An output.Image of (width, height) saving to ``filename`` with LiAR's
    defaults (linear tone mapping, sRGB; ``fStops`` fixes the exposure)."""
    from .. import output
    from ..rgb_space import sRGB

    if not filename:
        raise ValueError("no display output is available: give a filename")
    image = output.Image(filename, (width, height))
    if fStops is not None:
        image.exposureStops = fStops
        image.autoExposure = False
    image.toneMapping = toneMapping or output.Image.ToneMapping.Linear
    image.rgbSpace = rgbSpace or sRGB
    return image


def renderOptions(argv=None, **defaults):
    """Parses ``--name value`` switches for every keyword (underscores as
    dashes), typed after its default; returns an object with those
    attributes."""
    import argparse

    parser = argparse.ArgumentParser()
    for name, default in defaults.items():
        kind = int if isinstance(default, int) else float if isinstance(default, float) else str
        parser.add_argument("--" + name.replace("_", "-"), dest=name, type=kind, default=default,
                            help=f"[default={default}]")
    options, _ = parser.parse_known_args(argv)
    return options
