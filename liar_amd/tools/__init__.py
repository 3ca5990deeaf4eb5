"""Scene tools mirroring ``liar.tools`` (src/liar/tools/): spectral data files
(``spd``), the Cornell box (``cornell_box``), script helpers (``scripting``),
and the Wavefront OBJ / pbrt-v3 importers."""
