"""Scene import helpers mirroring ``liar.tools`` (src/liar/tools/)."""
