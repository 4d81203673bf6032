"""The spherical region of interest and its .bild file (ROI.ml of the reference).

A .bild ROI file holds one ``.sphere x y z r`` line among Chimera directives (other
lines are ignored); ``from_bild`` reads it as ROI.from_bild does (src/ROI.ml:20-32: the
lines starting with ".sphere ", exactly one of them, else an error), ``to_bild`` writes
the three lines of ROI.to_bild (src/ROI.ml:41-45; coordinates with "%g").  The file is in
the receptor's original coordinates: callers that move the receptor (the reference
centres it in its simulation box) move the centre by the same offset.
"""
import numpy as np


def from_bild(fn):
    """(centre (3,), radius) of the single ``.sphere`` line of a .bild file."""
    with open(fn) as f:
        lines = [l for l in f.read().splitlines() if l.startswith(".sphere ")]
    if len(lines) != 1:
        raise ValueError("%s: expected exactly one '.sphere x y z r' line, found %d"
                         % (fn, len(lines)))
    fields = lines[0].split()[1:]
    if len(fields) < 4:
        raise ValueError("%s: '.sphere' line without x y z r: %r" % (fn, lines[0]))
    x, y, z, r = (float(v) for v in fields[:4])
    return np.array([x, y, z]), r


def to_bild(fn, center, radius):
    c = np.asarray(center, dtype=np.float64)
    with open(fn, "w") as f:
        f.write(".transparency 0.8\n.color blue\n.sphere %g %g %g %g\n" % (c[0], c[1], c[2], radius))


def is_inside(center, radius, p):
    """ROI.is_inside: squared distance to the centre below radius^2 (src/ROI.ml:65-66)."""
    d = np.asarray(p, dtype=np.float64) - np.asarray(center, dtype=np.float64)
    return float(d @ d) < radius * radius
