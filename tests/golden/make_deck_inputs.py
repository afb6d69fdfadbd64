"""Write the m.in of the golden deck (tests/golden/deck_fm): m along z tilted
by up to 0.3 rad per cell, seeded, in the reference's per-cell text format
(inoutput.cpp:11-27 reads it; the lines follow output_m's layout,
inoutput.cpp:337-344).  Formatting is done here in Python, independently of
the library's writer.  Run from the repository root."""
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "deck_fm" / "m.in"


def main():
    shape = (8, 8, 4)
    rng = np.random.default_rng(20)
    th = 0.3 * rng.random(shape)
    ph = 2 * np.pi * rng.random(shape)
    m = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    lines = []
    for i in range(shape[0]):
        for j in range(shape[1]):
            for k in range(shape[2]):
                lines.append("%d  %d  %d  %.16e  %.16e  %.16e\n" % (i + 1, j + 1, k + 1, *m[:, i, j, k]))
    OUT.write_text("".join(lines))


if __name__ == "__main__":
    main()
