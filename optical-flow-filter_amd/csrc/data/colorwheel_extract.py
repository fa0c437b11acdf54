"""Writes colorwheel_rgba.bin, the flow colour-coding wheel of FlowToColor.

The wheel is a data table of the reference: 256 x 256 RGBA texels, rows top
to bottom, the bytes of COLOR_WHEEL_D in src/colorwheel.cpp:12 (the same
image as python/flowfilter/rsc/colorWheel.png). FlowToColor's colours are
texels of this table, so they are identical to the reference's only with
these exact bytes: the table follows the Middlebury flow colour coding
(Baker et al., "A Database and Evaluation Methodology for Optical Flow",
2007) but cannot be regenerated from it bit for bit (its rim is
antialiased). This script reads the table's hexadecimal bytes from the
reference source and writes them unchanged as a raw binary file, which the
library embeds at build time.

    python colorwheel_extract.py [/root/reference]
"""
import hashlib
import os
import re
import sys

HEIGHT, WIDTH, DEPTH = 256, 256, 4
SHA256 = "966163576b67b0a8dfee44fd06e692f64572f7c92291fb4c2eae56c353ddea03"


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = open(os.path.join(ref, "src", "colorwheel.cpp")).read()
    start = src.index("COLOR_WHEEL_D[]")
    end = src.index("};", start)
    data = bytes(int(h, 16) for h in re.findall(r"0x([0-9A-Fa-f]{2})", src[start:end]))
    if len(data) != HEIGHT * WIDTH * DEPTH:
        raise SystemExit("unexpected table size %d" % len(data))
    if hashlib.sha256(data).hexdigest() != SHA256:
        raise SystemExit("table checksum differs from the one this build was made with")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "colorwheel_rgba.bin")
    with open(out, "wb") as fh:
        fh.write(data)
    print("wrote %s (%d bytes)" % (out, len(data)))


if __name__ == "__main__":
    main()
