"""Writes simple_lut.txt: the reference's 256-entry simple-point table
(st_spix/csrc/bist/seg_utils.h:15-27), packed as 32 bytes (bit j of byte k
is entry 8k+j) in hex.  Run in a container that holds the reference
checkout; the fixture is the data, this script is how it was made."""
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/st_spix/csrc/bist/seg_utils.h"
src = open(ref).read()
body = re.search(r"bass_lut\[256\]\s*=\s*\{([^}]*)\}", src).group(1)
vals = [int(v) for v in body.replace("\n", " ").split(",") if v.strip()]
assert len(vals) == 256
packed = bytes(sum(vals[8 * k + j] << j for j in range(8)) for k in range(32)).hex()
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "simple_lut.txt")
with open(out, "w") as f:
    f.write(packed + "\n")
print(packed)
