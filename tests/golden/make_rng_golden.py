"""Write tests/golden/rng_kat.json from the reference's CLI regression test.

test/testmcx.sh:155 greps the .jnii written by `mcx --bench cube60 -d 0
--json '{"Shapes":[{"Grid":{"Tag":0,"Size":[2,2,2]}}]}' -F jnii -D R` for a
prefix of its _ArrayZipData_ string.  That prefix is a complete base64 zlib
stream of 8 float32 values: the first 8 draws of lane 0's xorshift128+
stream seeded from glibc rand() at its default state (src/mcx_core.cu:
3056-3066, :1731-1742).  The values appear in row-major order of the
2x2x2 column-major array, i.e. stream positions [0,4,2,6,1,5,3,7].

Reads the reference's test script as text; run in the build container only.
"""
import base64
import json
import os
import re
import zlib

import numpy as np

REF = "/root/reference/test/testmcx.sh"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    text = open(REF).read()
    m = re.search(r"grep '(eJx[^']+)' testrng222\.jnii", text)
    prefix = m.group(1).replace("\\+", "+").replace("\\/", "/")
    raw = zlib.decompress(base64.b64decode(prefix + "=" * (-len(prefix) % 4)))
    disk = np.frombuffer(raw, dtype="<f4")
    perm = [0, 4, 2, 6, 1, 5, 3, 7]
    stream = np.empty(8, dtype=np.float32)
    stream[perm] = disk
    out = {
        "source": "test/testmcx.sh:155 (_ArrayZipData_ prefix of testrng222.jnii)",
        "zipdata_prefix": prefix,
        "disk_order": [float(x) for x in disk],
        "stream_order": [float(x) for x in stream],
        "disk_from_stream_permutation": perm,
        "seed_state": "glibc rand() default state (no srand)",
    }
    with open(os.path.join(HERE, "rng_kat.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
