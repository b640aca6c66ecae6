"""Writes the golden fixtures under tests/golden/ from the reference tree.

Run in the development container (the reference is not present on GPU hosts):

    python tests/golden/make_golden.py /root/reference

* decode_tables.json — SHA-256 of gpu_cuda/decode_tables/{gammaInA,zetaInA,
  gamma8192}.txt, their entry counts, and every 257th entry of each table
  (data vectors; the full tables are re-derived from the closed form in the
  tests and compared through the hashes).
* cbtComm-sample.txt — the 1000-contact ChronoGraph test fixture
  (third_party/ChronoGraph/src/test/resources/cbtComm-sample.txt), copied as
  data.
"""
import hashlib
import json
import os
import re
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref: str) -> None:
    tables = {}
    for name in ("gammaInA", "zetaInA", "gamma8192"):
        path = os.path.join(ref, "gpu_cuda", "decode_tables", name + ".txt")
        text = open(path, "rb").read()
        vals = [int(t) for t in re.split(rb"[,\s]+", text) if t]
        tables[name] = {
            "sha256": hashlib.sha256(text).hexdigest(),
            "entries": len(vals),
            "sample_stride": 257,
            "sample": vals[::257],
            # bytes after the last value (",", or ",\r\n" in zetaInA.txt)
            "terminator": text[len(text.rstrip(b",\r\n")):].decode(),
        }
    with open(os.path.join(HERE, "decode_tables.json"), "w") as f:
        json.dump(tables, f)
    shutil.copyfile(os.path.join(ref, "third_party", "ChronoGraph", "src", "test", "resources",
                                 "cbtComm-sample.txt"), os.path.join(HERE, "cbtComm-sample.txt"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
