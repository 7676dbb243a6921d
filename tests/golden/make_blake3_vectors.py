"""Writes blake3_vectors.json beside this script: the input lengths of BLAKE3's official test vectors (vendored in the
reference at liblizard/vendor/blake3/test_vectors/test_vectors.json; inputs are the bytes i mod 251) and the first 32
bytes of each one's hash, as hex. Run once where the reference checkout is present:
    python3 tests/golden/make_blake3_vectors.py [reference root, default /root/reference]"""
import json
import os
import sys

root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = os.path.join(root, "liblizard", "vendor", "blake3", "test_vectors", "test_vectors.json")
cases = json.load(open(src))["cases"]
out = [{"input_len": c["input_len"], "hash32": c["hash"][:64]} for c in cases]
json.dump({"input": "byte i of an input is i mod 251", "cases": out},
          open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "blake3_vectors.json"), "w"), indent=0)
